"""The sfci command line's checks that end a run before any GPU work (no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sfci_amd", "sfci")


def run(tmp_path, lines):
    if not os.path.exists(EXE):
        pytest.skip("sfci executable not built")
    (tmp_path / "sfci.input").write_text("\n".join(lines) + "\n")
    return subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)


def test_more_than_128_spin_orbitals_refused(tmp_path):
    r = run(tmp_path, ["task=ground", "num_orb=130", "num_e=20", "filename=x_IN"])
    assert r.returncode == 1 and "larger than 128" in r.stderr


def test_wide_runs_only_ground(tmp_path):
    r = run(tmp_path, ["task=excited", "num_orb=100", "num_e=20", "filename=x_IN", "init_hf=1048575"])
    assert r.returncode == 1 and "task=ground" in r.stderr


def test_init_hf_must_be_decimal(tmp_path):
    r = run(tmp_path, ["task=ground", "num_orb=100", "num_e=20", "filename=x_IN", "init_hf=12a"])
    assert r.returncode == 1 and "init_hf" in r.stderr
    # a 129-bit value does not fit
    r = run(tmp_path, ["task=ground", "num_orb=100", "num_e=20", "filename=x_IN", "init_hf=%d" % (1 << 128)])
    assert r.returncode == 1 and "init_hf" in r.stderr


def test_init_hf_beyond_64_bits_needs_more_spin_orbitals(tmp_path):
    r = run(tmp_path, ["task=ground", "num_orb=52", "num_e=14", "filename=x_IN", "init_hf=%d" % (1 << 64)])
    assert r.returncode == 1 and "init_hf" in r.stderr

