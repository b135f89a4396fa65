"""The `sfci` command-line driver: reads ./sfci.input (src/main.cpp:128-151) and runs
task=ground (the truncated Lanczos solve) or task=excited (CISD state 2 + one
shift-invert step) on the GPU; results must agree with the library calls."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sfci_amd", "sfci")


def run_cli(workdir, lines):
    with open(os.path.join(workdir, "sfci.input"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("cli: %s" % " ".join(lines), flush=True)
    r = subprocess.run([EXE], cwd=workdir, capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stderr
    return r.stdout


def field(out, name):
    vals = [ln.split(name, 1)[1] for ln in out.splitlines() if ln.startswith(name)]
    return float(vals[-1])


def test_cli_ground_and_excited(make_system, tmp_path):
    s = make_system(8, 6, 3)
    base = ["num_orb=16", "num_e=6", "filename=%s" % s["in_path"], "init_hf=%d" % s["init_hf"],
            "Nd=3136", "Nl=6", "num_restart=2", "num_thread=8", "bucket_unit=4"]
    out = run_cli(str(tmp_path), ["task=ground"] + base)
    e_hf = field(out, "E_hf:")
    orc = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    assert abs(e_hf - orc.diag(s["init_hf"])) < 1e-9
    assert "SFCI time:" in out
    target = e_hf + 0.5
    out = run_cli(str(tmp_path), ["task=excited", "target_E=%.17g" % target] + base)
    vk, vv, oe, oec = orc.excited(s["init_hf"], target, 3136, 6, 2, cut=1e-5)
    assert abs(field(out, "From CISD: ") - oec) < 1e-9
    assert abs(field(out, "energy after 0th SI:") - oe) < 1e-8
