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
    # the solve's final energy against the oracle's solve with the CLI's parameters (the
    # reference's 0.0 start of the stopping test, HPsiCap 6e7, cut 1e-5)
    o = orc.lanczos(s["init_hf"], 3136, 6, 2, hpsi_cap=60_000_000, cut=1e-5, e_prev_init=0.0)
    assert abs(field(out, "final e:") - o["e_final"]) < 1e-9
    target = e_hf + 0.5
    out = run_cli(str(tmp_path), ["task=excited", "target_E=%.17g" % target] + base)
    vk, vv, oe, oec = orc.excited(s["init_hf"], target, 3136, 6, 2, cut=1e-5)
    assert abs(field(out, "From CISD: ") - oec) < 1e-9
    assert abs(field(out, "energy after 0th SI:") - oe) < 1e-8


def test_cli_ground_128bit(make_system, tmp_path):
    """task=ground beyond 64 spin orbitals (the 128-bit-key branch): 66 spin orbitals, one
    electron of each spin, started from the determinant with both electrons in the top
    spatial orbital (init_hf = 3 * 2^64, above 64 bits).  The printed final energy equals
    the library solve with the same parameters, and with Nd above the full space (1089
    determinants) and no cut it is the lowest eigenvalue of the dense Hamiltonian."""
    from sfci_amd import gpu
    s = make_system(33, 2, 5)
    init = 3 << 64
    base = ["num_orb=66", "num_e=2", "filename=%s" % s["in_path"], "init_hf=%d" % init, "Nd=2000", "Nl=20",
            "num_restart=8", "hpsi_cap=20000", "coeff_cut=0"]
    out = run_cli(str(tmp_path), ["task=ground"] + base)
    e_cli = field(out, "final e:")
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    assert abs(field(out, "E_hf:") - o.diag_wide(init)) < 1e-9
    ctx = gpu.WideContext(s["in_path"], s["n_so"], s["n_elec"])
    try:
        ctx.set_cut(0.0)
        r = ctx.ground_state(init, 2000, 20, 8, hpsi_cap=20_000, e_prev_init=0.0)
    finally:
        ctx.close()
    print("cli 128-bit: final e %.12f, library %.12f" % (e_cli, r["e_final"]), flush=True)
    assert abs(e_cli - r["e_final"]) < 1e-9
    dets = sorted({(1 << (2 * i)) | (1 << (2 * j + 1)) for i in range(33) for j in range(33)})
    idx = {d: t for t, d in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        lo, hi, val = o.connections_wide(d)
        for a, b, x in zip(lo, hi, val):
            M[idx[(int(b) << 64) | int(a)], j] += x
    if e_cli > 0:
        # the reference's stopping test from 0.0 ends a positive-energy solve after one
        # restart, so only the variational bound holds
        assert e_cli >= np.linalg.eigvalsh(M)[0] - 1e-9
    else:
        assert abs(e_cli - np.linalg.eigvalsh(M)[0]) < 1e-9
