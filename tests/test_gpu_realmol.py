"""GPU parity on a real molecule: the reference's C2/cc-pVDZ example (example/C2,
symmetry-sparse integrals, 52 spin orbitals, 8 correlated electrons, init_hf=202125315
as in example/sfci.input), converted by this repository's FCIDUMP -> _IN converter."""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from sfci_amd import integrals as I

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HF = 202125315


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def c2(tmp_path_factory):
    d = tmp_path_factory.mktemp("c2")
    inp = I.convert_rhf(os.path.join(GOLDEN, "c2-cc-pvdz-hf_fcidump"), str(d / "c2_IN"))
    from sfci_amd import gpu
    print("c2: loading", inp, flush=True)
    ctx = gpu.Context(inp, 52, 8)
    yield ctx, O.Oracle(inp, 52, 8)
    ctx.close()


def test_c2_hv_bit_exact(c2):
    ctx, orc = c2
    k, v = np.array([HF], np.uint64), np.array([1.0])
    for step in range(3):
        vin = ctx.vector(len(k), k, v)
        out = ctx.vector(2_000_000)
        gd = ctx.hv(vin, out, len(k), 0.0, 1.0)
        gk, gv = out.download()
        ok, ov, od, _ = orc.hv(k, v, cut=1e-5, out_cap=2_000_000)
        print("c2 step %d: %d -> %d dets" % (step, len(k), len(gk)), flush=True)
        assert np.array_equal(gk, ok)
        assert np.array_equal(bits(gv), bits(ov))
        assert bits([gd])[0] == bits([od])[0]
        keep = np.argsort(-np.abs(ov), kind="stable")[:20000]
        keep = np.sort(keep)
        k, v = ok[keep], ov[keep] / np.linalg.norm(ov[keep])


def test_c2_lanczos_matches_oracle(c2):
    ctx, orc = c2
    g = ctx.ground_state(HF, 4000, 8, 2, hpsi_cap=2_000_000)
    o = orc.lanczos(HF, 4000, 8, 2, hpsi_cap=2_000_000)
    print("c2 lanczos gpu %r oracle %r" % (g["ritz"], o["ritz"]), flush=True)
    assert g["restarts"] == o["restarts"]
    assert np.allclose(g["ritz"], o["ritz"], rtol=0, atol=1e-9)
    assert abs(g["e_final"] - o["e_final"]) < 1e-9
    assert g["e_final"] < ctx.diag_element(HF)
