"""GPU parity of the generator's two ways of evaluating same-spin doubles.

Mode 1: every source enumerates and evaluates its own same-spin doubles (colex rank,
Slater-Condon element with both sign factors).  Mode 2: each distinct source spin string's
doubles are evaluated once into a table (excited rank, element with the other spin empty,
other-spin parity mask) and every source with that string streams it.  Both must give the
oracle's H.v bit for bit, on one target pass and on several, in FP64 and mixed precision,
for spin-free and spin-dependent (UHF) integrals.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_same_vec(k1, v1, k2, v2):
    assert len(k1) == len(k2), (len(k1), len(k2))
    assert np.array_equal(k1, k2)
    diff = np.nonzero(bits(v1) != bits(v2))[0]
    assert len(diff) == 0, "first mismatch at %d: %r vs %r" % (diff[0], v1[diff[0]], v2[diff[0]])


def grown_input(orc, hf, n, seed):
    """A vector of about n determinants grown from HF by oracle H.v and top-n: many sources
    share a spin-0 or spin-1 string, as in a solver's vectors; values perturbed by seed."""
    k, v = np.array([hf], np.uint64), np.array([1.0])
    for _ in range(6):
        if len(k) >= n:
            break
        k, v, _, _ = orc.hv(k, v, nd=n, cut=1e-7, out_cap=n)
    rng = np.random.default_rng(seed)
    return k, v * (1.0 + 0.1 * rng.normal(size=len(v)))


def run_modes(ctx, orc, k, v, lam, coeff, mixed, multipass):
    ok, ov, od, _ = orc.hv(k, v, nd=len(k), lam=lam, coeff=coeff, cut=1e-5, mode=2 if mixed else 0)
    for mode in (1, 2):
        ctx.set_gen_tables(mode)
        ctx.set_precision(mixed)
        if multipass:
            ctx.set_workspace(1 << 26)
        try:
            vin = ctx.vector(len(k), k, v)
            vout = ctx.vector(60_000_000)
            gd = ctx.hv(vin, vout, len(k), lam, coeff)
            gk, gv = vout.download()
            st = ctx.stats()
        finally:
            ctx.set_gen_tables(0)
            ctx.set_precision(False)
            ctx.set_workspace(8 << 30)
        log("gen tables mode %d mixed %s multipass %s: passes %d strings %d/%d out %d" % (
            mode, mixed, multipass, st["n_passes"], st["n_str0"], st["n_str1"], len(gk)))
        if mode == 2:
            assert st["n_str0"] > 0 and st["n_str1"] > 0
        else:
            assert st["n_str0"] == 0 and st["n_str1"] == 0
        if multipass:
            assert st["n_passes"] > 1
        assert_same_vec(gk, gv, ok, ov)
        assert bits([gd])[0] == bits([od])[0], (gd, od)


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("multipass", [False, True])
def test_gen_tables_rhf(make_system, mixed, multipass):
    from sfci_amd import gpu
    s = make_system(14, 10, 5)
    orc = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    k, v = grown_input(orc, s["init_hf"], 6000, 1)
    ctx = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    try:
        run_modes(ctx, orc, k, v, 0.13, -0.9, mixed, multipass)
    finally:
        ctx.close()


@pytest.mark.parametrize("multipass", [False, True])
def test_gen_tables_uhf(tmp_path, multipass):
    # spin-dependent integrals: the spin-0 and spin-1 tables read different blocks
    from sfci_amd import gpu, integrals as I
    h_a, h_b, g_aa, g_bb, g_ab, ec = I.synth_uhf_integrals(11, 9)
    fc = str(tmp_path / "u11_fcidump")
    I.write_fcidump_uhf(fc, h_a, h_b, g_aa, g_bb, g_ab, 10, ec)
    inp = I.convert_uhf(fc, str(tmp_path / "u11_IN"))
    orc = O.Oracle(inp, 22, 10)
    k, v = grown_input(orc, (1 << 10) - 1, 4000, 2)
    ctx = gpu.Context(inp, 22, 10)
    try:
        for mixed in (False, True):
            run_modes(ctx, orc, k, v, 0.2, 1.1, mixed, multipass)
    finally:
        ctx.close()


def test_gen_tables_c2_example(request):
    # the reference's own C2 integrals (52 spin orbitals, 8 electrons), several passes
    from sfci_amd import gpu
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    fc = os.path.join(here, "golden", "c2-cc-pvdz-hf_fcidump")
    from sfci_amd import integrals as I
    tmp = request.getfixturevalue("tmp_path")
    inp = I.convert_rhf(fc, str(tmp / "c2_IN"))
    orc = O.Oracle(inp, 52, 8)
    k, v = grown_input(orc, (1 << 8) - 1, 3000, 3)
    ctx = gpu.Context(inp, 52, 8)
    try:
        run_modes(ctx, orc, k, v, 0.0, 1.0, False, True)
    finally:
        ctx.close()
