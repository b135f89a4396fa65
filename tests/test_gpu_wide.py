"""GPU parity of the 128-bit-key H.v (more than 64 spin orbitals) against the oracle.

Values are compared through their IEEE bit patterns, keys word by word: both sides form
every sum exactly and round once.
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


def assert_same(lo1, hi1, v1, lo2, hi2, v2):
    assert len(lo1) == len(lo2), (len(lo1), len(lo2))
    assert np.array_equal(lo1, lo2) and np.array_equal(hi1, hi2)
    diff = np.nonzero(bits(v1) != bits(v2))[0]
    assert len(diff) == 0, "first mismatch at %d: %r vs %r" % (diff[0], v1[diff[0]], v2[diff[0]])


@pytest.fixture(scope="module")
def wsys(make_system):
    return make_system(33, 8, 7)          # 66 spin orbitals, 4 + 4 electrons


@pytest.fixture(scope="module")
def worc(wsys):
    return O.Oracle(wsys["in_path"], wsys["n_so"], wsys["n_elec"])


@pytest.fixture(scope="module")
def wctx(wsys):
    from sfci_amd import gpu
    log("wctx: loading %s" % wsys["in_path"])
    c = gpu.WideContext(wsys["in_path"], wsys["n_so"], wsys["n_elec"])
    yield c
    c.close()


@pytest.fixture(scope="module")
def winput(wsys, worc):
    """HF -> H.v -> top-600 -> H.v (oracle, top 2500): a vector reaching spin orbitals 64 and 65."""
    hf = wsys["init_hf"]
    lo, hi, v, _ = worc.hv_wide(np.array([hf], np.uint64), np.zeros(1, np.uint64), np.array([1.0]))
    lo, hi, v, _ = worc.hv_wide(lo, hi, v, nd=600, out_cap=2500)
    assert hi.any()
    return lo, hi, v


def gpu_wide_hv(ctx, lo, hi, v, nd, lam, coeff, cap):
    vin = ctx.vector(max(len(lo), 1), lo, hi, v)
    vout = ctx.vector(cap)
    d = ctx.hv(vin, vout, nd, lam, coeff)
    ol, oh, ov = vout.download()
    return ol, oh, ov, d


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("nd,cap,lam,coeff", [(2500, 60_000_000, 0.0, 1.0), (300, 5000, 0.37, -1.25)])
def test_wide_hv_bit_exact(wctx, worc, winput, mixed, nd, cap, lam, coeff):
    lo, hi, v = winput
    log("wide H.v: n_in=%d nd=%d cap=%d mixed=%s" % (len(lo), nd, cap, mixed))
    wctx.set_precision(mixed)
    try:
        gl, gh, gv, gd = gpu_wide_hv(wctx, lo, hi, v, nd, lam, coeff, cap)
    finally:
        wctx.set_precision(False)
    ol, oh, ov, od = worc.hv_wide(lo, hi, v, nd=nd, lam=lam, coeff=coeff, out_cap=cap, mode=2 if mixed else 0)
    assert_same(gl, gh, gv, ol, oh, ov)
    assert bits([gd])[0] == bits([od])[0]
    st = wctx.stats()
    assert st["n_src"] == min(nd, len(lo)) and st["n_out"] == len(ol)


def test_wide_hv_multipass(wctx, worc, winput):
    lo, hi, v = winput
    wctx.set_workspace(104 * 2_000_000)    # a few million records per pass
    try:
        gl, gh, gv, gd = gpu_wide_hv(wctx, lo, hi, v, 2500, 0.2, 1.0, 60_000_000)
        st = wctx.stats()
    finally:
        wctx.set_workspace(1 << 36)
    log("multipass: %d passes" % st["n_passes"])
    assert st["n_passes"] > 1
    ol, oh, ov, od = worc.hv_wide(lo, hi, v, lam=0.2)
    assert_same(gl, gh, gv, ol, oh, ov)
    assert bits([gd])[0] == bits([od])[0]


def test_wide_hv_shared_group_hashes(wctx, worc, winput):
    """With 16 grouping bits many distinct keys share a group: they must still merge exactly."""
    lo, hi, v = winput
    wctx.set_group_bits(16)
    try:
        gl, gh, gv, gd = gpu_wide_hv(wctx, lo, hi, v, 300, 0.1, 1.0, 60_000_000)
        st = wctx.stats()
    finally:
        wctx.set_group_bits(48)
    ol, oh, ov, od = worc.hv_wide(lo, hi, v, nd=300, lam=0.1)
    assert st["n_unique"] > 4 * 65536     # far more distinct targets than groups
    assert_same(gl, gh, gv, ol, oh, ov)
    assert bits([gd])[0] == bits([od])[0]


def test_wide_path_matches_64bit_path(make_system):
    """Below 65 spin orbitals the 128-bit path must reproduce the 64-bit sfci_hv."""
    from sfci_amd import gpu
    s = make_system(12, 8, 0)
    c64 = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    cw = gpu.WideContext(s["in_path"], s["n_so"], s["n_elec"])
    try:
        v = c64.vector(10, np.array([s["init_hf"]], np.uint64), np.array([1.0]))
        o1 = c64.vector(100_000)
        c64.hv(v, o1, 10)
        k, x = o1.download()
        for mixed in (False, True):
            c64.set_precision(mixed)
            cw.set_precision(mixed)
            a = c64.vector(len(k), k, x)
            b = c64.vector(200_000)
            d64 = c64.hv(a, b, 900, 0.3, 0.7)
            k64, v64 = b.download()
            wl, wh, wv, dw = gpu_wide_hv(cw, k, np.zeros(len(k), np.uint64), x, 900, 0.3, 0.7, 200_000)
            assert_same(wl, wh, wv, k64, np.zeros(len(k64), np.uint64), v64)
            assert bits([dw])[0] == bits([d64])[0]
    finally:
        c64.close()
        cw.close()


def test_wide_rejects_bad_input(wctx, wsys):
    from sfci_amd import gpu
    # keys out of order: refused on upload
    with pytest.raises(gpu.SfciError):
        wctx.vector(4, np.array([5, 3], np.uint64), np.zeros(2, np.uint64), np.ones(2))
    # wrong electron count: reported by the generation kernel's check, nothing emitted
    bad = np.array([(1 << 9) - 1], np.uint64)
    with pytest.raises(gpu.SfciError) as e:
        gpu_wide_hv(wctx, bad, np.zeros(1, np.uint64), np.ones(1), 1, 0.0, 1.0, 100)
    assert e.value.code == -1
    # the context stays usable
    hf = np.array([wsys["init_hf"]], np.uint64)
    gl, gh, gv, gd = gpu_wide_hv(wctx, hf, np.zeros(1, np.uint64), np.ones(1), 1, 0.0, 1.0, 100_000)
    assert len(gl) > 0
