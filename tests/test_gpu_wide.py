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
    # about 3.5 million records per pass (52 B each above the fixed partition slack)
    wctx.set_workspace(600_000_000)
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


@pytest.mark.parametrize("hash_bits,fill", [(9, 0), (9, 179), (16, 0), (12, 40)])
def test_wide_hv_merge_classes(wctx, worc, winput, hash_bits, fill):
    """Few buckets with a small table fill: buckets are merged in many hash classes (each
    rereading the bucket); many buckets: most hold a few targets.  Same result either way."""
    lo, hi, v = winput
    wctx.set_merge(hash_bits, fill)
    try:
        gl, gh, gv, gd = gpu_wide_hv(wctx, lo, hi, v, 2500, 0.1, 1.0, 60_000_000)
        st = wctx.stats()
    finally:
        wctx.set_merge(0, 0)
    log("hash bits %d fill %d: %d distinct targets, %d class splits" % (hash_bits, fill, st["n_unique"],
                                                                       st["n_split"]))
    assert st["hash_bits"] == hash_bits and st["n_fallback"] == 0
    if fill:
        assert st["n_split"] >= (1 << hash_bits)
    ol, oh, ov, od = worc.hv_wide(lo, hi, v, nd=2500, lam=0.1)
    assert_same(gl, gh, gv, ol, oh, ov)
    assert bits([gd])[0] == bits([od])[0]


def test_wide_energy_pull(wctx, worc, winput):
    """<v|H|v> by pull on 128-bit keys against the oracle's exact <v|H v> (the two round at
    different points: per source against per target)."""
    lo, hi, v = winput
    e = wctx.energy(wctx.vector(len(lo), lo, hi, v))
    _, _, _, od = worc.hv_wide(lo, hi, v, lam=0.0, coeff=1.0)
    log("pull %.15f, oracle <v|Hv> %.15f" % (e, od))
    assert abs(e - od) < 1e-12 * max(1.0, abs(od))


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
    # wrong electron count: refused on upload (the generation kernel checks it again)
    bad = np.array([(1 << 9) - 1], np.uint64)
    with pytest.raises(gpu.SfciError) as e:
        gpu_wide_hv(wctx, bad, np.zeros(1, np.uint64), np.ones(1), 1, 0.0, 1.0, 100)
    assert e.value.code == -1
    # the context stays usable
    hf = np.array([wsys["init_hf"]], np.uint64)
    gl, gh, gv, gd = gpu_wide_hv(wctx, hf, np.zeros(1, np.uint64), np.ones(1), 1, 0.0, 1.0, 100_000)
    assert len(gl) > 0


def _pair_contexts(make_system):
    from sfci_amd import gpu
    s = make_system(12, 8, 0)
    return s, gpu.Context(s["in_path"], s["n_so"], s["n_elec"]), gpu.WideContext(s["in_path"], s["n_so"], s["n_elec"])


def test_wide_vector_ops_match_64bit(make_system):
    """dot, norm, scale, addTwo, shrinkFrom and recombination on two-word keys reproduce the
    64-bit operations bit for bit (truncation ties included)."""
    s, c64, cw = _pair_contexts(make_system)
    try:
        rng = np.random.default_rng(5)
        v0 = c64.vector(10, np.array([s["init_hf"]], np.uint64), np.array([1.0]))
        big = c64.vector(100_000)
        c64.hv(v0, big, 10)
        k, x = big.download()
        z = np.zeros(len(k), np.uint64)

        def pick(n, seed, ties=False):
            r = np.random.default_rng(seed)
            i = np.sort(r.choice(len(k), size=n, replace=False))
            val = r.choice([0.5, -0.25, 0.125], size=n) if ties else r.normal(size=n)
            return k[i], val

        a_k, a_v = pick(900, 1)
        b_k, b_v = pick(700, 2, ties=True)
        A64, B64 = c64.vector(2000, a_k, a_v), c64.vector(2000, b_k, b_v)
        Aw = cw.vector(2000, a_k, np.zeros(len(a_k), np.uint64), a_v)
        Bw = cw.vector(2000, b_k, np.zeros(len(b_k), np.uint64), b_v)
        assert bits([c64.dot(A64, B64)])[0] == bits([cw.dot(Aw, Bw)])[0]
        assert bits([c64.norm(A64)])[0] == bits([cw.norm(Aw)])[0]
        c64.scale(A64, 0.37)
        cw.scale(Aw, 0.37)
        for cap in (2000, 1000):
            V64, Vw = c64.vector(cap), cw.vector(cap)
            V64.copy_from(A64)
            cw.copy(Vw, Aw)
            c64.add_two(V64, B64, -1.5)
            cw.add_two(Vw, Bw, -1.5)
            k1, x1 = V64.download()
            l2, h2, x2 = Vw.download()
            assert_same(l2, h2, x2, k1, np.zeros(len(k1), np.uint64), x1)
            S64, Sw = c64.vector(cap // 3), cw.vector(cap // 3)
            c64.shrink_from(S64, V64)
            cw.shrink_from(Sw, Vw)
            k1, x1 = S64.download()
            l2, h2, x2 = Sw.download()
            assert_same(l2, h2, x2, k1, np.zeros(len(k1), np.uint64), x1)
        coef = rng.normal(size=2)
        L64, Lw = c64.vector(800), cw.vector(800)
        c64.lincomb([A64, B64], coef, L64)
        cw.lincomb([Aw, Bw], coef, Lw)
        k1, x1 = L64.download()
        l2, h2, x2 = Lw.download()
        assert_same(l2, h2, x2, k1, np.zeros(len(k1), np.uint64), x1)
        # the same pull with the same rounding (per source) on both paths
        assert bits([c64.energy(L64)])[0] == bits([cw.energy(Lw)])[0]
    finally:
        c64.close()
        cw.close()


def test_wide_ground_state_matches_64bit(make_system):
    """Below 65 spin orbitals the 128-bit driver follows the 64-bit driver step for step."""
    s, c64, cw = _pair_contexts(make_system)
    try:
        r64 = c64.ground_state(s["init_hf"], 3000, 8, 3, hpsi_cap=200_000, e_prev_init=float("inf"))
        rw = cw.ground_state(s["init_hf"], 3000, 8, 3, hpsi_cap=200_000, e_prev_init=float("inf"))
        log("64-bit %s / 128-bit %s" % (r64["ritz"], rw["ritz"]))
        assert r64["restarts"] == rw["restarts"]
        assert np.array_equal(bits(r64["ritz"]), bits(rw["ritz"]))
        assert np.abs(r64["e_trunc"] - rw["e_trunc"]).max() < 1e-12 * abs(r64["e_final"])
        assert abs(r64["e_final"] - rw["e_final"]) < 1e-12 * abs(r64["e_final"])
    finally:
        c64.close()
        cw.close()


def test_wide_ground_state_full_space_fci(make_system):
    """66 spin orbitals, 1 + 1 electrons: with Nd above the full space (1089 determinants) and
    no cut, the 128-bit driver reaches the lowest eigenvalue of the dense Hamiltonian."""
    from sfci_amd import gpu
    s = make_system(33, 2, 5)
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    n_orb = s["n_orb"]
    dets = sorted({(1 << (2 * i)) | (1 << (2 * j + 1)) for i in range(n_orb) for j in range(n_orb)})
    idx = {d: t for t, d in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        lo, hi, val = o.connections_wide(d)
        for a, b, x in zip(lo, hi, val):
            M[idx[(int(b) << 64) | int(a)], j] += x
    e0 = np.linalg.eigvalsh(M)[0]
    ctx = gpu.WideContext(s["in_path"], s["n_so"], s["n_elec"])
    try:
        ctx.set_cut(0.0)
        r = ctx.ground_state(s["init_hf"], 2000, 20, 8, hpsi_cap=20_000, e_prev_init=float("inf"))
    finally:
        ctx.close()
    log("FCI %.12f, 128-bit solve %.12f after %d restarts" % (e0, r["e_final"], r["restarts"]))
    assert abs(r["e_final"] - e0) < 1e-9

