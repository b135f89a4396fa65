"""GPU parity: the HIP H.v and sparse-vector operations against the oracle, bit for bit.

All values are compared through their IEEE bit patterns: the oracle and the GPU
library both form every sum exactly (128-bit fixed point) and round once, and neither
fuses multiply-adds, so any difference is a bug.
"""
import itertools
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


def all_dets(n_orb, k):
    strings = []
    for occ in itertools.combinations(range(n_orb), k):
        s = 0
        for i in occ:
            s |= 1 << (2 * i)
        strings.append(s)
    a = np.array(strings, dtype=np.uint64)
    d = (a[:, None] | (a[None, :] << np.uint64(1))).ravel()
    return np.sort(d)


@pytest.fixture(scope="module")
def sys1(make_system):
    return make_system(12, 8, 0)          # the 12-orbital / 8-electron workload, seed 0


@pytest.fixture(scope="module")
def orc1(sys1):
    return O.Oracle(sys1["in_path"], sys1["n_so"], sys1["n_elec"])


@pytest.fixture(scope="module")
def ctx1(sys1):
    from sfci_amd import gpu
    log("ctx1: loading %s" % sys1["in_path"])
    c = gpu.Context(sys1["in_path"], sys1["n_so"], sys1["n_elec"])
    yield c
    c.close()


def random_input(sys, n, seed, ties=False):
    rng = np.random.default_rng(seed)
    dets = all_dets(sys["n_orb"], sys["n_elec"] // 2)
    pick = np.sort(rng.choice(len(dets), size=n, replace=False))
    k = dets[pick]
    if ties:
        v = rng.choice([-0.25, 0.125, 0.5, -0.5, 0.03125], size=n)
    else:
        v = rng.normal(size=n) * np.exp(-rng.uniform(0, 8, size=n))
    return k, v


def gpu_hv(ctx, k, v, nd, lam, coeff, cap):
    vin = ctx.vector(max(len(k), 1), k, v)
    vout = ctx.vector(cap)
    d = ctx.hv(vin, vout, nd, lam, coeff)
    ok, ov = vout.download()
    return ok, ov, d


@pytest.mark.parametrize("case", ["hf", "rand", "rand_shift", "cap", "nd_ties", "cut0"])
def test_hv_bit_exact(case, sys1, orc1, ctx1):
    log("hv case %s" % case)
    hf = sys1["init_hf"]
    cut, cap, lam, coeff = 1e-5, 60_000_000, 0.0, 1.0
    if case == "hf":
        k, v = np.array([hf], np.uint64), np.array([1.0])
    elif case == "rand":
        k, v = random_input(sys1, 3000, 1)
    elif case == "rand_shift":
        k, v = random_input(sys1, 3000, 2)
        lam, coeff = 0.37, -0.6
    elif case == "cap":
        k, v = random_input(sys1, 2000, 3, ties=True)
        cap = 5000
    elif case == "nd_ties":
        k, v = random_input(sys1, 2000, 4, ties=True)
    else:
        k, v = random_input(sys1, 1000, 5)
        cut = 0.0
    nd = 700 if case == "nd_ties" else len(k)
    ctx1.set_cut(cut)
    try:
        gk, gv, gd = gpu_hv(ctx1, k, v, nd, lam, coeff, cap)
    finally:
        ctx1.set_cut(1e-5)
    ok, ov, od, _ = orc1.hv(k, v, nd=nd, lam=lam, coeff=coeff, cut=cut, out_cap=cap)
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0], (gd, od)
    st = ctx1.stats()
    assert st["n_out"] == len(ok)


@pytest.mark.parametrize("hbits", [1, 2])
def test_hv_bucket_tags_list_and_search(make_system, hbits):
    # the whole 10-orbital / 4-electron space as input: every target is an input
    # determinant.  With 2 buckets each holds about 1000 of them, more than the
    # per-bucket tag list, so the rest are finished through the search; with 4 buckets
    # the list holds them all.  No bucket overflows in either case.
    from sfci_amd import gpu
    sysS = make_system(10, 4, 7)
    orcS = O.Oracle(sysS["in_path"], sysS["n_so"], sysS["n_elec"])
    ctx = gpu.Context(sysS["in_path"], sysS["n_so"], sysS["n_elec"])
    try:
        k = all_dets(sysS["n_orb"], 2)
        v = np.random.default_rng(11).normal(size=len(k))
        ctx.set_merge_mode(2, hbits)
        gk, gv, gd = gpu_hv(ctx, k, v, len(k), 0.3, 0.7, 60_000_000)
        st = ctx.stats()
    finally:
        ctx.close()
    log("tags: %d dets, hash bits %d, fallback %d" % (len(k), st["hash_bits"], st["n_fallback"]))
    assert st["n_fallback"] == 0 and st["hash_bits"] == hbits
    ok, ov, od, _ = orcS.hv(k, v, nd=len(k), lam=0.3, coeff=0.7, cut=1e-5)
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0]


def test_hv_multipass_matches(sys1, orc1, ctx1):
    k, v = random_input(sys1, 4000, 7)
    ctx1.set_workspace(1 << 26)            # 64 MB: forces several target-row passes
    try:
        gk, gv, gd = gpu_hv(ctx1, k, v, len(k), 0.1, 1.0, 60_000_000)
        st = ctx1.stats()
    finally:
        ctx1.set_workspace(8 << 30)
    log("multipass: %d passes" % st["n_passes"])
    assert st["n_passes"] > 1
    ok, ov, od, _ = orc1.hv(k, v, nd=len(k), lam=0.1, coeff=1.0, cut=1e-5)
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0]


@pytest.mark.parametrize("workspace", [8 << 30, 1 << 30])
def test_hv_many_sources_matches(sys1, orc1, ctx1, workspace):
    # 140,000 sources (2.0e8 candidates): at least 64 per generator workgroup, so the
    # generators claim their sources from the global counter in runs of eight; one target
    # pass, and several with a 1 GiB workspace
    k, v = random_input(sys1, 140_000, 11)
    ctx1.set_workspace(workspace)
    try:
        gk, gv, gd = gpu_hv(ctx1, k, v, len(k), 0.1, 1.0, 60_000_000)
        st = ctx1.stats()
    finally:
        ctx1.set_workspace(8 << 30)
    log("many sources: %d candidates, %d passes, %d out" % (st["n_candidates"], st["n_passes"], len(gk)))
    assert st["n_src"] == len(k)
    if workspace < (8 << 30):
        assert st["n_passes"] > 1
    ok, ov, od, _ = orc1.hv(k, v, nd=len(k), lam=0.1, coeff=1.0, cut=1e-5)
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0]


@pytest.mark.parametrize("mode,hbits", [(1, 0), (2, 3), (2, 1), (2, 14), (2, 10), (4, 10), (3, 4), (3, 10)])
def test_hv_merge_modes_identical(sys1, orc1, ctx1, mode, hbits):
    # full sort, hash buckets (compact records above 8 bits, full-key records in mode 4),
    # hash buckets too few to hold the targets (each bucket is split into hash classes
    # merged one at a time), and tables counted full at 230 occupied slots (classes
    # split again and again) give the same bits
    k, v = random_input(sys1, 3000, 31)
    ctx1.set_merge_mode(mode, hbits)
    try:
        gk, gv, gd = gpu_hv(ctx1, k, v, len(k), 0.05, 1.0, 60_000_000)
        st = ctx1.stats()
    finally:
        ctx1.set_merge_mode(0, 0)
    log("merge mode %d bits %d: fallback %d split %d hash_bits %d" % (
        mode, hbits, st["n_fallback"], st["n_split"], st["hash_bits"]))
    assert st["n_fallback"] == 0
    if (mode == 2 and hbits == 1) or mode == 3:
        assert st["n_split"] >= 2
    if mode == 3 and hbits == 4:
        # nested: more splits than buckets
        assert st["n_split"] > (1 << hbits)
    ok, ov, od, _ = orc1.hv(k, v, nd=len(k), lam=0.05, coeff=1.0, cut=1e-5)
    assert_same_vec(gk, gv, ok, ov)
    assert bits_of(gd) == bits_of(od)


def bits_of(x):
    return bits([x])[0]


@pytest.mark.parametrize("mixed", [False, True])
def test_hv_bucket_overflow_split(make_system, mixed):
    """Hash buckets too few for a pass's distinct targets (9 bits forced, 3.2e6 targets for
    512 tables of 4608 slots): each bucket is split into hash classes merged one at a time,
    with the same bits as the oracle."""
    from sfci_amd import gpu
    s = make_system(16, 8, 3)
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    dets = all_dets(16, 4)
    rng = np.random.default_rng(7)
    pick = np.sort(rng.choice(len(dets), size=4000, replace=False))
    k, v = dets[pick], rng.normal(size=4000)
    ctx = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    try:
        ctx.set_precision(mixed)
        ctx.set_merge_mode(2, 9)
        gk, gv, gd = gpu_hv(ctx, k, v, len(k), 0.1, 1.0, 60_000_000)
        st = ctx.stats()
    finally:
        ctx.close()
    log("overflow split: fallback %d split %d hash_bits %d unique %d" % (
        st["n_fallback"], st["n_split"], st["hash_bits"], st["n_unique"]))
    assert st["n_fallback"] == 0 and st["n_split"] >= 512 and st["hash_bits"] == 9
    ok, ov, od, _ = o.hv(k, v, nd=len(k), lam=0.1, coeff=1.0, cut=1e-5, mode=2 if mixed else 0)
    assert_same_vec(gk, gv, ok, ov)
    assert bits_of(gd) == bits_of(od)


def test_hv_empty(ctx1):
    vin, vout = ctx1.vector(10), ctx1.vector(10)
    assert ctx1.hv(vin, vout, 10) == 0.0
    assert vout.len == 0


def test_energy_and_diag(sys1, orc1, ctx1):
    k, v = random_input(sys1, 1500, 11)
    vec = ctx1.vector(len(k), k, v)
    assert bits([ctx1.energy(vec)])[0] == bits([orc1.energy(k, v)])[0]
    for d in [sys1["init_hf"], int(k[0]), int(k[-1])]:
        assert bits([ctx1.diag_element(d)])[0] == bits([orc1.diag(d)])[0]


def test_vector_ops_bit_exact(sys1, ctx1):
    rng = np.random.default_rng(3)
    ak, av = random_input(sys1, 5000, 21, ties=True)
    bk, bv = random_input(sys1, 7000, 22)
    av[::17] = 0.0                                   # zeros are kept by add_two
    A, B = ctx1.vector(20000, ak, av), ctx1.vector(20000, bk, bv)
    assert bits([ctx1.dot(A, B)])[0] == bits([O.dot(ak, av, bk, bv)])[0]
    assert bits([ctx1.norm(B)])[0] == bits([O.norm(bv)])[0]
    for cap in [20000, 6000, 1234]:
        V = ctx1.vector(cap, ak, av) if cap >= len(ak) else ctx1.vector(20000, ak, av)
        W = ctx1.vector(cap)
        if cap >= len(ak):
            ctx1.add_two(V, B, -0.7)
            gk, gv = V.download()
            ok, ov = O.add_two(ak, av, bk, bv, -0.7, cap)
        else:
            M = ctx1.vector(20000, bk, bv)
            ctx1.add_two(M, A, 0.3)
            ctx1.shrink_from(W, M)
            gk, gv = W.download()
            mk, mv = O.add_two(bk, bv, ak, av, 0.3, 20000)
            ok, ov = O.shrink(mk, mv, cap)
        assert_same_vec(gk, gv, ok, ov)
    ck, cv = random_input(sys1, 3000, 23)
    Cv = ctx1.vector(3000, ck, cv)
    coef = rng.normal(size=3)
    out = ctx1.vector(4000)
    ctx1.lincomb([A, B, Cv], coef, out)
    gk, gv = out.download()
    ok, ov = O.lincomb([(ak, av), (bk, bv), (ck, cv)], coef, 4000)
    # the oracle's lincomb also normalizes; undo on the GPU side by the same steps
    n = ctx1.norm(out)
    ctx1.scale(out, 1.0 / n)
    gk, gv = out.download()
    assert_same_vec(gk, gv, ok, ov)
    S = ctx1.vector(5000, ak, av)
    ctx1.scale(S, -1.75)
    sk, sv = S.download()
    assert np.array_equal(bits(sv), bits(av * -1.75))


def test_lanczos_matches_oracle(sys1, orc1, ctx1):
    nd, nl, nr = 3000, 8, 3
    log("lanczos gpu")
    g = ctx1.ground_state(sys1["init_hf"], nd, nl, nr, hpsi_cap=200_000, e_prev_init=float("inf"))
    log("lanczos oracle")
    o = orc1.lanczos(sys1["init_hf"], nd, nl, nr, hpsi_cap=200_000, e_prev_init=float("inf"))
    assert g["restarts"] >= 2
    assert g["restarts"] == o["restarts"]
    assert np.allclose(g["ritz"], o["ritz"], rtol=0, atol=1e-9)
    assert abs(g["e_final"] - o["e_final"]) < 1e-9
    log("ritz gpu %r oracle %r" % (g["ritz"], o["ritz"]))


@pytest.mark.parametrize("case", ["hf", "rand_shift", "multipass", "sort_merge", "hash_fallback"])
def test_hv_mixed_bit_exact(case, sys1, orc1, ctx1):
    # FP32 candidate products, exact sums: the GPU matches the oracle's mixed mode bit
    # for bit on every merge path
    log("mixed hv case %s" % case)
    lam, coeff = 0.0, 1.0
    if case == "hf":
        k, v = np.array([sys1["init_hf"]], np.uint64), np.array([1.0])
    else:
        k, v = random_input(sys1, 3000, 41)
        lam, coeff = 0.21, -0.8
    ctx1.set_precision(True)
    if case == "multipass":
        ctx1.set_workspace(1 << 26)
    elif case == "sort_merge":
        ctx1.set_merge_mode(1, 0)
    elif case == "hash_fallback":
        ctx1.set_merge_mode(2, 1)
    try:
        gk, gv, gd = gpu_hv(ctx1, k, v, len(k), lam, coeff, 60_000_000)
        st = ctx1.stats()
    finally:
        ctx1.set_precision(False)
        ctx1.set_workspace(8 << 30)
        ctx1.set_merge_mode(0, 0)
    if case == "multipass":
        assert st["n_passes"] > 1
    if case == "hash_fallback":
        assert st["n_split"] >= 1 and st["n_fallback"] == 0
    ok, ov, od, _ = orc1.hv(k, v, nd=len(k), lam=lam, coeff=coeff, cut=1e-5, mode=2)
    assert_same_vec(gk, gv, ok, ov)
    assert bits_of(gd) == bits_of(od)


def test_energy_is_fp64_under_mixed(sys1, orc1, ctx1):
    k, v = random_input(sys1, 1500, 12)
    vec = ctx1.vector(len(k), k, v)
    ctx1.set_precision(True)
    try:
        e = ctx1.energy(vec)
    finally:
        ctx1.set_precision(False)
    assert bits_of(e) == bits_of(orc1.energy(k, v))


def test_lanczos_mixed_energy_error(sys1, ctx1):
    # mixed-precision H.v inside the FP64 Lanczos recurrences: the final FP64 energy
    # stays within 1e-5 Ha of the FP64 solve
    nd, nl, nr = 3000, 8, 3
    g64 = ctx1.ground_state(sys1["init_hf"], nd, nl, nr, hpsi_cap=200_000, e_prev_init=float("inf"))
    ctx1.set_precision(True)
    try:
        g32 = ctx1.ground_state(sys1["init_hf"], nd, nl, nr, hpsi_cap=200_000, e_prev_init=float("inf"))
    finally:
        ctx1.set_precision(False)
    log("lanczos e64 %.12f mixed %.12f" % (g64["e_final"], g32["e_final"]))
    assert abs(g64["e_final"] - g32["e_final"]) < 1e-5


def test_hv_uhf_bit_exact(tmp_path):
    # spin-dependent integrals through the UHF converter: all four two-electron spin
    # blocks and both one-electron blocks differ
    from sfci_amd import gpu, integrals as I
    h_a, h_b, g_aa, g_bb, g_ab, ec = I.synth_uhf_integrals(10, 5)
    fc = str(tmp_path / "u10_fcidump")
    I.write_fcidump_uhf(fc, h_a, h_b, g_aa, g_bb, g_ab, 8, ec)
    inp = I.convert_uhf(fc, str(tmp_path / "u10_IN"))
    orc = O.Oracle(inp, 20, 8)
    ctx = gpu.Context(inp, 20, 8)
    try:
        sysu = dict(n_orb=10, n_elec=8)
        k, v = random_input(sysu, 2000, 51)
        for mixed in (False, True):
            log("uhf hv mixed=%s" % mixed)
            ctx.set_precision(mixed)
            gk, gv, gd = gpu_hv(ctx, k, v, len(k), 0.15, 0.9, 60_000_000)
            ok, ov, od, _ = orc.hv(k, v, nd=len(k), lam=0.15, coeff=0.9, cut=1e-5, mode=2 if mixed else 0)
            assert_same_vec(gk, gv, ok, ov)
            assert bits_of(gd) == bits_of(od)
    finally:
        ctx.close()


def test_energy_pull_matches_oracle_larger(sys1, orc1, ctx1):
    # 20000 determinants: most connections miss v, many rows are empty
    for n, seed in [(20000, 61), (1, 62)]:
        k, v = random_input(sys1, n, seed)
        vec = ctx1.vector(n, k, v)
        e = ctx1.energy(vec)
        oe = orc1.energy(k, v)
        log("energy n=%d gpu %.15f oracle %.15f" % (n, e, oe))
        assert bits_of(e) == bits_of(oe)
    assert ctx1.energy(ctx1.vector(4)) == 0.0


def test_hv_multipass_heavy_row_group(make_system):
    # every source sits in the four lowest spin-0 rows (18 orbitals, 5 electrons per spin:
    # 8568 rows, planned in groups of 4 rows): the first group alone exceeds the pass
    # capacity, so planning refines it to single rows
    from sfci_amd import gpu
    s = make_system(18, 10, 0)
    strings = []
    for occ in itertools.combinations(range(18), 5):
        v = 0
        for i in occ:
            v |= 1 << i
        strings.append(v)

    def colex(v):
        r, t = 0, 1
        pos = 0
        from math import comb
        while v:
            if v & 1:
                r += comb(pos, t)
                t += 1
            v >>= 1
            pos += 1
        return r

    by_rank = sorted(strings, key=colex)
    spread = lambda v: sum(1 << (2 * i) for i in range(18) if (v >> i) & 1)
    rows = [spread(v) for v in by_rank[:4]]
    cols = [spread(v) << 1 for v in by_rank[::2]]
    keys = np.sort(np.array([a | b for a in rows for b in cols], dtype=np.uint64))
    rng = np.random.default_rng(77)
    vals = rng.normal(size=len(keys)) * np.exp(-rng.uniform(0, 6, len(keys)))
    log("heavy row group: %d sources" % len(keys))
    orc = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    ctx = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    try:
        ctx.set_workspace(320 << 20)
        gk, gv, gd = gpu_hv(ctx, keys, vals, len(keys), 0.0, 1.0, 60_000_000)
        st = ctx.stats()
    finally:
        ctx.close()
    log("heavy row group: %d passes" % st["n_passes"])
    assert st["n_passes"] > 1
    ok, ov, od, _ = orc.hv(keys, vals, nd=len(keys), lam=0.0, coeff=1.0, cut=1e-5)
    assert_same_vec(gk, gv, ok, ov)
    assert bits_of(gd) == bits_of(od)
