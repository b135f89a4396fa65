"""Pinning the oracle (CPU restatement) before it is trusted as the GPU's checker.

Independent routes used here (none re-types the oracle's own formula):
  * second-quantized brute force with Jordan-Wigner signs built straight from the
    spatial integrals of the FCIDUMP (tests/sq_bruteforce.py)  -> Slater-Condon, the
    FCIDUMP -> _IN converter and the _IN loader together;
  * the reference's literal emission (one candidate per integral-list entry, summed in
    doubles) against the unique-emission exact mode            -> H.v within 1e-12;
  * closed forms: connection counts U and G, HF energy, Brillouin/Fock single elements;
  * symmetries: Hermiticity, spin flip;
  * numpy / brute-force restatements of dot, norm, addTwo, shrinkFrom and recombination
    with the declared tie rule;
  * residuals of the dense generalized eigensolver;
  * full-space Lanczos (cut 0, Nd >= dim) against the exact lowest eigenvalue.
"""
import itertools
from math import comb

import numpy as np
import pytest

import sq_bruteforce as B
from oracle import oracle as O
from sfci_amd import integrals as I


def orc_of(s):
    return O.Oracle(s["in_path"], s["n_so"], s["n_elec"])


@pytest.mark.parametrize("n_orb,n_elec,seed", [(4, 4, 0), (5, 4, 7), (4, 6, 3), (5, 6, 2)])
def test_slater_condon_vs_second_quantized(make_system, n_orb, n_elec, seed):
    s = make_system(n_orb, n_elec, seed)
    o = orc_of(s)
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    dets, M = B.dense_h(h, eri, ec, n_orb, n_elec // 2)
    idx = {d: i for i, d in enumerate(dets)}
    for j, d in enumerate(dets):
        k, v = o.connections(d)
        col = np.zeros(len(dets))
        for kk, vv in zip(k, v):
            col[idx[int(kk)]] += vv
        assert np.abs(col - M[:, j]).max() < 1e-12
        # literal emission + literal diagonal
        k2, v2 = o.offdiag_literal(d)
        col2 = np.zeros(len(dets))
        col2[j] = o.diag_literal(d)
        for kk, vv in zip(k2, v2):
            col2[idx[int(kk)]] += vv
        assert np.abs(col2 - M[:, j]).max() < 1e-12
        assert abs(o.diag(d) - M[j, j]) < 1e-12


def test_connection_counts_closed_form(make_system):
    s = make_system(12, 8, 0)
    o = orc_of(s)
    k, v = 4, 8
    S, Dss, Dos = 2 * k * v, 2 * comb(k, 2) * comb(v, 2), (k * v) ** 2
    U = S + Dss + Dos
    G = S * (3 * k - 1) + 2 * Dss + Dos + 1
    assert U == 1424 and G == 2401
    for d in [s["init_hf"], 0b101101011001011010, 0b11110000110000110000]:
        if bin(d & 0x555555).count("1") != 4 or bin(d & 0xAAAAAA).count("1") != 4:
            continue
        kk, _ = o.connections(d)
        assert len(kk) == U + 1
        assert len(np.unique(kk)) == U + 1
        kl, _ = o.offdiag_literal(d)
        assert len(kl) + 1 == G


def test_hf_energy_and_fock_closed_form(make_system):
    s = make_system(8, 6, 4)
    o = orc_of(s)
    h, eri, ec, n = I.read_fcidump(s["fcidump_path"])
    occ = range(3)
    e_hf = ec + 2 * sum(h[i, i] for i in occ) + sum(2 * eri[i, i, j, j] - eri[i, j, j, i] for i in occ for j in occ)
    assert abs(o.diag(s["init_hf"]) - e_hf) < 1e-12
    # <HF|H|S_i^a> = f_ai = h_ai + sum_j [2 (ai|jj) - (aj|ji)]  (single on spin 0)
    hf = s["init_hf"]
    k, v = o.connections(hf)
    amp = dict(zip(k.tolist(), v.tolist()))
    for i in occ:
        for a in range(3, n):
            f = h[a, i] + sum(2 * eri[a, i, j, j] - eri[a, j, j, i] for j in occ)
            t = hf - (1 << (2 * i)) + (1 << (2 * a))
            sign = -1 if bin(hf & ((1 << (2 * a)) - 1) & ~((1 << (2 * i + 1)) - 1)).count("1") & 1 else 1
            assert abs(amp[t] - sign * f) < 1e-12


def test_hermitian_and_spin_flip(make_system):
    s = make_system(12, 8, 0)
    o = orc_of(s)
    rng = np.random.default_rng(0)
    hf = s["init_hf"]
    k, _ = o.connections(hf)
    sample = [hf] + [int(x) for x in rng.choice(k, 6, replace=False)]

    def flip(d):
        a = d & 0x555555555555
        b = (d >> 1) & 0x555555555555
        return (a << 1) | b

    for d in sample:
        kd, vd = o.connections(d)
        row = dict(zip(kd.tolist(), vd.tolist()))
        for t in kd[:: max(1, len(kd) // 40)]:
            kt, vt = o.connections(int(t))
            back = dict(zip(kt.tolist(), vt.tolist()))
            assert abs(back[d] - row[int(t)]) < 1e-13
        kf, vf = o.connections(flip(d))
        fl = dict(zip(kf.tolist(), vf.tolist()))
        for t, val in list(row.items())[::50]:
            assert abs(abs(fl[flip(t)]) - abs(val)) < 1e-13


def _rand_vec(s, n, seed):
    rng = np.random.default_rng(seed)
    o = orc_of(s)
    hf = s["init_hf"]
    k1, _ = o.connections(hf)
    pool = set(k1.tolist())
    for t in k1[::97]:
        pool.update(o.connections(int(t))[0].tolist())
    pool = np.array(sorted(pool), dtype=np.uint64)
    pick = np.sort(rng.choice(len(pool), size=min(n, len(pool)), replace=False))
    return pool[pick], rng.normal(size=len(pick)) * np.exp(-rng.uniform(0, 6, len(pick)))


def test_hv_exact_vs_literal_reference_arithmetic(make_system):
    s = make_system(12, 8, 0)
    o = orc_of(s)
    k, v = _rand_vec(s, 1500, 1)
    ok, ov, od, nce = o.hv(k, v, lam=0.2, coeff=0.9, cut=1e-5, mode=0)
    lk, lv, ld, ncl = o.hv(k, v, lam=0.2, coeff=0.9, cut=1e-5, mode=1)
    assert nce == len(k) * 1425 and ncl == len(k) * 2401
    # sets agree except for values inside the 1e-12 band around the cut
    common = np.intersect1d(ok, lk)
    only = np.setxor1d(ok, lk)
    allk = np.concatenate([ok, lk])
    allv = np.concatenate([ov, lv])
    for t in only:
        val = np.abs(allv[np.nonzero(allk == t)[0][0]])
        assert abs(val - 1e-5) <= 1e-12 * 1e-5 * 10
    ia, ib = np.searchsorted(ok, common), np.searchsorted(lk, common)
    rel = np.abs(ov[ia] - lv[ib]) / np.abs(lv[ib])
    assert rel.max() < 1e-12
    assert abs(od - ld) <= 1e-12 * abs(ld)


def test_energy_readout_equals_hv_diag(make_system):
    """The reference's two energy paths agree: calc_energy's per-source pull
    (src/functions.cpp:1272-1313) and the diag output of H.v with lambda 0, coeff 1 and all
    of v kept (src/functions.cpp:674-677); both sums are exact, rounded differently."""
    s = make_system(12, 8, 0)
    o = orc_of(s)
    k1, v1, _, _ = o.hv(np.array([s["init_hf"]], np.uint64), np.array([1.0]))
    k2, v2, _, _ = o.hv(k1, v1)
    v2 = v2 / np.sqrt(v2 @ v2)
    e = o.energy(k2, v2)
    _, _, d, _ = o.hv(k2, v2, nd=len(k2), lam=0.0, coeff=1.0, cut=0.0)
    assert abs(e - d) <= 4e-15 * abs(e)


def test_hv_matches_dense_tiny(make_system):
    s = make_system(5, 4, 7)
    o = orc_of(s)
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    dets, M = B.dense_h(h, eri, ec, 5, 2)
    rng = np.random.default_rng(2)
    vals = rng.normal(size=len(dets))
    keys = np.array(dets, dtype=np.uint64)
    ok, ov, od, _ = o.hv(keys, vals, lam=0.3, coeff=-0.7, cut=0.0)
    y = -0.7 * (M @ vals - 0.3 * vals)
    got = dict(zip(ok.tolist(), ov.tolist()))
    for i, d in enumerate(dets):
        assert abs(got.get(d, 0.0) - y[i]) < 1e-12
    assert abs(od - vals @ y) < 1e-10
    assert abs(o.energy(keys, vals) - vals @ (M @ vals)) < 1e-10


def test_hv_topnd_input_and_cap_tie_rule(make_system):
    s = make_system(5, 4, 7)
    o = orc_of(s)
    dets = np.array(B.sz0_dets(5, 2), dtype=np.uint64)
    vals = np.array([0.5, -0.5, 0.25] * 40)[: len(dets)]
    nd = 37
    # the top-37 by (|c| desc, key asc): all |0.5| first (smallest keys among ties)
    order = sorted(range(len(dets)), key=lambda i: (-abs(vals[i]), int(dets[i])))[:nd]
    kept = np.sort(dets[order])
    kv = vals[np.searchsorted(dets, kept)]
    a = o.hv(dets, vals, nd=nd, cut=0.0)
    b = o.hv(kept, kv, cut=0.0)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
    full = o.hv(dets, vals, cut=0.0)
    cap = 20
    c = o.hv(dets, vals, cut=0.0, out_cap=cap)
    order = sorted(range(len(full[0])), key=lambda i: (-abs(full[1][i]), int(full[0][i])))[:cap]
    assert np.array_equal(c[0], np.sort(full[0][order]))


def _brute_top(k, v, cap):
    order = sorted(range(len(k)), key=lambda i: (-abs(v[i]), int(k[i])))[:cap]
    order = sorted(order, key=lambda i: int(k[i]))
    return np.array([k[i] for i in order], dtype=np.uint64), np.array([v[i] for i in order])


def test_vector_ops_vs_bruteforce():
    rng = np.random.default_rng(5)
    ak = np.sort(rng.choice(10_000, 600, replace=False)).astype(np.uint64)
    bk = np.sort(rng.choice(10_000, 900, replace=False)).astype(np.uint64)
    av = rng.choice([0.5, -0.5, 0.25, 0.0, 1.0], size=600)
    bv = rng.normal(size=900)
    da = dict(zip(ak.tolist(), av.tolist()))
    db = dict(zip(bk.tolist(), bv.tolist()))
    assert abs(O.dot(ak, av, bk, bv) - sum(da[t] * db[t] for t in da if t in db)) < 1e-13
    assert abs(O.norm(bv) - np.sqrt(np.sum(bv * bv))) < 1e-13
    for cap in [2000, 1000, 300]:
        merged = dict(da)
        for t, x in db.items():
            merged[t] = merged.get(t, 0.0) + (-0.3) * x if t in merged else (-0.3) * x
        mk = np.array(sorted(merged), dtype=np.uint64)
        mv = np.array([merged[int(t)] for t in mk])
        ek, ev = _brute_top(mk, mv, cap)
        gk, gv = O.add_two(ak, av, bk, bv, -0.3, cap)
        assert np.array_equal(gk, ek) and np.allclose(gv, ev, rtol=0, atol=0)
        sk, sv = O.shrink(ak, av, cap // 4)
        ek, ev = _brute_top(ak, av, cap // 4)
        assert np.array_equal(sk, ek) and np.array_equal(sv, ev)
    c = np.array([0.7, -1.3])
    lk, lv = O.lincomb([(ak, av), (bk, bv)], c, 800)
    comb = {}
    for (kk, vv), ci in zip([(ak, av), (bk, bv)], c):
        for t, x in zip(kk.tolist(), vv.tolist()):
            comb[t] = comb.get(t, 0.0) + ci * x
    ck = np.array(sorted(comb), dtype=np.uint64)
    cv = np.array([comb[int(t)] for t in ck])
    ek, ev = _brute_top(ck, cv, 800)
    ev = ev / np.sqrt(np.sum(ev * ev))
    assert np.array_equal(lk, ek) and np.allclose(lv, ev, rtol=1e-14, atol=1e-15)


def test_gsaes_residuals():
    rng = np.random.default_rng(9)
    for n in [1, 3, 20, 40]:
        A = rng.normal(size=(n, n))
        K = A + A.T
        Bm = rng.normal(size=(n, n))
        S = Bm @ Bm.T + n * np.eye(n)
        ev, X = O.gsaes(K, S)
        assert np.all(np.diff(ev) >= 0)
        R = K @ X - S @ X @ np.diag(ev)
        assert np.abs(R).max() < 1e-10 * max(1, np.abs(K).max())
        assert np.allclose(X.T @ S @ X, np.eye(n), atol=1e-10)
        import scipy.linalg
        assert np.allclose(ev, scipy.linalg.eigh(K, S, eigvals_only=True), atol=1e-10)


@pytest.mark.parametrize("n_orb,n_elec,seed", [(5, 4, 7), (6, 6, 1)])
def test_lanczos_full_space_reaches_fci(make_system, n_orb, n_elec, seed):
    s = make_system(n_orb, n_elec, seed)
    o = orc_of(s)
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    dets, M = B.dense_h(h, eri, ec, n_orb, n_elec // 2)
    e0 = np.linalg.eigvalsh(M)[0]
    r = o.lanczos(s["init_hf"], nd=len(dets), nl=12, restarts=6, cut=0.0)
    assert abs(r["e_final"] - e0) < 1e-9
    assert abs(r["ritz"][-1] - e0) < 1e-9


def _dyadic_system(tmp_path, n_orb, n_elec, seed):
    """A Hamiltonian whose integrals are multiples of 1/64: every Slater-Condon element
    is then a short dyadic number, exactly representable in single precision."""
    h, canon, ecore = I.synth_integrals(n_orb, seed)
    h = np.round(h * 64) / 64
    canon = [(round(v * 64) / 64, i, j, k, l) for v, i, j, k, l in canon]
    fc, inp = str(tmp_path / "dy_fcidump"), str(tmp_path / "dy_IN")
    I.write_fcidump(fc, h, canon, n_elec, round(ecore * 64) / 64)
    I.convert_rhf(fc, inp)
    return inp


def test_hv_mixed_mode_exact_on_single_precision_data(tmp_path):
    # When every H_ij and every coeff*c_j is exact in single precision, the mixed
    # candidates equal the FP64 ones and both modes give the same bits.
    inp = _dyadic_system(tmp_path, 8, 6, 5)
    o = O.Oracle(inp, 16, 6)
    dets = np.array(B.sz0_dets(8, 3), dtype=np.uint64)
    rng = np.random.default_rng(4)
    pick = np.sort(rng.choice(len(dets), size=600, replace=False))
    k = dets[pick]
    v = rng.integers(-16, 17, size=len(k)) / 16.0
    a = o.hv(k, v, lam=0.25, coeff=0.5, cut=1e-5, mode=0)
    b = o.hv(k, v, lam=0.25, coeff=0.5, cut=1e-5, mode=2)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
    assert a[2] == b[2]
    # a value that is not float-exact does change the mixed result
    v2 = v.copy()
    v2[0] = 0.1
    c = o.hv(k, v2, lam=0.25, coeff=0.5, cut=0.0, mode=0)
    d = o.hv(k, v2, lam=0.25, coeff=0.5, cut=0.0, mode=2)
    assert not np.array_equal(c[1], d[1])


def test_hv_mixed_mode_error_bound(make_system):
    # Each mixed candidate carries at most ~1.5 ulp(float) relative error, so every
    # output differs from FP64 by at most 2^-22 * sum_j |H_ij coeff c_j|.
    s = make_system(5, 4, 7)
    o = orc_of(s)
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    dets, M = B.dense_h(h, eri, ec, 5, 2)
    rng = np.random.default_rng(9)
    vals = rng.normal(size=len(dets))
    keys = np.array(dets, dtype=np.uint64)
    a = o.hv(keys, vals, lam=0.3, coeff=-0.7, cut=0.0, mode=0)
    b = o.hv(keys, vals, lam=0.3, coeff=-0.7, cut=0.0, mode=2)
    bound = 2.0 ** -22 * (np.abs(M - np.diag(np.diag(M))) @ np.abs(0.7 * vals))
    ya = dict(zip(a[0].tolist(), a[1].tolist()))
    yb = dict(zip(b[0].tolist(), b[1].tolist()))
    diff = [abs(ya.get(d, 0.0) - yb.get(d, 0.0)) - bound[i] for i, d in enumerate(dets)]
    assert max(diff) <= 1e-300
    assert any(ya[d] != yb[d] for d in ya)
