"""Pinning the oracle's 128-bit-key H.v (more than 64 spin orbitals).

The reference stops at 64 spin orbitals, so nothing of its own pins these functions.
Independent routes:
  * on a system of 24 spin orbitals the 128-bit functions must reproduce the 64-bit
    oracle (itself pinned by tests/test_oracle.py) bit for bit, in the exact and the mixed
    mode, with top-Nd and top-cap truncation;
  * on a system of 66 spin orbitals, the connections of determinants that straddle the
    64-bit word boundary against the second-quantized brute force (tests/sq_bruteforce.py,
    arbitrary-width Python integers), and a whole H.v against the brute-force sum
    sum_j c_j coeff (H - lambda) |d_j>.
"""
from math import comb

import numpy as np
import pytest

import sq_bruteforce as B
from oracle import oracle as O
from sfci_amd import gpu as G
from sfci_amd import integrals as I

MASK = 0xFFFFFFFFFFFFFFFF


@pytest.fixture(scope="module")
def wide_sys(make_system):
    # 33 spatial orbitals = 66 spin orbitals: spin orbitals 64 and 65 live in the high word
    return make_system(33, 4, 7)


def test_wide_oracle_equals_64bit_oracle(make_system):
    s = make_system(12, 8, 0)
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    k1, v1, _, _ = o.hv(np.array([s["init_hf"]], np.uint64), np.array([1.0]))
    k2, v2, _, _ = o.hv(k1, v1)
    zero = np.zeros(len(k2), np.uint64)
    for mode in (0, 2):
        for nd, cap, lam, coeff in [(len(k2), 60_000_000, 0.0, 1.0), (500, 3000, 0.37, -1.25)]:
            ok, ov, od, _ = o.hv(k2, v2, nd=nd, lam=lam, coeff=coeff, out_cap=cap, mode=mode)
            wl, wh, wv, wd = o.hv_wide(k2, zero, v2, nd=nd, lam=lam, coeff=coeff, out_cap=cap, mode=mode)
            assert np.array_equal(ok, wl) and not wh.any()
            assert np.array_equal(ov.view(np.uint64), wv.view(np.uint64))
            assert od == wd
    # connections, diagonal
    d = int(k2[7])
    ck, cv = o.connections(d)
    wl, wh, wv = o.connections_wide(d)
    assert np.array_equal(ck, wl) and not wh.any() and np.array_equal(cv, wv)
    assert o.diag(d) == o.diag_wide(d)


def _straddling_dets(n_orb, k, rng, count):
    """Determinants with k electrons per spin, each with an electron above spin orbital 63."""
    out = []
    while len(out) < count:
        a = sorted(rng.choice(n_orb, size=k, replace=False))
        b = sorted(rng.choice(n_orb, size=k, replace=False))
        d = 0
        for i in a:
            d |= 1 << (2 * int(i))
        for j in b:
            d |= 1 << (2 * int(j) + 1)
        if d >> 64 and d & MASK:
            out.append(d)
    return sorted(set(out))


def test_wide_connections_vs_second_quantized(wide_sys):
    s = wide_sys
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    n_orb, k = s["n_orb"], s["n_elec"] // 2
    v = n_orb - k
    n_conn = 1 + 2 * k * v + 2 * comb(k, 2) * comb(v, 2) + (k * v) ** 2
    rng = np.random.default_rng(3)
    dets = _straddling_dets(n_orb, k, rng, 6)
    # fixed cases: an electron in spin orbital 64 / 65, and one at 63 (the word boundary)
    dets += [1 | (1 << 64) | (1 << 3) | (1 << 65), (1 << 63) | (1 << 62) | (1 << 64) | (1 << 1)]
    for d in dets:
        lo, hi, val = o.connections_wide(d)
        assert len(lo) == n_conn
        got = {}
        for a, b, x in zip(lo, hi, val):
            key = (int(b) << 64) | int(a)
            assert key not in got
            got[key] = float(x)
        assert next(iter(got)) == d          # the diagonal comes first
        ref = B.apply_h(h, eri, ec, d, n_orb)
        for t in set(got) | set(ref):
            assert abs(got.get(t, 0.0) - ref.get(t, 0.0)) < 1e-12, (hex(d), hex(t))
        assert abs(o.diag_wide(d) - ref[d]) < 1e-12


def test_wide_hv_vs_bruteforce_sum(wide_sys):
    s = wide_sys
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    n_orb, k = s["n_orb"], s["n_elec"] // 2
    rng = np.random.default_rng(11)
    dets = _straddling_dets(n_orb, k, rng, 5) + [s["init_hf"]]
    dets = sorted(set(dets), key=lambda x: (x >> 64, x & MASK))
    c = rng.normal(size=len(dets))
    lam, coeff = 0.41, -0.8
    lo, hi = G.split_keys(dets)
    ol, oh, ov, od = o.hv_wide(lo, hi, c, lam=lam, coeff=coeff, cut=0.0)
    y = {}
    for d, cj in zip(dets, c):
        for t, hv in B.apply_h(h, eri, ec, d, n_orb).items():
            y[t] = y.get(t, 0.0) + coeff * cj * hv
        y[d] = y.get(d, 0.0) - coeff * cj * lam
    got = dict(zip(G.join_keys(ol, oh), ov))
    keys = G.join_keys(ol, oh)
    assert keys == sorted(keys)
    for t in set(got) | {t for t, x in y.items() if x != 0.0}:
        assert abs(got.get(t, 0.0) - y.get(t, 0.0)) < 1e-11
    diag = sum(cj * y[d] for d, cj in zip(dets, c))
    assert abs(od - diag) < 1e-10
