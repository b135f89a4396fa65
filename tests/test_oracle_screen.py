"""CPU checks of the screener's oracle (oracle/screen_oracle.py): each function against an
independent route (torch's bfloat16 conversion, log2, central finite differences, Adam's
closed-form first step, brute-force selection, the already-pinned mixed-precision H.v)."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import screen_oracle as S


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 1e-3, 2000) * 7, [0.0, -0.0, 1.0, 65504.0,
                                                                                  1.00390625, 1.01171875]])
    x = x.astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(S.bf16(x).view(np.uint32), ref.view(np.uint32))


def test_plog2_is_piecewise_linear_log2():
    v = np.float32(2.0) ** np.arange(-40, 20, dtype=np.float32)
    assert np.array_equal(S.plog2(v), np.arange(-40, 20, dtype=np.float32))     # exact at powers of two
    rng = np.random.default_rng(1)
    x = np.sort(np.abs(rng.normal(0, 1, 50000)).astype(np.float32) * np.float32(1e-3))
    p = S.plog2(x)
    assert np.all(np.diff(p) >= 0)                                               # monotone
    err = np.log2(x.astype(np.float64)) - p
    assert err.min() >= -1e-6 and err.max() <= 0.0861                          # max of log2(1+m) - m
    assert S.plog2(np.array([0.0], np.float32))[0] == -150.0
    # one IEEE rounding of exponent + fraction: same as the float32 sum done by hand
    b = x.view(np.uint32)
    e = (b >> 23).astype(np.int64) - 127
    m = (b & 0x7FFFFF).astype(np.float64) / 8388608.0
    assert np.array_equal(p, (e + m).astype(np.float32))


def test_features_layout():
    det = np.array([(1 << 0) | (1 << 3) | (1 << 63)], np.uint64)
    x = S.features(det, [2], [np.float32(-13.25)], 64, -10.0, 0.5)
    assert x.shape == (1, 66)
    occ = np.full(64, -1.0, np.float32)
    occ[[0, 3, 63]] = 1.0
    assert np.array_equal(x[0, :64], occ)
    assert x[0, 64] == 2.0 and x[0, 65] == np.float32(-1.625)
    x2 = S.features(det, [1], [0.0], 10, 0.0, 1.0)                              # n_so < 64: zeros above
    assert np.all(x2[0, 10:64] == 0.0)


def _tiny_params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.1, S.NPARAM).astype(np.float32)
    return p


def test_gradient_vs_finite_differences():
    """With bf16 rounding off, the analytic gradient equals central differences of the
    float64 loss (pins every term of the chain rule)."""
    rng = np.random.default_rng(2)
    p = _tiny_params(3).astype(np.float64)
    x = rng.normal(0, 1, (12, S.IN))
    t = rng.normal(0, 1, 12)
    g, _ = S.gradient(p, x, t, round_bf16=False)
    idx = np.concatenate([rng.integers(S.OFF_W1, S.OFF_B1, 12), rng.integers(S.OFF_B1, S.OFF_W2, 4),
                          rng.integers(S.OFF_W2, S.OFF_B2, 12), rng.integers(S.OFF_B2, S.OFF_W3, 4),
                          rng.integers(S.OFF_W3, S.OFF_B3, 6), [S.OFF_B3]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = np.mean((S.forward(pp, x, round_bf16=False) - t) ** 2)
        lm = np.mean((S.forward(pm, x, round_bf16=False) - t) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_forward_rounding_changes_little():
    rng = np.random.default_rng(4)
    p = _tiny_params(5)
    x = rng.normal(0, 1, (64, S.IN)).astype(np.float32)
    y0 = S.forward(p, x, round_bf16=False)
    y1 = S.forward(p, x, round_bf16=True)
    assert np.max(np.abs(y0 - y1)) < 0.05 * (np.max(np.abs(y0)) + 1)


def test_adam_first_step_closed_form():
    """After one step from zero moments, m/c1 = g and v/c2 = g^2: p1 = p0 - lr g / (|g| + eps)."""
    rng = np.random.default_rng(6)
    p0 = rng.normal(0, 1, 1000).astype(np.float32)
    g = rng.normal(0, 1, 1000).astype(np.float32)
    p1, m, v = S.adam(p0, np.zeros(1000), np.zeros(1000), g, 1, lr=1e-2)
    exp = p0.astype(np.float64) - 1e-2 * g / (np.abs(g) + 1e-8)
    assert np.allclose(p1, exp, rtol=0, atol=2e-6)
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g * g, rtol=1e-4)   # (1 - b) in float32


def test_keep_mask_brute_force():
    rng = np.random.default_rng(7)
    s = rng.integers(0, 50, 400).astype(np.float32)      # many ties
    for n_src, keep in [(3, 10), (5, 100), (0, 1), (10, 500), (7, 8)]:
        m = S.keep_mask(s, n_src, keep)
        room = keep - n_src
        if room >= len(s):
            assert m.all()
            continue
        thr = sorted(s, reverse=True)[room - 1]
        assert np.array_equal(m, s >= thr)
        assert m.sum() >= room
        assert (s[~m] < s[m].min()).all()


def test_screen_unscreened_equals_mixed_hv(make_system):
    """eps = 0 and nothing dropped: the screened merge is the mixed-precision H.v."""
    sy = make_system(8, 6, 0)
    orc = O.Oracle(sy["in_path"], sy["n_so"], sy["n_elec"])
    hf = np.array([sy["init_hf"]], np.uint64)
    k1, v1, _, _ = orc.hv(hf, np.array([1.0]), cut=0.0)
    for lam, coeff, nd in [(0.0, 1.0, len(k1)), (0.7, -1.3, 40)]:
        ok, ov, od, _ = orc.hv(k1, v1, nd=nd, lam=lam, coeff=coeff, cut=1e-5, mode=2)
        sk, sv, sd = S.screened_hv(orc, k1, v1, nd, lam, coeff, 0.0, 1e9, lambda t, v, c: np.abs(v))
        assert np.array_equal(sk, ok)
        assert np.array_equal(sv.view(np.uint64), ov.view(np.uint64))
        assert sd == od


def test_screen_candidates_prefilter_and_ranks(make_system):
    sy = make_system(8, 6, 0)
    orc = O.Oracle(sy["in_path"], sy["n_so"], sy["n_elec"])
    hf = sy["init_hf"]
    k1, v1, _, _ = orc.hv(np.array([hf], np.uint64), np.array([1.0]), cut=0.0)
    t, v, c = S.screen_candidates(orc, k1, v1, len(k1), 1.0, 1e-4)
    assert np.all(np.abs(v) > np.float32(1e-4))
    # brute force over the connections of every source
    exp = []
    for d, cj in zip(k1, v1):
        kk, hh = orc.connections(int(d))
        for tt, h in zip(kk[1:], hh[1:]):
            val = np.float32(h) * np.float32(cj)
            if abs(val) > np.float32(1e-4):
                exp.append((int(tt), val, bin(int(tt) ^ int(d)).count("1") // 2))
    assert len(exp) == len(t)
    assert [e[0] for e in exp] == [int(a) for a in t]
    assert np.array_equal(np.array([e[1] for e in exp], np.float32), v)
    assert [e[2] for e in exp] == [int(a) for a in c]
