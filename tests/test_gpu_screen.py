"""GPU checks of the learned candidate screener: the network's features, forward pass,
gradient and Adam step against the oracle (oracle/screen_oracle.py), training, and the
screened H.v bit for bit against the oracle's screened product."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import screen_oracle as S
from sfci_amd import gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sysS(make_system):
    return make_system(10, 6, 3)


@pytest.fixture(scope="module")
def orcS(sysS):
    return O.Oracle(sysS["in_path"], sysS["n_so"], sysS["n_elec"])


@pytest.fixture(scope="module")
def ctxS(sysS):
    c = gpu.Context(sysS["in_path"], sysS["n_so"], sysS["n_elec"])
    yield c
    c.set_screen(0)
    c.close()


def _rand_samples(n, n_so, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros(n, gpu.SAMPLE_DTYPE)
    s["det"] = rng.integers(0, 1 << n_so, n, dtype=np.uint64)
    s["cls"] = rng.integers(0, 3, n)
    s["logv"] = rng.uniform(-30, 0, n).astype(np.float32)
    s["label"] = rng.normal(-10, 3, n).astype(np.float32)
    return s


def _rand_x(n, seed):
    rng = np.random.default_rng(seed)
    x = np.where(rng.random((n, S.IN)) < 0.5, -1.0, 1.0).astype(np.float32)
    x[:, 64] = rng.integers(0, 3, n)
    x[:, 65] = rng.normal(0, 1, n)
    return x


def test_mlp_features_bit_exact(ctxS, sysS):
    net = gpu.Mlp(ctxS, seed=1)
    net.set_norm(-12.5, 0.25, -9.0, 0.5)
    smp = _rand_samples(3000, sysS["n_so"], 2)
    xg = net.sample_features(smp)
    xo = S.sample_features(smp, sysS["n_so"], -12.5, 0.25)
    assert np.array_equal(xg.view(np.uint32), xo.view(np.uint32))


def test_mlp_forward_matches_oracle(ctxS):
    net = gpu.Mlp(ctxS, seed=3)
    p = net.get_params()
    x = _rand_x(1000, 4)
    yg = net.forward(x)
    yo = S.forward(p, x)
    scale = 1.0 + np.max(np.abs(yo))
    print("forward: max |gpu - oracle| = %.3g (scale %.3g)" % (np.max(np.abs(yg - yo)), scale))
    assert np.max(np.abs(yg - yo)) <= 1e-3 * scale
    assert np.mean(np.abs(yg - yo)) <= 1e-4 * scale
    # samples take the same kernel path with in-kernel features
    smp = _rand_samples(500, 20, 5)
    net.set_norm(-15.0, 0.1, 0.0, 1.0)
    ys = net.predict_samples(smp)
    yo2 = S.forward(p, S.sample_features(smp, 20, -15.0, 0.1))
    assert np.max(np.abs(ys - yo2)) <= 5e-3 * (1.0 + np.max(np.abs(yo2)))


def test_mlp_gradient_matches_oracle(ctxS):
    net = gpu.Mlp(ctxS, seed=7)
    p = net.get_params()
    x = _rand_x(512, 8)
    t = np.random.default_rng(9).normal(0, 1, 512).astype(np.float32)
    gg, lg = net.gradient(x, t)
    go, lo = S.gradient(p, x, t)
    assert abs(lg - lo) <= 1e-3 * abs(lo)
    for name, a, b in [("W1", S.OFF_W1, S.OFF_B1), ("b1", S.OFF_B1, S.OFF_W2), ("W2", S.OFF_W2, S.OFF_B2),
                       ("b2", S.OFF_B2, S.OFF_W3), ("w3", S.OFF_W3, S.OFF_B3), ("b3", S.OFF_B3, S.NPARAM)]:
        rel = np.linalg.norm(gg[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        print("grad %s: relative L2 difference %.3g" % (name, rel))
        assert rel <= 2e-3, name


def test_mlp_adam_step_matches_oracle(ctxS):
    net = gpu.Mlp(ctxS, seed=11)
    p0 = net.get_params()
    g = np.random.default_rng(12).normal(0, 1e-2, S.NPARAM).astype(np.float32)
    net.adam_step(g, 1e-3)
    p1, m, v = S.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    net.adam_step(g * 0.5, 1e-3)
    p2, _, _ = S.adam(p1, m, v, g * 0.5, 2, 1e-3)
    assert np.allclose(net.get_params(), p2, rtol=2e-6, atol=1e-9)


def test_mlp_train_reduces_loss(ctxS, sysS):
    n_so = sysS["n_so"]
    smp = _rand_samples(20000, n_so, 13)
    x = S.sample_features(smp, n_so, -15.0, 0.1)
    smp["label"] = (-2.0 * x[:, 64] + 1.5 * x[:, 65] + x[:, 0] - x[:, 3] + 0.5 * x[:, 5] * x[:, 6]).astype(np.float32)
    net = gpu.Mlp(ctxS, seed=14)
    net.set_norm(-15.0, 0.1, 0.0, 1.0)
    curve = net.train(smp, steps=300, batch=2048, lr=3e-3, seed=15)
    print("training loss: first %.4f last %.4f" % (curve[0], curve[-5:].mean()))
    assert np.all(np.isfinite(curve))
    assert curve[-5:].mean() < 0.05 * curve[0]
    pred = net.predict_samples(smp[:2000])
    assert np.corrcoef(pred, smp["label"][:2000])[0, 1] > 0.98


def _vector(orc, hf, rounds=2):
    k, v = np.array([hf], np.uint64), np.array([1.0])
    for _ in range(rounds):
        k, v, _, _ = orc.hv(k, v, cut=1e-5)
        v = v / np.linalg.norm(v)
    return k, v


@pytest.mark.parametrize("case", [(0.0, 1.0, None, 1e-4, 1.0), (0.3, -1.7, 60, 1e-5, 20.0), (0.0, 1.0, None, 0.0, 0.5)])
def test_screened_hv_magnitude_bit_exact(case, sysS, orcS, ctxS):
    lam, coeff, nd, eps, keep = case
    k, v = _vector(orcS, sysS["init_hf"])
    nd = nd or len(k)
    ctxS.set_screen(1, None, keep, eps)
    try:
        vin = ctxS.vector(len(k), k, v)
        vout = ctxS.vector(200_000)
        d = ctxS.hv(vin, vout, nd, lam, coeff)
        st = ctxS.stats()
    finally:
        ctxS.set_screen(0)
    gk, gv = vout.download()
    ok, ov, od = S.screened_hv(orcS, k, v, nd, lam, coeff, eps, keep, lambda t, c, r: np.abs(c), cut=1e-5,
                               out_cap=200_000)
    print("screened: %d emitted, %d kept, %d out" % (st["n_candidates"], st["n_kept"], st["n_out"]))
    assert st["n_kept"] < st["n_candidates"]
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv.view(np.uint64), ov.view(np.uint64))
    assert d == od


def test_screened_hv_unscreened_equals_mixed(sysS, orcS, ctxS):
    k, v = _vector(orcS, sysS["init_hf"])
    vin = ctxS.vector(len(k), k, v)
    a, b = ctxS.vector(200_000), ctxS.vector(200_000)
    ctxS.set_precision(True)
    try:
        d0 = ctxS.hv(vin, a, len(k), 0.2, 1.1)
        ctxS.set_screen(1, None, 1e9, 0.0)
        d1 = ctxS.hv(vin, b, len(k), 0.2, 1.1)
    finally:
        ctxS.set_screen(0)
        ctxS.set_precision(False)
    ka, va = a.download()
    kb, vb = b.download()
    assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint64), vb.view(np.uint64)) and d0 == d1


def test_screened_hv_network_bit_exact(sysS, orcS, ctxS):
    """Network screening: the kept set is the one the network's own outputs select (the
    outputs taken from the same kernel on the oracle's candidates), merged exactly."""
    k, v = _vector(orcS, sysS["init_hf"])
    net = gpu.Mlp(ctxS, seed=21)
    net.set_norm(-14.0, 0.125, 0.0, 1.0)
    nd, eps, keep = len(k), 1e-5, 0.7
    ctxS.set_screen(2, net, keep, eps)
    try:
        vin = ctxS.vector(len(k), k, v)
        vout = ctxS.vector(200_000)
        d = ctxS.hv(vin, vout, nd, 0.0, 1.0)
        st = ctxS.stats()
    finally:
        ctxS.set_screen(0)

    def scores(t, c, r):
        smp = np.zeros(len(t), gpu.SAMPLE_DTYPE)
        smp["det"], smp["cls"], smp["logv"] = t, r, S.plog2(c)
        return net.predict_samples(smp)

    ok, ov, od = S.screened_hv(orcS, k, v, nd, 0.0, 1.0, eps, keep, scores, cut=1e-5, out_cap=200_000)
    gk, gv = vout.download()
    assert st["n_kept"] < st["n_candidates"]
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv.view(np.uint64), ov.view(np.uint64))
    assert d == od


def test_screen_samples_labels(sysS, orcS, ctxS):
    k, v = _vector(orcS, sysS["init_hf"])
    vin = ctxS.vector(len(k), k, v)
    s1 = ctxS.screen_samples(vin, 1e-5, 3000, seed=5)
    s2 = ctxS.screen_samples(vin, 1e-5, 3000, seed=5)
    assert 0 < len(s1) <= 3000 and np.array_equal(s1, s2)
    t, c, r = S.screen_candidates(orcS, k, v, len(k), 1.0, 1e-5)
    cand = set(zip(t.tolist(), S.plog2(c).view(np.uint32).tolist(), r.tolist()))
    assert all((int(a), int(b), int(cc)) in cand
               for a, b, cc in zip(s1["det"], s1["logv"].view(np.uint32), s1["cls"]))
    yk, yv, _, _ = orcS.hv(k, v, cut=-1.0)
    pos = np.searchsorted(yk, s1["det"])
    assert np.array_equal(yk[pos], s1["det"])
    exp = np.log(np.maximum(np.abs(yv[pos]), 1e-30)).astype(np.float32)
    assert np.allclose(s1["label"], exp, rtol=1e-6, atol=0)


@pytest.mark.parametrize("mode", [1, 2])
def test_screened_lanczos_judged(mode, sysS, ctxS):
    """The screened solve: every restart judged by the unscreened energy readout, which never
    rises over accepted restarts; a settled screened solve is confirmed by an unscreened
    restart (an untrained network's stall is caught there); the screening setting survives
    the call."""
    hf = sysS["init_hf"]
    r0 = ctxS.ground_state(hf, 300, 8, 6, hpsi_cap=200_000, e_prev_init=float("inf"))
    net = None
    if mode == 2:
        net = gpu.Mlp(ctxS, seed=31)
        net.set_norm(-14.0, 0.125, 0.0, 1.0)
    ctxS.set_screen(mode, net, 4.0, 1e-6)
    try:
        r1 = ctxS.ground_state(hf, 300, 8, 12, hpsi_cap=200_000, e_prev_init=float("inf"))
        assert ctxS.get_screen()["mode"] == mode
    finally:
        ctxS.set_screen(0)
    print("plain E %.10f (%d restarts); screened E %.10f (%d restarts, %d screened, %d rejected)" % (
        r0["e_final"], r0["restarts"], r1["e_final"], r1["restarts"], r1["n_screened"], r1["n_reverts"]))
    assert r1["n_screened"] >= 1
    et = r1["e_trunc"]
    accepted = et[: r1["n_screened"]]
    assert np.all(np.diff(accepted) <= 1e-9)
    assert r1["e_final"] <= et[0] + 1e-9
    assert r1["restarts"] > r1["n_screened"]          # at least the confirming unscreened restart
    assert abs(r1["e_final"] - r0["e_final"]) < 1e-4
