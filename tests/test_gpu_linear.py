"""GPU parity for the linear-solver path: the uncut projected product
(spmspvm_bucket_linear), the residual product (spmspvm_bucket_res_norm), addThree, the
cut-then-cap shrink, the restarted Krylov solver, shift-invert, MINRES and CISD, against
the oracle.  Products and vector operations are bit-exact; the CISD eigenvector comes
from different dense eigensolvers (rocSOLVER dsyevd on the GPU, Jacobi in the oracle) and
is compared within 1e-9."""
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


def same(k1, v1, k2, v2):
    assert len(k1) == len(k2), (len(k1), len(k2))
    assert np.array_equal(k1, k2)
    d = np.nonzero(bits(v1) != bits(v2))[0]
    assert len(d) == 0, "first mismatch at %d: %r vs %r" % (d[0], v1[d[0]], v2[d[0]])


@pytest.fixture(scope="module")
def sysA(make_system):
    return make_system(12, 8, 0)


@pytest.fixture(scope="module")
def orcA(sysA):
    return O.Oracle(sysA["in_path"], sysA["n_so"], sysA["n_elec"])


@pytest.fixture(scope="module")
def ctxA(sysA):
    from sfci_amd import gpu
    log("linear: ctx %s" % sysA["in_path"])
    c = gpu.Context(sysA["in_path"], sysA["n_so"], sysA["n_elec"])
    yield c
    c.close()


@pytest.fixture(scope="module")
def sysB(make_system):
    return make_system(8, 6, 3)


@pytest.fixture(scope="module")
def orcB(sysB):
    return O.Oracle(sysB["in_path"], sysB["n_so"], sysB["n_elec"])


@pytest.fixture(scope="module")
def ctxB(sysB):
    from sfci_amd import gpu
    c = gpu.Context(sysB["in_path"], sysB["n_so"], sysB["n_elec"])
    yield c
    c.close()


def krylov_like(orc, hf, n_vec, nd, seed):
    """Normalized vectors grown from HF by H.v (the shape Krylov vectors have)."""
    rng = np.random.default_rng(seed)
    k, v = np.array([hf], np.uint64), np.array([1.0])
    out = []
    for i in range(n_vec):
        ok, ov, _, _ = orc.hv(k, v, lam=0.0, coeff=1.0, cut=1e-5, out_cap=nd)
        ov = ov * (1.0 + 0.1 * rng.normal(size=len(ov)))
        ov /= np.linalg.norm(ov)
        out.append((ok, ov))
        k, v = ok, ov
    return out


def test_hv_linear_bit_exact(sysA, orcA, ctxA):
    basis = krylov_like(orcA, sysA["init_hf"], 4, 3000, 1)
    lam = -3.7
    dv = [ctxA.vector(3000, k, v) for k, v in basis]
    for cap in (3000, 500):
        log("hv_linear cap %d" % cap)
        out = ctxA.vector(cap)
        hij = ctxA.hv_linear(dv[-1], out, lam, dv)
        gk, gv = out.download()
        ok, ov, oh = orcA.hv_linear(*basis[-1], lam, basis, cut=1e-5, out_cap=cap)
        same(gk, gv, ok, ov)
        assert np.array_equal(bits(hij), bits(oh))
    # a single basis vector (the first product of a restart) and an empty input
    out = ctxA.vector(3000)
    hij = ctxA.hv_linear(dv[0], out, lam, dv[:1])
    ok, ov, oh = orcA.hv_linear(*basis[0], lam, basis[:1], cut=1e-5, out_cap=3000)
    same(*out.download(), ok, ov)
    assert bits(hij)[0] == bits(oh)[0]
    e = ctxA.vector(10)
    assert np.array_equal(ctxA.hv_linear(e, out, lam, [e]), [0.0]) and out.len == 0


def test_hv_residual_and_shrink_cut(sysA, orcA, ctxA):
    (xk, xv), (bk, bv) = krylov_like(orcA, sysA["init_hf"], 2, 2500, 2)
    X, Bv, out = ctxA.vector(2500, xk, xv), ctxA.vector(2500, bk, bv), ctxA.vector(800)
    r = ctxA.hv_residual(X, out, 0.2, Bv)
    ok, ov, orr = orcA.hv_residual(xk, xv, bk, bv, 0.2, cut=1e-5, out_cap=800)
    same(*out.download(), ok, ov)
    assert bits([r])[0] == bits([orr])[0]
    dst = ctxA.vector(300)
    ctxA.shrink_cut(dst, X, 0.01)
    keep = np.abs(xv) > 0.01
    sk, sv = O.shrink(xk[keep], xv[keep], 300)
    same(*dst.download(), sk, sv)


def test_add_three_bit_exact(sysA, orcA, ctxA):
    vs = krylov_like(orcA, sysA["init_hf"], 3, 2000, 3)
    D = [ctxA.vector(2000, k, v) for k, v in vs]
    for cap in (6000, 700):
        out = ctxA.vector(cap)
        ctxA.add_three(out, D[0], D[1], D[2], 0.5, -1.25, 3.0)
        ok, ov = O.add_three(vs[0], vs[1], vs[2], 0.5, -1.25, 3.0, cap)
        same(*out.download(), ok, ov)


def test_linear_solve_matches_oracle(sysA, orcA, ctxA):
    (bk, bv), = krylov_like(orcA, sysA["init_hf"], 1, 1500, 4)
    target = ctxA.diag_element(sysA["init_hf"]) - 0.5
    nd, m, nr = 3000, 6, 2
    log("linear_solve gpu")
    B = ctxA.vector(nd, bk, bv)
    X = ctxA.vector(nd)
    res = ctxA.linear_solve(B, target, nd, m, nr, X)
    log("linear_solve oracle")
    xk, xv, ores = orcA.linear_solve(bk, bv, target, nd, m, nr, cut=1e-5)
    gk, gv = X.download()
    log("residuals gpu %r oracle %r" % (res, ores))
    same(gk, gv, xk, xv)
    assert np.array_equal(bits(res), bits(ores))


def test_minres_matches_oracle(sysA, orcA, ctxA):
    (bk, bv), = krylov_like(orcA, sysA["init_hf"], 1, 1500, 5)
    lam = ctxA.diag_element(sysA["init_hf"]) - 0.3
    nd = 2500
    R, G, X = ctxA.vector(nd, bk, bv), ctxA.vector(nd), ctxA.vector(nd)
    it, r2 = ctxA.minres(R, G, X, nd, 8, 1e-12, lam)
    xk, xv, oit, or2 = orcA.minres(bk, bv, np.zeros(0, np.uint64), np.zeros(0), nd, 8, 1e-12, lam, cut=1e-5)
    log("minres gpu it %d r2 %g oracle it %d r2 %g" % (it, r2, oit, or2))
    assert it == oit
    same(*X.download(), xk, xv)
    assert bits([r2])[0] == bits([or2])[0]


def test_cisd_matches_oracle(sysB, orcB, ctxB):
    hf = sysB["init_hf"]
    for st in (0, 2):
        out = ctxB.vector(5000)
        e, n = ctxB.cisd(hf, st, out)
        gk, gv = out.download()
        ok, ov, oe = orcB.cisd(hf, st)
        log("cisd state %d: block %d, e gpu %.12f oracle %.12f" % (st, n, e, oe))
        assert n == len(ok)
        assert np.array_equal(gk, ok)
        assert abs(e - oe) < 1e-10
        assert np.abs(gv - ov).max() < 1e-9


def test_excited_matches_oracle(sysB, orcB, ctxB):
    hf = sysB["init_hf"]
    nd, m, nr = 3136, 6, 2
    target = ctxB.diag_element(hf) + 0.5
    V = ctxB.vector(nd)
    e, ec = ctxB.excited(hf, target, nd, m, nr, V)
    vk, vv, oe, oec = orcB.excited(hf, target, nd, m, nr, cut=1e-5)
    log("excited gpu e %.12f cisd %.12f oracle e %.12f cisd %.12f" % (e, ec, oe, oec))
    assert abs(ec - oec) < 1e-10
    assert abs(e - oe) < 1e-8
    gk, gv = V.download()
    common, ia, ib = np.intersect1d(gk, vk, return_indices=True)
    assert len(common) >= 0.99 * max(len(gk), len(vk))
    assert np.abs(gv[ia] - vv[ib]).max() < 1e-7
