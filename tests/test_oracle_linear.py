"""Oracle pins for the linear-solver path (spmspvm_bucket_linear / res_norm, addThree,
the restarted Krylov linear solver, shift-invert, MINRES, CISD).

Every check goes through an independent route: the dense Hamiltonian of a tiny system
built by second-quantized brute force (tests/sq_bruteforce.py), numpy's dense solvers,
and plain dictionary sums.  Plausible slips (a dropped projection, the wrong basis vector
subtracted, a missing cut, a transposed least-squares operand, a wrong MINRES rotation)
each move a result far outside these tolerances.
"""
import numpy as np
import pytest

import sq_bruteforce as B
from oracle import oracle as O
from sfci_amd import integrals as I


@pytest.fixture(scope="module")
def tiny(make_system):
    s = make_system(5, 4, 7)
    o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
    h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
    dets, M = B.dense_h(h, eri, ec, 5, 2)
    keys = np.array(dets, dtype=np.uint64)
    return s, o, keys, M


def _vec(keys, x):
    nz = np.nonzero(x)[0]
    return keys[nz], x[nz]


def _dense(keys, k, v):
    x = np.zeros(len(keys))
    x[np.searchsorted(keys, k)] = v
    return x


def _cut_top(keys, w, cut, cap):
    idx = [i for i in range(len(w)) if abs(w[i]) > cut]
    idx = sorted(idx, key=lambda i: (-abs(w[i]), int(keys[i])))[:cap]
    idx = sorted(idx, key=lambda i: int(keys[i]))
    return keys[idx], w[idx]


def test_hv_linear_vs_dense(tiny):
    s, o, keys, M = tiny
    rng = np.random.default_rng(1)
    lam = -0.4
    basis = []
    for j in range(4):
        x = rng.normal(size=len(keys)) * (rng.random(len(keys)) < 0.6)
        basis.append(_vec(keys, x / np.linalg.norm(x)))
    A = M - lam * np.eye(len(keys))
    xin = _dense(keys, *basis[-1])
    w = A @ xin
    hij_ref = np.array([_dense(keys, *b) @ w for b in basis])
    w2 = w - hij_ref[-1] * _dense(keys, *basis[-1]) - hij_ref[-2] * _dense(keys, *basis[-2])
    for cut, cap in [(1e-5, 10 ** 6), (0.05, 37)]:
        ok, ov, hij = o.hv_linear(*basis[-1], lam, basis, cut=cut, out_cap=cap)
        assert np.allclose(hij, hij_ref, atol=1e-12, rtol=0)
        rk, rv = _cut_top(keys, w2, cut, cap)
        assert np.array_equal(ok, rk)
        assert np.allclose(ov, rv, atol=1e-12, rtol=0)


def test_hv_residual_vs_dense(tiny):
    s, o, keys, M = tiny
    rng = np.random.default_rng(2)
    lam = 0.3
    x = _vec(keys, rng.normal(size=len(keys)) * (rng.random(len(keys)) < 0.5))
    b = _vec(keys, rng.normal(size=len(keys)) * (rng.random(len(keys)) < 0.5))
    r = (M - lam * np.eye(len(keys))) @ _dense(keys, *x) - _dense(keys, *b)
    ok, ov, res = o.hv_residual(*x, *b, lam, cut=0.02, out_cap=50)
    assert abs(res - np.linalg.norm(r)) < 1e-12
    rk, rv = _cut_top(keys, r, 0.02, 50)
    assert np.array_equal(ok, rk) and np.allclose(ov, rv, atol=1e-12, rtol=0)


def test_add_three_bruteforce():
    rng = np.random.default_rng(3)
    pool = np.sort(rng.choice(2 ** 40, size=400, replace=False).astype(np.uint64))
    vs = []
    for n in (120, 90, 60):
        k = np.sort(rng.choice(pool, size=n, replace=False))
        v = rng.normal(size=n)
        vs.append((k, v))
    a, b, c = 0.7, -1.3, 0.25
    # make one key cancel exactly and a few tiny
    vs[1][1][0] = 0.0
    tot = {}
    for (k, v), co in zip(vs, (a, b, c)):
        for kk, vv in zip(k.tolist(), v.tolist()):
            tot[kk] = tot.get(kk, 0.0) + co * vv
    for cap in (1000, 50):
        ok, ov = O.add_three(vs[0], vs[1], vs[2], a, b, c, cap)
        items = sorted(tot.items(), key=lambda kv: (-abs(kv[1]), kv[0]))[:cap]
        items = sorted([kv for kv in items if abs(kv[1]) > 1e-14])
        assert ok.tolist() == [kv[0] for kv in items]
        assert np.allclose(ov, [kv[1] for kv in items], atol=1e-14, rtol=1e-14)


def test_lstsq_vs_numpy():
    rng = np.random.default_rng(4)
    for m, n in [(9, 8), (31, 30), (5, 5)]:
        A = rng.normal(size=(m, n))
        b = rng.normal(size=m)
        x, r = O.lstsq(A, b)
        assert r == n
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.allclose(x, ref, atol=1e-10)
        assert np.abs(A.T @ (A @ x - b)).max() < 1e-10
    # rank deficient: two equal columns -> one of them gets zero weight
    A = rng.normal(size=(7, 4))
    A[:, 2] = A[:, 1]
    b = rng.normal(size=7)
    x, r = O.lstsq(A, b)
    assert r == 3 and (x[1] == 0.0 or x[2] == 0.0)
    assert np.abs(A.T @ (A @ x - b)).max() < 1e-10


def test_linear_solve_converges_to_dense_solution(tiny):
    s, o, keys, M = tiny
    n = len(keys)
    e0 = np.linalg.eigvalsh(M)[0]
    target = e0 - 0.7                       # A = H - target is positive definite
    rng = np.random.default_rng(5)
    bd = rng.normal(size=n)
    bd /= np.linalg.norm(bd)
    b = _vec(keys, bd)
    xk, xv, res = o.linear_solve(*b, target, nd=n, num_iter=12, num_restart=4, cut=0.0)
    x_ref = np.linalg.solve(M - target * np.eye(n), bd)
    x = _dense(keys, xk, xv)
    assert res[-1] < 1e-8 and res[-1] < res[0]
    assert np.abs(x - x_ref).max() < 1e-8
    # reported residual is ||(H - target) x - b|| of the returned x
    assert abs(res[-1] - np.linalg.norm((M - target * np.eye(n)) @ x - bd)) < 1e-12


def test_shift_invert_moves_toward_the_nearest_state(tiny):
    s, o, keys, M = tiny
    n = len(keys)
    ev, U = np.linalg.eigh(M)
    target = ev[1] - 0.02                   # nearest eigenvalue: ev[1]
    rng = np.random.default_rng(6)
    bd = U[:, 1] + 0.3 * rng.normal(size=n) / np.sqrt(n)
    bd /= np.linalg.norm(bd)
    vk, vv, e, res = o.shift_invert(*_vec(keys, bd), target, nd=n, num_iter=15, num_restart=3, cut=0.0)
    v = _dense(keys, vk, vv)
    assert abs(np.linalg.norm(v) - 1.0) < 1e-12
    assert abs(e - v @ M @ v) < 1e-10
    # one inverse-iteration step: the overlap with state 1 grows
    assert abs(v @ U[:, 1]) > abs(bd @ U[:, 1])


def test_minres_vs_dense(tiny):
    s, o, keys, M = tiny
    n = len(keys)
    ev = np.linalg.eigvalsh(M)
    lam = 0.5 * (ev[3] + ev[4])              # indefinite A = H - lam
    rng = np.random.default_rng(7)
    bd = rng.normal(size=n)
    xk, xv, it, r2 = o.minres(*_vec(keys, bd), np.zeros(0, np.uint64), np.zeros(0), nd=n, iters=400,
                              tol=1e-20, lam=lam, cut=0.0)
    x = _dense(keys, xk, xv)
    x_ref = np.linalg.solve(M - lam * np.eye(n), bd)
    assert np.abs(x - x_ref).max() < 1e-7 * np.abs(x_ref).max()
    assert r2 < 1e-20 or it == 400


def test_cisd_vs_dense_block(tiny, make_system):
    for (no, ne, seed) in [(5, 4, 7), (6, 6, 1)]:
        s = make_system(no, ne, seed)
        o = O.Oracle(s["in_path"], s["n_so"], s["n_elec"])
        h, eri, ec, _ = I.read_fcidump(s["fcidump_path"])
        dets, M = B.dense_h(h, eri, ec, no, ne // 2)
        keys = np.array(dets, dtype=np.uint64)
        hf = s["init_hf"]
        ih = int(np.searchsorted(keys, hf))
        blk = sorted(set([hf] + [int(keys[i]) for i in np.nonzero(M[ih])[0]]))
        idx = np.searchsorted(keys, np.array(blk, dtype=np.uint64))
        Mb = M[np.ix_(idx, idx)]
        ev, U = np.linalg.eigh(Mb)
        for st in (0, 2):
            ok, ov, e = o.cisd(hf, st)
            assert ok.tolist() == blk
            assert abs(e - ev[st]) < 1e-10
            u = U[:, st] * np.sign(U[np.argmax(np.abs(U[:, st])), st])
            assert np.abs(ov - u).max() < 1e-8
            assert abs(np.linalg.norm(ov) - 1.0) < 1e-12
