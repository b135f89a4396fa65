"""ctypes binding of the oracle library (test infrastructure only).

The oracle is the CPU restatement of the reference hot path (see ``sfci_oracle.cpp``).
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it,
and only as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the committed Makefile."""
    args = ["make", "-s", "-C", HERE]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_load.restype = C.c_void_p
        L.oracle_load.argtypes = [C.c_char_p, C.c_int, C.c_int]
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_constant.restype = C.c_double
        L.oracle_constant.argtypes = [C.c_void_p]
        L.oracle_integral.restype = C.c_double
        L.oracle_integral.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.oracle_one_body.restype = C.c_double
        L.oracle_one_body.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oracle_alpha_rank.restype = C.c_uint64
        L.oracle_alpha_rank.argtypes = [C.c_void_p, C.c_uint64]
        L.oracle_bucket_params.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint),
                                           C.POINTER(C.c_uint64)]
        L.oracle_diag.restype = C.c_double
        L.oracle_diag.argtypes = [C.c_void_p, C.c_uint64]
        L.oracle_diag_literal.restype = C.c_double
        L.oracle_diag_literal.argtypes = [C.c_void_p, C.c_uint64]
        L.oracle_connections.restype = C.c_int64
        L.oracle_connections.argtypes = [C.c_void_p, C.c_uint64, _u64p, _f64p, C.c_int64]
        L.oracle_offdiag_literal.restype = C.c_int64
        L.oracle_offdiag_literal.argtypes = [C.c_void_p, C.c_uint64, _u64p, _f64p, C.c_int64]
        L.oracle_hv.restype = C.c_int
        L.oracle_hv.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                C.c_double, C.c_int64, _u64p, _f64p, C.POINTER(C.c_int64),
                                C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.oracle_dot.restype = C.c_double
        L.oracle_dot.argtypes = [_u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64]
        L.oracle_norm.restype = C.c_double
        L.oracle_norm.argtypes = [_f64p, C.c_int64]
        L.oracle_add_two.restype = C.c_int64
        L.oracle_add_two.argtypes = [_u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64, C.c_double,
                                     C.c_int64, _u64p, _f64p]
        L.oracle_shrink.restype = C.c_int64
        L.oracle_shrink.argtypes = [_u64p, _f64p, C.c_int64, C.c_int64, _u64p, _f64p]
        L.oracle_lincomb.restype = C.c_int64
        L.oracle_lincomb.argtypes = [_u64p, _f64p, _i64p, C.c_int, _f64p, C.c_int64, _u64p, _f64p,
                                     C.POINTER(C.c_double)]
        L.oracle_energy.restype = C.c_double
        L.oracle_energy.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, C.c_int]
        L.oracle_gsaes.restype = C.c_int
        L.oracle_gsaes.argtypes = [C.c_int, _f64p, _f64p, _f64p, _f64p]
        L.oracle_lanczos.restype = C.c_int
        L.oracle_lanczos.argtypes = [C.c_void_p, C.c_uint64, C.c_int64, C.c_int, C.c_int, C.c_int64,
                                     C.c_double, C.c_int, C.c_double, _f64p, _f64p, C.POINTER(C.c_double), _u64p,
                                     _f64p, C.POINTER(C.c_int64)]
        L.oracle_lstsq.restype = C.c_int
        L.oracle_lstsq.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.oracle_n_conn.restype = C.c_uint64
        L.oracle_n_conn.argtypes = [C.c_void_p]
        L.oracle_hv_linear.restype = C.c_int
        L.oracle_hv_linear.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, C.c_double, _u64p, _f64p, _i64p, C.c_int,
                                       C.c_double, C.c_int64, _u64p, _f64p, C.POINTER(C.c_int64), _f64p]
        L.oracle_hv_residual.restype = C.c_int
        L.oracle_hv_residual.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64, C.c_double,
                                         C.c_double, C.c_int64, _u64p, _f64p, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double)]
        L.oracle_add_three.restype = C.c_int64
        L.oracle_add_three.argtypes = [_u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64,
                                       C.c_double, C.c_double, C.c_double, C.c_int64, _u64p, _f64p]
        L.oracle_linear_solve.restype = C.c_int
        L.oracle_linear_solve.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, C.c_double, C.c_int64, C.c_int,
                                          C.c_int, C.c_double, _u64p, _f64p, C.POINTER(C.c_int64), _f64p]
        L.oracle_shift_invert.restype = C.c_int
        L.oracle_shift_invert.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, C.c_double, C.c_int64, C.c_int,
                                          C.c_int, C.c_double, _u64p, _f64p, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_double), _f64p]
        L.oracle_cisd.restype = C.c_int
        L.oracle_cisd.argtypes = [C.c_void_p, C.c_uint64, C.c_int, _u64p, _f64p, C.c_int64, C.POINTER(C.c_int64),
                                  C.POINTER(C.c_double)]
        L.oracle_excited.restype = C.c_int
        L.oracle_excited.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_int64, C.c_int, C.c_int, C.c_double,
                                     _u64p, _f64p, C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.oracle_minres.restype = C.c_int
        L.oracle_minres.argtypes = [C.c_void_p, _u64p, _f64p, C.c_int64, _u64p, _f64p, C.c_int64, C.c_int64,
                                    C.c_int, C.c_double, C.c_double, C.c_double, _u64p, _f64p,
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.oracle_connections_wide.restype = C.c_int64
        L.oracle_connections_wide.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, _u64p, _u64p, _f64p, C.c_int64]
        L.oracle_diag_wide.restype = C.c_double
        L.oracle_diag_wide.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.oracle_hv_wide.restype = C.c_int
        L.oracle_hv_wide.argtypes = [C.c_void_p, _u64p, _u64p, _f64p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                     C.c_double, C.c_int64, _u64p, _u64p, _f64p, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_double), C.c_int, C.c_int]
        _lib = L
    return _lib


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One loaded ``_IN`` Hamiltonian."""

    def __init__(self, in_path: str, n_so: int, n_elec: int):
        self.n_so, self.n_elec = n_so, n_elec
        self._h = lib().oracle_load(in_path.encode(), n_so, n_elec)
        if not self._h:
            raise ValueError("oracle could not read %s" % in_path)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().oracle_free(self._h)
            self._h = None

    # --- Slater-Condon -------------------------------------------------------------
    def diag(self, d: int) -> float:
        return lib().oracle_diag(self._h, d)

    def diag_literal(self, d: int) -> float:
        return lib().oracle_diag_literal(self._h, d)

    def integral(self, m, n, p, q) -> float:
        return lib().oracle_integral(self._h, m, n, p, q)

    def one_body(self, m, n) -> float:
        return lib().oracle_one_body(self._h, m, n)

    def constant(self) -> float:
        return lib().oracle_constant(self._h)

    def alpha_rank(self, d: int) -> int:
        return lib().oracle_alpha_rank(self._h, d)

    def bucket_params(self):
        a, b, c = C.c_uint64(), C.c_uint(), C.c_uint64()
        lib().oracle_bucket_params(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def connections(self, d: int):
        n = lib().oracle_connections(self._h, d, np.zeros(1, np.uint64), np.zeros(1), 0)
        k, v = np.zeros(n, np.uint64), np.zeros(n)
        lib().oracle_connections(self._h, d, k, v, n)
        return k, v

    def offdiag_literal(self, d: int):
        n = lib().oracle_offdiag_literal(self._h, d, np.zeros(1, np.uint64), np.zeros(1), 0)
        k, v = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1))
        lib().oracle_offdiag_literal(self._h, d, k, v, n)
        return k[:n], v[:n]

    # --- H.v -----------------------------------------------------------------------
    def hv(self, keys, vals, nd=None, lam=0.0, coeff=1.0, cut=1e-5, out_cap=60_000_000,
           nthreads=0, mode=0):
        keys, vals = _u64(keys), _f64(vals)
        n = len(keys)
        nd = n if nd is None else nd
        # upper bound on the output: every candidate distinct
        bound = min(out_cap, max(1, n) * 400_000)
        ok, ov = np.zeros(bound, np.uint64), np.zeros(bound)
        olen, diag, ncand = C.c_int64(), C.c_double(), C.c_int64()
        rc = lib().oracle_hv(self._h, keys, vals, n, nd, lam, coeff, cut, bound, ok, ov,
                             C.byref(olen), C.byref(diag), nthreads, mode, C.byref(ncand))
        if rc:
            raise RuntimeError("oracle_hv failed (%d)" % rc)
        if olen.value == bound and bound < out_cap:
            raise RuntimeError("oracle_hv output bound too small")
        m = olen.value
        return ok[:m].copy(), ov[:m].copy(), diag.value, ncand.value

    # --- 128-bit keys (lo = spin orbitals 0-63, hi = 64-127) ---------------------------
    def connections_wide(self, d: int):
        lo, hi = d & 0xFFFFFFFFFFFFFFFF, d >> 64
        z = np.zeros(1, np.uint64)
        n = lib().oracle_connections_wide(self._h, lo, hi, z, z.copy(), np.zeros(1), 0)
        kl, kh, v = np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n)
        lib().oracle_connections_wide(self._h, lo, hi, kl, kh, v, n)
        return kl, kh, v

    def diag_wide(self, d: int) -> float:
        return lib().oracle_diag_wide(self._h, d & 0xFFFFFFFFFFFFFFFF, d >> 64)

    def hv_wide(self, lo, hi, vals, nd=None, lam=0.0, coeff=1.0, cut=1e-5, out_cap=60_000_000,
                nthreads=0, mode=0):
        """H.v on 128-bit keys given as (lo, hi) word arrays, ascending by hi:lo."""
        lo, hi, vals = _u64(lo), _u64(hi), _f64(vals)
        n = len(lo)
        nd = n if nd is None else nd
        bound = min(out_cap, max(1, min(n, nd)) * int(self.n_conn()))
        ol, oh, ov = np.zeros(bound, np.uint64), np.zeros(bound, np.uint64), np.zeros(bound)
        olen, diag = C.c_int64(), C.c_double()
        rc = lib().oracle_hv_wide(self._h, lo, hi, vals, n, nd, lam, coeff, cut, bound, ol, oh, ov,
                                  C.byref(olen), C.byref(diag), nthreads, mode)
        if rc:
            raise RuntimeError("oracle_hv_wide failed (%d)" % rc)
        m = olen.value
        return ol[:m].copy(), oh[:m].copy(), ov[:m].copy(), diag.value

    def energy(self, keys, vals, nthreads=0) -> float:
        keys, vals = _u64(keys), _f64(vals)
        return lib().oracle_energy(self._h, keys, vals, len(keys), nthreads)

    def lanczos(self, init_det: int, nd: int, nl: int, restarts: int, hpsi_cap=60_000_000,
                cut=1e-5, nthreads=0, e_prev_init=0.0):
        ritz, et = np.zeros(restarts), np.zeros(restarts)
        ef = C.c_double()
        ok, ov = np.zeros(nd, np.uint64), np.zeros(nd)
        olen = C.c_int64()
        r = lib().oracle_lanczos(self._h, init_det, nd, nl, restarts, hpsi_cap, cut, nthreads, e_prev_init,
                                 ritz, et, C.byref(ef), ok, ov, C.byref(olen))
        if r < 0:
            raise RuntimeError("oracle_lanczos failed (%d)" % r)
        return dict(restarts=r, ritz=ritz[:r].copy(), e_trunc=et[:r].copy(), e_final=ef.value,
                    keys=ok[:olen.value].copy(), vals=ov[:olen.value].copy())

    # --- linear-solver path (linear_oracle.cpp) ---------------------------------------
    def n_conn(self) -> int:
        return lib().oracle_n_conn(self._h)

    def hv_linear(self, k, v, lam, basis, cut=1e-5, out_cap=60_000_000):
        k, v = _u64(k), _f64(v)
        offs = np.zeros(len(basis) + 1, np.int64)
        for i, (bk, _) in enumerate(basis):
            offs[i + 1] = offs[i] + len(bk)
        bk = _u64(np.concatenate([b[0] for b in basis]) if offs[-1] else np.zeros(1, np.uint64))
        bv = _f64(np.concatenate([b[1] for b in basis]) if offs[-1] else np.zeros(1))
        bound = max(1, min(out_cap, len(k) * self.n_conn() + int(offs[-1])))
        ok, ov = np.zeros(bound, np.uint64), np.zeros(bound)
        olen = C.c_int64()
        hij = np.zeros(len(basis))
        rc = lib().oracle_hv_linear(self._h, k, v, len(k), lam, bk, bv, offs, len(basis), cut, out_cap, ok, ov,
                                    C.byref(olen), hij)
        if rc:
            raise RuntimeError("oracle_hv_linear failed (%d)" % rc)
        return ok[:olen.value].copy(), ov[:olen.value].copy(), hij

    def hv_residual(self, xk, xv, bk, bv, lam, cut=1e-5, out_cap=60_000_000):
        xk, xv, bk, bv = _u64(xk), _f64(xv), _u64(bk), _f64(bv)
        bound = max(1, min(out_cap, len(xk) * self.n_conn() + len(bk)))
        ok, ov = np.zeros(bound, np.uint64), np.zeros(bound)
        olen, res = C.c_int64(), C.c_double()
        rc = lib().oracle_hv_residual(self._h, xk, xv, len(xk), bk, bv, len(bk), lam, cut, out_cap, ok, ov,
                                      C.byref(olen), C.byref(res))
        if rc:
            raise RuntimeError("oracle_hv_residual failed (%d)" % rc)
        return ok[:olen.value].copy(), ov[:olen.value].copy(), res.value

    def linear_solve(self, bk, bv, target_e, nd, num_iter, num_restart, cut=1e-5):
        bk, bv = _u64(bk), _f64(bv)
        xk, xv = np.zeros(nd, np.uint64), np.zeros(nd)
        xl = C.c_int64()
        res = np.zeros(num_restart)
        rc = lib().oracle_linear_solve(self._h, bk, bv, len(bk), target_e, nd, num_iter, num_restart, cut, xk, xv,
                                       C.byref(xl), res)
        if rc:
            raise RuntimeError("oracle_linear_solve failed (%d)" % rc)
        return xk[:xl.value].copy(), xv[:xl.value].copy(), res

    def shift_invert(self, bk, bv, target_e, nd, num_iter, num_restart, cut=1e-5):
        bk, bv = _u64(bk), _f64(bv)
        xk, xv = np.zeros(nd, np.uint64), np.zeros(nd)
        xl, e = C.c_int64(), C.c_double()
        res = np.zeros(num_restart)
        rc = lib().oracle_shift_invert(self._h, bk, bv, len(bk), target_e, nd, num_iter, num_restart, cut, xk, xv,
                                       C.byref(xl), C.byref(e), res)
        if rc:
            raise RuntimeError("oracle_shift_invert failed (%d)" % rc)
        return xk[:xl.value].copy(), xv[:xl.value].copy(), e.value, res

    def cisd(self, ref, state_idx):
        cap = self.n_conn()
        ok, ov = np.zeros(cap, np.uint64), np.zeros(cap)
        n, e = C.c_int64(), C.c_double()
        rc = lib().oracle_cisd(self._h, ref, state_idx, ok, ov, cap, C.byref(n), C.byref(e))
        if rc:
            raise RuntimeError("oracle_cisd failed (%d)" % rc)
        return ok[:n.value].copy(), ov[:n.value].copy(), e.value

    def excited(self, init_det, target_e, nd, num_iter, num_restart, cut=1e-5):
        xk, xv = np.zeros(nd, np.uint64), np.zeros(nd)
        xl, e, ec = C.c_int64(), C.c_double(), C.c_double()
        rc = lib().oracle_excited(self._h, init_det, target_e, nd, num_iter, num_restart, cut, xk, xv,
                                  C.byref(xl), C.byref(e), C.byref(ec))
        if rc:
            raise RuntimeError("oracle_excited failed (%d)" % rc)
        return xk[:xl.value].copy(), xv[:xl.value].copy(), e.value, ec.value

    def minres(self, rk, rv, gk, gv, nd, iters, tol, lam, cut=1e-5):
        rk, rv, gk, gv = _u64(rk), _f64(rv), _u64(gk), _f64(gv)
        xk, xv = np.zeros(nd, np.uint64), np.zeros(nd)
        xl, it, r2 = C.c_int64(), C.c_int(), C.c_double()
        rc = lib().oracle_minres(self._h, rk, rv, len(rk), gk, gv, len(gk), nd, iters, tol, lam, cut, xk, xv,
                                 C.byref(xl), C.byref(it), C.byref(r2))
        if rc:
            raise RuntimeError("oracle_minres failed (%d)" % rc)
        return xk[:xl.value].copy(), xv[:xl.value].copy(), it.value, r2.value


def add_three(x, y, z, a, b, c, cap):
    (xk, xv), (yk, yv), (zk, zv) = [(_u64(k), _f64(v)) for k, v in (x, y, z)]
    n = max(1, len(xk) + len(yk) + len(zk))
    ok, ov = np.zeros(n, np.uint64), np.zeros(n)
    m = lib().oracle_add_three(xk, xv, len(xk), yk, yv, len(yk), zk, zv, len(zk), a, b, c, cap, ok, ov)
    return ok[:m].copy(), ov[:m].copy()


# --- sparse vector ops -------------------------------------------------------------
def dot(ak, av, bk, bv) -> float:
    ak, av, bk, bv = _u64(ak), _f64(av), _u64(bk), _f64(bv)
    return lib().oracle_dot(ak, av, len(ak), bk, bv, len(bk))


def norm(v) -> float:
    v = _f64(v)
    return lib().oracle_norm(v, len(v))


def add_two(vk, vv, wk, wv, a, cap):
    vk, vv, wk, wv = _u64(vk), _f64(vv), _u64(wk), _f64(wv)
    n = len(vk) + len(wk)
    ok, ov = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1))
    m = lib().oracle_add_two(vk, vv, len(vk), wk, wv, len(wk), a, cap, ok, ov)
    return ok[:m].copy(), ov[:m].copy()


def shrink(sk, sv, cap):
    sk, sv = _u64(sk), _f64(sv)
    ok, ov = np.zeros(max(len(sk), 1), np.uint64), np.zeros(max(len(sk), 1))
    m = lib().oracle_shrink(sk, sv, len(sk), cap, ok, ov)
    return ok[:m].copy(), ov[:m].copy()


def lincomb(vecs, coeffs, cap):
    offs = np.zeros(len(vecs) + 1, np.int64)
    for i, (k, _) in enumerate(vecs):
        offs[i + 1] = offs[i] + len(k)
    kk = _u64(np.concatenate([k for k, _ in vecs])) if vecs else np.zeros(0, np.uint64)
    kv = _f64(np.concatenate([v for _, v in vecs])) if vecs else np.zeros(0)
    n = max(int(offs[-1]), 1)
    ok, ov = np.zeros(n, np.uint64), np.zeros(n)
    nrm = C.c_double()
    m = lib().oracle_lincomb(kk, kv, offs, len(vecs), _f64(coeffs), cap, ok, ov, C.byref(nrm))
    return ok[:m].copy(), ov[:m].copy()


def lstsq(A, b):
    m, n = A.shape
    x = np.zeros(n)
    r = lib().oracle_lstsq(m, n, _f64(A.ravel()), _f64(b), x)
    return x, r


def gsaes(K, S):
    n = K.shape[0]
    ev, vec = np.zeros(n), np.zeros(n * n)
    rc = lib().oracle_gsaes(n, _f64(K.ravel()), _f64(S.ravel()), ev, vec)
    if rc:
        raise RuntimeError("gsaes: S not positive definite")
    return ev, vec.reshape(n, n)
