"""ctypes binding of libsfci_hip.so (the C-ABI declared in include/sfci_gpu.h).

The shared library is built in-tree by ``sfci_amd/Makefile``.  Loading fails loudly when
it is missing: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsfci_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sfci_gpu.h")

SFCI_OK = 0
ERRORS = {-1: "SFCI_E_ARG", -2: "SFCI_E_HIP", -3: "SFCI_E_OVERFLOW", -4: "SFCI_E_CAPACITY",
          -5: "SFCI_E_RANGE", -6: "SFCI_E_IO"}


class SfciError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class InfoT(C.Structure):
    _fields_ = [("n_so", C.c_int), ("n_elec", C.c_int), ("string_dim", C.c_uint64),
                ("n_conn", C.c_uint64), ("e_core", C.c_double)]


class StatsT(C.Structure):
    _fields_ = [("n_src", C.c_uint64), ("n_candidates", C.c_uint64), ("n_unique", C.c_uint64),
                ("n_out", C.c_uint64), ("n_passes", C.c_uint64), ("ms_generate", C.c_double),
                ("ms_sort", C.c_double), ("ms_reduce", C.c_double), ("ms_select", C.c_double),
                ("ms_total", C.c_double), ("n_fallback", C.c_uint64), ("hash_bits", C.c_int),
                ("n_kept", C.c_uint64), ("screen_eps", C.c_double), ("ms_score", C.c_double),
                ("n_split", C.c_uint64), ("n_str0", C.c_uint64), ("n_str1", C.c_uint64)]


KERNEL_COUNT = 18   # SFCI_KERNEL_COUNT


class KernelStatT(C.Structure):
    _fields_ = [("ms", C.c_double), ("launches", C.c_uint64), ("bytes", C.c_double)]


class LanczosParamsT(C.Structure):
    _fields_ = [("nd", C.c_int64), ("nl", C.c_int), ("num_restart", C.c_int), ("hpsi_cap", C.c_int64),
                ("delta_e", C.c_double), ("verbose", C.c_int), ("e_prev_init", C.c_double),
                ("ms_restart", C.c_void_p)]


class LinearParamsT(C.Structure):
    _fields_ = [("nd", C.c_int64), ("num_iter", C.c_int), ("num_restart", C.c_int), ("verbose", C.c_int)]


class LanczosResultT(C.Structure):
    _fields_ = [("n_restarts", C.c_int), ("e_final", C.c_double), ("ms_total", C.c_double),
                ("ms_hv", C.c_double), ("n_hv", C.c_int64), ("n_screened", C.c_int), ("n_reverts", C.c_int)]


def build(force: bool = False) -> str:
    args = ["make", "-s", "-C", HERE, "-j4"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def declared_symbols() -> list:
    """Function names declared in include/sfci_gpu.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(sfci_[a-z_0-9]+)\s*\(", txt)))


_lib = None


def lib():
    """Load libsfci_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsfci_hip.so is not built (run `make -C sfci_amd` or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i64, u64, dbl, ip = C.c_void_p, C.c_int64, C.c_uint64, C.c_double, C.c_int
    pvp = C.POINTER(C.c_void_p)
    sig = {
        "sfci_create": (ip, [C.c_char_p, ip, ip, ip, pvp]),
        "sfci_create_from_entries": (ip, [ip, ip, i64, vp, vp, ip, pvp]),
        "sfci_destroy": (None, [vp]),
        "sfci_last_error": (C.c_char_p, [vp]),
        "sfci_info": (ip, [vp, C.POINTER(InfoT)]),
        "sfci_set_cut": (ip, [vp, dbl]),
        "sfci_set_workspace": (ip, [vp, u64]),
        "sfci_hv_stats": (ip, [vp, C.POINTER(StatsT)]),
        "sfci_set_merge_mode": (ip, [vp, ip, ip]),
        "sfci_set_gen_tables": (ip, [vp, ip]),
        "sfci_set_precision": (ip, [vp, ip]),
        "sfci_sync": (ip, [vp]),
        "sfci_kernel_stats": (ip, [vp, C.POINTER(KernelStatT), ip, ip]),
        "sfci_kernel_name": (C.c_char_p, [ip]),
        "sfci_vec_create": (ip, [vp, i64, pvp]),
        "sfci_vec_destroy": (None, [vp]),
        "sfci_vec_len": (i64, [vp]),
        "sfci_vec_cap": (i64, [vp]),
        "sfci_vec_upload": (ip, [vp, vp, vp, i64]),
        "sfci_vec_download": (ip, [vp, vp, vp, i64, C.POINTER(C.c_int64)]),
        "sfci_vec_device_ptrs": (ip, [vp, pvp, pvp]),
        "sfci_vec_set_len": (ip, [vp, i64]),
        "sfci_vec_copy": (ip, [vp, vp]),
        "sfci_vec_clear": (ip, [vp]),
        "sfci_hv": (ip, [vp, vp, vp, i64, dbl, dbl, C.POINTER(C.c_double)]),
        "sfci_energy": (ip, [vp, vp, C.POINTER(C.c_double)]),
        "sfci_diag_element": (ip, [vp, u64, C.POINTER(C.c_double)]),
        "sfci_dot": (ip, [vp, vp, vp, C.POINTER(C.c_double)]),
        "sfci_norm": (ip, [vp, vp, C.POINTER(C.c_double)]),
        "sfci_scale": (ip, [vp, vp, dbl]),
        "sfci_add_two": (ip, [vp, vp, vp, dbl]),
        "sfci_shrink_from": (ip, [vp, vp, vp]),
        "sfci_shrink_from_noalloc": (ip, [vp, vp, vp]),
        "sfci_lincomb": (ip, [vp, vp, vp, ip, vp]),
        "sfci_ground_state": (ip, [vp, u64, C.POINTER(LanczosParamsT), vp, vp, C.POINTER(LanczosResultT), vp]),
        "sfci_shrink_cut": (ip, [vp, vp, vp, dbl]),
        "sfci_hv_linear": (ip, [vp, vp, vp, dbl, vp, ip, vp]),
        "sfci_hv_residual": (ip, [vp, vp, vp, dbl, vp, C.POINTER(C.c_double)]),
        "sfci_add_three": (ip, [vp, vp, vp, vp, vp, dbl, dbl, dbl]),
        "sfci_linear_solve": (ip, [vp, vp, dbl, C.POINTER(LinearParamsT), vp, vp]),
        "sfci_shift_invert": (ip, [vp, vp, dbl, C.POINTER(LinearParamsT), vp, C.POINTER(C.c_double), vp]),
        "sfci_minres": (ip, [vp, vp, vp, vp, i64, ip, dbl, dbl, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
        "sfci_cisd": (ip, [vp, u64, ip, vp, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "sfci_excited": (ip, [vp, u64, dbl, C.POINTER(LinearParamsT), vp, C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
        "sfci_mlp_create": (ip, [vp, u64, pvp]),
        "sfci_mlp_destroy": (None, [vp]),
        "sfci_mlp_set_params": (ip, [vp, vp]),
        "sfci_mlp_get_params": (ip, [vp, vp]),
        "sfci_mlp_set_norm": (ip, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
        "sfci_mlp_forward": (ip, [vp, vp, i64, vp]),
        "sfci_mlp_gradient": (ip, [vp, vp, vp, i64, vp, C.POINTER(C.c_float)]),
        "sfci_mlp_adam_step": (ip, [vp, vp, C.c_float, C.c_float, C.c_float, C.c_float]),
        "sfci_mlp_train": (ip, [vp, vp, i64, ip, ip, C.c_float, u64, vp]),
        "sfci_mlp_predict_samples": (ip, [vp, vp, i64, vp]),
        "sfci_mlp_sample_features": (ip, [vp, vp, i64, vp]),
        "sfci_set_screen": (ip, [vp, ip, vp, dbl, dbl]),
        "sfci_screen_samples": (ip, [vp, vp, dbl, i64, u64, vp, C.POINTER(C.c_int64)]),
        "sfci_get_screen": (ip, [vp, C.POINTER(C.c_int), pvp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "sfci_wide_create": (ip, [C.c_char_p, ip, ip, ip, pvp]),
        "sfci_wide_destroy": (None, [vp]),
        "sfci_wide_last_error": (C.c_char_p, [vp]),
        "sfci_wide_info": (ip, [vp, C.POINTER(InfoT)]),
        "sfci_wide_set_cut": (ip, [vp, dbl]),
        "sfci_wide_set_precision": (ip, [vp, ip]),
        "sfci_wide_set_workspace": (ip, [vp, u64]),
        "sfci_wide_hv_stats": (ip, [vp, C.POINTER(StatsT)]),
        "sfci_wide_set_merge": (ip, [vp, ip, ip]),
        "sfci_wide_vec_copy": (ip, [vp, vp]),
        "sfci_wide_vec_clear": (ip, [vp]),
        "sfci_wide_dot": (ip, [vp, vp, vp, C.POINTER(C.c_double)]),
        "sfci_wide_norm": (ip, [vp, vp, C.POINTER(C.c_double)]),
        "sfci_wide_scale": (ip, [vp, vp, dbl]),
        "sfci_wide_add_two": (ip, [vp, vp, vp, dbl]),
        "sfci_wide_shrink_from": (ip, [vp, vp, vp]),
        "sfci_wide_lincomb": (ip, [vp, vp, vp, ip, vp]),
        "sfci_wide_energy": (ip, [vp, vp, C.POINTER(C.c_double)]),
        "sfci_wide_sync": (ip, [vp]),
        "sfci_wide_ground_state": (ip, [vp, vp, C.POINTER(LanczosParamsT), vp, vp, C.POINTER(LanczosResultT), vp]),
        "sfci_wide_vec_create": (ip, [vp, i64, pvp]),
        "sfci_wide_vec_destroy": (None, [vp]),
        "sfci_wide_vec_len": (i64, [vp]),
        "sfci_wide_vec_upload": (ip, [vp, vp, vp, i64]),
        "sfci_wide_vec_download": (ip, [vp, vp, vp, i64, C.POINTER(C.c_int64)]),
        "sfci_wide_hv": (ip, [vp, vp, vp, i64, dbl, dbl, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Vector:
    """Device sparse vector (dynVec) of fixed capacity."""

    def __init__(self, ctx: "Context", cap: int):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().sfci_vec_create(ctx._h, int(cap), C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            lib().sfci_vec_destroy(self._h)
            self._h = None

    @property
    def len(self) -> int:
        return lib().sfci_vec_len(self._h)

    @property
    def cap(self) -> int:
        return lib().sfci_vec_cap(self._h)

    def upload(self, keys, vals) -> "Vector":
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self.ctx._check(lib().sfci_vec_upload(self._h, _ptr(k), _ptr(v), len(k)))
        return self

    def download(self):
        n = self.len
        k, v = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1))
        got = C.c_int64()
        self.ctx._check(lib().sfci_vec_download(self._h, _ptr(k), _ptr(v), n, C.byref(got)))
        return k[:got.value].copy(), v[:got.value].copy()

    def copy_from(self, other: "Vector") -> "Vector":
        self.ctx._check(lib().sfci_vec_copy(self._h, other._h))
        return self

    def clear(self):
        lib().sfci_vec_clear(self._h)


class Context:
    """One loaded Hamiltonian on one GPU."""

    def __init__(self, in_path: str, n_so: int, n_elec: int, device: int = 0):
        h = C.c_void_p()
        rc = lib().sfci_create(in_path.encode(), n_so, n_elec, device, C.byref(h))
        if rc:
            raise SfciError(rc, "sfci_create(%s)" % in_path)
        self._h = h
        self.n_so, self.n_elec = n_so, n_elec

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            lib().sfci_destroy(self._h)
            self._h = None

    def _check(self, rc: int):
        if rc != SFCI_OK:
            msg = lib().sfci_last_error(self._h) if self._h else b""
            raise SfciError(rc, (msg or b"").decode())

    def info(self) -> dict:
        i = InfoT()
        self._check(lib().sfci_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in InfoT._fields_}

    def set_cut(self, cut: float):
        self._check(lib().sfci_set_cut(self._h, cut))

    def set_workspace(self, nbytes: int):
        self._check(lib().sfci_set_workspace(self._h, int(nbytes)))

    def set_precision(self, mixed: bool):
        self._check(lib().sfci_set_precision(self._h, 1 if mixed else 0))

    def set_merge_mode(self, mode: int, hash_bits: int = 0):
        self._check(lib().sfci_set_merge_mode(self._h, mode, hash_bits))

    def set_gen_tables(self, mode: int):
        """0 automatic, 1 every source evaluates its own same-spin doubles, 2 per-string tables"""
        self._check(lib().sfci_set_gen_tables(self._h, int(mode)))

    def stats(self) -> dict:
        s = StatsT()
        self._check(lib().sfci_hv_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in StatsT._fields_}

    def kernel_stats(self, reset: bool = False) -> dict:
        n = KERNEL_COUNT
        arr = (KernelStatT * n)()
        self._check(lib().sfci_kernel_stats(self._h, arr, n, 1 if reset else 0))
        return {lib().sfci_kernel_name(i).decode(): dict(ms=arr[i].ms, launches=arr[i].launches,
                                                           bytes=arr[i].bytes) for i in range(n)}

    def sync(self):
        self._check(lib().sfci_sync(self._h))

    def vector(self, cap: int, keys=None, vals=None) -> Vector:
        v = Vector(self, cap)
        if keys is not None:
            v.upload(keys, vals)
        return v

    def hv(self, vin: Vector, vout: Vector, nd: int, lam: float = 0.0, coeff: float = 1.0) -> float:
        d = C.c_double()
        self._check(lib().sfci_hv(self._h, vin._h, vout._h, int(nd), lam, coeff, C.byref(d)))
        return d.value

    def energy(self, v: Vector) -> float:
        d = C.c_double()
        self._check(lib().sfci_energy(self._h, v._h, C.byref(d)))
        return d.value

    def diag_element(self, det: int) -> float:
        d = C.c_double()
        self._check(lib().sfci_diag_element(self._h, det, C.byref(d)))
        return d.value

    def dot(self, a: Vector, b: Vector) -> float:
        d = C.c_double()
        self._check(lib().sfci_dot(self._h, a._h, b._h, C.byref(d)))
        return d.value

    def norm(self, a: Vector) -> float:
        d = C.c_double()
        self._check(lib().sfci_norm(self._h, a._h, C.byref(d)))
        return d.value

    def scale(self, a: Vector, s: float):
        self._check(lib().sfci_scale(self._h, a._h, s))

    def add_two(self, v: Vector, w: Vector, a: float):
        self._check(lib().sfci_add_two(self._h, v._h, w._h, a))

    def shrink_from(self, dst: Vector, src: Vector):
        self._check(lib().sfci_shrink_from(self._h, dst._h, src._h))

    def shrink_from_noalloc(self, dst: Vector, src: Vector):
        """shrink_from with no device allocation (a fixed scratch array per GPU)."""
        self._check(lib().sfci_shrink_from_noalloc(self._h, dst._h, src._h))

    def lincomb(self, vecs: Sequence[Vector], coeffs, out: Vector):
        arr = (C.c_void_p * len(vecs))(*[v._h.value for v in vecs])
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        self._check(lib().sfci_lincomb(self._h, C.cast(arr, C.c_void_p), _ptr(c), len(vecs), out._h))

    def ground_state(self, init_det: int, nd: int, nl: int, num_restart: int, hpsi_cap: int = 60_000_000,
                     delta_e: float = 1e-5, verbose: bool = False, v_out: Optional[Vector] = None,
                     e_prev_init: float = 0.0) -> dict:
        ritz, et, ms = np.zeros(num_restart), np.zeros(num_restart), np.zeros(num_restart)
        p = LanczosParamsT(nd, nl, num_restart, hpsi_cap, delta_e, 1 if verbose else 0, e_prev_init,
                           ms.ctypes.data)
        r = LanczosResultT()
        self._check(lib().sfci_ground_state(self._h, init_det, C.byref(p), _ptr(ritz), _ptr(et), C.byref(r),
                                            v_out._h if v_out is not None else None))
        n = r.n_restarts
        return dict(restarts=n, ritz=ritz[:n].copy(), e_trunc=et[:n].copy(), ms_restart=ms[:n].copy(),
                    e_final=r.e_final,
                    ms_total=r.ms_total, ms_hv=r.ms_hv, n_hv=r.n_hv, n_screened=r.n_screened,
                    n_reverts=r.n_reverts)

    # --- candidate screening ---------------------------------------------------------
    def set_screen(self, mode: int, net: Optional["Mlp"] = None, keep_factor: float = 3.0, eps_pre: float = 1e-7):
        """mode 0 off, 1 rank by |H_ij c_j|, 2 rank by the network, 3 by the network's predicted
        ln|y| of the target plus ln|H_ij c_j| (SFCI_SCREEN_*)."""
        self._check(lib().sfci_set_screen(self._h, int(mode), net._h if net is not None else None,
                                          float(keep_factor), float(eps_pre)))
        self._screen_net = net   # keep the network alive while the context uses it

    def get_screen(self) -> dict:
        m, n, k, e = C.c_int(), C.c_void_p(), C.c_double(), C.c_double()
        self._check(lib().sfci_get_screen(self._h, C.byref(m), C.byref(n), C.byref(k), C.byref(e)))
        return dict(mode=m.value, net=n.value, keep_factor=k.value, eps_pre=e.value)

    def screen_samples(self, v: Vector, eps_pre: float, max_samples: int, seed: int = 0) -> np.ndarray:
        out = np.zeros(max(int(max_samples), 1), SAMPLE_DTYPE)
        n = C.c_int64()
        self._check(lib().sfci_screen_samples(self._h, v._h, float(eps_pre), int(max_samples), int(seed), _ptr(out),
                                              C.byref(n)))
        return out[:n.value].copy()

    # --- linear-solver path --------------------------------------------------------
    def shrink_cut(self, dst: Vector, src: Vector, cut: float):
        self._check(lib().sfci_shrink_cut(self._h, dst._h, src._h, cut))

    def hv_linear(self, vin: Vector, vout: Vector, lam: float, basis: Sequence[Vector]) -> np.ndarray:
        arr = (C.c_void_p * len(basis))(*[b._h.value for b in basis])
        hij = np.zeros(len(basis))
        self._check(lib().sfci_hv_linear(self._h, vin._h, vout._h, lam, C.cast(arr, C.c_void_p), len(basis),
                                         _ptr(hij)))
        return hij

    def hv_residual(self, x: Vector, out: Vector, lam: float, b: Vector) -> float:
        r = C.c_double()
        self._check(lib().sfci_hv_residual(self._h, x._h, out._h, lam, b._h, C.byref(r)))
        return r.value

    def add_three(self, out: Vector, x: Vector, y: Vector, z: Vector, a: float, b: float, c: float):
        self._check(lib().sfci_add_three(self._h, out._h, x._h, y._h, z._h, a, b, c))

    def linear_solve(self, b: Vector, target_e: float, nd: int, num_iter: int, num_restart: int,
                     x_out: Vector, verbose: bool = False) -> np.ndarray:
        p = LinearParamsT(nd, num_iter, num_restart, 1 if verbose else 0)
        res = np.zeros(num_restart)
        self._check(lib().sfci_linear_solve(self._h, b._h, target_e, C.byref(p), x_out._h, _ptr(res)))
        return res

    def shift_invert(self, b: Vector, target_e: float, nd: int, num_iter: int, num_restart: int,
                     v_out: Vector, verbose: bool = False):
        p = LinearParamsT(nd, num_iter, num_restart, 1 if verbose else 0)
        res = np.zeros(num_restart)
        e = C.c_double()
        self._check(lib().sfci_shift_invert(self._h, b._h, target_e, C.byref(p), v_out._h, C.byref(e), _ptr(res)))
        return e.value, res

    def minres(self, rhs: Vector, guess: Vector, result: Vector, nd: int, iters: int, tol: float, lam: float):
        it, r2 = C.c_int(), C.c_double()
        self._check(lib().sfci_minres(self._h, rhs._h, guess._h, result._h, int(nd), iters, tol, lam,
                                      C.byref(it), C.byref(r2)))
        return it.value, r2.value

    def cisd(self, ref_det: int, state_idx: int, out: Vector):
        e, n = C.c_double(), C.c_int64()
        self._check(lib().sfci_cisd(self._h, ref_det, state_idx, out._h, C.byref(e), C.byref(n)))
        return e.value, n.value

    def excited(self, init_det: int, target_e: float, nd: int, num_iter: int, num_restart: int, v_out: Vector,
                verbose: bool = False):
        p = LinearParamsT(nd, num_iter, num_restart, 1 if verbose else 0)
        e, ec = C.c_double(), C.c_double()
        self._check(lib().sfci_excited(self._h, init_det, target_e, C.byref(p), v_out._h, C.byref(e), C.byref(ec)))
        return e.value, ec.value


# ---- screening network ---------------------------------------------------------------
SAMPLE_DTYPE = np.dtype([("det", "<u8"), ("logv", "<f4"), ("label", "<f4"), ("cls", "<u4"), ("pad", "<u4")])
MLP_IN, MLP_HID = 66, 256
MLP_NPARAM = MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_HID + 1


class Mlp:
    """The screening network (sfci_mlp) on a context's GPU."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().sfci_mlp_create(ctx._h, int(seed), C.byref(h)))
        self._h = h

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None) and self.ctx._h:
            lib().sfci_mlp_destroy(self._h)
        self._h = None

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.size == MLP_NPARAM
        self.ctx._check(lib().sfci_mlp_set_params(self._h, _ptr(a)))

    def get_params(self) -> np.ndarray:
        a = np.zeros(MLP_NPARAM, np.float32)
        self.ctx._check(lib().sfci_mlp_get_params(self._h, _ptr(a)))
        return a

    def set_norm(self, x_shift: float, x_scale: float, y_shift: float, y_scale: float):
        self.ctx._check(lib().sfci_mlp_set_norm(self._h, x_shift, x_scale, y_shift, y_scale))

    def forward(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        y = np.zeros(len(a), np.float32)
        self.ctx._check(lib().sfci_mlp_forward(self._h, _ptr(a), len(a), _ptr(y)))
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray):
        a = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
        tt = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float()
        self.ctx._check(lib().sfci_mlp_gradient(self._h, _ptr(a), _ptr(tt), len(a), _ptr(g), C.byref(loss)))
        return g, loss.value

    def adam_step(self, grad: np.ndarray, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        self.ctx._check(lib().sfci_mlp_adam_step(self._h, _ptr(g), lr, beta1, beta2, eps))

    def train(self, samples: np.ndarray, steps: int, batch: int = 65536, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        s = np.ascontiguousarray(samples, dtype=SAMPLE_DTYPE)
        curve = np.zeros(max(steps, 1), np.float32)
        self.ctx._check(lib().sfci_mlp_train(self._h, _ptr(s), len(s), int(steps), int(batch), float(lr), int(seed),
                                             _ptr(curve)))
        return curve[:steps]

    def predict_samples(self, samples: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(samples, dtype=SAMPLE_DTYPE)
        y = np.zeros(len(s), np.float32)
        self.ctx._check(lib().sfci_mlp_predict_samples(self._h, _ptr(s), len(s), _ptr(y)))
        return y

    def sample_features(self, samples: np.ndarray) -> np.ndarray:
        s = np.ascontiguousarray(samples, dtype=SAMPLE_DTYPE)
        x = np.zeros((len(s), MLP_IN), np.float32)
        self.ctx._check(lib().sfci_mlp_sample_features(self._h, _ptr(s), len(s), _ptr(x)))
        return x


# ---- 128-bit keys (more than 64 spin orbitals) ----------------------------------------
def split_keys(keys) -> tuple:
    """Python ints -> (lo, hi) uint64 arrays (lo = spin orbitals 0-63)."""
    lo = np.array([k & 0xFFFFFFFFFFFFFFFF for k in keys], np.uint64)
    hi = np.array([k >> 64 for k in keys], np.uint64)
    return lo, hi


def join_keys(lo, hi) -> list:
    return [(int(h) << 64) | int(l) for l, h in zip(lo, hi)]


class WideVector:
    """Device sparse vector with 128-bit keys (two words per entry, ascending as hi:lo)."""

    def __init__(self, ctx: "WideContext", cap: int):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._check(lib().sfci_wide_vec_create(ctx._h, int(cap), C.byref(h)))
        self._h = h
        self.cap = int(cap)

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            lib().sfci_wide_vec_destroy(self._h)
            self._h = None

    @property
    def len(self) -> int:
        return lib().sfci_wide_vec_len(self._h)

    def upload(self, lo, hi, vals) -> "WideVector":
        lo = np.ascontiguousarray(lo, dtype=np.uint64)
        hi = np.ascontiguousarray(hi, dtype=np.uint64)
        k = np.empty(2 * len(lo), np.uint64)
        k[0::2], k[1::2] = lo, hi
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self.ctx._check(lib().sfci_wide_vec_upload(self._h, _ptr(k), _ptr(v), len(lo)))
        return self

    def download(self):
        """-> (lo, hi, vals)"""
        n = self.len
        k, v = np.zeros(2 * max(n, 1), np.uint64), np.zeros(max(n, 1))
        got = C.c_int64()
        self.ctx._check(lib().sfci_wide_vec_download(self._h, _ptr(k), _ptr(v), n, C.byref(got)))
        m = got.value
        return k[0:2 * m:2].copy(), k[1:2 * m:2].copy(), v[:m].copy()


class WideContext:
    """One loaded Hamiltonian with more than 64 spin orbitals (128-bit keys) on one GPU."""

    def __init__(self, in_path: str, n_so: int, n_elec: int, device: int = 0):
        h = C.c_void_p()
        rc = lib().sfci_wide_create(in_path.encode(), n_so, n_elec, device, C.byref(h))
        if rc:
            raise SfciError(rc, "sfci_wide_create(%s)" % in_path)
        self._h = h
        self.n_so, self.n_elec = n_so, n_elec

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            lib().sfci_wide_destroy(self._h)
            self._h = None

    def _check(self, rc: int):
        if rc != SFCI_OK:
            msg = lib().sfci_wide_last_error(self._h) if self._h else b""
            raise SfciError(rc, (msg or b"").decode())

    def info(self) -> dict:
        i = InfoT()
        self._check(lib().sfci_wide_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in InfoT._fields_}

    def set_cut(self, cut: float):
        self._check(lib().sfci_wide_set_cut(self._h, cut))

    def set_precision(self, mixed: bool):
        self._check(lib().sfci_wide_set_precision(self._h, 1 if mixed else 0))

    def set_workspace(self, nbytes: int):
        self._check(lib().sfci_wide_set_workspace(self._h, int(nbytes)))

    def set_merge(self, hash_bits: int = 0, table_fill: int = 0):
        """Merge knobs (results identical for every value): bucket hash bits (9..16, 0 =
        automatic) and the LDS-table fill that splits a bucket into classes (0 = default)."""
        self._check(lib().sfci_wide_set_merge(self._h, int(hash_bits), int(table_fill)))

    def stats(self) -> dict:
        s = StatsT()
        self._check(lib().sfci_wide_hv_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in StatsT._fields_}

    def vector(self, cap: int, lo=None, hi=None, vals=None) -> WideVector:
        v = WideVector(self, cap)
        if lo is not None:
            v.upload(lo, hi, vals)
        return v

    def hv(self, vin: WideVector, vout: WideVector, nd: int, lam: float = 0.0, coeff: float = 1.0) -> float:
        d = C.c_double()
        self._check(lib().sfci_wide_hv(self._h, vin._h, vout._h, int(nd), lam, coeff, C.byref(d)))
        return d.value

    def copy(self, dst: WideVector, src: WideVector):
        self._check(lib().sfci_wide_vec_copy(dst._h, src._h))

    def dot(self, a: WideVector, b: WideVector) -> float:
        d = C.c_double()
        self._check(lib().sfci_wide_dot(self._h, a._h, b._h, C.byref(d)))
        return d.value

    def norm(self, a: WideVector) -> float:
        d = C.c_double()
        self._check(lib().sfci_wide_norm(self._h, a._h, C.byref(d)))
        return d.value

    def scale(self, a: WideVector, s: float):
        self._check(lib().sfci_wide_scale(self._h, a._h, s))

    def add_two(self, v: WideVector, w: WideVector, a: float):
        self._check(lib().sfci_wide_add_two(self._h, v._h, w._h, a))

    def shrink_from(self, dst: WideVector, src: WideVector):
        self._check(lib().sfci_wide_shrink_from(self._h, dst._h, src._h))

    def lincomb(self, vecs: Sequence[WideVector], coeffs, out: WideVector):
        arr = (C.c_void_p * len(vecs))(*[v._h.value for v in vecs])
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        self._check(lib().sfci_wide_lincomb(self._h, C.cast(arr, C.c_void_p), _ptr(c), len(vecs), out._h))

    def energy(self, v: WideVector) -> float:
        d = C.c_double()
        self._check(lib().sfci_wide_energy(self._h, v._h, C.byref(d)))
        return d.value

    def ground_state(self, init_det: int, nd: int, nl: int, num_restart: int, hpsi_cap: int = 60_000_000,
                     delta_e: float = 1e-5, verbose: bool = False, v_out: Optional[WideVector] = None,
                     e_prev_init: float = 0.0) -> dict:
        ritz, et, ms = np.zeros(num_restart), np.zeros(num_restart), np.zeros(num_restart)
        p = LanczosParamsT(nd, nl, num_restart, hpsi_cap, delta_e, 1 if verbose else 0, e_prev_init,
                           ms.ctypes.data)
        r = LanczosResultT()
        init = np.array([init_det & 0xFFFFFFFFFFFFFFFF, init_det >> 64], np.uint64)
        self._check(lib().sfci_wide_ground_state(self._h, _ptr(init), C.byref(p), _ptr(ritz), _ptr(et), C.byref(r),
                                                 v_out._h if v_out is not None else None))
        n = r.n_restarts
        return dict(restarts=n, ritz=ritz[:n].copy(), e_trunc=et[:n].copy(), ms_restart=ms[:n].copy(),
                    e_final=r.e_final, ms_total=r.ms_total, ms_hv=r.ms_hv, n_hv=r.n_hv)
