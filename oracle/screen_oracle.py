"""CPU restatement of the learned candidate screener (test infrastructure only).

No reference counterpart: the reference keeps every candidate of H.v
(src/functions.cpp:623-639).  This module states, in numpy, what the GPU library computes
for the screening path, so that the tests can check it:

  features      +-1 occupation of 64 spin orbitals (0 above n_so), excitation rank,
                (plog2|v| - x_shift) * x_scale in float32, where plog2 is the float's
                exponent plus its mantissa fraction (one IEEE rounding)
  network       z1 = X W1^T + b1, h1 = bf16(silu(z1)), z2 = h1 W2^T + b2,
                h2 = bf16(silu(z2)), y = h2 . bf16(w3) + b3; X and W1, W2 rounded to bf16
  gradient      mean squared error; dz2 = bf16(dy w3 silu'(z2)), dz1 = bf16(dh1 silu'(z1))
  Adam          bias-corrected, FP32 master parameters
  screened H.v  candidates of the mixed-precision product with |v| > eps, the kept subset
                (score >= the keep-th largest score, diagonal tags counted first), exact
                merge with every diagonal term, cut, cap, key order

Sums here run in float64 (the GPU accumulates in FP32), so network outputs agree to a
tolerance, not bit for bit.  `round_bf16=False` turns every bf16 rounding off, which is what
the finite-difference check of the gradient uses.  Only tests/ and bench.py's baseline leg
use this module.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

IN, HID = 66, 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAM = OFF_B3 + 1

SAMPLE_DTYPE = np.dtype([("det", "<u8"), ("logv", "<f4"), ("label", "<f4"), ("cls", "<u4"), ("pad", "<u4")])


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(a.shape)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def plog2(v) -> np.ndarray:
    """Exponent + mantissa fraction of |v| (float32): log2 piecewise linear; -150 for 0."""
    a = np.abs(np.ascontiguousarray(v, dtype=np.float32))
    b = a.view(np.uint32)
    e = ((b >> 23).astype(np.int32) - 127).astype(np.float32)
    f = (b & 0x7FFFFF).astype(np.float32) * np.float32(1.0 / 8388608.0)
    out = (e + f).astype(np.float32)
    out[b == 0] = np.float32(-150.0)
    return out


def features(det, cls, logv, n_so: int, x_shift: float, x_scale: float) -> np.ndarray:
    det = np.asarray(det, dtype=np.uint64)
    n = len(det)
    x = np.zeros((n, IN), np.float32)
    for p in range(min(n_so, 64)):
        x[:, p] = np.where((det >> np.uint64(p)) & np.uint64(1), 1.0, -1.0)
    x[:, 64] = np.asarray(cls, dtype=np.float32)
    lv = np.asarray(logv, dtype=np.float32)
    x[:, 65] = ((lv - np.float32(x_shift)).astype(np.float32) * np.float32(x_scale)).astype(np.float32)
    return x


def sample_features(samples: np.ndarray, n_so: int, x_shift=0.0, x_scale=1.0) -> np.ndarray:
    return features(samples["det"], samples["cls"], samples["logv"], n_so, x_shift, x_scale)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, round_bf16: bool = True, keep=False):
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    X = r(x).astype(np.float64)
    z1 = X @ r(W1).astype(np.float64).T + b1.astype(np.float64)
    h1 = r(_silu(z1).astype(np.float32) if round_bf16 else _silu(z1)).astype(np.float64)
    z2 = h1 @ r(W2).astype(np.float64).T + b2.astype(np.float64)
    h2 = r(_silu(z2).astype(np.float32) if round_bf16 else _silu(z2)).astype(np.float64)
    w3r = r(w3).astype(np.float64)
    y = h2 @ w3r + float(b3)
    if keep:
        return y, dict(X=X, z1=z1, h1=h1, z2=z2, h2=h2, w3r=w3r, W2r=r(W2).astype(np.float64))
    return y


def gradient(p, x, t, round_bf16: bool = True):
    """Flat gradient of mean((y - t)^2) and the loss."""
    y, a = forward(p, x, round_bf16, keep=True)
    t = np.asarray(t, dtype=np.float64)
    n = len(t)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / n
    rr = (lambda v: bf16(v.astype(np.float32)).astype(np.float64)) if round_bf16 else (lambda v: v)
    dz2 = rr(dy[:, None] * a["w3r"][None, :] * _silu_grad(a["z2"]))
    g = np.zeros(NPARAM, np.float64)
    g[OFF_W3:OFF_B3] = a["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    g[OFF_W2:OFF_B2] = (dz2.T @ a["h1"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dh1 = dz2 @ a["W2r"]
    dz1 = rr(dh1 * _silu_grad(a["z1"]))
    g[OFF_W1:OFF_B1] = (dz1.T @ a["X"]).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return g, loss


def adam(p, m, v, g, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step in float32; returns (p, m, v)."""
    f = np.float32
    p, m, v, g = (np.asarray(a, dtype=np.float32).copy() for a in (p, m, v, g))
    b1, b2 = f(b1), f(b2)
    m = (b1 * m + (f(1) - b1) * g).astype(np.float32)
    v = (b2 * v + (f(1) - b2) * g * g).astype(np.float32)
    c1 = f(1) - b1 ** f(step)
    c2 = f(1) - b2 ** f(step)
    p = (p - f(lr) * (m / c1) / (np.sqrt(v / c2) + f(eps))).astype(np.float32)
    return p, m, v


# ---- screened H.v ----------------------------------------------------------------------
def _bind(L):
    if getattr(L, "_screen_bound", False):
        return
    u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
    L.oracle_screen_candidates.restype = C.c_int64
    L.oracle_screen_candidates.argtypes = [C.c_void_p, u64p, f64p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                           u64p, f32p, u32p, C.c_int64]
    L.oracle_screen_merge.restype = C.c_int
    L.oracle_screen_merge.argtypes = [C.c_void_p, u64p, f64p, C.c_int64, C.c_int64, C.c_double, C.c_double,
                                      C.c_double, C.c_int64, u64p, f32p, C.c_int64, u64p, f64p,
                                      C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L._screen_bound = True


def screen_candidates(orc, keys, vals, nd, coeff, eps):
    """(targets, float32 values, excitation ranks) of the prefiltered candidates."""
    from oracle import oracle as O
    L = O.lib()
    _bind(L)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    z = np.zeros(1, np.uint64), np.zeros(1, np.float32), np.zeros(1, np.uint32)
    n = L.oracle_screen_candidates(orc._h, k, v, len(k), nd, coeff, eps, z[0], z[1], z[2], 0)
    t, cv, cl = np.zeros(max(n, 1), np.uint64), np.zeros(max(n, 1), np.float32), np.zeros(max(n, 1), np.uint32)
    L.oracle_screen_candidates(orc._h, k, v, len(k), nd, coeff, eps, t, cv, cl, n)
    return t[:n], cv[:n], cl[:n]


def keep_mask(scores, n_src: int, keep: int) -> np.ndarray:
    """Candidates whose score reaches the keep-th largest, the n_src diagonal tags counted
    first (their score is +inf); ties at the threshold kept."""
    s = np.asarray(scores, dtype=np.float32)
    room = keep - n_src
    if room >= len(s):
        return np.ones(len(s), bool)
    if room <= 0:
        thr = np.float32(np.inf)
    else:
        thr = np.sort(s)[::-1][room - 1]
    return s >= thr


def screen_merge(orc, keys, vals, nd, lam, coeff, cut, out_cap, ct, cv):
    from oracle import oracle as O
    L = O.lib()
    _bind(L)
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    ct = np.ascontiguousarray(ct, dtype=np.uint64)
    cv = np.ascontiguousarray(cv, dtype=np.float32)
    bound = max(1, len(ct) + min(len(k), nd))
    ok, ov = np.zeros(bound, np.uint64), np.zeros(bound)
    n, d = C.c_int64(), C.c_double()
    rc = L.oracle_screen_merge(orc._h, k, v, len(k), nd, lam, coeff, cut, out_cap, ct, cv, len(ct), ok, ov,
                               C.byref(n), C.byref(d))
    if rc:
        raise RuntimeError("oracle_screen_merge: a term left the fixed-point range")
    return ok[:n.value].copy(), ov[:n.value].copy(), d.value


def screened_hv(orc, keys, vals, nd, lam, coeff, eps, keep_factor, score_fn, cut=1e-5, out_cap=60_000_000):
    """The screened product: score_fn(targets, values, ranks) -> float32 scores (mode 1:
    |value|; mode 2: the network's outputs)."""
    n_src = min(len(keys), nd)
    ct, cv, cl = screen_candidates(orc, keys, vals, nd, coeff, eps)
    keep = max(int(keep_factor * nd), n_src + 1)
    if len(ct) + n_src > keep:
        m = keep_mask(score_fn(ct, cv, cl), n_src, keep)
        ct, cv = ct[m], cv[m]
    return screen_merge(orc, keys, vals, nd, lam, coeff, cut, out_cap, ct, cv)
