"""Second-quantized brute force for tiny systems (test helper, independent of the oracle).

H = sum_pq h_pq a+_p a_q + 1/2 sum_pqrs (pq|rs) a+_p a+_r a_s a_q over spin orbitals
``2*P + s``, with (pq|rs) = (PQ|RS) when spin(p) == spin(q) and spin(r) == spin(s), built
straight from the spatial integrals (not from an ``_IN`` file) and applied to occupation
bit strings with Jordan-Wigner signs (ascending orbital order).
"""
import itertools

import numpy as np


def _ann(d, q):
    if not (d >> q) & 1:
        return None, 0
    s = -1 if bin(d & ((1 << q) - 1)).count("1") & 1 else 1
    return d & ~(1 << q), s


def _cre(d, p):
    if (d >> p) & 1:
        return None, 0
    s = -1 if bin(d & ((1 << p) - 1)).count("1") & 1 else 1
    return d | (1 << p), s


def apply_h(h, eri, ecore, d, n_orb):
    """Return {det: <det|H|d>} for one determinant.

    ``h`` and ``eri`` are spatial arrays, or spin-resolved: ``h = {s: h_s}`` and
    ``eri = {(s_pq, s_rs): eri_block}`` with s the spin bit of the spin orbitals."""
    if not isinstance(h, dict):
        h = {0: h, 1: h}
    if not isinstance(eri, dict):
        eri = {(a, b): eri for a in (0, 1) for b in (0, 1)}
    out = {d: ecore}
    nso = 2 * n_orb
    for p in range(nso):
        for q in range(nso):
            if (p & 1) != (q & 1):
                continue
            v = h[p & 1][p >> 1, q >> 1]
            if v == 0:
                continue
            d1, s1 = _ann(d, q)
            if d1 is None:
                continue
            d2, s2 = _cre(d1, p)
            if d2 is None:
                continue
            out[d2] = out.get(d2, 0.0) + s1 * s2 * v
    occ = [i for i in range(nso) if (d >> i) & 1]
    for q in occ:
        for s in occ:
            if s == q:
                continue
            # a_s a_q |d>
            d1, s1 = _ann(d, q)
            d2, s2 = _ann(d1, s)
            for r in range(nso):
                if (r & 1) != (s & 1):
                    continue
                d3, s3 = _cre(d2, r)
                if d3 is None:
                    continue
                for p in range(nso):
                    if (p & 1) != (q & 1):
                        continue
                    v = eri[(p & 1, r & 1)][p >> 1, q >> 1, r >> 1, s >> 1]
                    if v == 0:
                        continue
                    d4, s4 = _cre(d3, p)
                    if d4 is None:
                        continue
                    out[d4] = out.get(d4, 0.0) + 0.5 * s1 * s2 * s3 * s4 * v
    return out


def sz0_dets(n_orb, n_alpha, n_beta=None):
    """All determinants with n_alpha even bits and n_beta odd bits set, ascending."""
    n_beta = n_alpha if n_beta is None else n_beta
    dets = []
    for a in itertools.combinations(range(n_orb), n_alpha):
        for b in itertools.combinations(range(n_orb), n_beta):
            d = 0
            for i in a:
                d |= 1 << (2 * i)
            for j in b:
                d |= 1 << (2 * j + 1)
            dets.append(d)
    return sorted(dets)


def dense_h(h, eri, ecore, n_orb, n_alpha):
    dets = sz0_dets(n_orb, n_alpha)
    idx = {d: i for i, d in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)))
    for j, d in enumerate(dets):
        for t, v in apply_h(h, eri, ecore, d, n_orb).items():
            M[idx[t], j] += v
    return dets, M
