"""Integral files: synthetic RHF-style integrals, FCIDUMP writer, FCIDUMP -> ``_IN`` converter.

The ``_IN`` format is the one the reference's main program reads
(``src/functions.cpp:297-395``, ``ReadElements``): seven blocks, each a line with an
entry count followed by ``coeff m n p q`` lines, spin orbitals numbered ``2*i + s`` with
``i`` the 1-based spatial orbital and ``s = 1`` for the "u" spin, ``s = 0`` for "d".
Block order: ``uu, dd, ud, du`` two-electron, ``u, d`` one-electron, then the constant.

``convert_rhf`` restates ``conv_rhf`` (``util/convert_fcidump.py:31-151``): the same
block order, the same eight index permutations per two-electron line with the same
``m != p and n != q`` / ``n != p and m != q`` guards, one dedupe dictionary shared by the
four two-electron blocks, one-electron lines emitted as ``(m, n)`` and, when ``m != n``,
``(n, m)`` without dedupe, and the coefficient text copied verbatim.  One deliberate
difference: the reference keeps its accumulator lists across files converted in one
call (``convert_fcidump.py:33-35``); here every call converts one file from scratch.

``convert_uhf`` restates ``conv_uhf`` (``util/convert_fcidump.py:153-276``) for the UHF
FCIDUMP layout: after the header, the ``aa``, ``bb`` and ``ab`` two-electron blocks, the
``a`` and ``b`` one-electron blocks and the core line, consecutive blocks separated by a
line whose four indices are 0 and whose value is exactly 0.0.  The ``ab`` block serves
both the ``ud`` and the ``du`` output blocks (``convert_fcidump.py:197-200``); the first
block is the "u" spin.

``synth_fcidump`` writes the synthetic integrals of the benchmark workloads.  Draw
order from ``numpy.random.default_rng(seed)``:

1. ``n*(n-1)/2`` draws ``normal(0, sigma_h)`` filling ``h[p, q] = h[q, p]`` for
   ``p = 1..n-1, q = 0..p-1`` in that (row-major lower-triangle) order; the diagonal is
   ``linspace(eps_lo, eps_hi, n)``.
2. Canonical two-electron entries ``(ij|kl)`` with ``i >= j``, ``k >= l``,
   ``ij >= kl`` (``ij = i*(i+1)/2 + j``), enumerated with ``ij`` ascending and ``kl``
   ascending inside it.  First, one ``uniform(0.3, 0.8)`` draw per entry of the form
   ``(pp|qq)`` in that order; then one ``normal(0, sigma_g)`` draw per remaining entry
   in that order, scaled by ``exp(-(|i-j| + |k-l|)/4)``.
3. The core energy is 0.
"""
from __future__ import annotations

import math
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

__all__ = [
    "synth_integrals",
    "write_fcidump",
    "synth_fcidump",
    "convert_rhf",
    "convert_uhf",
    "write_fcidump_uhf",
    "synth_uhf_integrals",
    "read_fcidump",
    "WORKLOADS",
    "make_workload",
]


def synth_integrals(n_orb: int, seed: int, eps_lo: float = -2.0, eps_hi: float = 1.5,
                    sigma_h: float = 0.05, sigma_g: float = 0.02):
    """Return ``(h, canon, ecore)``: ``h`` is ``n_orb x n_orb``; ``canon`` is a list of
    ``(value, i, j, k, l)`` with 0-based indices in canonical order."""
    rng = np.random.default_rng(seed)
    n = n_orb
    h = np.zeros((n, n))
    h[np.arange(n), np.arange(n)] = np.linspace(eps_lo, eps_hi, n)
    off = rng.normal(0.0, sigma_h, size=n * (n - 1) // 2)
    t = 0
    for p in range(1, n):
        for q in range(p):
            h[p, q] = h[q, p] = off[t]
            t += 1
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    quads = []
    for a, (i, j) in enumerate(pairs):
        for b in range(a + 1):
            k, l = pairs[b]
            quads.append((i, j, k, l))
    q = np.array(quads, dtype=np.int64).reshape(-1, 4)
    is_j = (q[:, 0] == q[:, 1]) & (q[:, 2] == q[:, 3])
    vals = np.empty(len(q))
    vals[is_j] = rng.uniform(0.3, 0.8, size=int(is_j.sum()))
    rest = ~is_j
    decay = np.exp(-(np.abs(q[rest, 0] - q[rest, 1]) + np.abs(q[rest, 2] - q[rest, 3])) / 4.0)
    vals[rest] = rng.normal(0.0, sigma_g, size=int(rest.sum())) * decay
    canon = [(float(v), int(a), int(b), int(c), int(d)) for v, (a, b, c, d) in zip(vals, q)]
    return h, canon, 0.0


def _fmt(x: float) -> str:
    # 17 significant digits: parses back to the identical double with %lf / strtod.
    return "%.16E" % x


def write_fcidump(path: str, h: np.ndarray, canon, n_elec: int, ecore: float = 0.0,
                  ms2: int = 0) -> None:
    """FCIDUMP with a `` /`` terminator line (the converter starts reading after a line
    that is exactly ``' /\\n'``, ``convert_fcidump.py:77``)."""
    n = h.shape[0]
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2=%2d,\n" % (n, n_elec, ms2))
        f.write("  ORBSYM=" + ",".join(["1"] * n) + ",\n")
        f.write("  ISYM=1,\n")
        f.write(" /\n")
        for v, i, j, k, l in canon:
            f.write("%s %3d %3d %3d %3d\n" % (_fmt(v), i + 1, j + 1, k + 1, l + 1))
        for i in range(n):
            for j in range(i + 1):
                f.write("%s %3d %3d %3d %3d\n" % (_fmt(h[i, j]), i + 1, j + 1, 0, 0))
        f.write("%s %3d %3d %3d %3d\n" % (_fmt(ecore), 0, 0, 0, 0))


def read_fcidump(path: str):
    """Spatial integrals from an FCIDUMP body (after the `` /`` line):
    ``(h, eri, ecore, n_orb)`` with ``eri`` the full 8-fold-expanded 4-index array."""
    header, body = [], []
    with open(path) as f:
        lines = f.readlines()
    start = None
    for t, line in enumerate(lines):
        s = line.strip()
        if s in ("/", "&END", "$END"):
            start = t + 1
            break
        header.append(line)
    if start is None:
        raise ValueError("no header terminator in %s" % path)
    hdr = "".join(header).replace(" ", "").upper()
    n = int(hdr.split("NORB=")[1].split(",")[0])
    h = np.zeros((n, n))
    eri = np.zeros((n, n, n, n))
    ecore = 0.0
    for line in lines[start:]:
        tok = line.split()
        if len(tok) < 5:
            continue
        v = float(tok[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) for x in tok[1:5])
        if i == 0 and j == 0 and k == 0 and l == 0:
            ecore = v
        elif k == 0 and l == 0:
            h[i - 1, j - 1] = h[j - 1, i - 1] = v
        else:
            i, j, k, l = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                               (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
                eri[a, b, c, d] = v
    return h, eri, ecore, n


def convert_rhf(fcidump_path: str, out_path: Optional[str] = None) -> str:
    """FCIDUMP -> ``_IN`` with ``conv_rhf`` semantics (``util/convert_fcidump.py:31-151``).

    ``out_path`` defaults to the reference's naming: the text before the first ``_`` of
    the input path, plus ``_IN`` (``convert_fcidump.py:41``)."""
    if out_path is None:
        out_path = fcidump_path.split("_")[0] + "_IN"
    const: List[list] = []
    onebody: List[list] = []
    twobody: List[list] = []
    started = False
    with open(fcidump_path) as f:
        for line in f:
            if started:
                t = line.split()
                row = [t[0], int(t[1]), int(t[2]), int(t[3]), int(t[4])]
                if t[1] == "0" and t[2] == "0" and t[3] == "0" and t[4] == "0":
                    const.append(row)
                elif t[3] == "0" and t[4] == "0":
                    onebody.append(row)
                else:
                    twobody.append(row)
            if line == " /\n":
                started = True
    seen: Dict[Tuple[int, int, int, int], bool] = {}
    blocks = []

    def emit(out, coeff, m, n, p, q):
        key = (m, n, p, q)
        if key not in seen:
            out.append("%s %d %d %d %d\n" % (coeff, m, n, p, q))
            seen[key] = True

    for spin, rows in (("uu", twobody), ("dd", twobody), ("ud", twobody), ("du", twobody)):
        s1 = 1 if spin[0] == "u" else 0
        s2 = 1 if spin[1] == "u" else 0
        out: List[str] = []
        for c, i, j, k, l in rows:
            m, n, p, q = (i << 1) + s1, (j << 1) + s1, (k << 1) + s2, (l << 1) + s2
            if m != p and n != q:
                emit(out, c, m, n, p, q)
                emit(out, c, n, m, q, p)
                emit(out, c, p, q, m, n)
                emit(out, c, q, p, n, m)
            if n != p and m != q:
                emit(out, c, n, m, p, q)
                emit(out, c, p, q, n, m)
                emit(out, c, m, n, q, p)
                emit(out, c, q, p, m, n)
        blocks.append(out)
    for spin in ("u", "d"):
        s = 1 if spin == "u" else 0
        out = []
        for c, i, j, _, _ in onebody:
            m, n = (i << 1) + s, (j << 1) + s
            out.append("%s %d %d 0 0\n" % (c, m, n))
            if m != n:
                out.append("%s %d %d 0 0\n" % (c, n, m))
        blocks.append(out)
    blocks.append(["%s 0 0 0 0\n" % c for c, *_ in const])
    with open(out_path, "w") as f:
        for out in blocks:
            f.write("%d\n" % len(out))
            f.writelines(out)
    return out_path


def convert_uhf(fcidump_path: str, out_path: Optional[str] = None) -> str:
    """UHF FCIDUMP -> ``_IN`` with ``conv_uhf`` semantics (``util/convert_fcidump.py:153-276``)."""
    if out_path is None:
        out_path = fcidump_path.split("_")[0] + "_IN"
    groups: List[list] = []
    raw: List[list] = []
    started = False
    with open(fcidump_path) as f:
        for line in f:
            if started:
                t = line.split()
                if t[1] == "0" and t[2] == "0" and t[3] == "0" and t[4] == "0" and float(t[0]) == 0.0:
                    groups.append(raw)
                    raw = []
                else:
                    raw.append([t[0], int(t[1]), int(t[2]), int(t[3]), int(t[4])])
            if line == " /\n":
                started = True
    groups.append(raw)
    groups = groups[:3] + [groups[2]] + groups[3:]
    labels = ["uu", "dd", "ud", "du", "u", "d", "c"]
    seen: Dict[Tuple[int, int, int, int], bool] = {}

    def emit(out, coeff, m, n, p, q):
        key = (m, n, p, q)
        if key not in seen:
            out.append("%s %d %d %d %d\n" % (coeff, m, n, p, q))
            seen[key] = True

    blocks = []
    for spin, rows in zip(labels, groups):
        out: List[str] = []
        if len(spin) == 2:
            s1 = 1 if spin[0] == "u" else 0
            s2 = 1 if spin[1] == "u" else 0
            for c, i, j, k, l in rows:
                m, n, p, q = (i << 1) + s1, (j << 1) + s1, (k << 1) + s2, (l << 1) + s2
                if m != p and n != q:
                    emit(out, c, m, n, p, q)
                    emit(out, c, n, m, q, p)
                    emit(out, c, p, q, m, n)
                    emit(out, c, q, p, n, m)
                if n != p and m != q:
                    emit(out, c, n, m, p, q)
                    emit(out, c, p, q, n, m)
                    emit(out, c, m, n, q, p)
                    emit(out, c, q, p, m, n)
        elif spin == "c":
            out = ["%s 0 0 0 0\n" % r[0] for r in rows]
        else:
            sp = 1 if spin == "u" else 0
            for c, i, j, _, _ in rows:
                m, n = (i << 1) + sp, (j << 1) + sp
                out.append("%s %d %d 0 0\n" % (c, m, n))
                if m != n:
                    out.append("%s %d %d 0 0\n" % (c, n, m))
        blocks.append(out)
    with open(out_path, "w") as f:
        for out in blocks:
            f.write("%d\n" % len(out))
            f.writelines(out)
    return out_path


def synth_uhf_integrals(n_orb: int, seed: int, delta: float = 0.02, **kw):
    """UHF-style integrals: the RHF-style set of ``synth_integrals(n_orb, seed, **kw)``
    plus spin-dependent perturbations drawn from ``default_rng(seed + 1000)``.
    Returns ``(h_a, h_b, g_aa, g_bb, g_ab, ecore)`` with dense arrays; ``g_aa``/``g_bb``
    keep the 8-fold symmetry, ``g_ab[i,j,k,l]`` only ``i <-> j`` and ``k <-> l``."""
    h, canon, _ = synth_integrals(n_orb, seed, **kw)
    n = n_orb
    g = np.zeros((n, n, n, n))
    for v, i, j, k, l in canon:
        for a, b, c, d in ((i, j, k, l), (j, i, k, l), (i, j, l, k), (j, i, l, k),
                           (k, l, i, j), (l, k, i, j), (k, l, j, i), (l, k, j, i)):
            g[a, b, c, d] = v
    rng = np.random.default_rng(seed + 1000)

    def sym2(x):
        return 0.5 * (x + x.T)

    def sym8(x):
        y = x + x.transpose(1, 0, 2, 3)
        y = y + y.transpose(0, 1, 3, 2)
        y = y + y.transpose(2, 3, 0, 1)
        return y / 8.0

    def sym4(x):
        y = x + x.transpose(1, 0, 2, 3)
        y = y + y.transpose(0, 1, 3, 2)
        return y / 4.0

    h_a = h + delta * sym2(rng.normal(size=(n, n)))
    h_b = h + delta * sym2(rng.normal(size=(n, n)))
    g_aa = g + delta * sym8(rng.normal(size=(n, n, n, n)))
    g_bb = g + delta * sym8(rng.normal(size=(n, n, n, n)))
    g_ab = g + delta * sym4(rng.normal(size=(n, n, n, n)))
    return h_a, h_b, g_aa, g_bb, g_ab, 1.25


def write_fcidump_uhf(path: str, h_a, h_b, g_aa, g_bb, g_ab, n_elec: int, ecore: float) -> str:
    """UHF FCIDUMP: ``aa`` and ``bb`` blocks over canonical ``i>=j, k>=l, ij>=kl``, the
    ``ab`` block over every ``i>=j`` (a) and ``k>=l`` (b), then the ``a`` and ``b``
    one-electron blocks (``i>=j``) and the core line, separated by zero lines."""
    n = h_a.shape[0]
    sep = "%s 0 0 0 0\n" % _fmt(0.0)
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    with open(path, "w") as f:
        f.write(" &FCI NORB=%3d,NELEC=%3d,MS2= 0,\n" % (n, n_elec))
        f.write("  ORBSYM=" + ",".join(["1"] * n) + ",\n")
        f.write("  ISYM=1,IUHF=1,\n")
        f.write(" /\n")
        for g in (g_aa, g_bb):
            for a, (i, j) in enumerate(pairs):
                for b in range(a + 1):
                    k, l = pairs[b]
                    f.write("%s %3d %3d %3d %3d\n" % (_fmt(g[i, j, k, l]), i + 1, j + 1, k + 1, l + 1))
            f.write(sep)
        for (i, j) in pairs:
            for (k, l) in pairs:
                f.write("%s %3d %3d %3d %3d\n" % (_fmt(g_ab[i, j, k, l]), i + 1, j + 1, k + 1, l + 1))
        f.write(sep)
        for h in (h_a, h_b):
            for (i, j) in pairs:
                f.write("%s %3d %3d   0   0\n" % (_fmt(h[i, j]), i + 1, j + 1))
            f.write(sep)
        f.write("%s   0   0   0   0\n" % _fmt(ecore))
    return path


def synth_fcidump(path: str, n_orb: int, n_elec: int, seed: int, **kw) -> str:
    h, canon, ecore = synth_integrals(n_orb, seed, **kw)
    write_fcidump(path, h, canon, n_elec, ecore)
    return path


# Benchmark workloads: spatial orbitals, electrons, seed, Nd, Nl, restarts, generator args.
WORKLOADS = {
    "c1": dict(n_orb=12, n_elec=8, seed=0, nd=20000, nl=20, restarts=5, gen={}),
    "c2": dict(n_orb=26, n_elec=14, seed=1, nd=1000000, nl=30, restarts=8, gen={}),
    "c3": dict(n_orb=32, n_elec=16, seed=2, nd=4000000, nl=30, restarts=8, gen={}),
    "c4": dict(n_orb=50, n_elec=20, seed=3, nd=10000000, nl=40, restarts=10,
               gen=dict(eps_lo=-1.0, eps_hi=0.5, sigma_g=0.04)),
    "c5": dict(n_orb=32, n_elec=16, seed=4, nd=4000000, nl=30, restarts=8, gen={}),
}


def make_workload(name: str, workdir: str, n_orb: Optional[int] = None,
                  n_elec: Optional[int] = None, seed: Optional[int] = None) -> dict:
    """Write ``<workdir>/<name>_fcidump`` and its ``_IN``; return the workload dict with
    ``in_path``, ``n_so``, ``n_elec`` and ``init_hf`` (lowest ``n_elec`` spin orbitals)."""
    w = dict(WORKLOADS[name])
    if n_orb is not None:
        w["n_orb"] = n_orb
    if n_elec is not None:
        w["n_elec"] = n_elec
    if seed is not None:
        w["seed"] = seed
    os.makedirs(workdir, exist_ok=True)
    tag = "%s-o%d-e%d-s%d" % (name, w["n_orb"], w["n_elec"], w["seed"])
    fcid = os.path.join(workdir, tag + "_fcidump")
    inp = os.path.join(workdir, tag + "_IN")
    if not os.path.exists(inp):
        synth_fcidump(fcid, w["n_orb"], w["n_elec"], w["seed"], **w["gen"])
        convert_rhf(fcid, inp)
    w["fcidump_path"] = fcid
    w["in_path"] = inp
    w["n_so"] = 2 * w["n_orb"]
    w["init_hf"] = (1 << w["n_elec"]) - 1
    return w
