"""sfci_shrink_from_noalloc (no device allocation, a fixed scratch array per GPU) against
sfci_shrink_from (src/all.h:281-299): the same kept entries, bit for bit, on calls in a row
whose inputs shrink, with ties at the truncation threshold."""
import ctypes as C
import itertools
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def all_dets(n_orb, k):
    strings = []
    for occ in itertools.combinations(range(n_orb), k):
        s = 0
        for i in occ:
            s |= 1 << (2 * i)
        strings.append(s)
    a = np.array(strings, dtype=np.uint64)
    return np.sort((a[:, None] | (a[None, :] << np.uint64(1))).ravel())


def both(ctx, k, v, cap):
    from sfci_amd import gpu
    src = ctx.vector(len(k), k, v)
    a, b = ctx.vector(cap), ctx.vector(cap)
    ctx.shrink_from(a, src)
    rc = gpu.lib().sfci_shrink_from_noalloc(ctx._h, b._h, src._h)
    assert rc == 0, gpu.lib().sfci_last_error(ctx._h)
    return a.download(), b.download()


def test_shrink_noalloc_matches_shrink_from(make_system):
    from sfci_amd import gpu
    s = make_system(12, 8, 0)
    dets = all_dets(12, 4)
    rng = np.random.default_rng(5)
    ctx = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    try:
        cases = [
            (200_000, 50_000, "ties"),     # many entries share |c| at the threshold
            (30_000, 7_000, "ties"),       # second call, smaller input: stale scratch must not matter
            (20_000, 19_999, "normal"),
            (5_000, 6_000, "normal"),      # src fits: a copy
            (70_000, 1, "ties"),
            (4_000, 3_000, "equal"),       # every |c| equal: the key order decides
        ]
        for n, cap, kind in cases:
            pick = np.sort(rng.choice(len(dets), size=n, replace=False))
            k = dets[pick]
            if kind == "ties":
                v = rng.choice([-0.5, 0.25, -0.125, 0.0625, 1e-3, -1e-3], size=n)
                m = rng.random(n) < 0.3
                v[m] = rng.normal(size=int(m.sum()))
            elif kind == "equal":
                v = np.where(rng.random(n) < 0.5, -0.01, 0.01)
            else:
                v = rng.normal(size=n) * np.exp(-rng.uniform(0, 10, size=n))
            (ak, av), (bk, bv) = both(ctx, k, v, cap)
            log("noalloc n=%d cap=%d %s: kept %d / %d" % (n, cap, kind, len(bk), len(ak)))
            assert len(ak) == min(n, cap)
            assert np.array_equal(ak, bk)
            assert np.array_equal(av.view(np.uint64), bv.view(np.uint64))
    finally:
        ctx.close()


def test_shrink_noalloc_rejects_same_vector(make_system):
    from sfci_amd import gpu
    s = make_system(12, 8, 0)
    ctx = gpu.Context(s["in_path"], s["n_so"], s["n_elec"])
    try:
        v = ctx.vector(10)
        assert gpu.lib().sfci_shrink_from_noalloc(ctx._h, v._h, v._h) == -1
    finally:
        ctx.close()
