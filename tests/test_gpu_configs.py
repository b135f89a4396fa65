"""GPU parity at the orbital spaces of the five benchmark workloads.

Each test builds the workload's own integrals (sfci_amd.integrals.make_workload: the same
generator, seed and sizes as the benchmark), grows an input vector from the Hartree-Fock
determinant, and compares the HIP H.v (src/functions.cpp:511-804) with the oracle's on a
sample of sources the oracle finishes in seconds: keys exactly, values and diag through
their IEEE bit patterns.

  26 spatial / 52 spin orbitals, 14 electrons, seed 1   2000 sources, several target
                                                         passes, the two-level hash merge
  32 spatial / 64 spin orbitals, 16 electrons, seeds 2, 4   FP64 and mixed precision, the
                                                         screened product (seed 4)
  50 spatial / 100 spin orbitals, 20 electrons, seed 3   the 128-bit-key H.v
  12 spatial / 24 spin orbitals, 8 electrons, seed 0     the whole ground-state solve at
                                                         Nd 2e4, Nl 20, 5 restarts
"""
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


def assert_same_vec(k1, v1, k2, v2):
    assert len(k1) == len(k2), (len(k1), len(k2))
    assert np.array_equal(k1, k2)
    diff = np.nonzero(bits(v1) != bits(v2))[0]
    assert len(diff) == 0, "first mismatch at %d: %r vs %r" % (diff[0], v1[diff[0]], v2[diff[0]])


@pytest.fixture(scope="module")
def wdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("workloads"))


def workload(wdir, name):
    from sfci_amd import integrals as I
    log("workload %s: integrals" % name)
    return I.make_workload(name, wdir)


def grow(ctx, hf, sizes, cap=60_000_000):
    """HF -> (H.v -> top-n -> normalize) for each n in sizes; returns host (keys, values)."""
    v = ctx.vector(1, np.array([hf], np.uint64), np.array([1.0]))
    out = ctx.vector(cap)
    for n in sizes:
        ctx.hv(v, out, max(v.len, 1))
        nxt = ctx.vector(n)
        ctx.shrink_from(nxt, out)
        ctx.scale(nxt, 1.0 / ctx.norm(nxt))
        v = nxt
    return v.download()


def gpu_hv(ctx, k, v, nd, lam, coeff, cap):
    vin = ctx.vector(max(len(k), 1), k, v)
    vout = ctx.vector(cap)
    d = ctx.hv(vin, vout, nd, lam, coeff)
    ok, ov = vout.download()
    return ok, ov, d


# --- 26 spatial orbitals, 14 electrons, seed 1 ------------------------------------------

@pytest.fixture(scope="module")
def c2(wdir):
    from sfci_amd import gpu
    w = workload(wdir, "c2")
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
    k, v = grow(ctx, w["init_hf"], [2000, 2000])
    log("c2 input: %d determinants" % len(k))
    yield w, ctx, orc, k, v
    ctx.close()


@pytest.mark.parametrize("case", ["single_pass", "multipass_auto", "multipass_12bits", "multipass_16bits",
                                  "mixed_multipass", "nd_shift", "split_9bits", "nested_split_12bits",
                                  "full_key_records_16bits"])
def test_c2_sample_hv_bit_exact(c2, case):
    w, ctx, orc, k, v = c2
    nd, lam, coeff, mode = len(k), 0.0, 1.0, 0
    ctx.set_merge_mode(0, 0)
    if case != "single_pass":
        ctx.set_workspace(1 << 28)            # 256 MB: about 7 target passes of 2000 sources
    if case == "multipass_12bits":
        ctx.set_merge_mode(2, 12)
    elif case == "multipass_16bits":
        ctx.set_merge_mode(2, 16)
    elif case == "mixed_multipass":
        ctx.set_precision(True)
        mode = 2
    elif case == "nd_shift":
        nd, lam, coeff = 1500, 0.41, -0.75
    elif case == "split_9bits":
        ctx.set_workspace(8 << 30)            # one pass: tens of thousands of targets per bucket
        ctx.set_merge_mode(2, 9)
    elif case == "nested_split_12bits":
        ctx.set_merge_mode(3, 12)             # tables counted full at 230 slots: classes split again
    elif case == "full_key_records_16bits":
        ctx.set_merge_mode(4, 16)             # 16-byte records with the dense target (no compaction)
    log("c2 hv case %s" % case)
    try:
        gk, gv, gd = gpu_hv(ctx, k, v, nd, lam, coeff, 60_000_000)
        st = ctx.stats()
    finally:
        ctx.set_workspace(8 << 30)
        ctx.set_merge_mode(0, 0)
        ctx.set_precision(False)
    log("c2 %s: %d passes, hash bits %d, %d candidates, %d unique, %d out, fallback %d" % (
        case, st["n_passes"], st["hash_bits"], st["n_candidates"], st["n_unique"], st["n_out"], st["n_fallback"]))
    ok, ov, od, nc = orc.hv(k, v, nd=nd, lam=lam, coeff=coeff, cut=1e-5, mode=mode)
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0], (gd, od)
    assert st["n_candidates"] == min(nd, len(k)) * ctx.info()["n_conn"]
    if case not in ("single_pass", "split_9bits"):
        assert st["n_passes"] > 1
    assert st["n_fallback"] == 0
    if case == "split_9bits":
        assert st["n_split"] >= 512 and st["hash_bits"] == 9
    if case == "nested_split_12bits":
        assert st["n_split"] > 4096 and st["hash_bits"] == 12
    if case == "multipass_12bits":
        assert st["hash_bits"] == 12
    if case in ("multipass_16bits", "full_key_records_16bits"):
        assert st["hash_bits"] == 16


def test_c2_energy_pull_bit_exact(c2):
    w, ctx, orc, k, v = c2
    vec = ctx.vector(len(k), k, v)
    assert bits([ctx.energy(vec)])[0] == bits([orc.energy(k, v)])[0]


# --- 32 spatial orbitals (64 spin orbitals), 16 electrons, seeds 2 and 4 -----------------

@pytest.fixture(scope="module")
def pool64(wdir):
    """Contexts, oracles and grown inputs of the 64-spin-orbital workloads, made on first use."""
    from sfci_amd import gpu
    made = {}

    def get(name):
        if name not in made:
            w = workload(wdir, name)
            ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
            orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
            k, v = grow(ctx, w["init_hf"], [1000, 1000])
            log("%s input: %d determinants" % (name, len(k)))
            made[name] = (name, w, ctx, orc, k, v)
        return made[name]

    yield get
    for m in made.values():
        m[2].close()


@pytest.fixture(params=["c3", "c5"])
def c64(pool64, request):
    return pool64(request.param)


@pytest.mark.parametrize("mixed", [False, True])
def test_64so_hv_bit_exact(c64, mixed):
    name, w, ctx, orc, k, v = c64
    ctx.set_precision(mixed)
    ctx.set_workspace(1 << 28)
    try:
        gk, gv, gd = gpu_hv(ctx, k, v, len(k), 0.2, 1.0, 60_000_000)
        st = ctx.stats()
    finally:
        ctx.set_precision(False)
        ctx.set_workspace(8 << 30)
    log("%s mixed=%s: %d passes, hash bits %d, %d unique, %d out" % (
        name, mixed, st["n_passes"], st["hash_bits"], st["n_unique"], st["n_out"]))
    ok, ov, od, _ = orc.hv(k, v, nd=len(k), lam=0.2, coeff=1.0, cut=1e-5, mode=2 if mixed else 0)
    assert (gk >> np.uint64(63)).any(), "the product should reach spin orbital 63 (the key's top bit)"
    assert_same_vec(gk, gv, ok, ov)
    assert bits([gd])[0] == bits([od])[0]


def test_64so_mixed_error_bound(c64):
    """Mixed precision against FP64 on the same input.  A mixed off-diagonal term is
    (float)H_ij * (float)c_j, within 3 * 2^-24 of H_ij c_j relative, and the sums are exact,
    so every common output differs by at most 2^-22 * sum_j |H_ij c_j| (plus one FP64
    rounding), which is below 2^-22 * |c|_1 when every |H_ij| < 1."""
    name, w, ctx, orc, k, v = c64
    gk64, gv64, _ = gpu_hv(ctx, k, v, len(k), 0.0, 1.0, 60_000_000)
    ctx.set_precision(True)
    try:
        gk32, gv32, _ = gpu_hv(ctx, k, v, len(k), 0.0, 1.0, 60_000_000)
    finally:
        ctx.set_precision(False)
    common, i64, i32 = np.intersect1d(gk64, gk32, return_indices=True)
    bound = 2.0 ** -22 * np.abs(v).sum()
    err = np.abs(gv64[i64] - gv32[i32]).max()
    log("%s: mixed - fp64 max %.3e (bound %.3e), %d common of %d / %d" % (
        name, err, bound, len(common), len(gk64), len(gk32)))
    assert err <= bound
    assert len(common) >= 0.99 * min(len(gk64), len(gk32))


def test_c5_screened_hv_bit_exact(pool64):
    """The screened product at the learned-screener workload's orbital space: magnitude
    ranking, and an (untrained) network's ranking, each against the oracle's screened
    product with the same kept set."""
    from oracle import screen_oracle as S
    from sfci_amd import gpu
    name, w, ctx, orc, k, v = pool64("c5")
    k, v = k[:300], v[:300] / np.linalg.norm(v[:300])
    nd, eps, keep = len(k), 1e-6, 3.0
    net = gpu.Mlp(ctx, seed=4)
    net.set_norm(-14.0, 0.125, 0.0, 1.0)

    def scores(t, c, r):
        smp = np.zeros(len(t), gpu.SAMPLE_DTYPE)
        smp["det"], smp["cls"], smp["logv"] = t, r, S.plog2(c)
        return net.predict_samples(smp)

    for mode, fn in ((1, lambda t, c, r: np.abs(c)), (2, scores)):
        ctx.set_precision(True)
        ctx.set_screen(mode, net if mode == 2 else None, keep, eps)
        try:
            gk, gv, gd = gpu_hv(ctx, k, v, nd, 0.0, 1.0, 1_000_000)
            st = ctx.stats()
        finally:
            ctx.set_screen(0)
            ctx.set_precision(False)
        log("c5 screened mode %d: %d emitted, %d kept, %d out" % (mode, st["n_candidates"], st["n_kept"], st["n_out"]))
        assert st["n_kept"] < st["n_candidates"]
        ok, ov, od = S.screened_hv(orc, k, v, nd, 0.0, 1.0, eps, keep, fn, cut=1e-5, out_cap=1_000_000)
        assert_same_vec(gk, gv, ok, ov)
        assert bits([gd])[0] == bits([od])[0]


# --- 50 spatial orbitals (100 spin orbitals), 20 electrons, seed 3: 128-bit keys ---------

@pytest.fixture(scope="module")
def c4(wdir):
    from sfci_amd import gpu
    w = workload(wdir, "c4")
    log("c4: loading the 100-spin-orbital tables")
    ctx = gpu.WideContext(w["in_path"], w["n_so"], w["n_elec"])
    orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
    lo, hi = gpu.split_keys([w["init_hf"]])
    v = ctx.vector(1, lo, hi, np.array([1.0]))
    out = ctx.vector(2_000_000)
    ctx.hv(v, out, 1)
    s = ctx.vector(200)
    ctx.shrink_from(s, out)
    ctx.scale(s, 1.0 / ctx.norm(s))
    sl, sh, sv = s.download()
    log("c4 input: %d determinants, %d above spin orbital 63" % (len(sl), int((sh != 0).sum())))
    yield w, ctx, orc, sl, sh, sv
    ctx.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_c4_wide_hv_bit_exact(c4, mixed):
    w, ctx, orc, lo, hi, v = c4
    assert (hi != 0).any()
    ctx.set_precision(mixed)
    try:
        vin = ctx.vector(len(lo), lo, hi, v)
        vout = ctx.vector(60_000_000)
        gd = ctx.hv(vin, vout, len(lo), 0.1, 1.0)
        gl, gh, gv = vout.download()
    finally:
        ctx.set_precision(False)
    ol, oh, ov, od = orc.hv_wide(lo, hi, v, nd=len(lo), lam=0.1, coeff=1.0, mode=2 if mixed else 0)
    log("c4 wide mixed=%s: %d out" % (mixed, len(gl)))
    assert len(gl) == len(ol)
    assert np.array_equal(gl, ol) and np.array_equal(gh, oh)
    assert np.array_equal(bits(gv), bits(ov))
    assert bits([gd])[0] == bits([od])[0]


# --- 12 spatial orbitals, 8 electrons, seed 0: the whole solve ----------------------------

def test_c1_ground_state_matches_oracle(wdir):
    """The reference's ground-state solve at the CPU-runnable workload's full size
    (Nd 2e4, Nl 20, 5 restarts, HPsiCap 6e7) on the GPU and in the oracle: the same number
    of restarts, Ritz values and final energy within 1e-9 Ha."""
    from sfci_amd import gpu
    w = workload(wdir, "c1")
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    try:
        g = ctx.ground_state(w["init_hf"], w["nd"], w["nl"], w["restarts"], e_prev_init=float("inf"))
    finally:
        ctx.close()
    log("c1 gpu: %d restarts, E %.12f" % (g["restarts"], g["e_final"]))
    orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
    o = orc.lanczos(w["init_hf"], w["nd"], w["nl"], w["restarts"], e_prev_init=float("inf"))
    log("c1 oracle: %d restarts, E %.12f" % (o["restarts"], o["e_final"]))
    assert g["restarts"] == o["restarts"]
    assert np.allclose(g["ritz"], o["ritz"], rtol=0, atol=1e-9)
    assert abs(g["e_final"] - o["e_final"]) < 1e-9
