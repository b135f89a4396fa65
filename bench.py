#!/usr/bin/env python3
"""H.v throughput of the SFCI hot path on one MI355X.

One step = one call of the GPU H.v (SpMVMtply_bucket, src/functions.cpp:511-804) in FP64
on an input vector already resident in HBM.  Workload (default): 26 spatial = 52 spin
orbitals, 14 electrons, synthetic integrals seed 1, Nd = 1,000,000, output capacity
60,000,000 (the reference's HPsiCap).  The input vector is built untimed on the GPU
from the Hartree-Fock determinant by repeated H.v, top-Nd truncation and normalization
(power iteration under the reference's 1e-5 output cut) until it holds Nd determinants.

The printed JSON line carries:
  value          connections/s = n_src * (U + 1) / t_step   (U = connections per det)
  roofline       the dominant kernel by time: its algorithmic bytes per launch over its
                 average launch duration (HIP events on the library's stream)
  cpu_baseline   the oracle (CPU restatement, OpenMP) on a bounded sample of the same
                 input, timed on this host
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "H\u00b7v connections/s + time-to-E within 1e-5 Ha, NN/FP32 vs plain FP64, 1 GPU"
HBM_PEAK_GBS = 8000.0

def _plain(o):
    """numpy scalars and arrays as JSON values"""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError("not JSON serializable: %r" % type(o))


def jdump(res):
    return json.dumps(res, default=_plain)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_input(ctx, w, hf, nd, cap, max_rounds=12):
    """HF -> (H.v -> top-Nd -> normalize) repeated until the vector holds Nd determinants
    (or stops growing); returns the device vector."""
    v = ctx.vector(nd, np.array([hf], np.uint64), np.array([1.0]))
    out = ctx.vector(cap)
    for it in range(max_rounds):
        if v.len >= nd:
            break
        ctx.hv(v, out, nd, 0.0, 1.0)
        nxt = ctx.vector(nd)
        ctx.shrink_from(nxt, out)
        n = ctx.norm(nxt)
        ctx.scale(nxt, 1.0 / n)
        v = nxt
        log("input build %d: len %d" % (it, v.len))
    return v


def cpu_baseline(w, keys, vals, n_conn, target_s=15.0, max_threads=16):
    from oracle import oracle as O
    orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    # calibrate on a small sample, then size the timed sample for ~target_s of CPU work
    s0 = max(1, min(len(keys), int(2e6 // n_conn) + 1))
    t = time.perf_counter()
    orc.hv(keys[:s0], vals[:s0], cut=1e-5, nthreads=threads)
    dt = time.perf_counter() - t
    rate = s0 / max(dt, 1e-6)
    s = int(min(len(keys), max(s0, rate * target_s), 1.2e8 // n_conn + 1))
    t = time.perf_counter()
    orc.hv(keys[:s], vals[:s], cut=1e-5, nthreads=threads)
    dt = time.perf_counter() - t
    return dict(value=s * n_conn / dt, unit="connections/s", cores=threads, kind="port",
                sample="first %d of the %d input determinants (key order), all %d connections each, "
                       "oracle H.v with %d OpenMP threads: %.2f s" % (s, len(keys), n_conn, threads, dt))


def fixed_vector_precision(ctx, v, nd, cap):
    """One H.v of the same vector in FP64 and in mixed precision: diag = <v|H v> of each
    (the energy for a normalized v) and how far the mixed product's entries lie from the
    FP64 ones.  No truncation path is involved: the difference is the FP32 rounding of the
    candidate products alone."""
    out = ctx.vector(cap)
    res = {}
    vals = {}
    for tag, mixed in (("fp64", False), ("mixed", True)):
        ctx.set_precision(mixed)
        try:
            d = ctx.hv(v, out, nd, 0.0, 1.0)
        finally:
            ctx.set_precision(False)
        k, y = out.download()
        res["energy_" + tag] = d
        vals[tag] = (k, y)
    (k64, y64), (k32, y32) = vals["fp64"], vals["mixed"]
    common, i64, i32 = np.intersect1d(k64, k32, assume_unique=True, return_indices=True)
    dy = np.abs(y32[i32] - y64[i64])
    res.update(energy_mixed_minus_fp64=res["energy_mixed"] - res["energy_fp64"],
               n_fp64=int(len(k64)), n_mixed=int(len(k32)), n_common=int(len(common)),
               max_abs_entry_diff=float(dy.max()) if len(dy) else 0.0,
               max_rel_entry_diff=float((dy / np.abs(y64[i64])).max()) if len(dy) else 0.0)
    return res


def time_to_e(ctx, w, nd, args):
    """The reference's ground-state solve (truncated Lanczos, restarts until the Ritz
    value drops by less than 1e-5 Ha) on the GPU; optionally the CPU oracle's solve."""
    mixed = args.precision == "mixed"
    w = dict(w, nl=args.nl or w["nl"], restarts=args.restarts or w["restarts"])
    fp64_ref = None
    fixed = None
    if mixed and args.fp64_energy is not None:
        # the plain FP64 solve's energy from a separate run (a solve pair too long for one job)
        fp64_ref = dict(e_final=args.fp64_energy, seconds=None, restarts=None, given=True)
    elif mixed:
        # the plain FP64 solve first: the mixed solve's energy is judged against it
        t = time.perf_counter()
        r64 = ctx.ground_state(w["init_hf"], nd, w["nl"], w["restarts"], hpsi_cap=args.cap,
                               e_prev_init=float("inf"), verbose=args.verbose)
        fp64_ref = dict(seconds=time.perf_counter() - t, e_final=r64["e_final"], restarts=r64["restarts"])
        log("fp64 solve: %.2f s, E = %.12f" % (fp64_ref["seconds"], r64["e_final"]))
    if mixed:
        ctx.set_precision(True)
    v_fin = ctx.vector(nd) if args.fixed_vector else None
    t = time.perf_counter()
    # the synthetic workloads have positive energies; starting the stopping test from
    # +inf (instead of the reference's 0.0) makes the solve run until two restarts agree
    r = ctx.ground_state(w["init_hf"], nd, w["nl"], w["restarts"], hpsi_cap=args.cap, e_prev_init=float("inf"),
                         verbose=args.verbose, v_out=v_fin)
    dt = time.perf_counter() - t
    ctx.set_precision(False)
    if v_fin is not None:
        fixed = fixed_vector_precision(ctx, v_fin, nd, args.cap)
        log("fixed vector: %s" % fixed)
    res = dict(metric="time-to-E (truncated Lanczos until two restarts' Ritz values differ by < 1e-5 Ha)",
               value=dt, unit="s",
               higher_is_better=False, dtype="f32/f64" if mixed else "f64", data="synthetic",
               config=dict(workload="%d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d, Nl=%d, restarts<=%d"
                           % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd, w["nl"], w["restarts"])),
               restarts=r["restarts"], ritz=list(r["ritz"]), e_final=r["e_final"], n_hv=r["n_hv"],
               ms_hv_total=r["ms_hv"])
    res["ms_restart"] = [float(x) for x in r["ms_restart"]]
    res["e_trunc"] = [float(x) for x in r["e_trunc"]]
    if fixed:
        res["fixed_vector"] = fixed
    if fp64_ref:
        res["fp64_solve"] = fp64_ref
        res["energy_error_vs_fp64"] = r["e_final"] - fp64_ref["e_final"]
        res["speedup_vs_fp64"] = fp64_ref["seconds"] / dt if fp64_ref["seconds"] else None
    if args.cpu:
        from oracle import oracle as O
        orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
        threads = max(1, min(16, os.cpu_count() or 1))
        t = time.perf_counter()
        o = orc.lanczos(w["init_hf"], nd, w["nl"], w["restarts"], hpsi_cap=args.cap, nthreads=threads,
                        e_prev_init=float("inf"))
        tc = time.perf_counter() - t
        res["cpu_baseline"] = dict(value=tc, unit="s", cores=threads, kind="port", e_final=o["e_final"],
                                   restarts=o["restarts"], sample="the same solve with the oracle")
        res["energy_error_vs_oracle"] = r["e_final"] - o["e_final"]
        res["cpu_over_gpu"] = tc / dt
    print(jdump(res))


def train_network(ctx, name, data, seeds, nd_tr, nl_tr, restarts_tr, eps, samples_per_seed, steps, batch, lr, cap):
    """Training samples from plain FP64 solves of the named workload's orbital space on other
    seeds (the last seed held out), and a network trained on them on the GPU.  Returns
    (network, quality on the held-out seed, data seconds, training seconds)."""
    from sfci_amd import gpu, integrals as I
    samples, t_data = [], time.perf_counter()
    for seed in seeds:
        w = I.make_workload(name, data, seed=seed)
        c = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
        v = c.vector(nd_tr)
        r = c.ground_state(w["init_hf"], nd_tr, nl_tr, restarts_tr, hpsi_cap=cap, v_out=v, e_prev_init=float("inf"))
        smp = c.screen_samples(v, eps, samples_per_seed, seed=seed)
        samples.append(smp)
        log("training data seed %d: E %.8f after %d restarts, %d samples" % (seed, r["e_final"], r["restarts"],
                                                                            len(smp)))
        c.close()
    t_data = time.perf_counter() - t_data
    held = samples[-1]
    train = np.concatenate(samples[:-1])
    net = gpu.Mlp(ctx, seed=0)
    net.set_norm(float(train["logv"].mean()), 1.0 / float(train["logv"].std()), float(train["label"].mean()),
                 1.0 / float(train["label"].std()))
    t = time.perf_counter()
    curve = net.train(train, steps=steps, batch=batch, lr=lr, seed=1)
    t_train = time.perf_counter() - t
    pred = net.predict_samples(held)

    def rank_corr(a, b):
        ra, rb = np.argsort(np.argsort(a)), np.argsort(np.argsort(b))
        return float(np.corrcoef(ra, rb)[0, 1])

    def top_recall(score, frac=0.1):
        k = max(1, int(frac * len(held)))
        top_true = set(np.argsort(-held["label"])[:k].tolist())
        return len(top_true & set(np.argsort(-score)[:k].tolist())) / k

    quality = dict(spearman_network=rank_corr(pred, held["label"]),
                   spearman_magnitude=rank_corr(held["logv"], held["label"]),
                   top10pct_recall_network=top_recall(pred), top10pct_recall_magnitude=top_recall(held["logv"]),
                   loss_first=float(curve[0]), loss_last=float(np.mean(curve[-20:])), train_samples=int(len(train)))
    return net, quality, t_data, t_train


def time_to_e_block(args):
    """Time to the ground-state energy at the CPU-runnable workload (12 spatial / 24 spin
    orbitals, 8 electrons, seed 0, Nd 2e4, Nl 20, at most 5 restarts): the reference's solve
    (src/sparse_solver.cpp:246-548, timed as src/main.cpp:267-276) in plain FP64, in mixed
    precision, and in mixed precision with the learned screener (network trained here on seeds
    10-18 of the same orbital space, 19 held out; candidates ranked by the network's predicted
    ln|y| of their target plus ln|H_ij c_j|, by the network alone, and by |H_ij c_j| alone),
    each on the GPU, plus the oracle's FP64 solve on the host cores.  Every energy is compared
    with the oracle's."""
    from oracle import oracle as O
    from sfci_amd import gpu, integrals as I

    data = args.data or os.path.join(tempfile.gettempdir(), "sfci_bench")
    w = I.make_workload("c1", data)
    nd, nl, nr = w["nd"], w["nl"], w["restarts"]
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    net, quality, t_data, t_train = train_network(ctx, "c1", data, range(10, 20), nd, nl, nr, args.eps, 400_000,
                                                  args.train_steps, args.batch, args.lr, args.cap)

    def solve(mode, mixed):
        ctx.set_precision(mixed)
        ctx.set_screen(mode, net if mode >= 2 else None, 3.0, args.eps)
        try:
            ctx.ground_state(w["init_hf"], nd, nl, nr, hpsi_cap=args.cap, e_prev_init=float("inf"))   # warm
            ctx.sync()
            t = time.perf_counter()
            r = ctx.ground_state(w["init_hf"], nd, nl, nr, hpsi_cap=args.cap, e_prev_init=float("inf"))
            dt = time.perf_counter() - t
        finally:
            ctx.set_screen(0)
            ctx.set_precision(False)
        return dict(seconds=dt, e_final=r["e_final"], restarts=r["restarts"], n_hv=r["n_hv"],
                    e_trunc=[float(x) for x in r["e_trunc"]], s_restart=[x / 1e3 for x in r["ms_restart"]],
                    screened_restarts=r["n_screened"], rejected_restarts=r["n_reverts"])

    # screened runs keep 3 Nd candidates per product, ranked by: the network's predicted ln|y|
    # of the target plus ln|H_ij c_j| (mode 3), the network alone (2), |H_ij c_j| alone (1)
    runs = dict(fp64=solve(0, False), mixed=solve(0, True), network_magnitude_screened_mixed=solve(3, True),
                network_screened_mixed=solve(2, True), magnitude_screened_mixed=solve(1, True))
    ctx.close()
    orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
    threads = max(1, min(16, os.cpu_count() or 1))
    t = time.perf_counter()
    o = orc.lanczos(w["init_hf"], nd, nl, nr, hpsi_cap=args.cap, nthreads=threads, e_prev_init=float("inf"))
    t_cpu = time.perf_counter() - t
    e64 = runs["fp64"]["e_final"]
    for tag, x in runs.items():
        x["energy_error_vs_oracle"] = x["e_final"] - o["e_final"]
        # time to E: the end of the first restart whose energy is within 1e-5 Ha of the plain
        # FP64 GPU solve's final energy
        hit = [t for e, t in zip(x["e_trunc"], x["s_restart"]) if e <= e64 + 1e-5]
        x["time_to_e_s"] = hit[0] if hit else None
    t64 = runs["fp64"]["time_to_e_s"]
    for tag, x in runs.items():
        x["time_to_e_speedup_vs_fp64"] = t64 / x["time_to_e_s"] if x["time_to_e_s"] else None
        log("time to E, %s: %.3f s, E %.12f, error vs oracle %.2e" % (tag, x["seconds"], x["e_final"],
                                                                      x["energy_error_vs_oracle"]))
    return dict(workload="12 spatial / 24 spin orbitals, 8 electrons, seed 0, Nd=%d, Nl=%d, restarts<=%d, "
                         "HPsiCap=%d" % (nd, nl, nr, args.cap),
                stop="two restarts' Ritz values differ by < 1e-5 Ha (first comparison against +inf)",
                **runs,
                oracle_fp64=dict(seconds=t_cpu, e_final=o["e_final"], restarts=o["restarts"], cores=threads,
                                 kind="port"),
                gpu_fp64_over_oracle=t_cpu / runs["fp64"]["seconds"],
                screener=dict(keep_factor=3.0, eps_pre=args.eps, train_seeds="10-18 (19 held out)",
                              data_seconds=t_data, train_seconds=t_train, steps=args.train_steps, batch=args.batch,
                              **quality))


def screener(args):
    """Learned-screener workload (32 spatial / 64 spin orbitals, 16 electrons): training
    samples from converged plain FP64 solves on seeds 10-19, the network trained on the GPU,
    then three solves of seed 4 timed from the HF determinant: plain FP64, network-screened
    mixed precision, and the same screening ranked by |H_ij c_j| alone (no network)."""
    from sfci_amd import gpu, integrals as I

    data = args.data or os.path.join(tempfile.gettempdir(), "sfci_bench")
    nd, nl, nd_tr = args.nd or 200_000, args.nl or 30, args.nd_train
    w = I.make_workload("c5", data)
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    ctx.kernel_stats(reset=True)
    net, quality, t_data, t_train = train_network(ctx, "c5", data, range(10, 20), nd_tr, 10, 3, args.eps,
                                                  args.samples_per_seed, args.train_steps, args.batch, args.lr,
                                                  args.cap)
    ks = ctx.kernel_stats()["mlp_train_step"]
    log("network: %s" % quality)

    keeps = [float(k) for k in str(args.keep).split(",")]

    def solve(tag, mode, mixed, keep=keeps[0], delta_e=1e-5):
        ctx.set_precision(mixed)
        ctx.set_screen(mode, net if mode >= 2 else None, keep, args.eps)
        ctx.sync()
        t0 = time.perf_counter()
        r = ctx.ground_state(w["init_hf"], nd, nl, args.restarts or 8, hpsi_cap=args.cap, e_prev_init=float("inf"),
                             verbose=True, delta_e=delta_e)
        dt = time.perf_counter() - t0
        ctx.set_screen(0)
        ctx.set_precision(False)
        out = dict(seconds=dt, e_final=r["e_final"], restarts=r["restarts"], n_hv=r["n_hv"], ms_hv=r["ms_hv"],
                   screened_restarts=r["n_screened"], rejected=r["n_reverts"], e_trunc=list(r["e_trunc"]),
                   s_restart=[x / 1e3 for x in r["ms_restart"]])
        log("%s: %.2f s, E %.10f, %d restarts (%d screened, %d rejected)" % (
            tag, dt, r["e_final"], r["restarts"], r["n_screened"], r["n_reverts"]))
        return out

    # the plain FP64 reference runs every restart (no early stop: at a truncated Nd the Ritz
    # value can rise before it settles); its energy is the lowest restart energy
    plain = solve("plain fp64 (all restarts)", 0, False, delta_e=-float("inf"))
    e_ref = float(min(plain["e_trunc"]))
    plain["e_reference"] = e_ref
    hit = [t for e, t in zip(plain["e_trunc"], plain["s_restart"]) if e <= e_ref + 1e-5]
    plain["time_to_reference_s"] = hit[0]
    # one product of a converged-solve-sized vector in each mode (diagnostics)
    vin = ctx.vector(nd)
    ctx.ground_state(w["init_hf"], nd, 4, 1, hpsi_cap=args.cap, v_out=vin, e_prev_init=float("inf"))
    out = ctx.vector(args.cap)
    single = {}
    for tag, mode, mixed in (("plain_fp64", 0, False), ("plain_mixed", 0, True), ("magnitude", 1, True),
                             ("network", 2, True), ("network_magnitude", 3, True)):
        ctx.set_precision(mixed)
        ctx.set_screen(mode, net if mode >= 2 else None, keeps[0], args.eps)
        ctx.hv(vin, out, nd)
        ctx.hv(vin, out, nd)
        st = ctx.stats()
        ctx.set_screen(0)
        ctx.set_precision(False)
        single[tag] = {k: st[k] for k in ("ms_total", "ms_generate", "ms_score", "ms_select", "ms_sort", "ms_reduce",
                                          "n_candidates", "n_kept", "n_unique", "n_out", "screen_eps")}
        log("single H.v %s: %s" % (tag, single[tag]))
    del vin, out
    def annotate(x):
        """time to E: the end of the first restart whose (judged) energy is at or below the
        plain FP64 reference energy + 1e-5 Ha, against the plain solve's own time to it."""
        x["energy_error_vs_plain_fp64"] = x["e_final"] - e_ref
        x["within_1e-5"] = bool(x["e_final"] <= e_ref + 1e-5)
        hit = [t for e, t in zip(x["e_trunc"], x["s_restart"]) if e <= e_ref + 1e-5]
        x["time_to_plain_energy_s"] = hit[0] if hit else None
        x["speedup_to_plain_energy"] = plain["time_to_reference_s"] / hit[0] if hit else None

    runs = {}
    abl = [float(k) for k in str(args.ablation_keep).split(",") if k]
    for kf in keeps:
        runs[kf] = dict(network_magnitude=solve("network x magnitude screened mixed, keep %g Nd" % kf, 3, True, kf),
                        network=solve("network-screened mixed, keep %g Nd" % kf, 2, True, kf))
        if kf in abl:
            runs[kf]["magnitude"] = solve("magnitude-screened mixed, keep %g Nd" % kf, 1, True, kf)
        for x in runs[kf].values():
            x["keep_factor"] = kf
            annotate(x)
    # the prefilter alone: every candidate with |H_ij c_j| > eps merged, nothing ranked
    pre = solve("prefilter only (eps %g, nothing dropped by rank)" % args.eps, 1, True, 1e12)
    annotate(pre)
    nn, mag = runs[keeps[0]]["network_magnitude"], runs[keeps[0]].get("magnitude")
    res = dict(metric="time-to-E within 1e-5 Ha of the plain FP64 GPU solve (learned screener)",
               value=nn["time_to_plain_energy_s"] if nn["time_to_plain_energy_s"] is not None else nn["seconds"],
               unit="s", higher_is_better=False, dtype="bf16 network / f32 products / f64",
               data="synthetic",
               config=dict(workload="%d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d, Nl=%d (reduced Nd)"
                           % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd, nl),
                           keep_factor=keeps[0], eps_pre=args.eps, train_seeds="10-18 (19 held out)",
                           train_nd=nd_tr, train_samples=quality["train_samples"], train_steps=args.train_steps,
                           batch=args.batch),
               plain_fp64=plain, network_magnitude_screened=nn, network_only_screened=runs[keeps[0]]["network"],
               magnitude_screened=mag, network_quality=quality,
               by_keep_factor={str(k): v for k, v in runs.items()}, prefilter_only=pre,
               single_hv=single,
               training=dict(seconds=t_train, data_seconds=t_data, ms_per_step=ks["ms"] / max(1, ks["launches"])))
    print(jdump(res))
    ctx.close()


def copy_probe():
    """A plain copy of known size through the library (k_copy: 2^26 keys and 2^26 values,
    1 GiB read and 1 GiB written), so that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE can be
    calibrated against it."""
    from sfci_amd import gpu, integrals as I
    import ctypes as C

    w = I.make_workload("c1", os.path.join(tempfile.gettempdir(), "sfci_bench"))
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    n = 1 << 26
    a, b = ctx.vector(n), ctx.vector(n)
    ctx._check(gpu.lib().sfci_vec_set_len(a._h, n))
    b.copy_from(a)
    ctx.sync()
    print(jdump(dict(copy_bytes_read=16 * n, copy_bytes_written=16 * n)))
    ctx.close()


def screen_hv(args):
    """One screened H.v of a grown input on the learned-screener workload (seed 4), in each
    scoring mode, after a short training of the network on samples of the same product:
    reports the network kernel's MFMA throughput (165 kFLOP per scored candidate)."""
    from sfci_amd import gpu, integrals as I

    data = args.data or os.path.join(tempfile.gettempdir(), "sfci_bench")
    w = I.make_workload("c5", data)
    nd = args.nd or 200_000
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    v = build_input(ctx, w, w["init_hf"], nd, args.cap)
    smp = ctx.screen_samples(v, args.eps, 400_000, seed=1)
    net = gpu.Mlp(ctx, seed=0)
    net.set_norm(float(smp["logv"].mean()), 1.0 / float(smp["logv"].std()), float(smp["label"].mean()),
                 1.0 / float(smp["label"].std()))
    net.train(smp, steps=200, batch=args.batch, lr=args.lr, seed=1)
    out = ctx.vector(args.cap)
    res = {}
    for tag, mode in (("plain_mixed", 0), ("magnitude", 1), ("network", 2), ("network_magnitude", 3)):
        ctx.set_precision(True)
        ctx.set_screen(mode, net if mode >= 2 else None, float(str(args.keep).split(",")[0]), args.eps)
        ctx.hv(v, out, nd)
        ctx.sync()
        ctx.kernel_stats(reset=True)
        t = time.perf_counter()
        for _ in range(args.steps):
            ctx.hv(v, out, nd)
        ctx.sync()
        dt = (time.perf_counter() - t) / args.steps
        st = ctx.stats()
        ks = ctx.kernel_stats()
        ctx.set_screen(0)
        ctx.set_precision(False)
        r = dict(ms=dt * 1e3, emitted=st["n_candidates"], kept=st["n_kept"], out=st["n_out"],
                 ms_generate=st["ms_generate"], ms_score=st["ms_score"], ms_select=st["ms_select"])
        if mode == 2:
            k = ks["k_mlp_fwd"]
            per = k["ms"] / max(1, k["launches"])
            flops = 2.0 * (80 * 256 + 256 * 256) + 2.0 * 256
            r["k_mlp_fwd"] = dict(avg_launch_ms=per, rows=st["n_candidates"],
                                  tflops=flops * st["n_candidates"] / (per * 1e-3) / 1e12,
                                  frac_of_bf16_peak=flops * st["n_candidates"] / (per * 1e-3) / 2.5e15)
        res[tag] = r
        log("%s: %s" % (tag, r))
    print(jdump(dict(metric="screened H.v (learned-screener workload)", unit="ms", config=dict(
        workload="%d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d" % (w["n_orb"], w["n_so"], w["n_elec"],
                                                                               w["seed"], nd),
        eps_pre=args.eps, keep_factor=args.keep), modes=res)))
    ctx.close()


def wide_hv(args):
    """H.v on 128-bit keys: the strong-correlation workload (50 spatial / 100 spin orbitals,
    20 electrons, seed 3) at a reduced Nd (default 1e4: its full Nd of 1e7 would be 2.3e12
    connections per H.v).  The input is grown from HF by H.v with output capacity Nd,
    normalized on the host between rounds (not timed)."""
    from sfci_amd import gpu, integrals as I

    data = args.data or os.path.join(tempfile.gettempdir(), "sfci_bench")
    t0 = time.perf_counter()
    w = I.make_workload("c4", data)
    nd = args.nd or 10_000
    log("integrals ready (%.1f s): %s" % (time.perf_counter() - t0, w["in_path"]))
    t0 = time.perf_counter()
    ctx = gpu.WideContext(w["in_path"], w["n_so"], w["n_elec"])
    log("context ready (%.1f s)" % (time.perf_counter() - t0))
    n_conn = ctx.info()["n_conn"]
    mixed = args.precision == "mixed"
    hf = w["init_hf"]
    if args.time_to_e:
        # the ground-state solve on 128-bit keys, FP64 first, then (with --precision mixed)
        # mixed precision judged against it
        nl = args.nl or w["nl"]
        restarts = args.restarts or w["restarts"]
        res = dict(metric="time-to-E, 128-bit keys (truncated Lanczos until two restarts' Ritz values differ by "
                          "< 1e-5 Ha)", unit="s", higher_is_better=False, data="synthetic",
                   config=dict(workload="%d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d (reduced), "
                                        "Nl=%d, restarts<=%d" % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd,
                                                                  nl, restarts)))
        for tag, mx in (("fp64", False), ("mixed", True)) if mixed else (("fp64", False),):
            ctx.set_precision(mx)
            t = time.perf_counter()
            r = ctx.ground_state(hf, nd, nl, restarts, hpsi_cap=args.cap, e_prev_init=float("inf"),
                                 verbose=args.verbose)
            dt = time.perf_counter() - t
            res[tag] = dict(seconds=dt, restarts=r["restarts"], ritz=list(r["ritz"]), e_trunc=list(r["e_trunc"]),
                            e_final=r["e_final"], n_hv=r["n_hv"], ms_hv_total=r["ms_hv"])
            log("%s solve: %.2f s, E = %.12f" % (tag, dt, r["e_final"]))
        ctx.set_precision(False)
        res["value"] = res["fp64"]["seconds"]
        if mixed:
            res["energy_error_mixed_vs_fp64"] = res["mixed"]["e_final"] - res["fp64"]["e_final"]
            res["speedup_mixed_vs_fp64"] = res["fp64"]["seconds"] / res["mixed"]["seconds"]
        print(jdump(res))
        ctx.close()
        return
    lo, hi = gpu.split_keys([hf])
    v = ctx.vector(nd, lo, hi, np.array([1.0]))
    out = ctx.vector(nd)
    for it in range(6):
        if v.len >= nd:
            break
        ctx.hv(v, out, nd)
        ol, oh, ov = out.download()
        v = ctx.vector(nd, ol, oh, ov / np.sqrt(np.dot(ov, ov)))
        log("grow %d: %d determinants" % (it, v.len))
    ctx.set_precision(mixed)
    big = ctx.vector(args.cap)
    for i in range(args.warmup):
        ctx.hv(v, big, nd)
        log("warmup %d: %s" % (i, ctx.stats()))
    t = time.perf_counter()
    sts = []
    for i in range(args.steps):
        ctx.hv(v, big, nd)
        sts.append(ctx.stats())
        log("step %d: %s" % (i, sts[-1]))
    dt = (time.perf_counter() - t) / args.steps
    n_in = v.len
    conn = n_in * n_conn
    avg = {k: sum(x[k] for x in sts) / len(sts) for k in ("ms_generate", "ms_sort", "ms_reduce", "ms_select",
                                                          "ms_total", "n_passes", "n_unique", "n_out", "n_split",
                                                          "n_fallback")}
    res = dict(metric="H.v connections/s, 128-bit keys", value=conn / dt, unit="connections/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=dt * 1e3, higher_is_better=True,
               dtype="f32/f64" if mixed else "f64", data="synthetic",
               config=dict(workload="H.v, %d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d (reduced), "
                           "cap=%d" % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd, args.cap),
                           n_in=n_in, n_conn=n_conn, candidates=conn),
               phases_ms=avg,
               generate_write_gbs=26.0 * conn / (avg["ms_generate"] * 1e-3) / 1e9 if avg["ms_generate"] else None)
    if not args.no_cpu:
        from oracle import oracle as O
        orc = O.Oracle(w["in_path"], w["n_so"], w["n_elec"])
        vl, vh, vv = v.download()
        threads = max(1, min(16, os.cpu_count() or 1))
        m = min(len(vl), 1000)
        t = time.perf_counter()
        orc.hv_wide(vl[:m], vh[:m], vv[:m], nthreads=threads)
        tc = time.perf_counter() - t
        res["cpu_baseline"] = dict(value=m * n_conn / tc, unit="connections/s", cores=threads, kind="port",
                                   sample="first %d of the %d input determinants (key order), oracle H.v on "
                                          "128-bit keys with %d OpenMP threads: %.2f s" % (m, n_in, threads, tc))
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    print(jdump(res))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", help="c1..c5 (spatial orbitals / electrons / seed / Nd)")
    ap.add_argument("--nd", type=int, default=None, help="override Nd")
    ap.add_argument("--cap", type=int, default=60_000_000, help="H.v output capacity (HPsiCap)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tte", action="store_true", help="leave out the time-to-E block of the default run")
    ap.add_argument("--full-key-records", action="store_true",
                    help="merge with 16-byte full-key records and 16 hash bits (no compact records)")
    ap.add_argument("--workspace", type=float, default=None,
                    help="GiB the H.v candidate buffers of one target pass may use (fewer: more passes)")
    ap.add_argument("--data", default=None, help="directory for the generated integral files")
    ap.add_argument("--time-to-e", action="store_true",
                    help="run the whole ground-state solve instead (GPU, and the CPU oracle with --cpu)")
    ap.add_argument("--cpu", action="store_true", help="with --time-to-e: also time the CPU oracle's solve")
    ap.add_argument("--verbose", action="store_true", help="with --time-to-e: print the solver's per-step lines")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="H.v candidate precision: fp64, or mixed (FP32 products, exact sums, FP64 vectors)")
    ap.add_argument("--screener", action="store_true",
                    help="learned-screener workload: train on seeds 10-19, time plain / network / magnitude solves")
    ap.add_argument("--screen-hv", action="store_true",
                    help="time one screened H.v per scoring mode on the learned-screener workload")
    ap.add_argument("--wide", action="store_true",
                    help="H.v on 128-bit keys (100 spin orbitals, 20 electrons, seed 3; reduced Nd)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one plain device copy of 1 GiB read + 1 GiB written (counter calibration)")
    ap.add_argument("--nl", type=int, default=None)
    ap.add_argument("--fp64-energy", type=float, default=None,
                    help="with --time-to-e --precision mixed: the plain FP64 solve's energy from a separate run")
    ap.add_argument("--fixed-vector", action="store_true",
                    help="with --time-to-e: FP64 and mixed H.v of the solve's final vector (precision alone)")
    ap.add_argument("--restarts", type=int, default=None)
    ap.add_argument("--nd-train", type=int, default=50_000)
    ap.add_argument("--samples-per-seed", type=int, default=400_000)
    ap.add_argument("--eps", type=float, default=1e-6, help="screening prefilter |H_ij c_j| > eps")
    ap.add_argument("--keep", default="3", help="screening keeps keep * Nd candidates (comma list: one solve each)")
    ap.add_argument("--ablation-keep", default="3", help="keep factors also solved with |H_ij c_j| ranking")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: sharding H.v across GPUs (by target row range) is a design, not built",
              file=sys.stderr)
        sys.exit(2)
    if args.screener:
        screener(args)
        return
    if args.screen_hv:
        screen_hv(args)
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.wide:
        wide_hv(args)
        return
    from sfci_amd import gpu, integrals as I

    data = args.data or os.path.join(tempfile.gettempdir(), "sfci_bench")
    t0 = time.perf_counter()
    w = I.make_workload(args.workload, data)
    nd = args.nd or w["nd"]
    log("integrals ready (%.1f s): %s" % (time.perf_counter() - t0, w["in_path"]))
    ctx = gpu.Context(w["in_path"], w["n_so"], w["n_elec"])
    if args.time_to_e:
        time_to_e(ctx, w, nd, args)
        return
    info = ctx.info()
    n_conn = info["n_conn"]
    v = build_input(ctx, w, w["init_hf"], nd, args.cap)
    out = ctx.vector(args.cap)
    mixed = args.precision == "mixed"
    ctx.set_precision(mixed)
    if args.full_key_records:
        ctx.set_merge_mode(4, 16)
    if args.workspace:
        ctx.set_workspace(int(args.workspace * (1 << 30)))
    rec = 12.0 if mixed else 16.0        # bytes of one candidate record (key + value)
    for i in range(args.warmup):
        ctx.hv(v, out, nd, 0.0, 1.0)
        log("warmup %d: %s" % (i, ctx.stats()))
    ctx.sync()
    ctx.kernel_stats(reset=True)
    t = time.perf_counter()
    tot_cand = tot_out = tot_uniq = 0
    for i in range(args.steps):
        ctx.hv(v, out, nd, 0.0, 1.0)
        st = ctx.stats()
        tot_cand += st["n_candidates"]
        tot_out += st["n_out"]
        tot_uniq += st["n_unique"]
        log("step %d: %.1f ms (%d passes, %d candidates, %d unique, %d out)" % (
            i, st["ms_total"], st["n_passes"], st["n_candidates"], st["n_unique"], st["n_out"]))
    ctx.sync()
    dt = time.perf_counter() - t
    ks = ctx.kernel_stats()
    n_in = v.len
    conn = n_in * n_conn
    ms_step = dt * 1e3 / args.steps
    dom = max(ks.items(), key=lambda kv: kv[1]["ms"])
    dom_name, d = dom
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    # HBM bytes per launch need rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE runs of their own (see
    # profiles/); they are not collected in this run
    traffic = None
    roofline = dict(bound="hbm", kernel=dom_name, achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=achieved / HBM_PEAK_GBS, traffic=traffic,
                    avg_launch_ms=d["ms"] / max(1, d["launches"]),
                    bytes_per_launch=d["bytes"] / max(1, d["launches"]))
    # whole-H.v bytes of an ideal single-pass design: N_in and M_out at 16 bytes, the
    # M_c candidate records written once and read once
    m_c = tot_cand / args.steps
    m_out = tot_out / args.steps
    b_hv = 16.0 * (n_in + m_out) + 2 * rec * m_c
    # sort + merge time per step (radix passes, their scans, and the exact reduce)
    t_sm = sum(x["ms"] for k, x in ks.items()
               if k.startswith(("k_radix_", "k_part_", "scan", "k_bucket_", "k_reduce"))) / args.steps
    res = dict(metric=METRIC, value=conn / (ms_step * 1e-3), unit="connections/s", n_gpus=1,
               steps=args.steps, warmup=args.warmup, ms_per_step=ms_step, higher_is_better=True,
               scaling="weak", vs_baseline=None, dtype="f32/f64" if mixed else "f64", data="synthetic",
               config=dict(workload="H.v, %d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d, cap=%d"
                           % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd, args.cap),
                           n_in=n_in, n_conn=n_conn, candidates=int(m_c), unique=int(tot_uniq / args.steps),
                           n_out=int(m_out)),
               roofline=roofline,
               determinants_merged_per_s=m_c / (t_sm * 1e-3 + 1e-30),
               sort_merge_roofline_frac=(rec * m_c + 16.0 * m_out) / (t_sm * 1e-3 + 1e-30) / 1e9 / HBM_PEAK_GBS,
               hv_ideal_bytes_frac=b_hv / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
               kernels={k: dict(ms_per_step=x["ms"] / args.steps, launches=x["launches"],
                                gbs=(x["bytes"] / (x["ms"] * 1e-3) / 1e9) if x["ms"] > 0 else 0.0)
                        for k, x in ks.items() if x["launches"]})
    if mixed:
        res["precision"] = "FP32 candidate products, exact 128-bit fixed-point sums, FP64 vectors"
    # energy readout <v|H|v> (calc_energy) on the same vector, outside the timed steps
    ctx.sync()
    t = time.perf_counter()
    e = ctx.energy(v)
    ctx.sync()
    res["energy_readout"] = dict(ms=(time.perf_counter() - t) * 1e3, energy=e)
    if not args.no_cpu:
        keys, vals = v.download()
        res["cpu_baseline"] = cpu_baseline(w, keys, vals, n_conn)
        res["gpu_over_cpu"] = res["value"] / res["cpu_baseline"]["value"]
    else:
        res["cpu_baseline"] = None
    del v, out
    ctx.close()
    if not args.no_tte and args.nd is None and args.workload == "c2" and not mixed:
        res["time_to_e"] = time_to_e_block(args)
    print(jdump(res))


if __name__ == "__main__":
    main()
