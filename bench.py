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

METRIC = json.load(open(os.path.join(ROOT, "BASELINE.json")))["metric"] if os.path.exists(
    os.path.join(ROOT, "BASELINE.json")) else "H.v connections/s"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (fetch + write) measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on the default workload in FP64, scaled by the counters' reading of a plain
# copy of known size (summary in profiles/); other workloads and precisions report null
PMC_TRAFFIC = {
    ("c2", "fp64"): {
        "k_generate": 1.008e11,
        "k_part_scatter_cur<DigitHash>": 1.679e11,
        "k_part_scatter<DigitHash>": 1.694e11,
        "k_part_hist<DigitHash>": 8.18e10,
        "k_bucket_reduce": 8.19e10,
    },
}


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


def time_to_e(ctx, w, nd, args):
    """The reference's ground-state solve (truncated Lanczos, restarts until the Ritz
    value drops by less than 1e-5 Ha) on the GPU; optionally the CPU oracle's solve."""
    mixed = args.precision == "mixed"
    fp64_ref = None
    if mixed:
        # the plain FP64 solve first: the mixed solve's energy is judged against it
        t = time.perf_counter()
        r64 = ctx.ground_state(w["init_hf"], nd, w["nl"], w["restarts"], hpsi_cap=args.cap,
                               e_prev_init=float("inf"), verbose=args.verbose)
        fp64_ref = dict(seconds=time.perf_counter() - t, e_final=r64["e_final"], restarts=r64["restarts"])
        log("fp64 solve: %.2f s, E = %.12f" % (fp64_ref["seconds"], r64["e_final"]))
        ctx.set_precision(True)
    t = time.perf_counter()
    # the synthetic workloads have positive energies; starting the stopping test from
    # +inf (instead of the reference's 0.0) makes the solve run until two restarts agree
    r = ctx.ground_state(w["init_hf"], nd, w["nl"], w["restarts"], hpsi_cap=args.cap, e_prev_init=float("inf"),
                         verbose=args.verbose)
    dt = time.perf_counter() - t
    ctx.set_precision(False)
    res = dict(metric="time-to-E (truncated Lanczos until two restarts' Ritz values differ by < 1e-5 Ha)",
               value=dt, unit="s",
               higher_is_better=False, dtype="f32/f64" if mixed else "f64", data="synthetic",
               config=dict(workload="%d spatial / %d spin orbitals, %d electrons, seed %d, Nd=%d, Nl=%d, restarts<=%d"
                           % (w["n_orb"], w["n_so"], w["n_elec"], w["seed"], nd, w["nl"], w["restarts"])),
               restarts=r["restarts"], ritz=list(r["ritz"]), e_final=r["e_final"], n_hv=r["n_hv"],
               ms_hv_total=r["ms_hv"])
    if fp64_ref:
        res["fp64_solve"] = fp64_ref
        res["energy_error_vs_fp64"] = r["e_final"] - fp64_ref["e_final"]
        res["speedup_vs_fp64"] = fp64_ref["seconds"] / dt
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
    print(json.dumps(res))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", help="c1..c5 (spatial orbitals / electrons / seed / Nd)")
    ap.add_argument("--nd", type=int, default=None, help="override Nd")
    ap.add_argument("--cap", type=int, default=60_000_000, help="H.v output capacity (HPsiCap)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--data", default=None, help="directory for the generated integral files")
    ap.add_argument("--time-to-e", action="store_true",
                    help="run the whole ground-state solve instead (GPU, and the CPU oracle with --cpu)")
    ap.add_argument("--cpu", action="store_true", help="with --time-to-e: also time the CPU oracle's solve")
    ap.add_argument("--verbose", action="store_true", help="with --time-to-e: print the solver's per-step lines")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="H.v candidate precision: fp64, or mixed (FP32 products, exact sums, FP64 vectors)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: multi-GPU sharding of H.v is described in DESIGN.md and not built",
              file=sys.stderr)
        sys.exit(2)
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
    traffic = PMC_TRAFFIC.get((args.workload, args.precision), {}).get(dom_name) if args.nd is None else None
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
    print(json.dumps(res))
    ctx.close()


if __name__ == "__main__":
    main()
