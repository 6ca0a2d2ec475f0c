#!/usr/bin/env python3
"""Benchmark: exhaustive third-order K2 scan, triplets x samples per second.

Workload (one "step" = one full scan of one cohort, inputs resident in HBM):
4096 SNPs x 16384 samples (8192 cases / 8192 controls), Hardy-Weinberg
genotypes with per-SNP MAF ~ U[0.05, 0.5], seed 3; C(4096,3) = 1.14e10
triplets.  Prints one JSON line with the metric, the roofline of the scan
kernel and the reference CPU baseline timed on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                  [--no-batch] [--no-serve]

More legs are reported in the same line (skippable); their cohorts are the
ones cross-dpc-episdet_amd/workloads.py defines, and where the reference
program's result for a cohort is recorded (tests/golden/
workload_reference.json) the line counts how many results print the same:
  configs: 256 SNPs x 1024 samples (HWE) and 2048 x 4096 with a planted
    triplet.
  batched: 64 cohorts (SNPs ~ U[128,1024], samples ~ U[512,8192], case
    fraction ~ U[0.3,0.7]) in one graph-captured persistent launch, against
    64 separate plain scans; per-cohort results must be identical.
  serving: 512 requests (SNPs log-U[64,2048], samples log-U[256,16384], case
    fraction U[0.3,0.7]).  First all submitted at once, which measures the
    server's capacity on this mix; then Poisson arrivals at 0.8 x that
    capacity, against one-request-at-a-time plain scans (FIFO) under the
    same arrival times, each request's service its cohort's creation and
    scan (`plain_fifo`; the scans alone in `plain_fifo_scan_only`); p50/p99
    latency for all, identical results.

The roofline's `frac` counts the operations the kernel executes: 8 cells
per triplet-sample are multiplied (MFMA) or popcounted (VALU), the other 19
derived from margins.  The reference counts all 27; the rate at its 54
ops per triplet-sample is reported apart as `reference_equivalent`.
"""
import argparse
import importlib.util
import threading
import json
import math
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cross-dpc-episdet_amd")
ORACLE = os.path.join(ROOT, "oracle")

# MI355X integer VALU issue rate: 256 CU x 64 lanes x 2.4 GHz (lane-ops/s)
VALU_PEAK_OPS = 256 * 64 * 2.4e9
# MI355X dense int8 and FP4 matrix-core rates (ops/s)
MFMA_I8_PEAK_OPS = 5.0e15
MFMA_FP4_PEAK_OPS = 10.1e15
# Work per triplet-sample.  Executed (the roofline's unit): the kernels
# multiply (or popcount) only the 8 cells without genotype 2 and derive the
# other 19 from pair and single margins: 8 MACs = 16 ops per triplet-sample
# on the matrix cores, or 8 AND + 8 popcount lane-ops per 32 samples on the
# VALU.  The reference's 27-cell count is 54 ops (27 MACs, or 27 AND +
# 27 popcount per 32 samples); the rate at that count is reported apart.
WORK = {
    1: {"bound": "valu", "peak": VALU_PEAK_OPS, "executed_ops": 16 / 32, "reference_ops": 54 / 32,
        "unit_text": "16 int32 lane-ops (8 AND + 8 popcount) per (triplet x 32 samples)"},
    2: {"bound": "mfma", "peak": MFMA_I8_PEAK_OPS, "executed_ops": 16.0, "reference_ops": 54.0,
        "unit_text": "16 int8 ops (8 MACs of 0/1 bytes) per triplet-sample"},
    3: {"bound": "mfma", "peak": MFMA_FP4_PEAK_OPS, "executed_ops": 16.0, "reference_ops": 54.0,
        "unit_text": "16 FP4 ops (8 MACs of 0/1 e2m1 values, exact FP32 sums) per "
                     "triplet-sample"},
}


KERNELS = {1: "scan_kernel (VALU AND + popcount)", 2: "mfma_scan_kernel (int8 MFMA)",
           3: "mfma_scan_kernel (FP4 MFMA)"}
DTYPES = {1: "u32", 2: "i8", 3: "fp4"}


def load_binding():
    spec = importlib.util.spec_from_file_location("episdet", os.path.join(PKG, "episdet.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["episdet"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tsc_hz():
    sys.path.insert(0, ORACLE)
    import oracle
    t0, c0 = time.perf_counter(), oracle.rdtsc()
    time.sleep(0.5)
    t1, c1 = time.perf_counter(), oracle.rdtsc()
    return (c1 - c0) / (t1 - t0)


def host_cpu():
    """CPU model, physical cores of the host (distinct socket/core pairs in
    /proc/cpuinfo) and the logical CPUs this process may use."""
    model, cores, cur = None, set(), {}
    with open("/proc/cpuinfo") as f:
        for line in f.read().splitlines() + [""]:
            if not line.strip():
                if "core id" in cur:
                    cores.add((cur.get("physical id", "0"), cur["core id"]))
                cur = {}
                continue
            k, _, v = line.partition(":")
            cur[k.strip()] = v.strip()
            if k.strip() == "model name" and model is None:
                model = v.strip()
    return {"model": model, "physical_cores": len(cores) or os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0))}


# Threads the reference CPU scan runs with: the CPU share a one-GPU host
# gives one job (16 threads), or fewer where fewer CPUs are usable.  The
# reference fixes its thread count at compile time (CPU/Makefile:4); builds
# for 8 and 16 threads exist under oracle/_ref.
CPU_THREADS_MAX = 16


def cpu_baseline(n_samples, n_snps_sample, n_snps_full):
    """Reference CPU v4 (AVX-512 VPOPCNTDQ build, or the AVX2 build when the
    host lacks it), compiled from the reference sources into oracle/_ref with
    FREQ=1 so it prints TSC cycles; converted with a measured TSC rate.  The
    sample scans n_snps_sample SNPs; the rate (triplets x samples per second)
    is taken to hold at n_snps_full, whose scan time is extrapolated."""
    flags = open("/proc/cpuinfo").read()
    cpu = host_cpu()
    threads = CPU_THREADS_MAX if cpu["usable_cpus"] >= CPU_THREADS_MAX else 8
    if "avx512_vpopcntdq" in flags:
        exe, extra = f"v4_popc_avx512_t{threads}", ["128", "16"]
    else:
        exe, extra = f"v4_avx_t{threads}", ["64", "32"]
    path = os.path.join(ORACLE, "_ref", exe)
    if not os.path.exists(path):
        return None
    hz = tsc_hz()
    log(f"cpu baseline: {exe} {n_samples} {n_snps_sample} (TSC {hz / 1e9:.3f} GHz)")
    out = subprocess.run([path, str(n_samples), str(n_snps_sample)] + extra,
                         capture_output=True, text=True, timeout=600, check=True).stdout
    cycles = float(out.split("Time:")[1].split()[0])
    secs = cycles / hz
    rate = trip(n_snps_sample) * n_samples / secs
    return {"value": rate, "unit": "triplets*samples/s", "cores": threads,
            "kind": "reference",
            "sample": f"{exe} block_pac={extra[0]} block_snp={extra[1]} on N={n_samples}, "
                      f"M={n_snps_sample} (reference generator), scan {secs:.2f} s",
            "cpu_model": cpu["model"], "host_physical_cores": cpu["physical_cores"],
            "usable_cpus": cpu["usable_cpus"],
            "threads_note": f"{threads} threads: the CPU share of one GPU's job on this host; "
                            "the reference's thread count is a compile-time constant",
            "per_thread_value": rate / threads,
            "all_cores_linear_estimate": rate / threads * cpu["physical_cores"],
            "extrapolation": {"snps": n_snps_full,
                              "scan_s": trip(n_snps_full) * n_samples / rate,
                              "assumes": "the same triplets x samples/s at "
                                         f"{n_snps_full} SNPs as at {n_snps_sample}"}}


def same_result(a, b):
    return ((a.i, a.j, a.k) == (b.i, b.j, b.k) and
            [(c.i, c.j, c.k, c.score) for c in a.top] == [(c.i, c.j, c.k, c.score) for c in b.top])


def trip(m):
    return m * (m - 1) * (m - 2) // 6


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, max(0, math.ceil(q * len(xs)) - 1))]


def load_workloads():
    sys.path.insert(0, PKG)
    import workloads
    return workloads


def reference_results(group):
    """The reference program's recorded results for a workload group
    (best triplet, printed score, spec), or None."""
    path = os.path.join(ROOT, "tests", "golden", "workload_reference.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["groups"].get(group)


def reference_agreement(results, specs, group):
    """How many results print the reference's score and name its triplet
    (a differing triplet counts when its score prints the same: a tie)."""
    ref = reference_results(group)
    if ref is None:
        return None
    n = same = 0
    for r, s, e in zip(results, specs, ref):
        if e["spec"] != s:
            continue
        n += 1
        same += "%f" % r.score == e["score"]
    return {"compared": n, "score_equal": same}


def fp_error(results):
    """FP32 vs FP64 scoring of the top candidates over several scans."""
    return {"max_abs_err_fp32": max(r.max_abs_err_fp32 for r in results),
            "fp64_order_differs": sum(bool(r.fp64_order_differs) for r in results),
            "scans": len(results)}


def batched_leg(epi, ctx):
    """64 uneven cohorts: one graph-captured launch vs 64 plain scans."""
    wl = load_workloads()
    specs = wl.batch_64()
    cohorts, work = [], 0
    for s in specs:
        g, p, _ = wl.materialize(s, epi)
        cohorts.append(ctx.cohort(g, p))
        work += wl.triplets_x_samples(s)
    n = len(specs)
    try:
        plain, t_plain = [], 0.0
        for c in cohorts:
            t0 = time.perf_counter()
            plain.append(ctx.scan(c))
            t_plain += time.perf_counter() - t0
        ctx.scan_batch(cohorts)  # capture + first replay
        reps, t0 = 3, time.perf_counter()
        for _ in range(reps):
            batch = ctx.scan_batch(cohorts)
        t_batch = (time.perf_counter() - t0) / reps
    finally:
        for c in cohorts:
            c.close()
    ok = all(same_result(a, b) for a, b in zip(plain, batch))
    return {"cohorts": n, "triplets_x_samples": work,
            "value": work / t_batch, "ms": t_batch * 1e3,
            "sequential_value": work / t_plain, "sequential_ms": t_plain * 1e3,
            "speedup_vs_sequential": t_plain / t_batch, "identical_to_plain": ok,
            "fp_error": fp_error(batch),
            "reference": reference_agreement(batch, specs, "batch_64")}


def configs_leg(epi, ctx):
    """The two smaller configurations: 256 SNPs x 1024 samples (HWE) and
    2048 SNPs x 4096 samples with one planted 3-way interaction; FP32 vs
    FP64 score error of the top 64 and recovery of the planted triplet."""
    wl = load_workloads()
    out = []
    for spec, group in ((wl.scan_256x1024(), None), (wl.planted_2048x4096(), "planted_2048x4096")):
        M, N = spec["n_snps"], spec["n_samples"]
        g, p, pl = wl.materialize(spec, epi)
        with ctx.cohort(g, p) as co:
            ctx.scan(co)  # warm-up (row table, expanded planes)
            r = ctx.scan(co)
        ref = reference_results(group) if group else None
        out.append({"config": spec["name"], "snps": M, "samples": N, "scan_ms": r.scan_ms,
                    "triplets_x_samples_per_s": trip(M) * N / (r.scan_ms / 1e3),
                    "best": [r.i, r.j, r.k], "score": r.score, "score_fp64": r.score_fp64,
                    "max_abs_err_fp32_top64": r.max_abs_err_fp32,
                    "fp64_order_differs": bool(r.fp64_order_differs),
                    "planted": list(pl) if pl else None,
                    "planted_recovered": (list(pl) == [r.i, r.j, r.k]) if pl else None,
                    "reference": ({"best": ref[0]["best"], "score": ref[0]["score"],
                                   "same": ref[0]["best"] == [r.i, r.j, r.k] and
                                   ref[0]["score"] == "%f" % r.score} if ref else None)})
        log(f"{spec['name']}: best {out[-1]['best']} err {r.max_abs_err_fp32:.2e}")
    return out


def serve_stream(ctx, reqs, arrivals, round_ms, max_active):
    """Submits request r at arrivals[r] seconds (from a feeder thread) and
    polls every result; returns (results, makespan s, server stats)."""
    n = len(reqs)
    results, ids = [None] * n, [None] * n
    with ctx.server(max_active=max_active, max_samples=16384, round_ms=round_ms) as srv:
        t0 = time.perf_counter()

        def feeder():
            for r in range(n):
                delay = arrivals[r] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                ids[r] = srv.submit(reqs[r][0], reqs[r][1])

        th = threading.Thread(target=feeder)
        th.start()
        th.join()
        for r in range(n):
            results[r] = srv.poll(ids[r], 600000)
        t_end = time.perf_counter() - t0
        st = srv.stats()
    return results, t_end, st


def serving_leg(epi, ctx, n=512, load=0.8):
    """The request stream through the server: its capacity on this mix
    (every request submitted at once), then Poisson arrivals at `load` x
    that capacity, against FIFO plain scans on the same arrival times."""
    import numpy as np
    wl = load_workloads()
    specs, rng = wl.stream(n)
    reqs = []
    for s in specs:
        g, p, _ = wl.materialize(s, epi)
        reqs.append((g, p, wl.triplets_x_samples(s)))
    total = sum(w for _, _, w in reqs)
    round_ms, max_active = 3.5, 64
    # capacity: all requests queued at t = 0
    sat, t_sat, st_sat = serve_stream(ctx, reqs, np.zeros(n), round_ms, max_active)
    capacity = total / t_sat
    lam = load * capacity / (total / n)  # requests per second
    gaps = rng.exponential(1.0 / lam, size=n)
    arrivals = np.cumsum(gaps) - gaps[0]
    log(f"serving: capacity {capacity:.3e} triplets*samples/s; "
        f"{n} requests at {lam:.1f}/s")
    results, t_end, st = serve_stream(ctx, reqs, arrivals, round_ms, max_active)
    lat = [res.latency_ms for res in results]
    # plain one-at-a-time baseline: real service times, FIFO queue on the
    # same arrival times.  A request arrives as host arrays, as it does at
    # the server, so its service is the cohort's creation (upload, packing,
    # margins, seed) and its scan; the scan alone is reported beside it.
    service, service_scan, plain = [], [], []
    for g, p, _ in reqs:
        t0 = time.perf_counter()
        with ctx.cohort(g, p) as co:
            t1 = time.perf_counter()
            plain.append(ctx.scan(co))
            t2 = time.perf_counter()
        service.append(t2 - t0)
        service_scan.append(t2 - t1)

    def fifo(svc):
        done, lat_q = 0.0, []
        for r in range(n):
            done = max(done, arrivals[r]) + svc[r]
            lat_q.append((done - arrivals[r]) * 1e3)
        return {"value": total / done, "makespan_ms": done * 1e3,
                "service_ms_total": sum(svc) * 1e3,
                "latency_ms": {"p50": pct(lat_q, 0.5), "p99": pct(lat_q, 0.99)}}
    ok = all(same_result(a, b) for a, b in zip(plain, results)) and \
        all(same_result(a, b) for a, b in zip(plain, sat))
    return {"requests": n, "latency_samples": len(lat), "offered_load": load,
            "offered_load_of": "the server's measured capacity on this mix",
            "arrival_rate_per_s": lam, "round_ms": round_ms, "max_active": max_active,
            "triplets_x_samples": total,
            "capacity": {"value": capacity, "makespan_ms": t_sat * 1e3,
                         "device_busy_frac": st_sat["device_ms"] / (t_sat * 1e3),
                         "rounds": st_sat["rounds"]},
            "value": total / t_end, "makespan_ms": t_end * 1e3,
            "arrivals_span_ms": float(arrivals[-1]) * 1e3,
            "device_busy_frac": st["device_ms"] / (t_end * 1e3),
            "rounds": st["rounds"],
            "latency_ms": {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99)},
            "plain_fifo": fifo(service),
            "plain_fifo_scan_only": fifo(service_scan),
            "identical_to_plain": ok,
            "fp_error": fp_error(results),
            "reference": reference_agreement(results, specs, "stream_64")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--snps", type=int, default=4096)
    ap.add_argument("--samples", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--snps-j", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-snps", type=int, default=1536)
    ap.add_argument("--variant", type=int, default=0, help="0 auto, 1 VALU, 2 int8 MFMA, 3 FP4 MFMA")
    ap.add_argument("--no-compare", action="store_true",
                    help="skip the one-scan comparison of the VALU and MFMA kernels")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 256x1024 and planted 2048x4096 configurations")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--requests", type=int, default=512,
                    help="requests in the serving leg (p99 needs more than 100)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    epi = load_binding()
    N, M, cases = args.samples, args.snps, args.samples // 2
    log(f"generating {N} x {M} HWE data (seed {args.seed})")
    g, p = epi.generate_hwe(args.seed, N, M, cases)
    tile = {"variant": args.variant, "snps_j": args.snps_j, "waves": args.waves}
    with epi.Context(0) as ctx:
        # the formulation the scans below run: the context's tile table
        # decides unless the command line names one
        if args.variant:
            variant = args.variant
        elif args.snps_j or args.waves:
            variant = 1
        else:
            variant = ctx.tune_get(N)["variant"]
        log(f"scan kernel: {KERNELS[variant]}")
        with ctx.cohort(g, p) as co:
            for w in range(args.warmup):
                r = ctx.scan(co, tile)
                log(f"warmup {w}: {r.scan_ms:.1f} ms, best {r.best} {r.score:.6f}")
            lat, kms = [], []
            t0 = time.perf_counter()
            for s in range(args.steps):
                r = ctx.scan(co, tile)
                lat.append(r.scan_ms)
                kms.append(r.kernel_ms)
                log(f"step {s}: {r.scan_ms:.1f} ms (scan kernel {r.kernel_ms:.1f} ms)")
            elapsed = time.perf_counter() - t0
            compare = None
            if not args.no_compare:
                # one scan with each formulation of the 8 popcounted cells
                compare = {}
                for name, var in (("valu", 1), ("mfma", 2), ("fp4", 3)):
                    rv = ctx.scan(co, {"variant": var})
                    kern = rv.kernel_ms / 1e3
                    compare[name] = {"kernel_ms": rv.kernel_ms, "scan_ms": rv.scan_ms,
                                     "triplets_x_samples_per_s": trip(M) * N / kern,
                                     "identical": same_result(rv, r)}
                    log(f"{name}: kernel {rv.kernel_ms:.1f} ms")
                # MFMA work executed: 8 MACs per triplet-sample (2 ops each)
                for name, peak in (("mfma", MFMA_I8_PEAK_OPS), ("fp4", MFMA_FP4_PEAK_OPS)):
                    tops = trip(M) * N * 16 / (compare[name]["kernel_ms"] / 1e3) / 1e12
                    compare[name]["executed_tops"] = tops
                    compare[name]["executed_frac"] = tops * 1e12 / peak
        device = ctx.device_name
        batched = serving = None
        others = None if args.no_configs else configs_leg(epi, ctx)
        if not args.no_batch:
            log("batched leg: 64 uneven cohorts")
            batched = batched_leg(epi, ctx)
            log(f"batched: {batched['ms']:.1f} ms vs {batched['sequential_ms']:.1f} ms sequential")
        if not args.no_serve:
            log("serving leg: Poisson stream")
            serving = serving_leg(epi, ctx, n=args.requests)
            log(f"serving: p50 {serving['latency_ms']['p50']:.1f} ms, "
                f"p99 {serving['latency_ms']['p99']:.1f} ms")
    ntrip = trip(M)
    value = ntrip * N * args.steps / elapsed
    kernel_s = statistics.mean(kms) / 1e3
    work = WORK[variant]
    units = ntrip * N  # triplet-samples per launch
    achieved = units * work["executed_ops"] / kernel_s / 1e12
    ref_equiv = units * work["reference_ops"] / kernel_s / 1e12
    peak = work["peak"] / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "scan_4096x16384_traffic.json")
    if os.path.exists(tpath) and (M, N) == (4096, 16384):
        with open(tpath) as f:
            for e in json.load(f).get("kernels", []):
                if e.get("variant") == variant:
                    traffic = e.get("hbm_bytes_per_launch")
    single = ntrip * N / kernel_s  # the scan kernel's own rate
    popcount_peak = VALU_PEAK_OPS * 32 / 54  # triplets x samples/s at 54 lane-ops per 32 samples
    if batched:
        batched["frac_of_single_scan"] = batched["value"] / value
    if serving:
        # the server's capacity on the mix against the single large scan
        serving["frac_of_single_scan"] = serving["capacity"]["value"] / value
        for key in ("plain_fifo", "plain_fifo_scan_only"):
            serving[key]["capacity"] = serving["triplets_x_samples"] / (
                serving[key]["service_ms_total"] / 1e3)
        # against the scans alone (the plain loop's creations excluded)
        serving["capacity_vs_plain"] = (serving["capacity"]["value"] /
                                        serving["plain_fifo_scan_only"]["capacity"])
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(N, args.cpu_sample_snps, M)
        except Exception as e:  # reported, never fatal for the GPU number
            log(f"cpu baseline failed: {e}")
        ref = reference_results("scan_4096x16384")
        if cpu and ref and (M, N) == (4096, 16384):
            # the reference program's whole scan of this workload, timed when
            # its result was recorded (another host: the one named there)
            e = ref[0]
            cpu["full_size_run"] = {"program": e["program"], "seconds": e["seconds"],
                                    "value": trip(M) * N / e["seconds"],
                                    "host": e.get("host", "the host that recorded the fixture"),
                                    "same_result": e["best"] == [r.i, r.j, r.k] and
                                    e["score"] == "%f" % r.score}
    out = {
        "metric": "triplets x samples/s at 4096 SNPs x 16384 samples; p50/p99 request latency",
        "value": value,
        "unit": "triplets*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[variant],
        "data": "synthetic",
        "config": {"workload": f"exhaustive 3-way K2 scan, {M} SNPs x {N} samples "
                               f"({cases} cases), HWE MAF U[0.05,0.5], seed {args.seed}",
                   "snps": M, "samples": N, "cases": cases, "triplets": ntrip,
                   "device": device},
        "roofline": {"bound": work["bound"], "achieved": achieved, "peak": peak,
                     "unit": "TOP/s", "frac": achieved / peak, "traffic": traffic,
                     "kernel": KERNELS[variant], "kernel_ms": kernel_s * 1e3,
                     "ops_per_unit": work["unit_text"],
                     "traffic_gbs": traffic / kernel_s / 1e9 if traffic else None},
        # Not roofline fractions: the rate at the reference's 27-cell count
        # (54 ops per triplet-sample) against the same peak, and the scan's
        # rate against the reference formulation's own bound on this GPU
        # (27 x (AND + popcount) lane-ops per 32 samples on 256 CU x 64
        # lanes x 2.4 GHz).  Both exceed 1 because 19 of the 27 cells are
        # derived instead of counted.
        "reference_equivalent": {"rate": ref_equiv, "unit": "TOP/s",
                                 "ops_per_unit": work["reference_ops"],
                                 "speedup_over_peak": ref_equiv / peak,
                                 "popcount_bound": popcount_peak,
                                 "popcount_bound_unit": "triplets*samples/s",
                                 "speedup_over_popcount_bound": single / popcount_peak},
        # request latency of the serving stream when it ran, else of the scans above
        "latency_ms": (serving["latency_ms"] if serving else
                       {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99)}),
        "scan_latency_ms": {"p50": pct(lat, 0.5), "p99": pct(lat, 0.99)},
        "best": [r.i, r.j, r.k],
        "score": r.score,
        "max_abs_err_fp32": r.max_abs_err_fp32,
        "cpu_baseline": cpu,
        "variants": compare,
        "batched": batched,
        "serving": serving,
        "other_configs": others,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
