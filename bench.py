#!/usr/bin/env python3
"""Benchmark: exhaustive third-order K2 scan, triplets x samples per second.

Workload (one "step" = one full scan of one cohort, inputs resident in HBM):
4096 SNPs x 16384 samples (8192 cases / 8192 controls), Hardy-Weinberg
genotypes with per-SNP MAF ~ U[0.05, 0.5], seed 3; C(4096,3) = 1.14e10
triplets.  Prints one JSON line with the metric, the roofline of the scan
kernel and the reference CPU baseline timed on this host.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import importlib.util
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
# reference work per (triplet x 32 samples): 27 cells x (3-input AND + popcount)
OPS_PER_TRIPLET_WORD = 54


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


def cpu_baseline(n_samples, n_snps_sample):
    """Reference CPU v4 (AVX-512 VPOPCNTDQ build, or the AVX2 build when the
    host lacks it), compiled from the reference sources into oracle/_ref with
    FREQ=1 so it prints TSC cycles; converted with a measured TSC rate."""
    flags = open("/proc/cpuinfo").read()
    threads = 16 if len(os.sched_getaffinity(0)) >= 16 else 8
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
    trip = n_snps_sample * (n_snps_sample - 1) * (n_snps_sample - 2) // 6
    return {"value": trip * n_samples / secs, "unit": "triplets*samples/s", "cores": threads,
            "kind": "reference",
            "sample": f"{exe} block_pac={extra[0]} block_snp={extra[1]} on N={n_samples}, "
                      f"M={n_snps_sample} (reference generator), scan {secs:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--snps", type=int, default=4096)
    ap.add_argument("--samples", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--snps-j", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-snps", type=int, default=1024)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    epi = load_binding()
    N, M, cases = args.samples, args.snps, args.samples // 2
    log(f"generating {N} x {M} HWE data (seed {args.seed})")
    g, p = epi.generate_hwe(args.seed, N, M, cases)
    tile = {"snps_j": args.snps_j, "waves": args.waves}
    with epi.Context(0) as ctx:
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
            device = ctx.device_name
    trip = M * (M - 1) * (M - 2) // 6
    value = trip * N * args.steps / elapsed
    kernel_s = statistics.mean(kms) / 1e3
    ops = trip * (N / 32.0) * OPS_PER_TRIPLET_WORD
    achieved = ops / kernel_s / 1e12
    peak = VALU_PEAK_OPS / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "scan_4096x16384_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(N, args.cpu_sample_snps)
        except Exception as e:  # reported, never fatal for the GPU number
            log(f"cpu baseline failed: {e}")
    lat_sorted = sorted(lat)
    out = {
        "metric": "triplets x samples/s at 4096 SNPs x 16384 samples",
        "value": value,
        "unit": "triplets*samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"exhaustive 3-way K2 scan, {M} SNPs x {N} samples "
                               f"({cases} cases), HWE MAF U[0.05,0.5], seed {args.seed}",
                   "snps": M, "samples": N, "cases": cases, "triplets": trip,
                   "device": device},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TOP/s",
                     "frac": achieved / peak, "traffic": traffic,
                     "ops_per_unit": "54 int32 lane-ops per (triplet x 32 samples)",
                     "kernel_ms": kernel_s * 1e3},
        "latency_ms": {"p50": lat_sorted[len(lat) // 2],
                       "p99": lat_sorted[min(len(lat) - 1, math.ceil(0.99 * len(lat)) - 1)]},
        "best": [r.i, r.j, r.k],
        "score": r.score,
        "max_abs_err_fp32": r.max_abs_err_fp32,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
