#!/usr/bin/env python3
"""Golden fixtures for the benchmark workloads, from the reference's own scan.

Run in the build container (needs oracle/_ref from `make -C oracle ref`).
For every cohort of the benchmark workloads (cross-dpc-episdet_amd/
workloads.py: the planted 2048 x 4096 scan, the 4096 x 16384 headline scan,
the 64 batch cohorts and the first 64 requests of the serving stream) the
genotypes are generated with the oracle's generator, written to files and
scanned by the reference program CPU/v4_popc_avx512.cpp through
oracle/ref_driver.cpp (the reference's transpose, packing, ln(n!) table,
blocked layout and scan; any case fraction).  The printed best triplet and
`%f` score are stored in workload_reference.json, written after every
cohort so that an interrupted run keeps what it finished.

  python tests/golden/make_workload_golden.py [--only NAME_PREFIX] [--threads 8]

The headline scan takes about half an hour on 8 cores.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cross-dpc-episdet_amd"))
import oracle  # noqa: E402
import workloads  # noqa: E402

LINE = re.compile(r"Best SNPs: (\d+), (\d+), (\d+) - Score: ([0-9.]+)")
OUT = os.path.join(HERE, "workload_reference.json")


def groups():
    stream, _ = workloads.stream(64)
    return [("planted_2048x4096", [workloads.planted_2048x4096()]),
            ("batch_64", workloads.batch_64()),
            ("stream_64", stream),
            ("scan_4096x16384", [workloads.scan_4096x16384()])]


def reference_best(spec, exe, td):
    A, B, planted = workloads.materialize(spec, oracle)
    N, M = A.shape
    gpath, ppath = os.path.join(td, "geno.bin"), os.path.join(td, "pheno.bin")
    A.tofile(gpath)
    B.tofile(ppath)
    block_pac = "128" if N > 8192 else "64"
    t0 = time.time()
    out = subprocess.run([exe, gpath, ppath, str(N), str(M), block_pac, "16"],
                         capture_output=True, text=True, check=True).stdout
    m = LINE.search(out)
    if not m:
        raise RuntimeError(f"unexpected output {out!r}")
    return {"best": [int(m.group(1)), int(m.group(2)), int(m.group(3))], "score": m.group(4),
            "planted": list(planted) if planted else None,
            "program": f"{os.path.basename(exe)} {block_pac} 16", "seconds": time.time() - t0,
            "host": host_name()}


def host_name():
    """CPU model and logical CPUs of the host that ran the program."""
    model = "unknown CPU"
    with open("/proc/cpuinfo") as f:
        for line in f:
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    return f"{model}, {len(os.sched_getaffinity(0))} logical CPUs"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="")
    ap.add_argument("--threads", type=int, default=8)
    args = ap.parse_args()
    exe = oracle.ref_binary(f"v4_popc_drive_t{args.threads}")
    if exe is None:
        sys.exit("oracle/_ref has no v4_popc_drive build: run make -C oracle ref")
    data = {"source": "reference CPU/v4_popc_avx512.cpp scan through oracle/ref_driver.cpp "
                      "on cohorts from the oracle generator", "groups": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    with tempfile.TemporaryDirectory() as td:
        for name, specs in groups():
            if args.only and not name.startswith(args.only):
                continue
            done = data["groups"].setdefault(name, [])
            for q, spec in enumerate(specs):
                if q < len(done) and done[q]["spec"] == spec:
                    continue
                r = reference_best(spec, exe, td)
                entry = dict(r, spec=spec)
                if q < len(done):
                    done[q] = entry
                else:
                    done.append(entry)
                print(name, q, spec["n_samples"], spec["n_snps"], r["best"], r["score"],
                      f"{r['seconds']:.1f}s", flush=True)
                with open(OUT + ".tmp", "w") as f:
                    json.dump(data, f, indent=1)
                os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
