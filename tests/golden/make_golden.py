#!/usr/bin/env python3
"""Regenerate the golden fixtures from the reference's own CPU programs.

Run in the build container (needs oracle/_ref, built by `make -C oracle ref`
from the reference sources).  Writes:
  kat_reference.json    best triplet + printed score of the reference
                        programs on their own srand(100) generator
                        (CPU/v2.cpp:559-599, CPU/v4_popc_avx512.cpp:4583-4630);
  ref_feed_vectors.json the same programs on seeded Hardy-Weinberg datasets,
                        fed through oracle/ref_feed.c (rand() replays the
                        dataset), including cell counts >= 748 so the
                        Stirling branch of addlog (CPU/v2.cpp:167-176) is
                        exercised;
  log_table.json        SHA-256 of the FP32 ln(n!) table bits for several
                        lengths, to detect libm drift between machines.
The datasets are regenerated in the tests from their seeds, so only seeds,
shapes and expected outputs are stored.
"""
import hashlib
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

RAND_MAX = 2147483647
LINE = re.compile(r"Best SNPs: (\d+), (\d+), (\d+) - Score: ([0-9.]+)")

# (N samples, M SNPs, programs) -- the known-answer list of the reference
KAT = [
    (1024, 256, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 32"]),
    (1000, 100, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 32"]),
    (333, 70, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 32"]),
    (16384, 64, ["v2_t1", "v4_popc_avx512_t8 64 16", "v4_popc_avx512_t8 256 16"]),
    (4096, 128, ["v2_t1", "v4_popc_avx512_t8 64 16"]),
    (777, 45, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 8"]),
    (100, 33, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 8"]),
    (2049, 97, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 8"]),
    (31, 40, ["v1_t1", "v2_t1", "v4_popc_avx512_t8 16 8"]),
    (16384, 512, ["v2_t8", "v4_popc_avx512_t8 128 16"]),
    (16384, 1024, ["v4_popc_avx512_t8 128 16"]),
    (4096, 2048, ["v4_popc_avx512_t8 64 16"]),
]

# seeded datasets fed to the reference (exactly N/2 cases, as it requires)
FEED = [
    {"seed": 11, "N": 1024, "M": 64, "maf": [0.05, 0.5]},
    {"seed": 12, "N": 1001, "M": 50, "maf": [0.05, 0.5]},
    {"seed": 13, "N": 4096, "M": 40, "maf": [0.01, 0.1]},
    {"seed": 14, "N": 16384, "M": 48, "maf": [0.05, 0.5]},
    {"seed": 15, "N": 3000, "M": 60, "maf": [0.02, 0.2]},
    {"seed": 16, "N": 65, "M": 30, "maf": [0.05, 0.5]},
]


def run(cmd, env=None):
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, check=True).stdout
    m = LINE.search(out)
    if not m:
        raise RuntimeError(f"{cmd}: unexpected output {out!r}")
    return [int(m.group(1)), int(m.group(2)), int(m.group(3))], m.group(4)


def feed_stream(A, B):
    N = A.shape[0]
    vals = [A.reshape(-1).astype(np.int32)]
    picks = []
    for idx in np.nonzero(B == 1)[0]:
        r = (int(idx) * RAND_MAX + N - 1) // N
        while int(N * (r / RAND_MAX)) > idx:
            r -= 1
        while int(N * (r / RAND_MAX)) < idx:
            r += 1
        assert int(N * (r / RAND_MAX)) == idx
        picks.append(r)
    vals.append(np.asarray(picks, np.int32))
    return np.concatenate(vals).astype("<i4")


def main():
    kat = []
    for N, M, progs in KAT:
        runs = []
        for p in progs:
            parts = p.split()
            exe = oracle.ref_binary(parts[0])
            best, score = run([exe, str(N), str(M)] + parts[1:])
            runs.append({"program": p, "best": best, "score": score})
            print(N, M, p, best, score, flush=True)
        kat.append({"N": N, "M": M, "runs": runs})
    with open(os.path.join(HERE, "kat_reference.json"), "w") as f:
        json.dump({"source": "reference CPU programs on their srand(100) generator",
                   "cases": kat}, f, indent=1)

    vec = []
    with tempfile.TemporaryDirectory() as td:
        for d in FEED:
            N, M = d["N"], d["M"]
            A, B, _ = oracle.hwe_generate(d["seed"], N, M, N // 2, *d["maf"])
            path = os.path.join(td, "feed.bin")
            feed_stream(A, B).tofile(path)
            env = dict(os.environ, EPISDET_FEED=path)
            runs = []
            for p in ["v2_feed", "v1_feed"]:
                best, score = run([oracle.ref_binary(p), str(N), str(M)], env=env)
                runs.append({"program": p, "best": best, "score": score})
            # largest cell total of the winning triplet: shows whether the
            # Stirling branch (n >= 748) is used
            pc = oracle.PackedCohort(A, B)
            c = pc.counts(*runs[0]["best"])
            big = int((c[:27] + c[27:] + 1).max())
            print(d, runs, "max n0+n1+1 =", big, flush=True)
            vec.append(dict(d, cases=N // 2, runs=runs, max_cell_plus1=big))
    with open(os.path.join(HERE, "ref_feed_vectors.json"), "w") as f:
        json.dump({"source": "reference CPU v1/v2 fed seeded HWE data through oracle/ref_feed.c",
                   "vectors": vec}, f, indent=1)

    tab = {}
    for n in [1024, 4098, 16386, 65538]:
        T = oracle.log_table(n - 1)
        tab[str(n)] = hashlib.sha256(T.tobytes()).hexdigest()
    T = oracle.log_table(20000)
    with open(os.path.join(HERE, "log_table.json"), "w") as f:
        json.dump({"sha256_float32_le": tab,
                   "samples": {str(n): float(T[n]).hex() for n in
                               [0, 1, 2, 3, 10, 100, 746, 747, 748, 749, 1000, 4097, 16385, 20000]}},
                  f, indent=1)


if __name__ == "__main__":
    main()
