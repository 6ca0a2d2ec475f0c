"""The five benchmark workloads as seeded cohort specifications.

One definition shared by the benchmark, the GPU tests and the script that
writes the golden fixtures, so that all three scan the same cohorts.  A
spec is a dict: seed, n_samples, n_snps, n_cases and the generator
("hwe": Hardy-Weinberg genotypes, per-SNP MAF ~ U[0.05, 0.5], exactly
n_cases cases; "planted": the same with one planted three-way
interaction).  `materialize` turns a spec into genotypes with a given
generator module, either this package's `episdet` binding or the CPU
oracle (both produce the same bytes).

  scan_256x1024     256 SNPs x 1024 samples, 512 cases
  planted_2048x4096 2048 SNPs x 4096 samples, 2048 cases, one planted triplet
  scan_4096x16384   4096 SNPs x 16384 samples, 8192 cases (the headline scan)
  batch_64          64 cohorts: SNPs ~ U{128..1024}, samples ~ U{512..8192},
                    case fraction ~ U[0.3, 0.7]
  stream(n)         Poisson request stream: SNPs log-U[64, 2048], samples
                    log-U[256, 16384], case fraction U[0.3, 0.7]; the first
                    64 requests do not depend on n
"""
from __future__ import annotations

import math

import numpy as np


def scan_256x1024():
    return {"name": "256x1024 HWE", "gen": "hwe", "seed": 1, "n_samples": 1024,
            "n_snps": 256, "n_cases": 512}


def planted_2048x4096():
    return {"name": "2048x4096 planted", "gen": "planted", "seed": 11, "n_samples": 4096,
            "n_snps": 2048, "n_cases": 2048}


def scan_4096x16384(seed: int = 3):
    return {"name": "4096x16384 HWE", "gen": "hwe", "seed": seed, "n_samples": 16384,
            "n_snps": 4096, "n_cases": 8192}


def batch_64(seed: int = 4, n: int = 64):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n):
        M = int(rng.integers(128, 1025))
        N = int(rng.integers(512, 8193))
        cases = int(round(N * rng.uniform(0.3, 0.7)))
        out.append({"name": f"batch cohort {b}", "gen": "hwe", "seed": seed * 1000 + b,
                    "n_samples": N, "n_snps": M, "n_cases": cases})
    return out


def stream(n: int = 64, seed: int = 5):
    """Request specs of the serving stream and the random generator that
    draws their arrival gaps afterwards."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        M = int(round(math.exp(rng.uniform(math.log(64), math.log(2048)))))
        N = int(round(math.exp(rng.uniform(math.log(256), math.log(16384)))))
        cases = int(round(N * rng.uniform(0.3, 0.7)))
        out.append({"name": f"request {r}", "gen": "hwe", "seed": seed * 1000 + r,
                    "n_samples": N, "n_snps": M, "n_cases": cases})
    return out, rng


def triplets_x_samples(spec) -> int:
    m = spec["n_snps"]
    return m * (m - 1) * (m - 2) // 6 * spec["n_samples"]


def materialize(spec, gen):
    """(genotypes [N][M] uint8, phenotypes [N] uint8, planted or None) from
    `gen`, a module with generate_hwe / generate_planted (the episdet
    binding) or hwe_generate / planted_generate (the oracle)."""
    N, M, cases, seed = spec["n_samples"], spec["n_snps"], spec["n_cases"], spec["seed"]
    if spec["gen"] == "planted":
        if hasattr(gen, "generate_planted"):
            g, p, pl = gen.generate_planted(seed, N, M, cases)
        else:
            g, p, pl = gen.planted_generate(seed, N, M, cases)
        return g, p, tuple(pl)
    if hasattr(gen, "generate_hwe"):
        g, p = gen.generate_hwe(seed, N, M, cases)
    else:
        g, p, _ = gen.hwe_generate(seed, N, M, cases)
    return g, p, None
