"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The product never imports it.
See oracle.c for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_i64 = ctypes.c_int64
_VP = ctypes.c_void_p


class Packed(ctypes.Structure):
    _fields_ = [("N", _i64), ("M", _i64), ("n", _i64 * 2), ("W", _i64 * 2),
                ("mask", ctypes.c_uint32 * 2), ("planes", _VP * 2)]


class Cand(ctypes.Structure):
    _fields_ = [("score", ctypes.c_float), ("i", ctypes.c_uint32), ("j", ctypes.c_uint32),
                ("k", ctypes.c_uint32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(LIB_PATH)
        L.orc_ref_generate.argtypes = [_i64, _i64, _VP, _VP]
        L.orc_hwe_generate.argtypes = [ctypes.c_uint64, _i64, _i64, ctypes.c_double,
                                       ctypes.c_double, _i64, _VP, _VP, _VP]
        L.orc_planted_generate.argtypes = [ctypes.c_uint64, _i64, _i64, ctypes.c_double,
                                           ctypes.c_double, _i64, _VP, _VP, _VP]
        L.orc_log_table.argtypes = [_i64, _VP]
        L.orc_log_table_f64.argtypes = [_i64, _VP]
        L.orc_count_brute.argtypes = [_VP, _VP, _i64, _i64, _i64, _i64, _i64, _VP]
        L.orc_pack.argtypes = [_VP, _VP, _i64, _i64, ctypes.POINTER(Packed)]
        L.orc_pack.restype = ctypes.c_int
        L.orc_packed_free.argtypes = [ctypes.POINTER(Packed)]
        L.orc_count_packed.argtypes = [ctypes.POINTER(Packed), _i64, _i64, _i64, _VP]
        L.orc_k2.argtypes = [_VP, _VP]
        L.orc_k2.restype = ctypes.c_float
        L.orc_k2_f64.argtypes = [_VP, _VP]
        L.orc_k2_f64.restype = ctypes.c_double
        L.orc_scan.argtypes = [ctypes.POINTER(Packed), _VP, ctypes.c_int, ctypes.c_int, _VP]
        L.orc_scan.restype = _i64
        L.orc_scan_all_scores.argtypes = [ctypes.POINTER(Packed), _VP, ctypes.c_int, _VP]
        L.orc_scan_all_scores.restype = _i64
        L.orc_rank3.argtypes = [_i64, _i64, _i64, _i64]
        L.orc_rank3.restype = _i64
        L.orc_rdtsc.restype = ctypes.c_uint64
        L.orc_bed_decode.argtypes = [_VP, _i64, _i64, _VP]
        L.orc_bed_encode.argtypes = [_VP, _i64, _i64, _VP]
        _lib = L
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def ref_generate(N, M):
    A = np.empty((N, M), np.uint8)
    B = np.empty(N, np.uint8)
    lib().orc_ref_generate(N, M, _p(A), _p(B))
    return A, B


def hwe_generate(seed, N, M, n_cases, maf_lo=0.05, maf_hi=0.5):
    A = np.empty((N, M), np.uint8)
    B = np.empty(N, np.uint8)
    maf = np.empty(M, np.float64)
    lib().orc_hwe_generate(seed, N, M, maf_lo, maf_hi, n_cases, _p(A), _p(B), _p(maf))
    return A, B, maf


def planted_generate(seed, N, M, n_cases, maf_lo=0.05, maf_hi=0.5):
    A = np.empty((N, M), np.uint8)
    B = np.empty(N, np.uint8)
    pl = np.zeros(3, np.int64)
    lib().orc_planted_generate(seed, N, M, maf_lo, maf_hi, n_cases, _p(A), _p(B), _p(pl))
    return A, B, tuple(int(x) for x in pl)


def log_table(n_max):
    T = np.empty(n_max + 1, np.float32)
    lib().orc_log_table(n_max, _p(T))
    return T


def log_table_f64(n_max):
    T = np.empty(n_max + 1, np.float64)
    lib().orc_log_table_f64(n_max, _p(T))
    return T


def count_brute(A, B, i, j, k):
    A = np.ascontiguousarray(A, np.uint8)
    B = np.ascontiguousarray(B, np.uint8)
    out = np.zeros(54, np.uint32)
    lib().orc_count_brute(_p(A), _p(B), A.shape[0], A.shape[1], i, j, k, _p(out))
    return out


class PackedCohort:
    """Class-split 32-bit planes (CPU/v2.cpp:103-164) plus the ln(n!) table."""

    def __init__(self, A, B):
        self.A = np.ascontiguousarray(A, np.uint8)
        self.B = np.ascontiguousarray(B, np.uint8)
        self.N, self.M = self.A.shape
        self.p = Packed()
        rc = lib().orc_pack(_p(self.A), _p(self.B), self.N, self.M, ctypes.byref(self.p))
        if rc != 0:
            raise MemoryError("orc_pack")
        self.T = log_table(self.N + 1)
        self.T64 = log_table_f64(self.N + 1)

    def __del__(self):
        try:
            lib().orc_packed_free(ctypes.byref(self.p))
        except Exception:
            pass

    def counts(self, i, j, k):
        out = np.zeros(54, np.uint32)
        lib().orc_count_packed(ctypes.byref(self.p), i, j, k, _p(out))
        return out

    def score(self, counts):
        c = np.ascontiguousarray(counts, np.uint32)
        return float(np.float32(lib().orc_k2(_p(c), _p(self.T))))

    def score_f64(self, counts):
        c = np.ascontiguousarray(counts, np.uint32)
        return float(lib().orc_k2_f64(_p(c), _p(self.T64)))

    def scan(self, topk=64, nthreads=8):
        top = (Cand * topk)()
        n = lib().orc_scan(ctypes.byref(self.p), _p(self.T), nthreads, topk, top)
        return n, [(float(np.float32(c.score)), c.i, c.j, c.k) for c in top]

    def all_scores(self, nthreads=8):
        M = self.M
        out = np.empty(M * (M - 1) * (M - 2) // 6, np.float32)
        lib().orc_scan_all_scores(ctypes.byref(self.p), _p(self.T), nthreads, _p(out))
        return out


def bed_decode(bed, N):
    """PLINK 1 .bed body [M][(N+3)//4] -> sample-major genotypes [N][M]
    (00 -> 2, 01 -> 3 missing, 10 -> 1, 11 -> 0)."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    M = bed.shape[0]
    assert bed.shape == (M, (N + 3) // 4)
    A = np.empty((N, M), np.uint8)
    lib().orc_bed_decode(_p(bed), N, M, _p(A))
    return A


def bed_encode(A):
    """Sample-major genotypes [N][M] -> PLINK 1 .bed body [M][(N+3)//4]."""
    A = np.ascontiguousarray(A, dtype=np.uint8)
    N, M = A.shape
    bed = np.empty((M, (N + 3) // 4), np.uint8)
    lib().orc_bed_encode(_p(A), N, M, _p(bed))
    return bed


def rank3(M, i, j, k):
    return int(lib().orc_rank3(M, i, j, k))


def rdtsc():
    return int(lib().orc_rdtsc())


def ref_binary(name):
    """Path of a reference program compiled into oracle/_ref (None if absent)."""
    p = os.path.join(REF_DIR, name)
    return p if os.path.exists(p) else None
