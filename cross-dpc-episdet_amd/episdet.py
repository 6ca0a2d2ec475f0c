"""ctypes binding of the episdet C ABI (include/episdet.h).

The reference (hiperbio/cross-dpc-episdet) is a command-line program,
``./v4 <num_samples> <num_snps> <block_size>`` (GPU/README.md:29-30); the
``bin/v4`` program built next to this file is its drop-in.  This module is
the same C ABI for Python callers (tests, the benchmark, notebooks): it
loads ``lib/libepisdet.so`` and nothing else.  There is no fallback: when the
library or a GPU is missing, calls raise ``EpiError``.
"""
from __future__ import annotations

import ctypes
import functools
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libepisdet.so")
CLI_PATH = os.path.join(HERE, "bin", "v4")

EPI_OK = 0
EPI_EINVAL = -1
EPI_ENOMEM = -2
EPI_EDEVICE = -3
EPI_EBUSY = -4
EPI_ECANCELED = -5
EPI_EAGAIN = -6
EPI_ENOENT = -7
EPI_EINTERNAL = -8
EPI_EIO = -9
EPI_EFORMAT = -10
EPI_TOPK = 64

_u32 = ctypes.c_uint32


class EpiCandidate(ctypes.Structure):
    _fields_ = [
        ("i", _u32), ("j", _u32), ("k", _u32),
        ("score", ctypes.c_float),
        ("score_fp64", ctypes.c_double),
        ("counts", _u32 * 54),
    ]


# numpy view of EpiCandidate (checked against the ctypes layout at import)
_CAND_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4"), ("k", "<u4"), ("score", "<f4"),
                        ("score_fp64", "<f8"), ("counts", "<u4", (54,))])
assert _CAND_DTYPE.itemsize == ctypes.sizeof(EpiCandidate)
assert all(_CAND_DTYPE.fields[f][1] == getattr(EpiCandidate, f).offset
           for f in ("i", "j", "k", "score", "score_fp64", "counts"))


class EpiResult(ctypes.Structure):
    _fields_ = [
        ("i", _u32), ("j", _u32), ("k", _u32),
        ("score", ctypes.c_float),
        ("score_fp64", ctypes.c_double),
        ("triplets", ctypes.c_uint64),
        ("n_samples", _u32), ("n_snps", _u32),
        ("n_controls", _u32), ("n_cases", _u32),
        ("scan_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
        ("queue_ms", ctypes.c_double),
        ("max_abs_err_fp32", ctypes.c_double),
        ("fp64_order_differs", ctypes.c_int32),
        ("n_top", _u32),
        ("top", EpiCandidate * EPI_TOPK),
    ]


class EpiTile(ctypes.Structure):
    _fields_ = [("variant", _u32), ("snps_j", _u32), ("waves", _u32), ("words", _u32)]


class EpiServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("max_active", _u32),
                ("max_samples", _u32), ("round_ms", ctypes.c_double)]


class EpiServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("rounds", ctypes.c_uint64), ("pending", _u32), ("active", _u32),
                ("bytes_in_use", ctypes.c_uint64), ("device_ms", ctypes.c_double)]


# Every entry point of include/episdet.h: (name, restype, argtypes).
_VP = ctypes.c_void_p
_PROTOS = [
    ("epi_version", ctypes.c_char_p, []),
    ("epi_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("epi_open", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_VP)]),
    ("epi_close", ctypes.c_int, [_VP]),
    ("epi_device_name", ctypes.c_int, [_VP, ctypes.c_char_p, ctypes.c_size_t]),
    ("epi_cohort_create", ctypes.c_int, [_VP, _VP, _VP, _u32, _u32, ctypes.POINTER(_VP)]),
    ("epi_cohort_destroy", ctypes.c_int, [_VP]),
    ("epi_cohort_bytes", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    ("epi_cohort_footprint", ctypes.c_int, [_u32, _u32, ctypes.POINTER(ctypes.c_uint64)]),
    ("epi_set_pair_budget", ctypes.c_int, [_VP, ctypes.c_uint64]),
    ("epi_set_snp_order", ctypes.c_int, [_VP, ctypes.c_int]),
    ("epi_scan", ctypes.c_int, [_VP, _VP, ctypes.POINTER(EpiTile), ctypes.POINTER(EpiResult)]),
    ("epi_scan_batch", ctypes.c_int, [_VP, ctypes.POINTER(_VP), _u32, ctypes.POINTER(EpiTile),
                                      ctypes.POINTER(EpiResult)]),
    ("epi_count", ctypes.c_int, [_VP, _VP, _VP, _u32, _VP, _VP, _VP]),
    ("epi_scan_dump", ctypes.c_int, [_VP, _VP, ctypes.POINTER(EpiTile), _VP, _VP]),
    ("epi_autotune", ctypes.c_int, [_VP, _u32, _u32, ctypes.POINTER(EpiTile),
                                    ctypes.POINTER(ctypes.c_double)]),
    ("epi_tune_set", ctypes.c_int, [_VP, _VP, ctypes.POINTER(EpiTile), _u32]),
    ("epi_tune_get", ctypes.c_int, [_VP, _u32, ctypes.POINTER(EpiTile)]),
    ("epi_server_create", ctypes.c_int, [_VP, ctypes.POINTER(EpiServerConfig),
                                         ctypes.POINTER(_VP)]),
    ("epi_server_destroy", ctypes.c_int, [_VP]),
    ("epi_submit", ctypes.c_int, [_VP, _VP, _VP, _u32, _u32, ctypes.POINTER(ctypes.c_uint64)]),
    ("epi_poll", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(EpiResult)]),
    ("epi_scan_noalloc", ctypes.c_int, [_VP, _VP, ctypes.POINTER(EpiTile),
                                        ctypes.POINTER(EpiResult)]),
    ("epi_submit_cohort", ctypes.c_int, [_VP, _VP, ctypes.POINTER(ctypes.c_uint64)]),
    ("epi_cancel", ctypes.c_int, [_VP, ctypes.c_uint64]),
    ("epi_server_get_stats", ctypes.c_int, [_VP, ctypes.POINTER(EpiServerStats)]),
    ("epi_cohort_create_bed", ctypes.c_int, [_VP, _VP, _VP, _u32, _u32, ctypes.POINTER(_VP)]),
    ("epi_read_plink", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_u32), ctypes.POINTER(_u32),
                                      _VP, _VP]),
    ("epi_write_plink", ctypes.c_int, [ctypes.c_char_p, _VP, _VP, _u32, _u32]),
    ("epi_read_raw", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _u32, ctypes.POINTER(_u32),
                                    _VP, _VP]),
    ("epi_generate_reference", ctypes.c_int, [_u32, _u32, _VP, _VP]),
    ("epi_generate_hwe", ctypes.c_int, [ctypes.c_uint64, _u32, _u32, ctypes.c_double,
                                        ctypes.c_double, _u32, _VP, _VP]),
    ("epi_generate_planted", ctypes.c_int, [ctypes.c_uint64, _u32, _u32, ctypes.c_double,
                                            ctypes.c_double, _u32, _VP, _VP, _VP]),
]


class EpiError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        super().__init__(f"{what}: {strerror(code)} ({code})")


_lib = None


def lib() -> ctypes.CDLL:
    """Load lib/libepisdet.so (built by `make` in this directory)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EpiError(EPI_EDEVICE, f"{LIB_PATH} is not built")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _PROTOS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def strerror(code: int) -> str:
    try:
        return lib().epi_strerror(code).decode()
    except EpiError:
        return f"error {code}"


def _check(code: int, what: str) -> None:
    if code != EPI_OK:
        raise EpiError(code, what)


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _tile(tile) -> Optional[EpiTile]:
    if tile is None:
        return None
    if isinstance(tile, EpiTile):
        return tile
    t = EpiTile()
    for k, v in dict(tile).items():
        setattr(t, k, int(v))
    return t


# ---------------------------------------------------------------- file input
def read_plink(prefix: str):
    """Reads <prefix>.bed/.bim/.fam (epi_read_plink): returns (bed, pheno),
    bed [n_snps][(n_samples+3)//4] uint8, pheno [n_samples] (0 control,
    1 case, 255 missing)."""
    n, m = _u32(), _u32()
    pre = os.fsencode(prefix)
    _check(lib().epi_read_plink(pre, ctypes.byref(n), ctypes.byref(m), None, None),
           "epi_read_plink")
    bed = np.empty((m.value, (n.value + 3) // 4), np.uint8)
    pheno = np.empty(n.value, np.uint8)
    _check(lib().epi_read_plink(pre, ctypes.byref(n), ctypes.byref(m), _ptr(bed), _ptr(pheno)),
           "epi_read_plink")
    return bed, pheno


def write_plink(prefix: str, geno: np.ndarray, pheno: np.ndarray) -> None:
    """Writes sample-major genotypes as <prefix>.bed/.bim/.fam (epi_write_plink)."""
    geno = np.ascontiguousarray(geno, dtype=np.uint8)
    pheno = np.ascontiguousarray(pheno, dtype=np.uint8)
    _check(lib().epi_write_plink(os.fsencode(prefix), _ptr(geno), _ptr(pheno), geno.shape[0],
                                 geno.shape[1]), "epi_write_plink")


def read_raw(geno_path: str, pheno_path: str, n_samples: int):
    """Reads raw sample-major genotype and phenotype bytes (epi_read_raw)."""
    m = _u32()
    gp, pp = os.fsencode(geno_path), os.fsencode(pheno_path)
    _check(lib().epi_read_raw(gp, pp, n_samples, ctypes.byref(m), None, None), "epi_read_raw")
    geno = np.empty((n_samples, m.value), np.uint8)
    pheno = np.empty(n_samples, np.uint8)
    _check(lib().epi_read_raw(gp, pp, n_samples, ctypes.byref(m), _ptr(geno), _ptr(pheno)),
           "epi_read_raw")
    return geno, pheno


# ---------------------------------------------------------------- generators
def cohort_footprint(n_samples: int, n_snps: int) -> int:
    """Device bytes of a cohort of this shape (epi_cohort_footprint); raises
    EpiError(EPI_EINVAL) for a shape no cohort may have.  Needs no GPU."""
    b = ctypes.c_uint64()
    _check(lib().epi_cohort_footprint(n_samples, n_snps, ctypes.byref(b)), "epi_cohort_footprint")
    return int(b.value)


def generate_reference(n_samples: int, n_snps: int):
    """The reference's srand(100) generator (GPU/v4.cpp:81-110)."""
    g = np.empty((n_samples, n_snps), np.uint8)
    p = np.empty(n_samples, np.uint8)
    _check(lib().epi_generate_reference(n_samples, n_snps, _ptr(g), _ptr(p)), "generate_reference")
    return g, p


def generate_hwe(seed: int, n_samples: int, n_snps: int, n_cases: int,
                 maf_lo: float = 0.05, maf_hi: float = 0.5):
    g = np.empty((n_samples, n_snps), np.uint8)
    p = np.empty(n_samples, np.uint8)
    _check(lib().epi_generate_hwe(seed, n_samples, n_snps, maf_lo, maf_hi, n_cases, _ptr(g),
                                  _ptr(p)), "generate_hwe")
    return g, p


def generate_planted(seed: int, n_samples: int, n_snps: int, n_cases: int,
                     maf_lo: float = 0.05, maf_hi: float = 0.5):
    g = np.empty((n_samples, n_snps), np.uint8)
    p = np.empty(n_samples, np.uint8)
    pl = np.zeros(3, np.uint32)
    _check(lib().epi_generate_planted(seed, n_samples, n_snps, maf_lo, maf_hi, n_cases, _ptr(g),
                                      _ptr(p), _ptr(pl)), "generate_planted")
    return g, p, tuple(int(x) for x in pl)


# ------------------------------------------------------------------ results
@dataclass
class Candidate:
    i: int
    j: int
    k: int
    score: float
    score_fp64: float
    counts: np.ndarray


@dataclass
class ScanResult:
    i: int
    j: int
    k: int
    score: float
    score_fp64: float
    triplets: int
    n_samples: int
    n_snps: int
    n_controls: int
    n_cases: int
    scan_ms: float
    kernel_ms: float
    latency_ms: float
    queue_ms: float
    max_abs_err_fp32: float
    fp64_order_differs: bool
    # the top candidates as a structured array (_CAND_DTYPE), turned into
    # Candidate objects on first access of `top`
    cand: np.ndarray = field(default=None, repr=False, compare=False)

    @property
    def best(self):
        return (self.i, self.j, self.k)

    @functools.cached_property
    def top(self) -> List[Candidate]:
        a = self.cand
        if a is None:
            return []
        counts = a["counts"]
        return [Candidate(i, j, k, s, s64, counts[e])
                for e, (i, j, k, s, s64) in enumerate(zip(a["i"].tolist(), a["j"].tolist(),
                                                          a["k"].tolist(), a["score"].tolist(),
                                                          a["score_fp64"].tolist()))]

    @classmethod
    def from_c(cls, r: EpiResult) -> "ScanResult":
        # one structured copy of the candidates (per-candidate ctypes access
        # costs microseconds each)
        top = np.frombuffer(r.top, dtype=_CAND_DTYPE, count=int(r.n_top)).copy()
        return cls(r.i, r.j, r.k, float(r.score), float(r.score_fp64), int(r.triplets),
                   r.n_samples, r.n_snps, r.n_controls, r.n_cases, float(r.scan_ms),
                   float(r.kernel_ms), float(r.latency_ms), float(r.queue_ms),
                   float(r.max_abs_err_fp32), bool(r.fp64_order_differs), top)


# ------------------------------------------------------------------ objects
class Context:
    """An opened device (epi_open / epi_close)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().epi_open(device, ctypes.byref(h)), f"epi_open({device})")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            lib().epi_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def device_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(lib().epi_device_name(self.handle, buf, 256), "epi_device_name")
        return buf.value.decode()

    def cohort(self, geno: np.ndarray, pheno: np.ndarray) -> "Cohort":
        return Cohort(self, geno, pheno)

    def cohort_bed(self, bed: np.ndarray, pheno: np.ndarray, n_samples: int) -> "Cohort":
        """Cohort from a PLINK 1 .bed body ([n_snps][(n_samples+3)//4] bytes)."""
        return Cohort(self, None, pheno, bed=bed, n_samples=n_samples)

    def scan(self, cohort: "Cohort", tile=None) -> ScanResult:
        r = EpiResult()
        t = _tile(tile)
        _check(lib().epi_scan(self.handle, cohort.handle, ctypes.byref(t) if t else None,
                              ctypes.byref(r)), "epi_scan")
        return ScanResult.from_c(r)

    def set_pair_budget(self, nbytes: int) -> None:
        """Pair-margin bytes cohorts created afterwards keep resident (0:
        the default); past it a cohort is scanned in strip rounds."""
        _check(lib().epi_set_pair_budget(self.handle, int(nbytes)), "epi_set_pair_budget")

    def set_snp_order(self, mode: int) -> None:
        """Device SNP order of cohorts created afterwards: 1 allele
        frequency, 0 original, 2 (the default) allele frequency from 512
        SNPs and 2048 samples on.  Results are those of the original order either way."""
        _check(lib().epi_set_snp_order(self.handle, int(mode)), "epi_set_snp_order")

    def scan_noalloc(self, cohort: "Cohort", tile=None) -> ScanResult:
        """epi_scan_noalloc: the same scan with no device allocation (a
        static workspace per GPU)."""
        r = EpiResult()
        t = _tile(tile)
        _check(lib().epi_scan_noalloc(self.handle, cohort.handle, ctypes.byref(t) if t else None,
                                      ctypes.byref(r)), "epi_scan_noalloc")
        return ScanResult.from_c(r)

    def scan_batch(self, cohorts: Sequence["Cohort"], tile=None) -> List[ScanResult]:
        n = len(cohorts)
        hs = (ctypes.c_void_p * n)(*[c.handle.value for c in cohorts])
        rs = (EpiResult * n)()
        t = _tile(tile)
        _check(lib().epi_scan_batch(self.handle, hs, n, ctypes.byref(t) if t else None, rs),
               "epi_scan_batch")
        return [ScanResult.from_c(rs[b]) for b in range(n)]

    def count(self, cohort: "Cohort", triplets):
        tr = np.ascontiguousarray(np.asarray(triplets, dtype=np.uint32).reshape(-1, 3))
        n = tr.shape[0]
        counts = np.zeros((n, 54), np.uint32)
        s32 = np.zeros(n, np.float32)
        s64 = np.zeros(n, np.float64)
        _check(lib().epi_count(self.handle, cohort.handle, _ptr(tr), n, _ptr(counts), _ptr(s32),
                               _ptr(s64)), "epi_count")
        return counts, s32, s64

    def autotune(self, n_samples: int, n_snps: int):
        t = EpiTile()
        ms = ctypes.c_double()
        _check(lib().epi_autotune(self.handle, n_samples, n_snps, ctypes.byref(t), ctypes.byref(ms)),
               "epi_autotune")
        return {"variant": t.variant, "snps_j": t.snps_j, "waves": t.waves}, float(ms.value)

    def tune_set(self, table):
        """table: [(max_samples, {"variant": V, "snps_j": J, "waves": W}), ...] in
        increasing order of max_samples (variant 1 = VALU, 2 = int8 MFMA)."""
        n = len(table)
        ms = np.array([int(m) for m, _ in table], np.uint32)
        tiles = (EpiTile * max(n, 1))()
        for e, (_, t) in enumerate(table):
            tiles[e] = _tile(t)
        _check(lib().epi_tune_set(self.handle, _ptr(ms), tiles, n), "epi_tune_set")

    def tune_get(self, n_samples: int):
        t = EpiTile()
        _check(lib().epi_tune_get(self.handle, n_samples, ctypes.byref(t)), "epi_tune_get")
        return {"variant": t.variant, "snps_j": t.snps_j, "waves": t.waves}

    def server(self, **config) -> "Server":
        return Server(self, **config)

    def scan_dump(self, cohort: "Cohort", tile=None, with_counts: bool = False):
        m = cohort.n_snps
        nt = m * (m - 1) * (m - 2) // 6
        scores = np.zeros(nt, np.float32)
        counts = np.zeros((nt, 54), np.uint32) if with_counts else None
        t = _tile(tile)
        _check(lib().epi_scan_dump(self.handle, cohort.handle, ctypes.byref(t) if t else None,
                                   _ptr(scores), _ptr(counts) if counts is not None else None),
               "epi_scan_dump")
        return scores, counts


class Cohort:
    """A bit-packed cohort resident in device memory (epi_cohort_create, or
    epi_cohort_create_bed with ``bed=``)."""

    def __init__(self, ctx: Context, geno: Optional[np.ndarray], pheno: np.ndarray,
                 bed: Optional[np.ndarray] = None, n_samples: int = 0):
        pheno = np.ascontiguousarray(pheno, dtype=np.uint8)
        h = ctypes.c_void_p()
        if bed is not None:
            bed = np.ascontiguousarray(bed, dtype=np.uint8)
            if (bed.ndim != 2 or pheno.shape != (n_samples,) or
                    bed.shape[1] != (n_samples + 3) // 4):
                raise EpiError(EPI_EINVAL, "bed must be [n_snps][(n_samples+3)//4], "
                                           "pheno [n_samples]")
            self.n_snps = bed.shape[0]
            self.n_samples = int(np.count_nonzero(pheno <= 1))  # others are left out
            _check(lib().epi_cohort_create_bed(ctx.handle, _ptr(bed), _ptr(pheno), n_samples,
                                               self.n_snps, ctypes.byref(h)),
                   "epi_cohort_create_bed")
        else:
            geno = np.ascontiguousarray(geno, dtype=np.uint8)
            if geno.ndim != 2 or pheno.shape != (geno.shape[0],):
                raise EpiError(EPI_EINVAL, "geno must be [n_samples][n_snps], pheno [n_samples]")
            self.n_samples, self.n_snps = geno.shape
            _check(lib().epi_cohort_create(ctx.handle, _ptr(geno), _ptr(pheno), self.n_samples,
                                           self.n_snps, ctypes.byref(h)), "epi_cohort_create")
        self.handle = h

    @property
    def device_bytes(self) -> int:
        b = ctypes.c_uint64()
        _check(lib().epi_cohort_bytes(self.handle, ctypes.byref(b)), "epi_cohort_bytes")
        return int(b.value)

    def close(self) -> None:
        """Frees the cohort; raises EpiError(EPI_EBUSY) and keeps it while a
        server request on it has not completed."""
        if self.handle:
            _check(lib().epi_cohort_destroy(self.handle), "epi_cohort_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Continuous-batching request loop (epi_server_*)."""

    def __init__(self, ctx: Context, memory_budget: int = 0, max_active: int = 0,
                 max_samples: int = 0, round_ms: float = 0.0):
        cfg = EpiServerConfig(memory_budget, max_active, max_samples, round_ms)
        h = ctypes.c_void_p()
        _check(lib().epi_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(h)),
               "epi_server_create")
        self.handle = h

    def submit(self, geno: np.ndarray, pheno: np.ndarray) -> int:
        geno = np.ascontiguousarray(geno, dtype=np.uint8)
        pheno = np.ascontiguousarray(pheno, dtype=np.uint8)
        rid = ctypes.c_uint64()
        _check(lib().epi_submit(self.handle, _ptr(geno), _ptr(pheno), geno.shape[0], geno.shape[1],
                                ctypes.byref(rid)), "epi_submit")
        return int(rid.value)

    def submit_cohort(self, cohort: "Cohort") -> int:
        """Queues a scan of a cohort already on the device (no upload or
        packing); the cohort cannot be closed until the request completes."""
        rid = ctypes.c_uint64()
        _check(lib().epi_submit_cohort(self.handle, cohort.handle, ctypes.byref(rid)),
               "epi_submit_cohort")
        return int(rid.value)

    def poll(self, rid: int, timeout_ms: int = -1) -> Optional[ScanResult]:
        """Result of request `rid`; None if not ready within timeout_ms."""
        r = EpiResult()
        code = lib().epi_poll(self.handle, rid, timeout_ms, ctypes.byref(r))
        if code == EPI_EAGAIN:
            return None
        _check(code, f"epi_poll({rid})")
        return ScanResult.from_c(r)

    def cancel(self, rid: int) -> None:
        _check(lib().epi_cancel(self.handle, rid), f"epi_cancel({rid})")

    def stats(self) -> dict:
        st = EpiServerStats()
        _check(lib().epi_server_get_stats(self.handle, ctypes.byref(st)), "epi_server_get_stats")
        return {k: getattr(st, k) for k, _ in EpiServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            lib().epi_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
