"""The C ABI library loads on a CPU-only host and exports every entry point
that include/episdet.h declares; host-side pieces that need no GPU (data
generators, argument checking, struct layout) behave as specified."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "episdet.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(epi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(epi):
    names = declared_functions()
    assert len(names) >= 15
    L = epi.lib()
    for n in names:
        assert hasattr(L, n), n
    # the Python binding covers the whole ABI too
    assert {p[0] for p in epi._PROTOS} == set(names)


def test_struct_layout_matches_c(epi):
    so = os.path.join(ROOT, "tests", "native", "libcells_host.so")
    if not os.path.exists(so):
        pytest.skip("tests/native/libcells_host.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(so)
    out = (ctypes.c_size_t * 6)()
    L.abi_sizes(out)
    assert out[0] == ctypes.sizeof(epi.EpiCandidate)
    assert out[1] == ctypes.sizeof(epi.EpiResult)
    assert out[2] == ctypes.sizeof(epi.EpiTile)
    assert out[3] == epi.EpiResult.top.offset
    assert out[4] == epi.EpiResult.n_top.offset
    assert out[5] == epi.EpiCandidate.counts.offset


def test_reference_generator_matches_oracle(epi, orc):
    for N, M in [(31, 40), (1024, 256), (2049, 97)]:
        g, p = epi.generate_reference(N, M)
        A, B = orc.ref_generate(N, M)
        assert np.array_equal(g, A) and np.array_equal(p, B)
        assert int(p.sum()) == N // 2


def test_hwe_and_planted_generators_match_oracle(epi, orc):
    g, p = epi.generate_hwe(5, 3000, 77, 1234)
    A, B, _ = orc.hwe_generate(5, 3000, 77, 1234)
    assert np.array_equal(g, A) and np.array_equal(p, B)
    g, p, pl = epi.generate_planted(8, 1500, 50, 700)
    A, B, pl2 = orc.planted_generate(8, 1500, 50, 700)
    assert np.array_equal(g, A) and np.array_equal(p, B) and pl == pl2


def test_generator_argument_checks(epi):
    with pytest.raises(epi.EpiError):
        epi.generate_hwe(1, 10, 5, 11)  # more cases than samples
    with pytest.raises(epi.EpiError):
        epi.generate_hwe(1, 10, 5, 5, 0.6, 0.5)  # empty MAF range
    with pytest.raises(epi.EpiError):
        epi.generate_planted(1, 10, 2, 5)  # fewer than 3 SNPs


@pytest.mark.parametrize("args", [[], ["10"], ["10", "20"], ["x", "20", "32"], ["10", "2", "32"],
                                  ["10", "20", "0"], ["0", "20", "32"], ["10", "20", "32", "--bogus"],
                                  ["10", "20", "32", "--cases", "11"],
                                  ["10", "20", "32", "--top", "65"],
                                  ["10", "20", "32", "--variant", "4"],
                                  ["--bfile"], ["--bfile", "x", "--hwe", "1"],
                                  ["--bfile", "x", "--cases", "3"],
                                  ["--raw", "g", "p"], ["--raw", "g", "p", "0"],
                                  ["--raw", "g", "p", "x"]])
def test_cli_rejects_malformed_arguments(epi, args):
    if not os.path.exists(epi.CLI_PATH):
        pytest.skip("bin/v4 not built")
    r = subprocess.run([epi.CLI_PATH] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_cli_reports_unreadable_input(epi, tmp_path):
    # input is read before any device work, so this runs without a GPU
    if not os.path.exists(epi.CLI_PATH):
        pytest.skip("bin/v4 not built")
    for args in (["--bfile", str(tmp_path / "absent")],
                 ["--raw", str(tmp_path / "g"), str(tmp_path / "p"), "10"]):
        r = subprocess.run([epi.CLI_PATH] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and "reading input failed" in r.stderr, r.stderr


def test_context_without_gpu_fails_loudly(epi):
    # On a host with no GPU the library reports a device error rather than
    # falling back to any CPU path.
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present")
    with pytest.raises(epi.EpiError) as e:
        epi.Context(0)
    assert e.value.code == epi.EPI_EDEVICE


def test_strerror_names_every_code(epi):
    for code in range(0, -11, -1):
        s = epi.strerror(code)
        assert s and "unknown" not in s


def test_cohort_shape_bounds(epi):
    """epi_cohort_footprint accepts exactly the shapes every kernel can
    launch and count exactly (no device needed): n_samples in [1, 2^22),
    n_snps in [3, 2^21 - 1], n_samples x n_snps <= 2^38."""
    ok = [(1, 3), (1024, 256), ((1 << 22) - 1, 3), (1, (1 << 21) - 1), (1 << 17, 1 << 20)]
    bad = [(0, 10), (10, 2), (10, 0), (1 << 22, 3), (1, 1 << 21), (1 << 21, (1 << 17) + 1)]
    for N, M in ok:
        b = epi.cohort_footprint(N, M)
        Mp = (M + 63) // 64 * 64
        assert b >= 32 * Mp * Mp  # the full pair margins of both classes
    for N, M in bad:
        with pytest.raises(epi.EpiError) as e:
            epi.cohort_footprint(N, M)
        assert e.value.code == epi.EPI_EINVAL, (N, M)
