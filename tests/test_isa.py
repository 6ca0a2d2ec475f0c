"""gfx950 assembly of the MFMA scan kernels (CPU only: hipcc cross-compiles).

The stage copies are inline asm that set M0, a register the compiler
reserves and does not expect asm to change; that is sound only while no
compiler-generated instruction of these kernels uses M0.  Also: the
production instances keep their stage and slot loops free of scratch
reloads (a reload's wait would also wait for the in-flight stage copies)."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "cross-dpc-episdet_amd", "csrc", "episdet_mfma.hip")
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("needs hipcc")
    out = tmp_path_factory.mktemp("isa") / "mfma.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "--cuda-device-only", "-S", "-O3",
                    "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-o", str(out), SRC], check=True, timeout=600,
                   capture_output=True)
    return out.read_text()


def kernels(text):
    """name -> body of every mfma_scan_kernel instance."""
    out = {}
    for m in re.finditer(r"^(_ZN3epi16mfma_scan_kernel\w+):", text, re.M):
        end = text.find("s_endpgm", m.end())
        out[m.group(1)] = text[m.end():end]
    return out


def test_m0_only_in_the_stage_copies(asm):
    ks = kernels(asm)
    assert len(ks) == 8
    for name, body in ks.items():
        lines = [l.split(";")[0].strip() for l in body.splitlines()]
        lines = [l for l in lines if l and not l.startswith((".", "#"))]
        for t, l in enumerate(lines):
            if re.search(r"\bm0\b", l):
                # the copy's own pair: s_mov_b32 m0, sN then global_load_lds
                assert l.startswith("s_mov_b32 m0,"), (name, l)
                assert lines[t + 1].startswith("global_load_lds_dwordx4"), (name, lines[t + 1])


def test_production_loops_without_scratch_reloads(asm):
    """Instances <TLDS, DUMP = false, FP4> (the four a scan runs): no
    scratch access in a block nested two loops deep (the stage loops and the
    epilogue's slot loop, inside the persistent item loop).  The dump-mode
    test instances do spill there."""
    ks = kernels(asm)
    checked = 0
    for name, body in ks.items():
        if not re.search(r"ILb[01]ELb0E", name):  # not dump mode
            continue
        checked += 1
        depth, notes = 0, False
        for l in body.splitlines():
            if re.match(r"^\.LBB\w+:", l):
                depth, notes = 0, True
            elif notes and not l.strip().startswith(";"):
                notes = False
            if notes:
                # the printer's loop notes: "in Loop: Header=... Depth=N",
                # "Parent Loop ... Depth=N", "Loop Header: Depth=N"
                depth = max([depth] + [int(d) for d in re.findall(r"Depth=(\d+)", l)])
            elif depth >= 2 and "scratch_" in l:
                raise AssertionError((name, l.strip()))
    assert checked == 4
