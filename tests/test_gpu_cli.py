"""The `v4` program, the drop-in for the reference's `./v4 <num_samples>
<num_snps> <block_size>` (GPU/README.md:29-30): the same four output lines
(GPU/v4.cpp:251, :610-612) with the reference's known answers, and the file
inputs and top-K lines the reference does not have."""
import json
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_v4(epi, args, timeout=300):
    say("v4", *args)
    r = subprocess.run([epi.CLI_PATH] + [str(a) for a in args], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    return r.stdout.splitlines()


def four_lines(lines):
    assert lines[0].startswith("Selected device: ") and len(lines[0]) > len("Selected device: ")
    assert lines[1].startswith("Time: ") and float(lines[1].split()[1]) > 0
    assert lines[2].startswith("Score: ")
    assert lines[3].startswith("Solution: ")
    return lines[2].split()[1], [int(x) for x in lines[3].split()[1:]]


@pytest.mark.parametrize("idx", [0, 2, 3, 6, 8, 11])
def test_cli_prints_reference_known_answers(epi, golden, idx):
    c = golden("kat_reference.json")["cases"][idx]
    lines = run_v4(epi, [c["N"], c["M"], 256])
    score, best = four_lines(lines)
    assert best == c["runs"][0]["best"]
    # the reference prints the float through std::cout: 6 significant digits
    assert score == "%.6g" % np.float32(float(c["runs"][0]["score"]))
    assert len(lines) == 4


def top_lines(lines, k):
    tops = [ln for ln in lines if ln.startswith("Top ")]
    assert len(tops) == k
    out = []
    for r, ln in enumerate(tops):
        head, rest = ln.split(":", 1)
        assert head == f"Top {r + 1}"
        f = rest.split()
        out.append((int(f[0]), int(f[1]), int(f[2]), np.float32(float(f[3]))))
    return out


def oracle_top(orc, geno, pheno, k):
    _, top = orc.PackedCohort(geno, pheno).scan(topk=k)
    return [(i, j, k_, np.float32(s)) for s, i, j, k_ in top]


def test_cli_plink_fileset_and_top_k(epi, orc, tmp_path):
    N, M = 1500, 90
    geno, pheno = epi.generate_hwe(31, N, M, 700)
    pre = str(tmp_path / "c")
    epi.write_plink(pre, geno, pheno)
    lines = run_v4(epi, ["--bfile", pre, "--top", 8, "--json"])
    _, best = four_lines(lines)
    want = oracle_top(orc, geno, pheno, 8)
    assert best == list(want[0][:3])
    assert top_lines(lines, 8) == want
    res = json.loads(lines[-1])
    assert (res["n_samples"], res["n_snps"], res["cases"]) == (N, M, 700)


def test_cli_raw_input(epi, orc, tmp_path):
    N, M = 640, 50
    geno, pheno = epi.generate_hwe(32, N, M, 320)
    gp, pp = tmp_path / "g.bin", tmp_path / "p.bin"
    geno.tofile(gp)
    pheno.tofile(pp)
    lines = run_v4(epi, ["--raw", gp, pp, N, "--top", 3, "--variant", 1])
    _, best = four_lines(lines)
    want = oracle_top(orc, geno, pheno, 3)
    assert best == list(want[0][:3])
    assert top_lines(lines, 3) == want


def test_cli_planted_triplet_is_found(epi):
    lines = run_v4(epi, [4096, 512, 256, "--planted", 7, "--json"])
    res = json.loads(lines[-1])
    assert [res["i"], res["j"], res["k"]] == res["planted"]
