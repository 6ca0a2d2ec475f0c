"""On-disk input (raw bytes and PLINK 1 filesets) through the C ABI, on the
CPU: the readers and the writer touch no GPU.  The .bed codec is pinned by
hand-decoded bytes of the published PLINK 1 format and by round trips
against the oracle's restatement of it."""
import ctypes
import os

import numpy as np
import pytest

# Five samples x three SNPs, decoded by hand from the PLINK 1 rules:
# codes 00 -> 2, 01 -> missing (3), 10 -> 1, 11 -> 0; sample s in bits
# 2(s%4)..2(s%4)+1 of byte s/4 of the SNP's row.
KAT_GENO = np.array([[0, 2, 1],
                     [1, 2, 0],
                     [2, 1, 0],
                     [3, 0, 2],
                     [0, 1, 3]], np.uint8)
# SNP0: 0b01_00_10_11 = 0x4B, then sample 4 -> 0x03
# SNP1: 0b11_10_00_00 = 0xE0, then 0x02
# SNP2: 0b00_11_11_10 = 0x3E, then 0x01
KAT_BED = np.array([[0x4B, 0x03], [0xE0, 0x02], [0x3E, 0x01]], np.uint8)
KAT_FAM = "a s0 0 0 1 1\na s1 0 0 2 2\na s2 0 0 1 0\nb s3 0 0 2 -9\nb s4 0 0 0 2\n"
KAT_PHENO = np.array([0, 1, 255, 255, 1], np.uint8)


def write_fileset(prefix, bed_rows, fam, n_bim, magic=b"\x6c\x1b\x01", extra=b""):
    with open(prefix + ".bed", "wb") as f:
        f.write(magic + np.ascontiguousarray(bed_rows).tobytes() + extra)
    with open(prefix + ".fam", "w") as f:
        f.write(fam)
    with open(prefix + ".bim", "w") as f:
        for j in range(n_bim):
            f.write(f"1 rs{j} 0 {100 + j} A G\n")


def test_bed_codec_known_answer(orc):
    assert np.array_equal(orc.bed_decode(KAT_BED, 5), KAT_GENO)
    assert np.array_equal(orc.bed_encode(KAT_GENO), KAT_BED)


def test_read_plink_known_answer(epi, tmp_path):
    pre = str(tmp_path / "kat")
    write_fileset(pre, KAT_BED, KAT_FAM, 3)
    bed, pheno = epi.read_plink(pre)
    assert np.array_equal(bed, KAT_BED)
    assert np.array_equal(pheno, KAT_PHENO)


def test_write_plink_round_trip(epi, orc, tmp_path):
    rng = np.random.default_rng(11)
    N, M = 37, 11  # N not a multiple of 4: the last byte of a row is partial
    geno = rng.choice(np.array([0, 1, 2, 7], np.uint8), size=(N, M), p=[0.4, 0.3, 0.2, 0.1])
    pheno = rng.choice(np.array([0, 1, 5], np.uint8), size=N)
    pre = str(tmp_path / "rt")
    epi.write_plink(pre, geno, pheno)
    bed, ph = epi.read_plink(pre)
    assert np.array_equal(bed, orc.bed_encode(geno))
    want = geno.copy()
    want[want > 2] = 3  # anything but 0/1/2 is written as missing
    assert np.array_equal(orc.bed_decode(bed, N), want)
    # phenotype 1 is a case (2 in .fam); everything else a control (1)
    assert np.array_equal(ph, (pheno == 1).astype(np.uint8))
    with open(pre + ".fam") as f:
        assert len(f.read().splitlines()) == N
    with open(pre + ".bim") as f:
        assert len(f.read().splitlines()) == M
    assert os.path.getsize(pre + ".bed") == 3 + M * ((N + 3) // 4)


@pytest.mark.parametrize("case", ["magic", "individual_major", "short", "long", "fam_cols",
                                  "few_snps"])
def test_read_plink_rejects_malformed(epi, tmp_path, case):
    pre = str(tmp_path / case)
    fam, nb, magic, rows, extra = KAT_FAM, 3, b"\x6c\x1b\x01", KAT_BED, b""
    if case == "magic":
        magic = b"\x00\x1b\x01"
    elif case == "individual_major":
        magic = b"\x6c\x1b\x00"
    elif case == "short":
        rows = KAT_BED[:2]
    elif case == "long":
        extra = b"\x00"
    elif case == "fam_cols":
        fam = "a s0 0 0 1\n" * 5
    elif case == "few_snps":
        rows, nb = KAT_BED[:2], 2
    write_fileset(pre, rows, fam, nb, magic, extra)
    with pytest.raises(epi.EpiError) as e:
        epi.read_plink(pre)
    assert e.value.code == epi.EPI_EFORMAT


def test_read_plink_missing_file_and_stale_counts(epi, tmp_path):
    with pytest.raises(epi.EpiError) as e:
        epi.read_plink(str(tmp_path / "absent"))
    assert e.value.code == epi.EPI_EIO
    pre = str(tmp_path / "kat")
    write_fileset(pre, KAT_BED, KAT_FAM, 3)
    n, m = ctypes.c_uint32(4), ctypes.c_uint32(3)  # one sample short of the fileset
    bed = np.zeros((3, 2), np.uint8)
    ph = np.zeros(5, np.uint8)
    rc = epi.lib().epi_read_plink(os.fsencode(pre), ctypes.byref(n), ctypes.byref(m),
                                  bed.ctypes.data, ph.ctypes.data)
    assert rc == epi.EPI_EINVAL
    # buffers must come together
    rc = epi.lib().epi_read_plink(os.fsencode(pre), ctypes.byref(n), ctypes.byref(m),
                                  bed.ctypes.data, None)
    assert rc == epi.EPI_EINVAL


def test_read_raw(epi, tmp_path):
    rng = np.random.default_rng(5)
    geno = rng.integers(0, 3, size=(21, 9), dtype=np.uint8)
    pheno = rng.integers(0, 2, size=21, dtype=np.uint8)
    gp, pp = str(tmp_path / "g.bin"), str(tmp_path / "p.bin")
    geno.tofile(gp)
    pheno.tofile(pp)
    g, p = epi.read_raw(gp, pp, 21)
    assert np.array_equal(g, geno) and np.array_equal(p, pheno)
    with pytest.raises(epi.EpiError) as e:
        epi.read_raw(gp, pp, 20)  # phenotype file holds 21 bytes
    assert e.value.code == epi.EPI_EFORMAT
    geno[:, :2].tofile(gp)
    with pytest.raises(epi.EpiError) as e:
        epi.read_raw(gp, pp, 21)  # two SNPs: fewer than a triplet
    assert e.value.code == epi.EPI_EFORMAT
    with pytest.raises(epi.EpiError) as e:
        epi.read_raw(str(tmp_path / "absent"), pp, 21)
    assert e.value.code == epi.EPI_EIO
