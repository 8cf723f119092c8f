"""PLINK binary decoding for the oracle (test infrastructure only).

Restates what the reference obtains from bed-reader 1.0.6 (not vendored; pinned
in the reference's uv.lock) through ``jamma.io.plink``:
  * ``stream_genotype_chunks`` src/jamma/io/plink.py:239-300 -- windowed reads of
    the SNP-major .bed into (n_samples, chunk) float arrays, NaN for missing;
  * ``get_plink_metadata`` src/jamma/io/plink.py:92-136;
  * ``validate_plink_dimensions`` src/jamma/io/plink.py:178-220.

bed-reader's default is ``count_A1=True``: the value is the number of copies of
allele 1 (the .bim 5th column). PLINK's 2-bit codes, low bits first, four
samples per byte: 00 -> homozygous A1 (2), 01 -> missing (NaN),
10 -> heterozygous (1), 11 -> homozygous A2 (0). The mapping is pinned by the
reference fixture ``tests/fixtures/gemma_synthetic``: af = mean/2 of the decoded
values reproduces the GEMMA ``af`` column exactly (tests/test_oracle.py).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

BED_MAGIC = bytes([0x6C, 0x1B, 0x01])

# code -> A1 count (NaN = missing)
_CODE_VALUE = np.array([2.0, np.nan, 1.0, 0.0])


def read_fam(prefix) -> list[list[str]]:
    with open(f"{prefix}.fam") as f:
        return [line.split() for line in f if line.strip()]


def read_bim(prefix) -> dict:
    chrom, rs, pos, a1, a0 = [], [], [], [], []
    with open(f"{prefix}.bim") as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            chrom.append(t[0])
            rs.append(t[1])
            pos.append(int(t[3]))
            a1.append(t[4])
            a0.append(t[5])
    return {"chromosome": np.array(chrom), "sid": np.array(rs),
            "bp_position": np.array(pos), "allele_1": np.array(a1),
            "allele_2": np.array(a0)}


def decode_bed_bytes(raw: bytes | np.ndarray, n_samples: int, n_snps: int) -> np.ndarray:
    """Decode SNP-major .bed payload (magic stripped) to (n_samples, n_snps) float64."""
    bps = (n_samples + 3) // 4
    buf = np.frombuffer(raw, dtype=np.uint8, count=bps * n_snps).reshape(n_snps, bps)
    codes = np.empty((n_snps, bps * 4), dtype=np.uint8)
    for k in range(4):
        codes[:, k::4] = (buf >> (2 * k)) & 3
    return _CODE_VALUE[codes[:, :n_samples]].T.copy()


def read_bed(prefix, snp_start: int = 0, snp_end: int | None = None,
             dtype=np.float64) -> np.ndarray:
    """Windowed read, as bed-reader ``read(index=np.s_[:, start:end])``."""
    n = len(read_fam(prefix))
    p = sum(1 for line in open(f"{prefix}.bim") if line.strip())
    snp_end = p if snp_end is None else min(snp_end, p)
    bps = (n + 3) // 4
    with open(f"{prefix}.bed", "rb") as f:
        if f.read(3) != BED_MAGIC:
            raise ValueError("not a SNP-major PLINK .bed file")
        f.seek(3 + snp_start * bps)
        raw = f.read(bps * (snp_end - snp_start))
    return decode_bed_bytes(raw, n, snp_end - snp_start).astype(dtype)


def encode_bed(genotypes: np.ndarray) -> bytes:
    """Inverse of decode: (n, p) A1 counts with NaN -> .bed bytes (with magic)."""
    n, p = genotypes.shape
    bps = (n + 3) // 4
    codes = np.full((p, bps * 4), 3, dtype=np.uint8)  # pad value irrelevant
    g = genotypes.T
    c = np.where(np.isnan(g), 1, np.where(g == 2, 0, np.where(g == 1, 2, 3)))
    codes[:, :n] = c
    out = np.zeros((p, bps), dtype=np.uint8)
    for k in range(4):
        out |= (codes[:, k::4] << (2 * k)).astype(np.uint8)
    return BED_MAGIC + out.tobytes()


def write_plink(prefix, genotypes: np.ndarray, phenotypes=None, chromosomes=None):
    """Write a PLINK fileset (used by tests to make synthetic inputs)."""
    n, p = genotypes.shape
    prefix = Path(prefix)
    prefix.parent.mkdir(parents=True, exist_ok=True)
    with open(f"{prefix}.bed", "wb") as f:
        f.write(encode_bed(genotypes))
    chromosomes = ["1"] * p if chromosomes is None else chromosomes
    with open(f"{prefix}.bim", "w") as f:
        for j in range(p):
            f.write(f"{chromosomes[j]}\trs{j:06d}\t0\t{1000 * (j + 1)}\tA\tG\n")
    with open(f"{prefix}.fam", "w") as f:
        for i in range(n):
            ph = "-9" if phenotypes is None or np.isnan(phenotypes[i]) else repr(float(phenotypes[i]))
            f.write(f"F{i}\tI{i}\t0\t0\t0\t{ph}\n")
