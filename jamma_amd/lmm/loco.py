"""Leave-one-chromosome-out association on the device.

``run_lmm_loco`` mirrors src/jamma/lmm/loco.py:54-252 (LOCO kinships from
per-chromosome int8 partial Grams, K_c = (S_full - S_c) / (p - p_c), one
eigendecomposition per chromosome, association of that chromosome's SNPs
with float64 statistics as in _run_lmm_for_chromosome :255-593).

``run_lmm_loco_multi`` adds what the reference lacks: several phenotypes with
the same valid-sample set share each chromosome's LOCO kinship,
eigendecomposition and genotype rotation; only the spectrum moments and the
lambda search run per phenotype.
"""

from __future__ import annotations

import logging
import time
from pathlib import Path

import numpy as np
import torch

from jamma_amd.device import require_gpu, round_up, to_dev
from jamma_amd.io.plink import BedFile, decode, get_chromosome_partitions, get_plink_metadata
from jamma_amd.kinship.compute import loco_kinship_streaming_device
from jamma_amd.kinship.io import write_kinship_matrix
from jamma_amd.lmm.eigen import eigh_device
from jamma_amd.lmm.engine import AssocProblem, Cohort
from jamma_amd.lmm.io import IncrementalAssocWriter
from jamma_amd.lmm.runner_streaming import (
    TILE_SNPS,
    TileSink,
    _TEST_TYPE,
    _meta_snp_info,
    default_precision,
    snp_filter,
)
from jamma_amd.lmm.stats import AssocResult

logger = logging.getLogger("jamma_amd")


def _valid_mask(phenotypes: np.ndarray, covariates) -> np.ndarray:
    ph = np.atleast_2d(np.asarray(phenotypes, dtype=np.float64).T).T
    valid = np.all(~np.isnan(ph) & (ph != -9.0), axis=1)
    if covariates is not None:
        valid &= np.all(~np.isnan(np.asarray(covariates, dtype=np.float64)), axis=1)
    if not valid.any():
        raise ValueError("No samples with valid phenotypes")
    return valid


def _chromosome_tiles(bed: BedFile, chr_idx: np.ndarray, sample_idx, n: int, maf: float,
                      miss: float, snps_indices):
    """Decode one chromosome's SNPs, float64 statistics + filter (loco.py:305-356)."""
    lo, hi = int(chr_idx.min()), int(chr_idx.max()) + 1
    raw = bed.window_to_device(lo, hi)
    _, c, _ = decode(raw, bed.bytes_per_snp, len(chr_idx), sample_idx=sample_idx, n_out=n,
                     snp_idx=(chr_idx - lo).astype(np.int32), want_g=False)
    counts = c[: len(chr_idx)].cpu().numpy()
    mask, means, af, miss_c = snp_filter(counts, n, maf, miss, np.float64)
    if snps_indices is not None:
        mask &= np.isin(chr_idx, np.asarray(snps_indices))
    local = np.where(mask)[0]
    return raw, lo, local, means, af, miss_c


def run_lmm_loco(
    bed_path: Path,
    phenotypes: np.ndarray,
    covariates: np.ndarray | None = None,
    maf_threshold: float = 0.01,
    miss_threshold: float = 0.05,
    lmm_mode: int = 1,
    output_path: Path | None = None,
    check_memory: bool = True,
    show_progress: bool = True,
    save_kinship: bool = False,
    kinship_output_dir: Path | None = None,
    kinship_output_prefix: str = "result",
    snps_indices: np.ndarray | None = None,
    ksnps_indices: np.ndarray | None = None,
    *,
    precision: str | None = None,
) -> list[AssocResult]:
    """LOCO association (loco.py:54-252); results in sorted-chromosome order."""
    del check_memory
    out = run_lmm_loco_multi(bed_path, np.asarray(phenotypes, dtype=np.float64)[:, None],
                             covariates=covariates, maf_threshold=maf_threshold,
                             miss_threshold=miss_threshold, lmm_mode=lmm_mode,
                             output_paths=[output_path] if output_path is not None else None,
                             show_progress=show_progress, save_kinship=save_kinship,
                             kinship_output_dir=kinship_output_dir,
                             kinship_output_prefix=kinship_output_prefix,
                             snps_indices=snps_indices, ksnps_indices=ksnps_indices,
                             precision=precision)
    return out[0]


def run_lmm_loco_multi(
    bed_path: Path,
    phenotypes: np.ndarray,
    covariates: np.ndarray | None = None,
    maf_threshold: float = 0.01,
    miss_threshold: float = 0.05,
    lmm_mode: int = 1,
    output_paths: list[Path] | None = None,
    show_progress: bool = True,
    save_kinship: bool = False,
    kinship_output_dir: Path | None = None,
    kinship_output_prefix: str = "result",
    snps_indices: np.ndarray | None = None,
    ksnps_indices: np.ndarray | None = None,
    *,
    precision: str | None = None,
) -> list[list[AssocResult]]:
    """LOCO for a (n_samples, n_pheno) phenotype matrix sharing one valid-sample set."""
    t0 = time.perf_counter()
    precision = precision or default_precision()
    if lmm_mode not in (1, 2, 3, 4):
        raise ValueError(f"lmm_mode must be 1..4, got {lmm_mode}")
    require_gpu()
    Y = np.asarray(phenotypes, dtype=np.float64)
    if Y.ndim == 1:
        Y = Y[:, None]
    meta = get_plink_metadata(bed_path)
    partitions = get_chromosome_partitions(bed_path)
    chrs = sorted(partitions.keys())
    if len(chrs) < 2:
        raise ValueError("LOCO requires SNPs on multiple chromosomes. "
                         f"Found only {len(chrs)} chromosome(s): {chrs}")
    valid = _valid_mask(Y, covariates)
    n = int(valid.sum())
    sample_idx = None if valid.all() else np.where(valid)[0].astype(np.int32)
    Yv = Y[valid]
    Wv = None if covariates is None else np.asarray(covariates, dtype=np.float64)[valid]
    snp_info = _meta_snp_info(meta)
    bed = BedFile(bed_path)
    n_ph = Yv.shape[1]
    writers = None
    if output_paths is not None:
        writers = [IncrementalAssocWriter(p, test_type=_TEST_TYPE[lmm_mode]) for p in output_paths]
        for w in writers:
            w.__enter__()
    results: list[list[AssocResult]] = [[] for _ in range(n_ph)]
    vdev = None if valid.all() else torch.as_tensor(np.where(valid)[0], device=require_gpu())
    ld = round_up(n, 64)
    try:
        for chr_name, K in loco_kinship_streaming_device(bed_path, maf_threshold=maf_threshold,
                                                         miss_threshold=miss_threshold,
                                                         ksnps_indices=ksnps_indices):
            if save_kinship and kinship_output_dir is not None:
                write_kinship_matrix(K.cpu().numpy(), Path(kinship_output_dir) /
                                     f"{kinship_output_prefix}.loco.cXX.chr{chr_name}.txt")
            Kv = K if vdev is None else K[vdev][:, vdev]
            del K
            d, U = eigh_device(Kv)
            del Kv
            cohort = Cohort(d, U, precision="both" if precision == "mixed" else "f64")
            del U
            probs = [AssocProblem(cohort, Yv[:, k], Wv) for k in range(n_ph)]
            chr_idx = partitions[chr_name]
            raw, lo, local, means, af, miss_c = _chromosome_tiles(
                bed, chr_idx, sample_idx, n, maf_threshold, miss_threshold, snps_indices)
            sinks = [TileSink(lmm_mode, snp_info, af, miss_c,
                              writers[k] if writers is not None else None, results[k])
                     for k in range(n_ph)]
            for t0i in range(0, len(local), TILE_SNPS):
                tl = local[t0i:t0i + TILE_SNPS]
                gsnp = chr_idx[tl]
                g, _, _ = decode(raw, bed.bytes_per_snp, len(tl), sample_idx=sample_idx, n_out=n,
                                 snp_idx=(gsnp - lo).astype(np.int32), want_counts=False, ld_g=ld)
                mdev = to_dev(means[tl], torch.float64)
                Z, f32, ldz = cohort.rotate_genotypes(g, ld, mdev, len(tl), precision)
                for k, pr in enumerate(probs):
                    xm = cohort.moments(Z, f32, ldz, len(tl), pr.V, pr.n_cvt + 1)
                    sinks[k].push(chr_idx[tl], tl, pr.assoc_moments(xm, len(tl), lmm_mode),
                                  (pr, g, ld, mdev) if precision == "mixed" else None)
                del Z
            for sk in sinks:
                sk.close()
            if show_progress:
                logger.info("LOCO chr %s: %d SNPs, T=%d", chr_name, len(local), cohort.n_terms)
            del cohort, probs
    finally:
        if writers is not None:
            for w in writers:
                w.__exit__(None, None, None)
    if show_progress:
        logger.info("LOCO association completed in %.2fs", time.perf_counter() - t0)
    return [[] for _ in range(n_ph)] if writers is not None else results
