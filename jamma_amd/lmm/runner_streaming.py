"""LMM association runners on the device engine.

``run_lmm_association_streaming`` keeps the signature and semantics of
src/jamma/lmm/runner_streaming.py:94-628 (two passes over the PLINK file:
pass-1 SNP statistics and filter, pass-2 association per file chunk) and
``run_lmm_association_jax`` those of the in-memory runner
src/jamma/lmm/runner_jax.py:52-481. Both feed one device pipeline per SNP
tile: on-device .bed decode -> rotation (FP64 or BF16x3 MFMA) -> spectrum
moments -> fused lambda search + Wald/LRT/Score.

The extra keyword ``precision`` selects the rotation: "f64" (default; the
reference's own precision) or "mixed" (BF16x3 MFMA, FP32 accumulation).
"""

from __future__ import annotations

import logging
import os
import time
from pathlib import Path

import numpy as np
import torch

from jamma_amd._lib import JM_NOUT, OUT
from jamma_amd.device import require_gpu, round_up, to_dev
from jamma_amd.io.plink import BedFile, decode, get_plink_metadata, sample_key, stats_from_counts
from jamma_amd.lmm.eigen import eigh_device
from jamma_amd.lmm.engine import AssocProblem, Cohort, guard_rows
from jamma_amd.lmm.io import IncrementalAssocWriter, snp_text_parts
from jamma_amd.lmm.stats import AssocResult

logger = logging.getLogger("jamma_amd")

TILE_SNPS = int(os.environ.get("JM_TILE_SNPS", "8192"))
_TEST_TYPE = {1: "wald", 2: "lrt", 3: "score", 4: "all"}


def default_precision() -> str:
    return os.environ.get("JM_PRECISION", "f64")


def hwe_pvalues(n_aa, n_ab, n_bb) -> np.ndarray:
    """Chi-square HWE p-values (core/snp_filter.py:94-150), chi2(1) sf = erfc(sqrt(x/2))."""
    from scipy.special import erfc

    n_aa, n_ab, n_bb = (np.asarray(v, dtype=np.float64) for v in (n_aa, n_ab, n_bb))
    n = n_aa + n_ab + n_bb
    with np.errstate(invalid="ignore", divide="ignore"):
        p = (2 * n_aa + n_ab) / (2 * n)
        q = 1.0 - p
        e_aa, e_ab, e_bb = n * p ** 2, 2 * n * p * q, n * q ** 2
        chi = (n_aa - e_aa) ** 2 / e_aa + (n_ab - e_ab) ** 2 / e_ab + (n_bb - e_bb) ** 2 / e_bb
    chi = np.where(np.isnan(chi), 0.0, chi)
    return erfc(np.sqrt(0.5 * np.maximum(chi, 0.0)))


def snp_filter(counts: np.ndarray, n_samples: int, maf_threshold: float, miss_threshold: float,
               stat_dtype=np.float32):
    """compute_snp_filter_mask (core/snp_filter.py:59-91) from genotype counts."""
    mean, miss, poly = stats_from_counts(counts, stat_dtype)
    af = mean / 2.0
    maf = np.minimum(af, 1.0 - af)
    mask = (maf >= maf_threshold) & (miss / n_samples <= miss_threshold) & poly
    return mask, mean, af, miss


def _meta_snp_info(meta) -> list[dict]:
    return [{"chr": str(meta["chromosome"][i]), "rs": str(meta["sid"][i]),
             "pos": int(meta["bp_position"][i]), "a1": str(meta["allele_1"][i]),
             "a0": str(meta["allele_2"][i])} for i in range(meta["n_snps"])]


def _make_result(mode: int, info: dict, af: float, n_miss: int, row: np.ndarray) -> AssocResult:
    """_yield_chunk_results (src/jamma/lmm/results.py:222-286)."""
    meta = dict(chr=info["chr"], rs=info["rs"], ps=info.get("pos", info.get("ps", 0)),
                n_miss=int(n_miss), allele1=info.get("a1", info.get("allele1", "")),
                allele0=info.get("a0", info.get("allele0", "")), af=float(af))
    g = lambda k: float(row[OUT[k]])  # noqa: E731
    if mode == 1:
        return AssocResult(**meta, beta=g("beta"), se=g("se"), logl_H1=g("logl_H1"),
                           l_remle=g("l_remle"), p_wald=g("p_wald"))
    if mode == 3:
        return AssocResult(**meta, beta=g("beta_score"), se=g("se_score"), p_score=g("p_score"))
    if mode == 2:
        return AssocResult(**meta, beta=float("nan"), se=float("nan"), l_mle=g("l_mle"),
                           p_lrt=g("p_lrt"))
    return AssocResult(**meta, beta=g("beta"), se=g("se"), logl_H1=g("logl_H1"),
                       l_remle=g("l_remle"), l_mle=g("l_mle"), p_wald=g("p_wald"),
                       p_lrt=g("p_lrt"), p_score=g("p_score"))


class TileSink:
    """Consumes the device result rows of SNP tiles one tile behind their launch.

    ``push`` starts a non-blocking copy of a tile's rows into pinned host memory and
    then finishes the previous tile: with a writer its rows go to the writer's
    background formatter (GEMMA text, no per-SNP Python objects), otherwise they
    become AssocResults (_yield_chunk_results, results.py:222-286). The device works
    on the next tile meanwhile. ``close`` finishes the last tile."""

    def __init__(self, mode: int, snp_info, af, miss, writer=None, results=None):
        self.mode, self.snp_info, self.af, self.miss = mode, snp_info, af, miss
        self.writer = writer
        self.results = results if results is not None else []
        self._pending = None
        self.retested = 0

    def push(self, snps: np.ndarray, stat_idx: np.ndarray, out_dev: torch.Tensor,
             guard=None) -> None:
        """snps: global SNP indices (snp_info rows); stat_idx: their rows of af / miss.
        guard (mixed precision): (problem, g, ld_g, means) of the tile, so that rows
        with a p-value within the guard band of a significance threshold are
        recomputed on the FP64 path before they are consumed."""
        k = len(snps)
        host = torch.empty((k, JM_NOUT), dtype=torch.float64, pin_memory=True)
        host.copy_(out_dev[:k], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        prev, self._pending = self._pending, (np.asarray(snps), np.asarray(stat_idx), host, ev,
                                              out_dev, guard)
        if prev is not None:
            self._finish(*prev)

    def close(self) -> None:
        prev, self._pending = self._pending, None
        if prev is not None:
            self._finish(*prev)

    def _finish(self, snps, stat_idx, host, ev, out_dev, guard) -> None:
        ev.synchronize()
        rows = host.numpy()
        if guard is not None:
            near = guard_rows(rows, self.mode)
            if len(near):
                prob, g, ld_g, means = guard
                prob.retest_f64(g, ld_g, means, near, self.mode, out_dev)
                rows[near] = out_dev[torch.as_tensor(near, device=out_dev.device)].cpu().numpy()
                self.retested += len(near)
        if self.writer is not None:
            parts = snp_text_parts([self.snp_info[int(s)] for s in snps])
            self.writer.write_rows(parts, self.miss[stat_idx], self.af[stat_idx], rows)
        else:
            for j, s in enumerate(snps):
                i = stat_idx[j]
                self.results.append(_make_result(self.mode, self.snp_info[int(s)], self.af[i],
                                                 self.miss[i], rows[j]))


def _prepare_problem(kinship, eigenvalues, eigenvectors, phenotypes, covariates, precision, l_min,
                     l_max, n_grid, n_refine):
    if eigenvalues is not None:
        d = np.asarray(eigenvalues, dtype=np.float64)
        U = eigenvectors
    else:
        d, U = eigh_device(kinship)
    cohort = Cohort(d, U, precision=precision)
    prob = AssocProblem(cohort, phenotypes, covariates, l_min=l_min, l_max=l_max, n_grid=n_grid,
                        n_refine=n_refine)
    return cohort, prob


def _check_eigen_args(kinship, eigenvalues, eigenvectors):
    if (eigenvalues is None) != (eigenvectors is None):
        raise ValueError("Must provide both eigenvalues and eigenvectors, or neither. "
                         f"Got eigenvalues={eigenvalues is not None}, "
                         f"eigenvectors={eigenvectors is not None}")
    if kinship is None and eigenvalues is None:
        raise ValueError("Either kinship or pre-computed eigendecomposition (eigenvalues + "
                         "eigenvectors) must be provided")


def _valid_subset(phenotypes, kinship, covariates):
    phenotypes = np.asarray(phenotypes, dtype=np.float64)
    valid = ~np.isnan(phenotypes) & (phenotypes != -9.0)
    if covariates is not None:
        covariates = np.asarray(covariates, dtype=np.float64)
        valid &= np.all(~np.isnan(covariates), axis=1)
    if not np.all(valid):
        phenotypes = phenotypes[valid]
        if kinship is not None:
            kinship = np.asarray(kinship)[np.ix_(valid, valid)]
        if covariates is not None:
            covariates = covariates[valid, :]
    return valid, phenotypes, kinship, covariates


def run_lmm_association_streaming(
    bed_path: Path,
    phenotypes: np.ndarray,
    kinship: np.ndarray | None = None,
    snp_info: list | None = None,
    covariates: np.ndarray | None = None,
    eigenvalues: np.ndarray | None = None,
    eigenvectors: np.ndarray | None = None,
    maf_threshold: float = 0.01,
    miss_threshold: float = 0.05,
    l_min: float = 1e-5,
    l_max: float = 1e5,
    n_grid: int = 50,
    n_refine: int = 10,
    chunk_size: int = 10_000,
    use_gpu: bool = False,
    check_memory: bool = True,
    show_progress: bool = True,
    output_path: Path | None = None,
    lmm_mode: int = 1,
    snps_indices: np.ndarray | None = None,
    hwe_threshold: float = 0.0,
    validate_genotypes: bool = True,
    *,
    precision: str | None = None,
) -> list[AssocResult]:
    """Two-pass streaming association (runner_streaming.py:94-628); returns []
    when ``output_path`` is set (results on disk). ``use_gpu`` and
    ``check_memory`` are accepted for compatibility: the device path is the
    only path."""
    del use_gpu, check_memory, validate_genotypes
    t0 = time.perf_counter()
    precision = precision or default_precision()
    if lmm_mode not in (1, 2, 3, 4):
        raise ValueError(f"lmm_mode must be 1..4, got {lmm_mode}")
    _check_eigen_args(kinship, eigenvalues, eigenvectors)
    require_gpu()
    meta = get_plink_metadata(bed_path)
    n_snps = meta["n_snps"]
    if snp_info is None:
        snp_info = _meta_snp_info(meta)
    valid, phenotypes, kinship, covariates = _valid_subset(phenotypes, kinship, covariates)
    n = phenotypes.shape[0]
    sample_idx = None if np.all(valid) else np.where(valid)[0].astype(np.int32)
    bed = BedFile(bed_path)

    # ---- pass 1: per-SNP counts on the device (float32 statistics); inside a
    # resident_bed block the kinship's pass over the same samples is reused ----
    skey = sample_key(sample_idx, bed.n_samples)
    counts = bed.cached_counts(skey)
    if counts is None:
        counts = np.zeros((n_snps, 4), dtype=np.int64)
        for start in range(0, n_snps, chunk_size):
            end = min(start + chunk_size, n_snps)
            raw = bed.window_to_device(start, end)
            _, c, _ = decode(raw, bed.bytes_per_snp, end - start, sample_idx=sample_idx, n_out=n,
                             want_g=False)
            counts[start:end] = c[: end - start].cpu().numpy()
        bed.cache_counts(skey, counts.copy())
    else:
        counts = counts.copy()
    mask, means, af, miss = snp_filter(counts, n, maf_threshold, miss_threshold, np.float32)
    if snps_indices is not None:
        idx = np.asarray(snps_indices, dtype=np.int64)
        if len(idx) and idx.max() >= n_snps:
            raise ValueError(f"SNP list filter index {idx.max()} out of range for {n_snps} SNPs")
        lm = np.zeros(n_snps, dtype=bool)
        lm[idx] = True
        mask &= lm
    if hwe_threshold > 0:
        mask &= hwe_pvalues(counts[:, 0], counts[:, 1], counts[:, 2]) >= hwe_threshold
    snp_indices = np.where(mask)[0]
    if len(snp_indices) == 0:
        return []

    cohort, prob = _prepare_problem(kinship, eigenvalues, eigenvectors, phenotypes, covariates,
                                    precision, l_min, l_max, n_grid, n_refine)
    means_dev_all = to_dev(means, torch.float64)
    writer = None
    if output_path is not None:
        writer = IncrementalAssocWriter(output_path, test_type=_TEST_TYPE[lmm_mode])
        writer.__enter__()
    results: list[AssocResult] = []
    sink = TileSink(lmm_mode, snp_info, af, miss, writer, results)
    try:
        ld = round_up(n, 64)
        for start in range(0, n_snps, chunk_size):
            end = min(start + chunk_size, n_snps)
            left = np.searchsorted(snp_indices, start, side="left")
            right = np.searchsorted(snp_indices, end, side="left")
            if left == right:
                continue
            sel = snp_indices[left:right]
            raw = bed.window_to_device(start, end)
            for t0i in range(0, len(sel), TILE_SNPS):
                tsel = sel[t0i:t0i + TILE_SNPS]
                g, _, _ = decode(raw, bed.bytes_per_snp, len(tsel), sample_idx=sample_idx, n_out=n,
                                 snp_idx=(tsel - start).astype(np.int32), want_counts=False,
                                 ld_g=ld)
                mdev = means_dev_all[torch.as_tensor(tsel, device=means_dev_all.device)]
                mdev = mdev.contiguous()
                out = prob.associate(g, ld, mdev, len(tsel), lmm_mode, precision)
                sink.push(tsel, tsel, out,
                          (prob, g, ld, mdev) if precision == "mixed" else None)
        sink.close()
    except BaseException as e:
        if writer is not None:
            writer.__exit__(type(e), e, None)
            writer = None
        raise
    finally:
        if writer is not None:
            writer.__exit__(None, None, None)
    if show_progress:
        logger.info("LMM association (%s, %d SNPs) completed in %.2fs", precision,
                    len(snp_indices), time.perf_counter() - t0)
    return [] if output_path is not None else results


def encode_genotypes_i8(genotypes: np.ndarray) -> tuple[np.ndarray, bool]:
    """(n, p) float with NaN -> SNP-major int8 rows (p, n), -1 missing.
    Returns (g, integral) where integral is False when a value is not 0/1/2."""
    g = np.asarray(genotypes)
    nanm = np.isnan(g)
    vals = np.where(nanm, 0, g)
    integral = bool(np.all((vals == 0) | (vals == 1) | (vals == 2)))
    out = np.where(nanm, -1, vals).astype(np.int8).T.copy()
    return out, integral


def counts_from_i8(gi8: np.ndarray) -> np.ndarray:
    """SNP-major int8 rows -> (p, 4) counts (#0, #1, #2, #missing)."""
    return np.stack([(gi8 == 0).sum(1), (gi8 == 1).sum(1), (gi8 == 2).sum(1), (gi8 < 0).sum(1)],
                    axis=1).astype(np.int64)


def run_lmm_association_jax(
    genotypes: np.ndarray,
    phenotypes: np.ndarray,
    kinship: np.ndarray,
    snp_info: list,
    covariates: np.ndarray | None = None,
    eigenvalues: np.ndarray | None = None,
    eigenvectors: np.ndarray | None = None,
    maf_threshold: float = 0.01,
    miss_threshold: float = 0.05,
    l_min: float = 1e-5,
    l_max: float = 1e5,
    n_grid: int = 50,
    n_refine: int = 10,
    use_gpu: bool = False,
    check_memory: bool = True,
    show_progress: bool = True,
    lmm_mode: int = 1,
    *,
    precision: str | None = None,
) -> list[AssocResult]:
    """In-memory runner (src/jamma/lmm/runner_jax.py:52-481) on the device engine.

    Hard calls 0/1/2 (NaN = missing; the PLINK case) take the int8 decode and either
    rotation; any other values (dosages) are mean-imputed on the device and rotated
    in FP64. SNP statistics follow the input dtype (float32 from a PLINK load, as in
    the reference's compute_snp_stats)."""
    del use_gpu, check_memory
    precision = precision or default_precision()
    if lmm_mode not in (1, 2, 3, 4):
        raise ValueError(f"lmm_mode must be 1..4, got {lmm_mode}")
    _check_eigen_args(kinship, eigenvalues, eigenvectors)
    require_gpu()
    genotypes = np.asarray(genotypes)
    stat_dtype = np.float32 if genotypes.dtype == np.float32 else np.float64
    valid, phenotypes, kinship, covariates = _valid_subset(phenotypes, kinship, covariates)
    if not np.all(valid):
        genotypes = genotypes[valid, :]
    n, p = genotypes.shape
    gi8, integral = encode_genotypes_i8(genotypes)
    if integral:
        counts = counts_from_i8(gi8)
        mask, means, af, miss = snp_filter(counts, n, maf_threshold, miss_threshold, stat_dtype)
    else:
        # dosages: the reference's filter on nanmean / missing / nanvar (input dtype),
        # FP64 rotation of the mean-imputed values (no int8 packing, no bf16 split)
        from jamma_amd.kinship.compute import snp_stats_dense

        means, miss, var = snp_stats_dense(genotypes)
        af = means / 2.0
        maf = np.minimum(af, 1.0 - af)
        mask = (maf >= maf_threshold) & (miss / n <= miss_threshold) & (var > 0)
        if precision != "f64":
            logger.info("dosage genotypes: FP64 rotation (the mixed path needs hard calls)")
        precision = "f64"
    snp_indices = np.where(mask)[0]
    if len(snp_indices) == 0:
        return []
    cohort, prob = _prepare_problem(kinship, eigenvalues, eigenvectors, phenotypes, covariates,
                                    precision, l_min, l_max, n_grid, n_refine)
    ld = round_up(n, 64)
    results = []
    sink = TileSink(lmm_mode, snp_info, af, miss, None, results)
    if not integral:
        for t0i in range(0, len(snp_indices), TILE_SNPS):
            tsel = snp_indices[t0i:t0i + TILE_SNPS]
            X = to_dev(np.ascontiguousarray(np.asarray(genotypes[:, tsel], dtype=np.float64).T),
                       torch.float64)
            mdev = to_dev(means[tsel], torch.float64)
            sink.push(tsel, tsel, prob.associate_dense(X, mdev, len(tsel), lmm_mode))
        sink.close()
        if show_progress:
            logger.info("LMM association (f64, dosages, in-memory, %d SNPs) done", len(snp_indices))
        return results
    for t0i in range(0, len(snp_indices), TILE_SNPS):
        tsel = snp_indices[t0i:t0i + TILE_SNPS]
        gt = np.zeros((len(tsel), ld), dtype=np.int8)
        gt[:, :n] = gi8[tsel]
        g = to_dev(gt, torch.int8)
        mdev = to_dev(means[tsel], torch.float64)
        sink.push(tsel, tsel, prob.associate(g, ld, mdev, len(tsel), lmm_mode, precision),
                  (prob, g, ld, mdev) if precision == "mixed" else None)
    sink.close()
    if show_progress:
        logger.info("LMM association (%s, in-memory, %d SNPs) done", precision, len(snp_indices))
    return results
