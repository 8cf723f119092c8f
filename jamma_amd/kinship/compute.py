"""Centred kinship (GEMMA -gk 1) on the device: exact integer GEMM on MFMA + FP64 corrections.

Drop-in for src/jamma/kinship/compute.py:
  * compute_centered_kinship   :93-226  (in-memory genotypes)
  * compute_kinship_streaming   :502-690 (two passes over the PLINK file)
  * compute_loco_kinship(_streaming) :363-499, :693-916 (LOCO by subtraction)
The 0/1/2 calls (missing as 0) are packed to 2 bits per genotype and X0 X0^T is
accumulated exactly into int32 S (jm_kinship_syrk_2bit: FP4 MFMA with integer sums in
FP32, or the int8 MFMA with JM_KINSHIP_SYRK_I8MFMA=1; JM_KINSHIP_INT8=1 selects the
int8-operand form jm_kinship_syrk_i8; all give the same S); mean imputation and
centring are applied as exact
FP64 low-rank / sparse corrections (jm_kinship_impute_terms,
jm_kinship_finalize). Per-SNP statistics come from jm_decode_bed counts.
"""

from __future__ import annotations

import logging
import os
import warnings
from collections.abc import Iterator
from pathlib import Path

import numpy as np
import torch

from jamma_amd import _lib
from jamma_amd.device import ptr, require_gpu, round_up, stream_ptr, to_dev
from jamma_amd.io.plink import (
    BedFile,
    decode,
    get_plink_metadata,
    sample_key,
    stats_from_counts,
)

logger = logging.getLogger("jamma_amd")


class KinshipAccumulator:
    """Device accumulators S (int32), u, C (float64) and s0 over SNP tiles.

    The GEMM operand of incoming tiles (sample-major, missing calls 0, rows padded
    to a multiple of 256) is staged in one buffer of up to ``stage_bytes``: 2-bit
    packed (jm_pack_gram_2bit; each tile starts on a 16-SNP boundary, the gap reads
    as zero genotypes) or, with JM_KINSHIP_INT8=1, int8. The GEMM runs once per full
    buffer and when ``S`` is read, so the reference's 10,000-SNP batches cost one
    large launch instead of many small ones. The imputation terms u and C are
    applied per tile."""

    def __init__(self, n: int, device=None, stage_bytes: int = 1 << 31,
                 max_snps: int | None = None):
        """max_snps (optional): the SNPs this accumulator will see at most; the staging
        buffer is not made wider than that."""
        self.device = device or require_gpu()
        self.n = n
        self.n_pad = round_up(n, 256)
        self._S = torch.zeros((n, n), dtype=torch.int32, device=self.device)
        self.u = torch.zeros((n,), dtype=torch.float64, device=self.device)
        self.C = None
        self.s0 = 0.0
        self.p = 0
        self.packed = os.environ.get("JM_KINSHIP_INT8", "0") != "1"
        # staged SNP columns: a multiple of 256 (packed: 64 bytes per 256 SNPs and row)
        per_snp = 0.25 if self.packed else 1.0
        cap = (1 << 19) if self.packed else (1 << 17)
        self.stage_cols = int(min(cap, max(256, int(stage_bytes / (self.n_pad * per_snp)) // 256 * 256)))
        if max_snps is not None:
            self.stage_cols = int(min(self.stage_cols, max(256, round_up(max_snps + 16, 256))))
        self._stage = None
        self._fill = 0
        self.last_flush_cols = 0  # SNP columns of the stage's last GEMM
        # optional list: each SYRK launch appends its (start, end) torch.cuda.Event pair
        self.syrk_timer = None

    @property
    def S(self) -> torch.Tensor:
        self.flush()
        return self._S

    @S.setter
    def S(self, value: torch.Tensor) -> None:
        self.flush()
        self._S = value

    def _syrk(self, op: torch.Tensor, ld: int, n_snps: int) -> None:
        ev = None
        if self.syrk_timer is not None:
            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        name = "jm_kinship_syrk_2bit" if self.packed else "jm_kinship_syrk_i8"
        _lib.call(name, ptr(op), ld, self.n, n_snps, ptr(self._S), stream_ptr())
        if ev is not None:
            ev[1].record()
            self.syrk_timer.append(ev)

    def _stage_ld(self) -> int:
        return self.stage_cols // 4 if self.packed else self.stage_cols

    def flush(self) -> None:
        """Run the GEMM over the staged SNPs (S += X0 X0^T)."""
        if self._fill == 0:
            return
        k, cols = self._fill, self.stage_cols
        if not self.packed:
            kp = min(round_up(k, 128), cols)
            if kp > k:  # columns left from an earlier, longer fill must read as zero
                self._stage[:, k:kp].zero_()
        self._syrk(self._stage, self._stage_ld(), k)
        self.last_flush_cols = k
        self._fill = 0

    def release(self) -> None:
        """Flush and free the staging buffer."""
        self.flush()
        self._stage = None

    def add_tile(self, g: torch.Tensor, ld_g: int, n_snps: int, mu: np.ndarray,
                 has_missing: bool, n_missing: int | None = None) -> None:
        """g: SNP-major int8 (n_snps x ld_g, -1 missing); mu: float64 nanmean per SNP;
        n_missing: the tile's number of missing calls (from the decode counts; counted
        on the device when not given)."""
        if n_snps == 0:
            return
        n = self.n
        ldt = round_up(n_snps, 128)
        st = stream_ptr()
        gT = torch.empty((n, ldt), dtype=torch.int8, device=self.device)
        if ldt > n_snps:
            gT[:, n_snps:].zero_()
        if self.packed:
            _lib.call("jm_transpose_i8", ptr(g), n_snps, n, ld_g, ptr(gT), ldt, None, 0, st)
            if n_snps + 16 > self.stage_cols:  # larger than the staging buffer: its own launch
                self.flush()
                ldp = round_up(n_snps, 256) // 4
                op = torch.zeros((self.n_pad, ldp), dtype=torch.uint8, device=self.device)
                _lib.call("jm_pack_gram_2bit", ptr(g), ld_g, n_snps, n, ptr(op), ldp, 0, st)
                self._syrk(op, ldp, n_snps)
                del op
            else:
                col0 = round_up(self._fill, 16)
                if col0 + n_snps > self.stage_cols:
                    self.flush()
                    col0 = 0
                if self._stage is None:  # padded rows stay zero
                    self._stage = torch.zeros((self.n_pad, self._stage_ld()), dtype=torch.uint8,
                                              device=self.device)
                _lib.call("jm_pack_gram_2bit", ptr(g), ld_g, n_snps, n, ptr(self._stage),
                          self._stage_ld(), col0, st)
                self._fill = col0 + n_snps
        elif n_snps > self.stage_cols:  # larger than the staging buffer: its own launch
            self.flush()
            gT0 = torch.zeros((self.n_pad, ldt), dtype=torch.int8, device=self.device)
            _lib.call("jm_transpose_i8", ptr(g), n_snps, n, ld_g, ptr(gT), ldt, ptr(gT0), ldt, st)
            self._syrk(gT0, ldt, n_snps)
            del gT0
        else:
            if self._fill + n_snps > self.stage_cols:
                self.flush()
            if self._stage is None:  # padded rows stay zero
                self._stage = torch.zeros((self.n_pad, self.stage_cols), dtype=torch.int8,
                                          device=self.device)
            dst = self._stage[:, self._fill:]
            _lib.call("jm_transpose_i8", ptr(g), n_snps, n, ld_g, ptr(gT), ldt, ptr(dst),
                      self.stage_cols, st)
            self._fill += n_snps
        mu_d = to_dev(np.asarray(mu, dtype=np.float64), torch.float64, self.device)
        if has_missing and self.C is None:
            self.C = torch.zeros((n, n), dtype=torch.float64, device=self.device)
        ws, ws_len = None, 0
        if has_missing:
            if n_missing is None:
                n_missing = int((g[:n_snps, :n] < 0).sum())
            ws_len = int(_lib.lib.jm_kinship_miss_ws_len(n, n_snps, int(n_missing)))
            ws = torch.empty((ws_len,), dtype=torch.int32, device=self.device)
        _lib.call("jm_kinship_impute_terms", ptr(g), ld_g, ptr(gT), ldt, n, n_snps, ptr(mu_d),
                  ptr(self.u), ptr(self.C) if has_missing else None, int(has_missing), ptr(ws),
                  ws_len, int(n_missing or 0), st)
        for m in np.asarray(mu, dtype=np.float64):
            self.s0 += m * m
        self.p += n_snps

    def finalize(self, p: float | None = None, sub: "KinshipAccumulator | None" = None) -> torch.Tensor:
        """K = (S - u 1' - 1 u' + s0 + C + C') / p, optionally minus ``sub`` (LOCO)."""
        p = float(self.p if p is None else p)
        self.release()
        if sub is not None:
            sub.release()
        K = torch.empty((self.n, self.n), dtype=torch.float64, device=self.device)
        _lib.call("jm_kinship_finalize", ptr(self.S), ptr(self.u), ptr(self.C), self.s0,
                  ptr(sub.S) if sub else None, ptr(sub.u) if sub else None,
                  ptr(sub.C) if sub else None, sub.s0 if sub else 0.0, self.n, p, ptr(K),
                  stream_ptr())
        return K


def _mu64(counts: np.ndarray) -> np.ndarray:
    """float64 nanmean of integer calls (kinship/missing.py:48-52): sum / count, 0 if none."""
    mean, _, _ = stats_from_counts(counts, np.float64)
    return mean


def _filter(counts, n, maf_threshold, miss_threshold, stat_dtype):
    mean, miss, poly = stats_from_counts(counts, stat_dtype)
    af = mean / 2.0
    maf = np.minimum(af, 1.0 - af)
    return (maf >= maf_threshold) & (miss / n <= miss_threshold) & poly


def _integral_i8(genotypes: np.ndarray) -> np.ndarray:
    g = np.asarray(genotypes)
    nanm = np.isnan(g)
    vals = np.where(nanm, 0, g)
    if not np.all((vals == 0) | (vals == 1) | (vals == 2)):
        raise ValueError("genotypes must be hard calls 0/1/2 (NaN = missing)")
    return np.where(nanm, -1, vals).astype(np.int8).T.copy()  # SNP-major


def _counts(gi8: np.ndarray) -> np.ndarray:
    return np.stack([(gi8 == 0).sum(1), (gi8 == 1).sum(1), (gi8 == 2).sum(1), (gi8 < 0).sum(1)],
                    axis=1).astype(np.int64)


def is_hard_calls(genotypes: np.ndarray) -> bool:
    """True when every non-NaN value is 0, 1 or 2 (the int8 path applies)."""
    g = np.asarray(genotypes)
    v = g[~np.isnan(g)]
    return bool(np.all((v == 0) | (v == 1) | (v == 2)))


def snp_stats_dense(genotypes: np.ndarray):
    """compute_snp_stats (core/snp_filter.py:36-56) on a dense matrix with NaN: per-SNP
    nanmean, missing count and nanvar in the input's dtype, NaN -> 0."""
    g = np.asarray(genotypes)
    miss = np.sum(np.isnan(g), axis=0)
    with np.errstate(invalid="ignore"), warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        mean = np.nanmean(g, axis=0)
        var = np.nanvar(g, axis=0)
    return (np.nan_to_num(mean, nan=0.0).astype(np.float64), miss.astype(np.int64),
            np.nan_to_num(var, nan=0.0).astype(np.float64))


def dosage_kinship_device(genotypes: np.ndarray, batch_size: int = 10000,
                          maf_threshold: float = 0.0, miss_threshold: float = 1.0) -> torch.Tensor:
    """compute_centered_kinship for non-integer genotypes (kinship/compute.py:93-226):
    the reference's filter on nanmean / missing count / nanvar, then per batch of SNPs
    impute_and_center on the device (jm_center_columns_f64) and K += Xc Xc^T on FP64
    MFMA (jm_dgemm), K /= p."""
    g = np.asarray(genotypes)
    n, p0 = g.shape
    mean, miss, var = snp_stats_dense(g)
    af = mean / 2.0
    maf = np.minimum(af, 1.0 - af)
    idx = np.where((maf >= maf_threshold) & (miss / n <= miss_threshold) & (var > 0))[0]
    if len(idx) == 0:
        raise ValueError(f"No SNPs passed filtering (maf>={maf_threshold}, miss<={miss_threshold},"
                         f" polymorphic). Original SNP count: {p0}")
    dev = require_gpu()
    K = torch.zeros((n, n), dtype=torch.float64, device=dev)
    st = stream_ptr()
    for s in range(0, len(idx), batch_size):
        sel = idx[s:s + batch_size]
        b = len(sel)
        X = to_dev(np.ascontiguousarray(g[:, sel], dtype=np.float64), torch.float64, dev)
        _lib.call("jm_center_columns_f64", ptr(X), n, b, b, None, st)
        # column-major view: X is (b x n) with ld b; K (n x n) += X^T X
        _lib.call("jm_dgemm", 1, 0, n, n, b, 1.0, ptr(X), b, ptr(X), b, 1.0, ptr(K), n, 0, st)
        del X
    K /= float(len(idx))
    return K


def centered_kinship_device(genotypes: np.ndarray, batch_size: int = 10000,
                            maf_threshold: float = 0.0, miss_threshold: float = 1.0) -> torch.Tensor:
    """compute_centered_kinship on the device -> device float64 K. Hard calls (0/1/2)
    take the int8 path; any other values the FP64 dosage path."""
    g = np.asarray(genotypes)
    if not is_hard_calls(g):
        return dosage_kinship_device(g, batch_size, maf_threshold, miss_threshold)
    n = g.shape[0]
    gi8 = _integral_i8(g)
    counts = _counts(gi8)
    stat_dtype = np.float32 if g.dtype == np.float32 else np.float64
    mask = _filter(counts, n, maf_threshold, miss_threshold, stat_dtype)
    idx = np.where(mask)[0]
    if len(idx) == 0:
        raise ValueError(f"No SNPs passed filtering (maf>={maf_threshold}, miss<={miss_threshold},"
                         f" polymorphic). Original SNP count: {g.shape[1]}")
    mu = _mu64(counts)
    acc = KinshipAccumulator(n)
    ld = round_up(n, 64)
    for s in range(0, len(idx), batch_size):
        sel = idx[s:s + batch_size]
        tile = np.zeros((len(sel), ld), dtype=np.int8)
        tile[:, :n] = gi8[sel]
        acc.add_tile(to_dev(tile, torch.int8), ld, len(sel), mu[sel], bool(counts[sel, 3].any()),
                     int(counts[sel, 3].sum()))
    return acc.finalize(len(idx))


def compute_centered_kinship(genotypes: np.ndarray, batch_size: int = 10000,
                             maf_threshold: float = 0.0, miss_threshold: float = 1.0,
                             check_memory: bool = True) -> np.ndarray:
    """Drop-in for kinship/compute.py:93-226 (returns host float64 K)."""
    del check_memory
    return centered_kinship_device(genotypes, batch_size, maf_threshold, miss_threshold).cpu().numpy()


def _bed_counts(bed: BedFile, chunk_size: int, sample_idx=None, n_out=None) -> np.ndarray:
    n_out = bed.n_samples if n_out is None else n_out
    key = sample_key(sample_idx, bed.n_samples)
    hit = bed.cached_counts(key)
    if hit is not None:
        return hit.copy()
    counts = np.zeros((bed.n_snps, 4), dtype=np.int64)
    for start in range(0, bed.n_snps, chunk_size):
        end = min(start + chunk_size, bed.n_snps)
        raw = bed.window_to_device(start, end)
        _, c, _ = decode(raw, bed.bytes_per_snp, end - start, sample_idx=sample_idx, n_out=n_out,
                         want_g=False)
        counts[start:end] = c[: end - start].cpu().numpy()
    bed.cache_counts(key, counts.copy())
    return counts


def _streaming_selection(bed, chunk_size, maf_threshold, miss_threshold, ksnps_indices):
    counts = _bed_counts(bed, chunk_size)
    mask = _filter(counts, bed.n_samples, maf_threshold, miss_threshold, np.float32)
    if ksnps_indices is not None:
        k = np.asarray(ksnps_indices, dtype=np.int64)
        if len(k) and k.max() >= bed.n_snps:
            raise ValueError(f"Kinship SNP list index {k.max()} out of range for {bed.n_snps} SNPs")
        lm = np.zeros(bed.n_snps, dtype=bool)
        lm[k] = True
        mask &= lm
    idx = np.where(mask)[0]
    if len(idx) == 0:
        raise ValueError(f"No SNPs passed filtering (maf>={maf_threshold}, miss<={miss_threshold}, "
                         f"polymorphic). Original SNP count: {bed.n_snps}")
    return counts, idx


def _accumulate_streaming(bed, idx, counts, chunk_size, acc_for):
    """Pass 2: decode selected SNPs per file chunk and feed acc_for(snp) groups."""
    n = bed.n_samples
    ld = round_up(n, 64)
    mu = _mu64(counts)
    for start in range(0, bed.n_snps, chunk_size):
        end = min(start + chunk_size, bed.n_snps)
        sel = idx[(idx >= start) & (idx < end)]
        if len(sel) == 0:
            continue
        raw = bed.window_to_device(start, end)
        for acc, part in acc_for(sel):
            g, _, _ = decode(raw, bed.bytes_per_snp, len(part), n_out=n,
                             snp_idx=(part - start).astype(np.int32), want_counts=False, ld_g=ld)
            acc.add_tile(g, ld, len(part), mu[part], bool(counts[part, 3].any()),
                         int(counts[part, 3].sum()))


def kinship_streaming_device(bed_path: Path, chunk_size: int = 10_000, maf_threshold: float = 0.0,
                             miss_threshold: float = 1.0, ksnps_indices=None) -> torch.Tensor:
    bed = BedFile(bed_path)
    counts, idx = _streaming_selection(bed, chunk_size, maf_threshold, miss_threshold,
                                       ksnps_indices)
    acc = KinshipAccumulator(bed.n_samples)
    _accumulate_streaming(bed, idx, counts, chunk_size, lambda sel: [(acc, sel)])
    return acc.finalize(len(idx))


def compute_kinship_streaming(bed_path: Path, chunk_size: int = 10_000, maf_threshold: float = 0.0,
                              miss_threshold: float = 1.0, check_memory: bool = True,
                              show_progress: bool = True,
                              ksnps_indices: np.ndarray | None = None) -> np.ndarray:
    """Drop-in for kinship/compute.py:502-690 (returns host float64 K)."""
    del check_memory, show_progress
    return kinship_streaming_device(bed_path, chunk_size, maf_threshold, miss_threshold,
                                    ksnps_indices).cpu().numpy()


def loco_kinship_streaming_device(bed_path: Path, chunk_size: int = 10_000,
                                  maf_threshold: float = 0.0, miss_threshold: float = 1.0,
                                  ksnps_indices=None) -> Iterator[tuple[str, torch.Tensor]]:
    """LOCO kinships K_c = (S_full - S_c) / (p - p_c) (kinship/compute.py:693-916),
    yielded one chromosome at a time in sorted order.

    The full accumulators (int32 S, FP64 u and C) are built in one pass; each
    chromosome's partial Gram is then recomputed from its own SNPs, subtracted and
    freed before the next, so at most two sets of n x n accumulators are live (at
    n = 20,000: 2 x (1.6 GB int32 + 3.2 GB FP64) instead of one set per chromosome).
    The extra work is one more int8 GEMM over the SNPs, the cheapest stage."""
    bed = BedFile(bed_path)
    chrom = get_plink_metadata(bed_path)["chromosome"]
    counts, idx = _streaming_selection(bed, chunk_size, maf_threshold, miss_threshold,
                                       ksnps_indices)
    names = sorted(set(chrom[idx].tolist()))
    full = KinshipAccumulator(bed.n_samples, max_snps=len(idx))
    _accumulate_streaming(bed, idx, counts, chunk_size, lambda sel: [(full, sel)])
    full.release()
    for c in names:
        sel_c = idx[chrom[idx] == c]
        p_loco = full.p - len(sel_c)
        if p_loco == 0:
            raise ValueError(f"Cannot compute LOCO kinship: all {full.p} filtered SNPs are on "
                             f"chromosome '{c}'.")
        acc = KinshipAccumulator(bed.n_samples, max_snps=len(sel_c))
        _accumulate_streaming(bed, sel_c, counts, chunk_size, lambda sel: [(acc, sel)])
        K = full.finalize(p_loco, sub=acc)
        del acc
        yield c, K
        del K


def compute_loco_kinship_streaming(bed_path: Path, chunk_size: int = 10_000,
                                   maf_threshold: float = 0.0, miss_threshold: float = 1.0,
                                   check_memory: bool = True, show_progress: bool = True,
                                   ksnps_indices: np.ndarray | None = None
                                   ) -> Iterator[tuple[str, np.ndarray]]:
    """Drop-in for kinship/compute.py:693-916: yields (chr, host K_loco)."""
    del check_memory, show_progress
    for c, K in loco_kinship_streaming_device(bed_path, chunk_size, maf_threshold, miss_threshold,
                                              ksnps_indices):
        yield c, K.cpu().numpy()


def compute_loco_kinship(genotypes: np.ndarray, chromosome_for_each_snp: np.ndarray,
                         batch_size: int = 10000, maf_threshold: float = 0.0,
                         miss_threshold: float = 1.0, check_memory: bool = True
                         ) -> Iterator[tuple[str, np.ndarray]]:
    """Drop-in for kinship/compute.py:363-499 (in-memory LOCO)."""
    del check_memory
    g = np.asarray(genotypes)
    n = g.shape[0]
    chrom = np.asarray(chromosome_for_each_snp)
    gi8 = _integral_i8(g)
    counts = _counts(gi8)
    stat_dtype = np.float32 if g.dtype == np.float32 else np.float64
    mask = _filter(counts, n, maf_threshold, miss_threshold, stat_dtype)
    idx = np.where(mask)[0]
    if len(idx) == 0:
        raise ValueError("No SNPs passed filtering")
    mu = _mu64(counts)
    ld = round_up(n, 64)
    names = sorted(set(chrom[idx].tolist()))

    def accumulate(sel_all):
        acc = KinshipAccumulator(n, max_snps=len(sel_all))
        for s in range(0, len(sel_all), batch_size):
            sel = sel_all[s:s + batch_size]
            tile = np.zeros((len(sel), ld), dtype=np.int8)
            tile[:, :n] = gi8[sel]
            acc.add_tile(to_dev(tile, torch.int8), ld, len(sel), mu[sel],
                         bool(counts[sel, 3].any()), int(counts[sel, 3].sum()))
        acc.release()
        return acc

    # the full accumulators once, then each chromosome's partial Gram on its own
    # (at most two sets of n x n accumulators live, as in the streaming form)
    full = accumulate(idx)
    for c in names:
        sel_c = idx[chrom[idx] == c]
        if full.p - len(sel_c) == 0:
            raise ValueError(f"Cannot compute LOCO kinship: all {full.p} filtered SNPs are on "
                             f"chromosome '{c}'. LOCO requires SNPs on multiple chromosomes.")
        acc = accumulate(sel_c)
        K = full.finalize(full.p - len(sel_c), sub=acc).cpu().numpy()
        del acc
        yield c, K
