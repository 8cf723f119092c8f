"""Centred kinship for the oracle (test infrastructure only).

Restates src/jamma/kinship/compute.py and src/jamma/kinship/missing.py of the
reference:
  * SNP stats / filter: core/snp_filter.py:36-91 (compute_snp_stats,
    compute_snp_filter_mask);
  * impute_and_center: kinship/missing.py:22-61 (float64 nanmean over all rows
    of the batch, NaN -> mean, subtract mean; all-missing column -> mean 0);
  * compute_centered_kinship: kinship/compute.py:93-226 (filter, batches of
    ``batch_size`` SNPs, K += Xc Xc^T, K / n_filtered);
  * compute_kinship_streaming: kinship/compute.py:502-690 (pass 1 on float32
    chunks for the var>0 / maf / miss filter, pass 2 on float64);
  * compute_loco_kinship(_streaming): kinship/compute.py:363-499, 693-916
    (global centring, K_c = (S_full - S_c) / (p - p_c), chromosomes sorted).
"""

from __future__ import annotations

import warnings

import numpy as np


def snp_stats(genotypes: np.ndarray):
    """core/snp_filter.py:36-56: (col_means, miss_counts, col_vars), NaN -> 0."""
    miss = np.sum(np.isnan(genotypes), axis=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        means = np.nanmean(genotypes, axis=0)
        var = np.nanvar(genotypes, axis=0)
    return np.nan_to_num(means, nan=0.0), miss, np.nan_to_num(var, nan=0.0)


def snp_filter_mask(means, miss, var, n_samples, maf_threshold, miss_threshold):
    """core/snp_filter.py:59-91."""
    miss_rates = miss / n_samples
    af = means / 2.0
    maf = np.minimum(af, 1.0 - af)
    mask = (maf >= maf_threshold) & (miss_rates <= miss_threshold) & (var > 0)
    return mask, af, maf


def impute_and_center(x: np.ndarray) -> np.ndarray:
    """kinship/missing.py:22-61 (float64)."""
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        mu = np.nanmean(x, axis=0, keepdims=True)
    mu = np.nan_to_num(mu, nan=0.0)
    return np.where(np.isnan(x), mu, x) - mu


def centered_kinship(genotypes, batch_size=10000, maf_threshold=0.0, miss_threshold=1.0):
    """kinship/compute.py:93-226 (stats in the input dtype, usually float64)."""
    n, _ = genotypes.shape
    means, miss, var = snp_stats(genotypes)
    mask, _, _ = snp_filter_mask(means, miss, var, n, maf_threshold, miss_threshold)
    x = np.asarray(genotypes[:, mask], dtype=np.float64)
    p = x.shape[1]
    if p == 0:
        raise ValueError("No SNPs passed filtering")
    K = np.zeros((n, n))
    for s in range(0, p, batch_size):
        xc = impute_and_center(x[:, s:s + batch_size])
        K += xc @ xc.T
    return K / p


def streaming_filter(genotypes, maf_threshold=0.0, miss_threshold=1.0):
    """Pass 1 of kinship/compute.py:586-620: float32 chunk statistics."""
    g32 = np.asarray(genotypes, dtype=np.float32)
    means, miss, var = snp_stats(g32)
    mask, _, _ = snp_filter_mask(means.astype(np.float64), miss, var.astype(np.float64),
                                 genotypes.shape[0], maf_threshold, miss_threshold)
    return mask


def kinship_streaming(genotypes, maf_threshold=0.0, miss_threshold=1.0, chunk_size=10000):
    """kinship/compute.py:502-690 on an in-memory (n, p) array with NaN."""
    n = genotypes.shape[0]
    mask = streaming_filter(genotypes, maf_threshold, miss_threshold)
    idx = np.where(mask)[0]
    if len(idx) == 0:
        raise ValueError("No SNPs passed filtering")
    K = np.zeros((n, n))
    for s in range(0, genotypes.shape[1], chunk_size):
        cols = idx[(idx >= s) & (idx < s + chunk_size)]
        if len(cols):
            xc = impute_and_center(np.asarray(genotypes[:, cols], dtype=np.float64))
            K += xc @ xc.T
    return K / len(idx)


def loco_kinship(genotypes, chromosomes, maf_threshold=0.0, miss_threshold=1.0):
    """kinship/compute.py:693-916 (streaming LOCO): {chr: K_loco}, chr sorted."""
    n = genotypes.shape[0]
    chromosomes = np.asarray(chromosomes)
    mask = streaming_filter(genotypes, maf_threshold, miss_threshold)
    idx = np.where(mask)[0]
    xc = impute_and_center(np.asarray(genotypes[:, idx], dtype=np.float64))
    chr_f = chromosomes[idx]
    s_full = xc @ xc.T
    out = {}
    for c in sorted(set(chr_f.tolist())):
        sel = chr_f == c
        p_loco = len(idx) - int(sel.sum())
        if p_loco == 0:
            raise ValueError("LOCO requires SNPs on multiple chromosomes")
        s_c = xc[:, sel] @ xc[:, sel].T
        out[c] = (s_full - s_c) / p_loco
    return out
