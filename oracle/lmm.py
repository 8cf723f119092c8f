"""LMM association oracle (test infrastructure only).

NumPy/SciPy FP64 restatement of the reference's per-SNP LMM math, vectorised
over SNPs. Reference files (relative to the reference repository root):
  * get_ab_index                 src/jamma/lmm/likelihood.py:25-43
  * build_index_table            src/jamma/lmm/likelihood_jax.py:40-109
  * Uab (products of rotated vectors)  likelihood_jax.py:112-145, :342-393
  * calc_pab (row 0 + recursion) likelihood_jax.py:148-186
  * REML with precomputed Iab    likelihood_jax.py:280-339
  * MLE                          likelihood_jax.py:189-228
  * grid + golden refinement     likelihood_jax.py:419-562, :565-598, :801-883
  * Wald / Score / LRT           likelihood_jax.py:601-675, :706-773, :231-251
  * null model MLE (host golden) likelihood.py:349-392, :395-474, :571-607
  * runner semantics             lmm/runner_streaming.py:188-566,
                                 lmm/prepare.py:41-191, lmm/eigen.py:28-128

Row 0 of Pab (Hi^T Uab) is formed as Hi-weighted dot products of the rotated
vectors instead of materialising Uab; this is the same sum in a different
order (differences at the 1e-16 relative level).
"""

from __future__ import annotations

import math
import warnings

import numpy as np
from scipy import special, stats

PHI = 0.6180339887498949  # likelihood_jax.py:449
P_YY_MIN = 1e-8           # likelihood_jax.py:333


# --------------------------------------------------------------------------
# Index tables
# --------------------------------------------------------------------------
def get_ab_index(a: int, b: int, n_cvt: int) -> int:
    """likelihood.py:25-43 (1-based a, b; packed upper triangle)."""
    cols = n_cvt + 2
    a1, b1 = (a, b) if a <= b else (b, a)
    return (2 * cols - a1 + 2) * (a1 - 1) // 2 + b1 - a1


def index_table(n_cvt: int) -> dict:
    """likelihood_jax.py:40-109."""
    n_index = (n_cvt + 3) * (n_cvt + 2) // 2
    pairs = []
    for a in range(1, n_cvt + 3):
        for b in range(a, n_cvt + 3):
            pairs.append((a - 1, b - 1, get_ab_index(a, b, n_cvt)))
    rec = {}
    for p in range(1, n_cvt + 2):
        rec[p] = [(get_ab_index(a, b, n_cvt), get_ab_index(a, p, n_cvt),
                   get_ab_index(b, p, n_cvt), get_ab_index(p, p, n_cvt))
                  for a in range(p + 1, n_cvt + 3) for b in range(a, n_cvt + 3)]
    return {
        "n_index": n_index,
        "idx_yy": get_ab_index(n_cvt + 2, n_cvt + 2, n_cvt),
        "idx_xx": get_ab_index(n_cvt + 1, n_cvt + 1, n_cvt),
        "idx_xy": get_ab_index(n_cvt + 2, n_cvt + 1, n_cvt),
        "pairs": pairs,
        "rec": rec,
        "logdet_diag": [(i, get_ab_index(i + 1, i + 1, n_cvt)) for i in range(n_cvt + 1)],
    }


# --------------------------------------------------------------------------
# Pab
# --------------------------------------------------------------------------
def pab_batch(n_cvt, Hi, UtW, Uty, Utx):
    """Pab for a batch of SNPs.

    Hi: (n,) shared weights or (s, n) per-SNP weights. Utx: (n, s) rotated
    genotypes, or None for the null model (genotype columns are zero,
    likelihood.py:80-142). Returns (s, n_cvt+2, n_index); s = 1 when Utx is
    None and Hi is shared.
    """
    t = index_table(n_cvt)
    ni = t["n_index"]
    UtW = UtW.reshape(UtW.shape[0], -1)
    Hi2 = Hi if Hi.ndim == 2 else Hi[None, :]
    if Utx is None:
        s = Hi2.shape[0]
        X = None
    else:
        s = Utx.shape[1] if Hi.ndim == 1 else Hi2.shape[0]
        X = Utx.T  # (s, n)
    vecs = [UtW[:, k] for k in range(n_cvt)] + [None, Uty]
    xcol = n_cvt
    P = np.zeros((s, n_cvt + 2, ni))
    HX = None if X is None else Hi2 * X  # (s, n)
    for a, b, idx in t["pairs"]:
        if a == xcol or b == xcol:
            if X is None:
                continue
            if a == xcol and b == xcol:
                P[:, 0, idx] = np.sum(HX * X, axis=1)
            else:
                other = vecs[b] if a == xcol else vecs[a]
                P[:, 0, idx] = HX @ other
        else:
            P[:, 0, idx] = Hi2 @ (vecs[a] * vecs[b])
    # recursion, likelihood_jax.py:176-184
    for p in range(1, n_cvt + 2):
        for i_ab, i_aw, i_bw, i_ww in t["rec"][p]:
            ps_ww = P[:, p - 1, i_ww]
            with np.errstate(divide="ignore", invalid="ignore"):
                inv = np.where(ps_ww != 0, 1.0 / ps_ww, 0.0)
            P[:, p, i_ab] = P[:, p - 1, i_ab] - P[:, p - 1, i_aw] * P[:, p - 1, i_bw] * inv
    return P


def _clamp_pyy(v):
    return np.where((v >= 0.0) & (v < P_YY_MIN), P_YY_MIN, v)


def reml_batch(n_cvt, lam, d, UtW, Uty, Utx, Iab):
    """likelihood_jax.py:280-339; lam scalar or (s,)."""
    n = d.shape[0]
    df = n - n_cvt - 1
    t = index_table(n_cvt)
    lam = np.asarray(lam, dtype=np.float64)
    v = lam[..., None] * d + 1.0 if lam.ndim else lam * d + 1.0
    with np.errstate(divide="ignore"):
        logdet_h = np.sum(np.log(np.abs(v)), axis=-1)
    P = pab_batch(n_cvt, 1.0 / v, UtW, Uty, Utx)
    logdet_hiw = 0.0
    with np.errstate(divide="ignore", invalid="ignore"):
        for row, col in t["logdet_diag"]:
            dp = P[:, row, col]
            di = Iab[:, row, col]
            logdet_hiw = logdet_hiw + np.where(dp > 0, np.log(np.where(dp > 0, dp, 1.0)), 0.0)
            logdet_hiw = logdet_hiw - np.where(di > 0, np.log(np.where(di > 0, di, 1.0)), 0.0)
        pyy = _clamp_pyy(P[:, n_cvt + 1, t["idx_yy"]])
        c = 0.5 * df * (np.log(df) - np.log(2 * np.pi) - 1.0)
        return c - 0.5 * logdet_h - 0.5 * logdet_hiw - 0.5 * df * np.log(pyy)


def mle_batch(n_cvt, lam, d, UtW, Uty, Utx):
    """likelihood_jax.py:189-228."""
    n = d.shape[0]
    t = index_table(n_cvt)
    lam = np.asarray(lam, dtype=np.float64)
    v = lam[..., None] * d + 1.0 if lam.ndim else lam * d + 1.0
    with np.errstate(divide="ignore"):
        logdet_h = np.sum(np.log(np.abs(v)), axis=-1)
    P = pab_batch(n_cvt, 1.0 / v, UtW, Uty, Utx)
    pyy = _clamp_pyy(P[:, n_cvt + 1, t["idx_yy"]])
    c = 0.5 * n * (np.log(n) - np.log(2 * np.pi) - 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        return c - 0.5 * logdet_h - 0.5 * n * np.log(pyy)


# --------------------------------------------------------------------------
# Lambda optimisation (batched JAX semantics)
# --------------------------------------------------------------------------
def grid_log_lambdas(l_min, l_max, n_grid):
    """likelihood_jax.py:540-543."""
    return np.linspace(np.log(l_min), np.log(l_max), n_grid)


def golden_refine(f, all_logls, log_lambdas, n_grid, n_iter):
    """likelihood_jax.py:419-488. f(log_lams (s,)) -> logls (s,)."""
    best = np.argmax(all_logls, axis=0)
    lo = np.maximum(best - 1, 0)
    hi = np.minimum(best + 1, n_grid - 1)
    a = log_lambdas[lo]
    b = log_lambdas[hi]
    c = b - PHI * (b - a)
    d = a + PHI * (b - a)
    fc = f(c)
    fd = f(d)
    for _ in range(n_iter):
        keep_left = fc > fd
        na = np.where(keep_left, a, c)
        nb = np.where(keep_left, d, b)
        nc = nb - PHI * (nb - na)
        nd = na + PHI * (nb - na)
        new = f(np.where(keep_left, nc, nd))
        nfc = np.where(keep_left, new, fd)
        nfd = np.where(keep_left, fc, new)
        a, b, c, d, fc, fd = na, nb, nc, nd, nfc, nfd
    log_opt = (a + b) / 2
    return np.exp(log_opt), f(log_opt)


def optimize_reml(n_cvt, d, UtW, Uty, Utx, l_min=1e-5, l_max=1e5, n_grid=50, n_iter=20,
                  return_evals=False):
    """golden_section_optimize_lambda, likelihood_jax.py:491-562."""
    ones = np.ones(d.shape[0])
    Iab = pab_batch(n_cvt, ones, UtW, Uty, Utx)
    ll = grid_log_lambdas(l_min, l_max, n_grid)
    lams = np.exp(ll)
    grid = np.stack([reml_batch(n_cvt, lam, d, UtW, Uty, Utx, Iab) for lam in lams])

    def f(log_lams):
        return reml_batch(n_cvt, np.exp(log_lams), d, UtW, Uty, Utx, Iab)

    lam, logl = golden_refine(f, grid, ll, n_grid, n_iter)
    return (lam, logl, grid) if return_evals else (lam, logl)


def optimize_mle(n_cvt, d, UtW, Uty, Utx, l_min=1e-5, l_max=1e5, n_grid=50, n_iter=20,
                 return_evals=False):
    """golden_section_optimize_lambda_mle, likelihood_jax.py:830-883."""
    ll = grid_log_lambdas(l_min, l_max, n_grid)
    lams = np.exp(ll)
    grid = np.stack([mle_batch(n_cvt, lam, d, UtW, Uty, Utx) for lam in lams])

    def f(log_lams):
        return mle_batch(n_cvt, np.exp(log_lams), d, UtW, Uty, Utx)

    lam, logl = golden_refine(f, grid, ll, n_grid, n_iter)
    return (lam, logl, grid) if return_evals else (lam, logl)


# --------------------------------------------------------------------------
# Test statistics
# --------------------------------------------------------------------------
def _f_pvalue(df, f_stat, valid):
    """F(1, df) survival through betainc, likelihood_jax.py:668-673."""
    with np.errstate(divide="ignore", invalid="ignore"):
        z = df / (df + np.maximum(f_stat, 1e-10))
    z = np.clip(z, 0.0, 1.0)
    p = special.betainc(df / 2.0, 0.5, z)
    p = np.where(f_stat <= 0, 1.0, p)
    return np.where(valid, p, np.nan)


def _beta_se(df, P_xx, P_xy, Px_yy, valid):
    with np.errstate(divide="ignore", invalid="ignore"):
        safe_xx = np.where(valid, P_xx, 1.0)
        beta = np.where(valid, P_xy / safe_xx, np.nan)
        tau = df / Px_yy
        var_beta = np.where(valid, 1.0 / (tau * safe_xx), np.nan)
        var_safe = np.where(np.abs(var_beta) < 0.001, np.abs(var_beta), var_beta)
        se = np.where(valid, np.sqrt(var_safe), np.nan)
    return beta, se, tau


def wald_batch(n_cvt, lam, d, UtW, Uty, Utx):
    """calc_wald_stats_jax, likelihood_jax.py:601-675."""
    n = d.shape[0]
    df = n - n_cvt - 1
    t = index_table(n_cvt)
    lam = np.asarray(lam, dtype=np.float64)
    Hi = 1.0 / (lam[:, None] * d + 1.0)
    P = pab_batch(n_cvt, Hi, UtW, Uty, Utx)
    P_xx = P[:, n_cvt, t["idx_xx"]]
    P_xy = P[:, n_cvt, t["idx_xy"]]
    P_yy = P[:, n_cvt, t["idx_yy"]]
    Px_yy = _clamp_pyy(P[:, n_cvt + 1, t["idx_yy"]])
    valid = P_xx > 0
    beta, se, tau = _beta_se(df, P_xx, P_xy, Px_yy, valid)
    with np.errstate(invalid="ignore"):
        f_stat = (P_yy - Px_yy) * tau
    return beta, se, _f_pvalue(df, f_stat, valid)


def score_batch(n_cvt, Hi_null, UtW, Uty, Utx):
    """batch_calc_score_stats / calc_score_stats_jax, likelihood_jax.py:706-798."""
    n = Hi_null.shape[0]
    df = n - n_cvt - 1
    t = index_table(n_cvt)
    P = pab_batch(n_cvt, Hi_null, UtW, Uty, Utx)
    P_yy = P[:, n_cvt, t["idx_yy"]]
    P_xx = P[:, n_cvt, t["idx_xx"]]
    P_xy = P[:, n_cvt, t["idx_xy"]]
    Px_yy = _clamp_pyy(P[:, n_cvt + 1, t["idx_yy"]])
    valid = P_xx > 0
    beta, se, _ = _beta_se(df, P_xx, P_xy, Px_yy, valid)
    with np.errstate(divide="ignore", invalid="ignore"):
        f_stat = n * (P_xy * P_xy) / (P_yy * np.where(valid, P_xx, 1.0))
    return beta, se, _f_pvalue(df, f_stat, valid)


def lrt_pvalue(logl_h1, logl_h0):
    """calc_lrt_pvalue_jax, likelihood_jax.py:231-251."""
    stat = np.maximum(2.0 * (np.asarray(logl_h1) - logl_h0), 0.0)
    return stats.chi2.sf(stat, df=1)


# --------------------------------------------------------------------------
# Null model (host scalar golden section)
# --------------------------------------------------------------------------
def golden_section_minimize(func, l_min=1e-5, l_max=1e5, n_grid=50, n_iter=20):
    """likelihood.py:395-474."""
    log_l_min = math.log(l_min)
    log_l_max = math.log(l_max)
    step = (log_l_max - log_l_min) / (n_grid - 1)
    log_lambdas = [log_l_min + i * step for i in range(n_grid)]
    best_idx = 0
    best_val = func(math.exp(log_lambdas[0]))
    for i in range(1, n_grid):
        val = func(math.exp(log_lambdas[i]))
        if val < best_val:
            best_val = val
            best_idx = i
    a = log_lambdas[max(best_idx - 1, 0)]
    b = log_lambdas[min(best_idx + 1, n_grid - 1)]
    c = b - PHI * (b - a)
    d = a + PHI * (b - a)
    fc = func(math.exp(c))
    fd = func(math.exp(d))
    for _ in range(n_iter):
        if fc < fd:
            b, d, fd = d, c, fc
            c = b - PHI * (b - a)
            fc = func(math.exp(c))
        else:
            a, c, fc = c, d, fd
            d = a + PHI * (b - a)
            fd = func(math.exp(d))
    log_opt = (a + b) / 2.0
    opt_lambda = math.exp(log_opt)
    return opt_lambda, -func(opt_lambda)


def mle_null(lam, d, UtW, Uty, n_cvt):
    """mle_log_likelihood_null, likelihood.py:349-392."""
    n = d.shape[0]
    t = index_table(n_cvt)
    v = lam * d + 1.0
    logdet_h = float(np.sum(np.log(np.abs(v))))
    P = pab_batch(n_cvt, 1.0 / v, UtW, Uty, None)[0]
    pyy = P[n_cvt, t["idx_yy"]]
    if 0.0 <= pyy < P_YY_MIN:
        pyy = P_YY_MIN
    c = 0.5 * n * (np.log(n) - np.log(2 * np.pi) - 1.0)
    return c - 0.5 * logdet_h - 0.5 * n * np.log(pyy)


def reml_null(lam, d, UtW, Uty, n_cvt):
    """reml_log_likelihood_null, likelihood.py:292-346."""
    n = d.shape[0]
    df = n - n_cvt
    t = index_table(n_cvt)
    v = lam * d + 1.0
    logdet_h = float(np.sum(np.log(np.abs(v))))
    P = pab_batch(n_cvt, 1.0 / v, UtW, Uty, None)[0]
    I = pab_batch(n_cvt, np.ones(n), UtW, Uty, None)[0]
    logdet_hiw = 0.0
    for i in range(n_cvt):
        idx = get_ab_index(i + 1, i + 1, n_cvt)
        if P[i, idx] > 0:
            logdet_hiw += np.log(P[i, idx])
        if I[i, idx] > 0:
            logdet_hiw -= np.log(I[i, idx])
    pyy = P[n_cvt, t["idx_yy"]]
    if 0.0 <= pyy < P_YY_MIN:
        pyy = P_YY_MIN
    c = 0.5 * df * (np.log(df) - np.log(2 * np.pi) - 1.0)
    return c - 0.5 * logdet_h - 0.5 * logdet_hiw - 0.5 * df * np.log(pyy)


def null_model_mle(d, UtW, Uty, n_cvt, l_min=1e-5, l_max=1e5):
    """compute_null_model_mle, likelihood.py:571-607 -> (lambda0, logl_H0)."""
    return golden_section_minimize(lambda lam: -mle_null(lam, d, UtW, Uty, n_cvt), l_min, l_max)


def null_model_reml(d, UtW, Uty, n_cvt, l_min=1e-5, l_max=1e5):
    """compute_null_model_lambda, likelihood.py:477-514."""
    return golden_section_minimize(lambda lam: -reml_null(lam, d, UtW, Uty, n_cvt), l_min, l_max)


# --------------------------------------------------------------------------
# Eigendecomposition and runner
# --------------------------------------------------------------------------
def eigendecompose(K, threshold=1e-10):
    """eigen.py:28-128 (numpy eigh; |d| < 1e-10 -> 0; ascending)."""
    d, U = np.linalg.eigh(K)
    d = np.where(np.abs(d) < threshold, 0.0, d)
    return d, U


def assoc_chunk(mode, n_cvt, d, UtW, Uty, Utx, Hi_null=None, logl_H0=None,
                l_min=1e-5, l_max=1e5, n_grid=50, n_refine=10):
    """The device compute of one chunk, runner_streaming.py:471-554."""
    n_iter = max(n_refine, 20)
    out = {}
    n = d.shape[0]
    if mode in (3, 4):
        b, s, p = score_batch(n_cvt, Hi_null, UtW, Uty, Utx)
        out["p_score"] = p
        if mode == 3:
            out["beta"], out["se"] = b, s
    if mode in (2, 4):
        lam_m, ll_m = optimize_mle(n_cvt, d, UtW, Uty, Utx, l_min, l_max, n_grid, n_iter)
        out["l_mle"] = lam_m
        out["logl_mle"] = ll_m
        out["p_lrt"] = lrt_pvalue(ll_m, logl_H0)
    if mode in (1, 4):
        lam, ll = optimize_reml(n_cvt, d, UtW, Uty, Utx, l_min, l_max, n_grid, n_iter)
        b, s, p = wald_batch(n_cvt, lam, d, UtW, Uty, Utx)
        out.update(l_remle=lam, logl_H1=ll, beta=b, se=s, p_wald=p)
    del n
    return out


def snp_pass1(genotypes_valid, dtype=np.float32):
    """Pass-1 statistics on a ``dtype`` chunk (float32 in runner_streaming.py:248-263,
    float64 in the LOCO per-chromosome runner, loco.py:305-324)."""
    g32 = np.asarray(genotypes_valid, dtype=dtype)
    miss = np.sum(np.isnan(g32), axis=0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        means = np.nanmean(g32, axis=0)
        var = np.nanvar(g32, axis=0)
    means = np.nan_to_num(means, nan=0.0).astype(np.float64)
    var = np.nan_to_num(var, nan=0.0).astype(np.float64)
    return means, miss.astype(np.int64), var


def run_lmm(genotypes, phenotypes, kinship=None, covariates=None, eigenvalues=None,
            eigenvectors=None, maf_threshold=0.01, miss_threshold=0.05, l_min=1e-5,
            l_max=1e5, n_grid=50, n_refine=10, lmm_mode=1, chunk=512, stat_dtype=np.float32):
    """run_lmm_association_streaming semantics (runner_streaming.py:94-628) on an
    in-memory (n_samples, n_snps) genotype array with NaN for missing.

    Returns a dict with ``snp_indices``, ``af``, ``n_miss`` and per-mode arrays.
    """
    phenotypes = np.asarray(phenotypes, dtype=np.float64)
    valid = ~np.isnan(phenotypes) & (phenotypes != -9.0)
    if covariates is not None:
        valid &= np.all(~np.isnan(covariates), axis=1)
    y = phenotypes[valid]
    if kinship is not None:
        kinship = kinship[np.ix_(valid, valid)]
    if covariates is not None:
        covariates = covariates[valid, :]
    G = genotypes[valid, :]
    n = y.shape[0]
    means, miss, var = snp_pass1(G, stat_dtype)
    af = means / 2.0
    maf = np.minimum(af, 1.0 - af)
    mask = (maf >= maf_threshold) & (miss / n <= miss_threshold) & (var > 0)
    idx = np.where(mask)[0]
    res = {"snp_indices": idx, "af": af[idx], "n_miss": miss[idx], "n_samples": n}
    if len(idx) == 0:
        return res
    if eigenvalues is None:
        d, U = eigendecompose(kinship)
    else:
        d, U = eigenvalues, eigenvectors
    W = np.ones((n, 1)) if covariates is None else covariates.astype(np.float64)
    n_cvt = W.shape[1]
    if n <= n_cvt + 1:
        raise ValueError("Over-parameterized model")
    UT = np.ascontiguousarray(U.T)
    UtW = UT @ W
    Uty = UT @ y
    logl_H0 = lam0 = Hi_null = None
    if lmm_mode in (2, 3, 4):
        lam0, logl_H0 = null_model_mle(d, UtW, Uty, n_cvt, l_min, l_max)
        Hi_null = 1.0 / (lam0 * d + 1.0)
    res.update(lambda_null_mle=lam0, logl_H0=logl_H0, eigenvalues=d, n_cvt=n_cvt)
    Gf = G[:, idx]
    fm = means[idx]
    Gf = np.where(np.isnan(Gf), fm[None, :], Gf)
    parts = []
    for s in range(0, len(idx), chunk):
        Utx = UT @ Gf[:, s:s + chunk]
        parts.append(assoc_chunk(lmm_mode, n_cvt, d, UtW, Uty, Utx, Hi_null, logl_H0,
                                 l_min, l_max, n_grid, n_refine))
    for k in parts[0]:
        res[k] = np.concatenate([pp[k] for pp in parts])
    return res


def run_lmm_loco(genotypes, chromosomes, phenotypes, covariates=None, maf_threshold=0.01,
                 miss_threshold=0.05, lmm_mode=1):
    """run_lmm_loco semantics (loco.py:54-593): LOCO kinships over all samples with the
    LMM thresholds, chromosomes in sorted order, float64 per-chromosome statistics.
    Returns {chr: run_lmm result dict (snp_indices relative to that chromosome)}."""
    from oracle import kinship as ok

    chromosomes = np.asarray(chromosomes)
    loco = ok.loco_kinship(genotypes, chromosomes, maf_threshold, miss_threshold)
    out = {}
    for c in sorted(set(chromosomes.tolist())):
        if c not in loco:
            continue
        sel = np.where(chromosomes == c)[0]
        r = run_lmm(genotypes[:, sel], phenotypes, kinship=loco[c], covariates=covariates,
                    maf_threshold=maf_threshold, miss_threshold=miss_threshold,
                    lmm_mode=lmm_mode, stat_dtype=np.float64)
        r["global_indices"] = sel[r["snp_indices"]]
        out[c] = r
    return out
