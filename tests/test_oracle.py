"""Pin the CPU oracle before trusting it (CPU only).

Checks, in order of strength:
  * the reference's GEMMA golden fixtures (tests/golden/, copied data files of
    the reference's tests/fixtures/): kinship, null-model log-likelihoods and
    the Wald / LRT / Score / all-tests association files, at the reference's
    own tolerances (src/jamma/validation/tolerances.py:89-96,
    tests/test_lmm_validation.py:32, validation/compare.py:684);
  * the reference's known-answer unit tests (missing-value kinship, GEMMA
    packed index);
  * closed forms: with K = I the likelihood is flat in lambda, Wald reduces to
    the OLS F test, LRT to n log(RSS0/RSS1), Score to n R^2 (partial);
  * brute force: the golden-section optimum against a dense lambda scan;
  * planted mistakes in the oracle must break the fixture comparison.
"""

from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import kinship as ok
from oracle import lmm as ol
from oracle import plink as opl
from tests.util import GOLDEN, fam_phenotypes, load_assoc, synthetic_prefix


@pytest.fixture(scope="module")
def synth():
    pre = synthetic_prefix()
    G = opl.read_bed(pre)
    y = fam_phenotypes(pre)
    K = np.loadtxt(GOLDEN / "gemma_synthetic" / "gemma_kinship.cXX.txt")
    C = np.loadtxt(GOLDEN / "gemma_covariate" / "covariates.txt")
    return G, y, K, C


# ---------------------------------------------------------------- decode / kinship
def test_decode_af_matches_gemma(synth):
    G, _, _, _ = synth
    ref = load_assoc(GOLDEN / "gemma_synthetic" / "gemma_assoc.assoc.txt")
    af = np.nanmean(G, axis=0) / 2
    assert np.max(np.abs(np.round(af, 3) - ref["af"])) < 1e-9


def test_bed_roundtrip_random():
    rng = np.random.default_rng(3)
    G = rng.integers(0, 3, size=(37, 11)).astype(float)
    G[rng.random(G.shape) < 0.1] = np.nan
    raw = opl.encode_bed(G)
    back = opl.decode_bed_bytes(raw[3:], 37, 11)
    assert np.array_equal(np.isnan(back), np.isnan(G))
    assert np.array_equal(np.nan_to_num(back, nan=-1), np.nan_to_num(G, nan=-1))


def test_kinship_matches_gemma_fixture(synth):
    G, _, K, _ = synth
    assert np.max(np.abs(ok.centered_kinship(G) - K)) < 1e-10  # tests/test_gemma_integration.py:79
    assert np.max(np.abs(ok.kinship_streaming(G, chunk_size=77) - K)) < 1e-10


def test_kinship_missing_known_answer():
    # tests/test_kinship_compute.py:112-121: [[0],[NaN],[2]] -> diag [1, 0, 1]
    X = np.array([[0.0], [np.nan], [2.0]])
    K = ok.centered_kinship(X)
    assert np.allclose(np.diag(K), [1.0, 0.0, 1.0])


def test_kinship_closed_form_with_missing():
    rng = np.random.default_rng(5)
    G = rng.integers(0, 3, size=(40, 300)).astype(float)
    G[rng.random(G.shape) < 0.05] = np.nan
    mu = np.nanmean(G, axis=0)
    Xc = np.where(np.isnan(G), 0.0, G - mu)
    keep = np.nanvar(G, axis=0) > 0
    K_ref = Xc[:, keep] @ Xc[:, keep].T / keep.sum()
    assert np.max(np.abs(ok.centered_kinship(G) - K_ref)) < 1e-13


def test_loco_subtraction_identity():
    rng = np.random.default_rng(7)
    G = rng.integers(0, 3, size=(30, 120)).astype(float)
    G[rng.random(G.shape) < 0.02] = np.nan
    chrom = np.array(["1"] * 50 + ["2"] * 40 + ["3"] * 30)
    loco = ok.loco_kinship(G, chrom)
    for c, Kc in loco.items():
        keep = chrom != c
        assert np.max(np.abs(Kc - ok.centered_kinship(G[:, keep]))) < 1e-12


# ---------------------------------------------------------------- LMM vs GEMMA
def test_get_ab_index_known_values():
    # tests/test_lmm_unit.py:478-501 (n_cvt = 1)
    got = [ol.get_ab_index(a, b, 1) for a, b in [(1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]]
    assert got == [0, 1, 2, 3, 4, 5]
    assert ol.get_ab_index(3, 2, 1) == ol.get_ab_index(2, 3, 1)


@pytest.mark.parametrize("cov,reml,mle", [(False, -140.636, -139.281), (True, -139.624, -138.905)])
def test_null_model_loglik_matches_gemma_log(synth, cov, reml, mle):
    _, y, K, C = synth
    d, U = ol.eigendecompose(K)
    W = C if cov else np.ones((100, 1))
    UtW, Uty = U.T @ W, U.T @ y
    _, l_reml = ol.null_model_reml(d, UtW, Uty, W.shape[1])
    _, l_mle = ol.null_model_mle(d, UtW, Uty, W.shape[1])
    assert abs(l_reml - reml) < 6e-4  # GEMMA prints 6 significant digits
    assert abs(l_mle - mle) < 6e-4


def _compare(res, ref, cols, lam_cols=()):
    tol = {"beta": 1e-2, "se": 1e-5, "p_wald": 1e-4, "p_score": 1e-4, "p_lrt": 5e-3,
           "l_remle": 5e-5, "l_mle": 5e-5}
    for k in cols:
        a, e = np.asarray(res[k]), ref[k]
        if k in lam_cols:  # validation/compare.py boundary handling
            m = (a > 1e-4) & (e > 1e-4) & (a < 1e4) & (e < 1e4)
            a, e = a[m], e[m]
        np.testing.assert_allclose(a, e, rtol=tol[k], atol=1e-12, err_msg=k)


@pytest.mark.parametrize("fixture,cov,mode,cols", [
    ("gemma_synthetic/gemma_assoc.assoc.txt", False, 1, ("beta", "se", "p_wald", "l_remle")),
    ("gemma_synthetic/gemma_lrt.assoc.txt", False, 2, ("p_lrt", "l_mle")),
    ("gemma_score/gemma_score.assoc.txt", False, 3, ("p_score",)),
    ("gemma_covariate/gemma_covariate.assoc.txt", True, 1, ("beta", "se", "p_wald", "l_remle")),
    ("gemma_covariate/gemma_covariate_lrt.assoc.txt", True, 2, ("p_lrt", "l_mle")),
    ("gemma_covariate/gemma_covariate_score.assoc.txt", True, 3, ("p_score",)),
    ("gemma_all_tests/gemma_all.assoc.txt", False, 4,
     ("beta", "se", "p_wald", "p_lrt", "p_score", "l_remle", "l_mle")),
    ("gemma_all_tests/gemma_all_covar.assoc.txt", True, 4,
     ("beta", "se", "p_wald", "p_lrt", "p_score", "l_remle", "l_mle")),
])
def test_oracle_matches_gemma(synth, fixture, cov, mode, cols):
    G, y, K, C = synth
    res = ol.run_lmm(G, y, kinship=K, covariates=C if cov else None, lmm_mode=mode)
    if mode == 3:
        res["beta"], res["se"] = res.get("beta"), res.get("se")
    ref = load_assoc(GOLDEN / fixture)
    assert len(res["snp_indices"]) == len(ref["rs"])
    _compare(res, ref, cols, lam_cols=("l_remle", "l_mle"))


def test_score_beta_se_match_gemma(synth):
    G, y, K, _ = synth
    res = ol.run_lmm(G, y, kinship=K, lmm_mode=3)
    ref = load_assoc(GOLDEN / "gemma_score" / "gemma_score.assoc.txt")
    np.testing.assert_allclose(res["beta"], ref["beta"], rtol=1e-2)
    np.testing.assert_allclose(res["se"], ref["se"], rtol=1e-4)


# ---------------------------------------------------------------- closed forms
def _ols_setup(seed=11, n=60, c=3, p=7):
    rng = np.random.default_rng(seed)
    W = np.column_stack([np.ones(n), rng.normal(size=(n, c - 1))])
    X = rng.normal(size=(n, p))
    y = X[:, 0] * 0.5 + W @ rng.normal(size=c) + rng.normal(size=n)
    return rng, W, X, y


def _rss(A, y):
    beta, *_ = np.linalg.lstsq(A, y, rcond=None)
    r = y - A @ beta
    return r @ r, beta


def test_identity_kinship_reduces_to_ols():
    _, W, X, y = _ols_setup()
    n, c = W.shape
    d = np.ones(n)
    U = np.linalg.qr(np.random.default_rng(2).normal(size=(n, n)))[0]  # any orthonormal basis
    UtW, Uty, Utx = U.T @ W, U.T @ y, U.T @ X
    lam, _ = ol.optimize_reml(c, d, UtW, Uty, Utx)
    beta, se, p = ol.wald_batch(c, lam, d, UtW, Uty, Utx)
    rss0, _ = _rss(W, y)
    df = n - c - 1
    for j in range(X.shape[1]):
        A = np.column_stack([W, X[:, j]])
        rss1, b = _rss(A, y)
        F = (rss0 - rss1) / (rss1 / df)
        cov = np.linalg.inv(A.T @ A) * rss1 / df
        assert abs(beta[j] - b[-1]) < 1e-10 * max(1, abs(b[-1]))
        assert abs(se[j] - np.sqrt(cov[-1, -1])) < 1e-10
        assert abs(p[j] - stats.f.sf(F, 1, df)) < 1e-10 * max(p[j], 1e-300) + 1e-15
        # MLE flat -> LRT = n log(RSS0/RSS1)
    _, ll1 = ol.optimize_mle(c, d, UtW, Uty, Utx)
    _, ll0 = ol.null_model_mle(d, UtW, Uty, c)
    for j in range(X.shape[1]):
        rss1, _ = _rss(np.column_stack([W, X[:, j]]), y)
        assert abs(2 * (ll1[j] - ll0) - n * np.log(rss0 / rss1)) < 1e-8


def test_identity_kinship_score_is_n_r2():
    _, W, X, y = _ols_setup(seed=13)
    n, c = W.shape
    d = np.ones(n)
    U = np.linalg.qr(np.random.default_rng(4).normal(size=(n, n)))[0]
    Hi = 1.0 / (0.7 * d + 1.0)
    _, _, p = ol.score_batch(c, Hi, U.T @ W, U.T @ y, U.T @ X)
    P = np.eye(n) - W @ np.linalg.solve(W.T @ W, W.T)
    for j in range(X.shape[1]):
        x = P @ X[:, j]
        r = P @ y
        F = n * (x @ r) ** 2 / ((r @ r) * (x @ x))
        assert abs(p[j] - stats.f.sf(F, 1, n - c - 1)) < 1e-10


def test_golden_section_vs_brute_force(synth):
    G, y, K, _ = synth
    d, U = ol.eigendecompose(K)
    Gi = np.where(np.isnan(G), np.nanmean(G, axis=0), G)[:, :20]
    UtW, Uty, Utx = U.T @ np.ones((100, 1)), U.T @ y, U.T @ Gi
    lam, ll = ol.optimize_reml(1, d, UtW, Uty, Utx)
    Iab = ol.pab_batch(1, np.ones(100), UtW, Uty, Utx)
    grid = np.exp(np.linspace(np.log(1e-5), np.log(1e5), 4001))
    dense = np.stack([ol.reml_batch(1, g, d, UtW, Uty, Utx, Iab) for g in grid])
    best = dense.max(axis=0)
    # the golden-section result is at least as good as a 4001-point scan
    # (up to the scan's resolution) ...
    assert np.all(ll >= best - 1e-6)
    # ... and is the maximum of a fine local scan around it (bracket 6.6e-5)
    for j in range(20):
        t = np.log(lam[j]) + np.linspace(-2e-4, 2e-4, 401)
        t = t[(t >= np.log(1e-5)) & (t <= np.log(1e5))]
        loc = np.array([ol.reml_batch(1, np.exp(tt), d, UtW, Uty, Utx[:, j:j + 1],
                                      Iab[j:j + 1])[0] for tt in t])
        assert ll[j] >= loc.max() - 1e-8


# ---------------------------------------------------------------- planted mistakes
def test_planted_mistake_in_recursion_is_caught(synth, monkeypatch):
    G, y, K, _ = synth
    real = ol.pab_batch

    def broken(n_cvt, Hi, UtW, Uty, Utx):
        P = real(n_cvt, Hi, UtW, Uty, Utx)
        P[:, n_cvt + 1] = P[:, n_cvt]  # "forget" to project out the genotype
        return P

    monkeypatch.setattr(ol, "pab_batch", broken)
    res = ol.run_lmm(G[:, :50], y, kinship=K, lmm_mode=1)
    ref = load_assoc(GOLDEN / "gemma_synthetic" / "gemma_assoc.assoc.txt")
    with pytest.raises(AssertionError):
        np.testing.assert_allclose(res["p_wald"], ref["p_wald"][:50], rtol=1e-4)


def test_planted_mistake_in_reml_constant_is_caught(synth, monkeypatch):
    _, y, K, _ = synth
    d, U = ol.eigendecompose(K)
    real = ol.reml_null

    def broken(lam, d, UtW, Uty, n_cvt):  # wrong df (n - n_cvt - 1 instead of n - n_cvt)
        return real(lam, d[:-1], UtW[:-1], Uty[:-1], n_cvt)

    monkeypatch.setattr(ol, "reml_null", broken)
    _, l_reml = ol.null_model_reml(d, U.T @ np.ones((100, 1)), U.T @ y, 1)
    assert abs(l_reml - (-140.636)) > 1e-2
