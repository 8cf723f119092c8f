"""Degenerate SNPs and phenotypes on the device: the reference's guards
(likelihood_jax.py:231-251, 601-675, 706-773; its property tests
tests/test_hypothesis.py:482-712 exercise the same cases on the CPU).

* a SNP that is zero in the rotated basis (all its spectrum moments are exactly 0,
  as for a SNP wholly explained by the covariates when the projection cancels
  exactly): P_xx = 0 -> beta, se, p_wald and p_score are NaN; the LRT compares two
  equal likelihoods (p ~ 1);
* a zero phenotype: P_yy = P_xy = 0, Px_yy is clamped to 1e-8, F <= 0 -> p_wald = 1
  and beta = 0; the Score statistic is 0/0 -> NaN;
* a phenotype of scale 1e-7: P_yy < 1e-8 at every level, the clamp decides
  Px_yy and F < 0 -> p_wald = 1, while Score stays finite.
Both search kernels (16 lanes per SNP, and one wavefront per SNP) are compared with
the oracle on the same rotated inputs, NaN for NaN.
"""

import numpy as np
import pytest
import torch

from oracle import lmm as ol

pytestmark = pytest.mark.gpu

COLS = {"beta": 0, "se": 1, "p_wald": 2, "l_remle": 3, "logl_H1": 4, "l_mle": 5,
        "p_lrt": 7, "p_score": 8}


def _setup(seed, n=300, p=12, ncov=2):
    rng = np.random.default_rng(seed)
    d = np.sort(rng.uniform(0.05, 3.0, n))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    W = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(ncov)])
    X = rng.binomial(2, 0.3, size=(n, p)).astype(np.float64)
    X[:, 0] = 0.0  # zero SNP
    return rng, d, Q, W, X


def _run(monkeypatch, d, Q, W, X, y):
    from jamma_amd.lmm.engine import AssocProblem, Cohort

    n, p = X.shape
    co = Cohort(d, torch.as_tensor(Q, device="cuda"), precision="f64")
    pb = AssocProblem(co, y, W)
    UT = Q.T
    Z = torch.as_tensor(np.ascontiguousarray(UT @ X), device="cuda")
    xm = co.moments(Z, False, p, p, pb.V, pb.n_cvt + 1)
    outs = []
    for wave in ("0", "1"):
        monkeypatch.setenv("JM_LMM_WAVE_PER_SNP", wave)
        outs.append(pb.assoc_moments(xm, p, 4).cpu().numpy())
    lam0, ll0 = ol.null_model_mle(d, UT @ W, UT @ y, W.shape[1])
    with np.errstate(all="ignore"):
        ref = ol.assoc_chunk(4, W.shape[1], d, UT @ W, UT @ y, UT @ X, 1.0 / (lam0 * d + 1.0),
                             ll0)
    return outs, ref


def _same_nan(got, want, what):
    assert np.array_equal(np.isnan(got), np.isnan(want)), (what, got, want)


def test_zero_snp_gives_nan_statistics(monkeypatch):
    rng, d, Q, W, X = _setup(5)
    y = Q @ (rng.normal(size=len(d)) * np.sqrt(d + 0.5)) + W @ rng.normal(size=W.shape[1])
    outs, ref = _run(monkeypatch, d, Q, W, X, y)
    for out in outs:
        for k in ("beta", "se", "p_wald", "p_score"):
            assert np.isnan(out[0, COLS[k]]), k
            _same_nan(out[:, COLS[k]], ref[k], k)
        # the other SNPs are ordinary
        for k in ("p_wald", "p_lrt", "p_score"):
            np.testing.assert_allclose(-np.log10(out[1:, COLS[k]]), -np.log10(ref[k][1:]),
                                       atol=1e-5, err_msg=k)
        # zero SNP: H1 = H0 up to the optimisers' rounding -> p_lrt ~ 1
        assert out[0, COLS["p_lrt"]] > 0.999 and ref["p_lrt"][0] > 0.999
        np.testing.assert_allclose(out[:, COLS["logl_H1"]], ref["logl_H1"], rtol=1e-9)


def test_zero_phenotype_clamps_and_f_guard(monkeypatch):
    rng, d, Q, W, X = _setup(6)
    y = np.zeros(len(d))
    outs, ref = _run(monkeypatch, d, Q, W, X, y)
    for out in outs:
        assert np.all(out[1:, COLS["p_wald"]] == 1.0) and np.all(ref["p_wald"][1:] == 1.0)
        assert np.all(out[1:, COLS["beta"]] == 0.0)
        for k in ("beta", "se", "p_wald", "p_score"):
            _same_nan(out[:, COLS[k]], ref[k], k)
        assert np.all(np.isnan(out[:, COLS["p_score"]]))  # 0 / 0 statistic
        ok = np.isfinite(ref["se"])
        np.testing.assert_allclose(out[ok, COLS["se"]], ref["se"][ok], rtol=1e-6)
        np.testing.assert_allclose(out[:, COLS["logl_H1"]], ref["logl_H1"], rtol=1e-9)
        np.testing.assert_allclose(out[:, COLS["l_remle"]], ref["l_remle"], rtol=5e-5)


def test_tiny_phenotype_pyy_clamp(monkeypatch):
    rng, d, Q, W, X = _setup(7)
    y = 1e-7 * rng.normal(size=len(d))
    outs, ref = _run(monkeypatch, d, Q, W, X, y)
    for out in outs:
        assert np.all(out[1:, COLS["p_wald"]] == 1.0) and np.all(ref["p_wald"][1:] == 1.0)
        for k in COLS:
            _same_nan(out[:, COLS[k]], ref[k], k)
        fin = np.isfinite(ref["p_score"])
        np.testing.assert_allclose(-np.log10(out[fin, COLS["p_score"]]),
                                   -np.log10(ref["p_score"][fin]), atol=1e-5)
        np.testing.assert_allclose(out[:, COLS["logl_H1"]], ref["logl_H1"], rtol=1e-9)
        np.testing.assert_allclose(out[:, COLS["l_remle"]], ref["l_remle"], rtol=5e-5)
        np.testing.assert_allclose(out[:, COLS["l_mle"]], ref["l_mle"], rtol=5e-5)
