"""Both lambda-search kernels on cohorts built from a chosen spectrum.

A random orthogonal basis and eigenvalues drawn from a uniform range give spectra
that compress into many clusters: 16 terms per lane with more clusters than lanes
in a group, 12 terms per lane, and SNP counts that leave idle lane groups. The
16-lanes-per-SNP kernel and the wavefront-per-SNP kernel (JM_LMM_WAVE_PER_SNP=1)
run on the same device moments and are checked against the oracle's chunk compute
(runner_streaming.py:471-554) on the same rotated inputs. The moments kernel is checked
against its defining sums with more partner rows than a block has wavefronts.
"""

import numpy as np
import pytest
import torch

from oracle import lmm as ol

pytestmark = pytest.mark.gpu

# (eigenvalue range, covariates besides the intercept, SNPs, terms range, seed)
CASES = [((0.01, 5.0), 0, 37, (193, 256), 11), ((0.3, 1.5), 0, 41, (129, 192), 12),
         ((0.5, 1.5), 3, 6, (129, 192), 13)]


def _spectrum(rng, lo, hi, n):
    return np.sort(rng.uniform(lo, hi, n))


@pytest.mark.parametrize("rng_d,ncov,p,t_range,seed", CASES)
def test_search_kernels_match_oracle(monkeypatch, rng_d, ncov, p, t_range, seed):
    from jamma_amd.lmm.engine import AssocProblem, Cohort

    rng = np.random.default_rng(seed)
    n = 400
    d = _spectrum(rng, *rng_d, n)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    W = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(ncov)])
    X = rng.normal(size=(n, p))
    y = Q @ (rng.normal(size=n) * np.sqrt(0.5 * d + 0.5)) + W @ rng.normal(size=W.shape[1])
    co = Cohort(d, torch.as_tensor(Q, device="cuda"), precision="f64")
    assert t_range[0] <= co.n_terms <= t_range[1], co.n_terms
    print(f"T = {co.n_terms}, clusters = {len(co.plan.cl_size)}", flush=True)
    pb = AssocProblem(co, y, W)
    UT = Q.T
    Z = torch.as_tensor(np.ascontiguousarray(UT @ X), device="cuda")
    xm = co.moments(Z, False, p, p, pb.V, pb.n_cvt + 1)

    def run():
        return pb.assoc_moments(xm, p, 4).cpu().numpy()

    monkeypatch.delenv("JM_LMM_WAVE_PER_SNP", raising=False)
    out16 = run()
    monkeypatch.setenv("JM_LMM_WAVE_PER_SNP", "1")
    out64 = run()
    lam0, ll0 = ol.null_model_mle(d, UT @ W, UT @ y, W.shape[1])
    ref = ol.assoc_chunk(4, W.shape[1], d, UT @ W, UT @ y, UT @ X, 1.0 / (lam0 * d + 1.0), ll0)
    for out in (out16, out64):
        # a golden-section tie taken the other way moves lambda by one bracket step
        # (rtol 5e-5) and -log10 p by up to ~1e-5
        np.testing.assert_allclose(out[:, 3], ref["l_remle"], rtol=5e-5)
        np.testing.assert_allclose(out[:, 5], ref["l_mle"], rtol=5e-5)
        for k, col in (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)):
            np.testing.assert_allclose(-np.log10(out[:, col]), -np.log10(ref[k]), atol=1e-5,
                                       err_msg=k)
        np.testing.assert_allclose(out[:, 4], ref["logl_H1"], rtol=1e-9)


@pytest.mark.parametrize("nv,z_f32", [(14, False), (5, True), (1, False)])
def test_moments_many_partner_rows(nv, z_f32):
    """The moments kernel against its definition, with more row groups than a block
    has wavefronts (nv = 14: 15 rows in 5 groups of 3 on 4 waves), a SNP count that
    leaves a partial 64-column block, FP32 and FP64 Z. The row groups take each of the
    kernel's paths: three V rows (nv = 14, 5), two V rows and the Z^2 row (nv = 14, 5)
    and a short last group (nv = 1):
      mom[s][b][t] = sum_{i in c(t)} Z[i][s] V[b][i] delta_i^m(t)   (b < nv)
      mom[s][nv][t] = sum_{i in c(t)} Z[i][s]^2 delta_i^m(t)."""
    from jamma_amd.lmm.engine import Cohort

    rng = np.random.default_rng(nv)
    n, p = 300, 45
    d = np.sort(rng.uniform(0.05, 3.0, n))
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    co = Cohort(d, torch.as_tensor(Q, device="cuda"), precision="f64")
    pl = co.plan
    Zh = rng.normal(size=(n, p))
    if z_f32:
        Zh = Zh.astype(np.float32).astype(np.float64)
    V = rng.normal(size=(nv, n))
    Zd = torch.as_tensor(Zh.astype(np.float32) if z_f32 else Zh, device="cuda")
    got = co.moments(Zd, z_f32, p, p, torch.as_tensor(V, device="cuda"), nv).cpu().numpy()
    rows = np.concatenate([V[None, :, :] * Zh.T[:, None, :], (Zh.T ** 2)[:, None, :]], axis=1)
    ref = np.zeros((p, nv + 1, pl.n_terms))
    scale = np.zeros_like(ref)
    for c in range(len(pl.cl_size)):
        sl = slice(pl.cl_start[c], pl.cl_start[c] + pl.cl_size[c])
        for m in range(pl.cl_terms[c]):
            w = pl.idx_delta[sl] ** m
            ref[:, :, pl.cl_toff[c] + m] = rows[:, :, sl] @ w
            scale[:, :, pl.cl_toff[c] + m] = np.abs(rows[:, :, sl]) @ np.abs(w)
    assert np.all(np.abs(got - ref) <= 1e-13 * scale + 1e-300)
