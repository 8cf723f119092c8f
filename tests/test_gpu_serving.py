"""Serving loop (GPU): continuous batching over resident cohorts.

Each request's result must be the same alone and in the mix (same device
instructions per SNP, so bit-identical), match the CPU oracle on the same
inputs, and admission must respect the HBM budget.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import kinship as ok
from oracle import lmm as ol
from tests.util import balding_nichols

pytestmark = pytest.mark.gpu


def _cohort(name, n, p, seed, precision="mixed"):
    from jamma_amd.device import round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.serve import ResidentCohort

    G = balding_nichols(n, p, miss=0.01, seed=seed)
    K = ok.centered_kinship(G)
    d, U = ol.eigendecompose(K)
    ld = round_up(n, 64)
    g8 = np.zeros((p, ld), np.int8)
    g8[:, :n] = np.where(np.isnan(G), -1, G).astype(np.int8).T
    counts = np.stack([(G == 0).sum(0), (G == 1).sum(0), (G == 2).sum(0), np.isnan(G).sum(0)], 1)
    mean32, _, _ = stats_from_counts(counts, np.float32)
    rc = ResidentCohort(name, to_dev(g8, torch.int8), ld, to_dev(mean32, torch.float64), d, U,
                        precision=precision)
    return rc, G, d, U, mean32


def _requests(rng, cohorts, k):
    from jamma_amd.serve import Request

    out = []
    for i in range(k):
        name, (rc, G, *_) = list(cohorts.items())[i % len(cohorts)]
        n = rc.n
        ncov = int(rng.integers(0, 4))
        W = np.column_stack([np.ones(n), rng.normal(size=(n, ncov))]) if ncov else None
        y = rng.normal(size=n) + 0.3 * np.nan_to_num(G[:, i % 50])
        length = int(rng.integers(20, 400))
        start = int(rng.integers(0, rc.p - length))
        out.append(Request(name, y, start, start + length, covariates=W, mode=4))
    return out


def _clone(r):
    from jamma_amd.serve import Request

    return Request(r.cohort, r.phenotype, r.snp_start, r.snp_end, r.covariates, r.mode)


@pytest.fixture(scope="module")
def cohorts():
    return {"a": _cohort("a", 300, 900, 1), "b": _cohort("b", 517, 700, 2)}


def test_alone_equals_mix_and_oracle(cohorts):
    from jamma_amd.serve import Server

    rng = np.random.default_rng(0)
    reqs = _requests(rng, cohorts, 8)
    alone = []
    for r in reqs:
        srv = Server(slice_snps=64, max_batch_snps=256)
        for rc, *_ in cohorts.values():
            srv.add_cohort(rc)
        with srv:
            alone.append(srv.submit(_clone(r)).result(timeout=120))
    # the mix also caps each cohort's steps by n^2 x SNPs: 128 SNPs at n = 300, one
    # 64-SNP slice at n = 517
    srv = Server(slice_snps=64, max_batch_snps=256, step_work=64 * 517 ** 2)
    for rc, *_ in cohorts.values():
        srv.add_cohort(rc)
    assert srv.step_snps == {"a": 128, "b": 64}
    with srv:
        srv.pause()  # every request prepared and queued first: the steps mix them
        futs = [srv.submit(_clone(r)) for r in reqs]
        srv.wait_prepared()
        srv.resume()
        mixed = [f.result(timeout=300) for f in futs]
        stats = srv.latency_stats()
    assert stats["max_concurrent"] >= 2, stats
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m, equal_nan=True)
    # oracle on the same inputs
    for r, res in zip(reqs, mixed):
        rc, G, d, U, mean32 = cohorts[r.cohort]
        X = G[:, r.snp_start:r.snp_end]
        X = np.where(np.isnan(X), mean32[r.snp_start:r.snp_end][None, :], X)
        W = r.covariates if r.covariates is not None else np.ones((rc.n, 1))
        UT = U.T
        UtW, Uty = UT @ W, UT @ r.phenotype
        lam0, ll0 = ol.null_model_mle(d, UtW, Uty, W.shape[1])
        ref = ol.assoc_chunk(4, W.shape[1], d, UtW, Uty, UT @ X, 1.0 / (lam0 * d + 1.0), ll0)
        for k, col in (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)):
            dl = np.abs(np.log10(res[:, col]) - np.log10(ref[k]))
            assert np.nanmax(dl) <= 1e-4, (r.cohort, k, np.nanmax(dl))


def test_admission_budget_queues_and_completes(cohorts):
    """An arena that holds the cohorts' step workspaces and 1.5 requests admits one
    request at a time; every request still completes. A request that can never fit
    is refused at submission, and a cohort whose workspace does not fit at add time."""
    from jamma_amd.serve import Request, Server
    from jamma_amd.serve.server import cohort_workspace_bytes, request_workspace_bytes, serve_config

    rng = np.random.default_rng(1)
    reqs = _requests(rng, cohorts, 5)
    rc = cohorts["a"][0]
    cfg = serve_config(slice_snps=64)
    ws = sum((cohort_workspace_bytes(c, cfg) + 255) // 256 * 256 for c, *_ in cohorts.values())
    one = max(request_workspace_bytes(cohorts[r.cohort][0], cfg,
                                      1 if r.covariates is None else r.covariates.shape[1])
              for r in reqs)
    srv = Server(hbm_budget_bytes=ws + int(one * 1.5), slice_snps=64)
    for c, *_ in cohorts.values():
        srv.add_cohort(c)
    with srv:
        futs = [srv.submit(r) for r in reqs]
        res = [f.result(timeout=300) for f in futs]
        stats = srv.latency_stats()
    assert stats["max_concurrent"] < len(reqs)  # the budget forces queueing
    assert all(np.isfinite(x[:, 2]).all() for x in res)
    ws_a = (cohort_workspace_bytes(rc, cfg) + 255) // 256 * 256
    tiny = Server(hbm_budget_bytes=ws_a + 1024, slice_snps=64)
    tiny.add_cohort(rc)
    with pytest.raises(MemoryError):
        tiny.submit(Request("a", np.zeros(rc.n), 0, rc.p, covariates=None))
    tiny.stop()
    tinier = Server(hbm_budget_bytes=1024, slice_snps=64)
    with pytest.raises(MemoryError):
        tinier.add_cohort(rc)
    tinier.stop()


def test_step_error_fails_its_cohort_alone(cohorts):
    """A step that fails inside the runtime (the failure hook: the cohort's next step
    fails on the host before any launch) fails the requests active on that cohort with
    the error, releases their arena space, and the loop keeps serving: a later request
    on the same cohort and one on another cohort complete and match their results
    alone."""
    from jamma_amd.serve import Request, Server

    rca, rcb = cohorts["a"][0], cohorts["b"][0]
    rng = np.random.default_rng(9)
    ya, yb, yc = (rng.normal(size=rc.n) for rc in (rca, rca, rcb))
    srv = Server(slice_snps=64)
    for c, *_ in cohorts.values():
        srv.add_cohort(c)
    ref = Server(slice_snps=64)
    for c, *_ in cohorts.values():
        ref.add_cohort(c)
    with ref:
        want_a2 = ref.submit(Request("a", yb, 100, 300)).result(timeout=120)
        want_b = ref.submit(Request("b", yc, 0, 150)).result(timeout=120)
    with srv:
        srv.inject_step_error("a")
        bad = srv.submit(Request("a", ya, 0, 200))
        with pytest.raises(RuntimeError, match="injected step error"):
            bad.result(timeout=120)
        good_a = srv.submit(Request("a", yb, 100, 300))
        good_b = srv.submit(Request("b", yc, 0, 150))
        assert np.array_equal(good_a.result(timeout=120), want_a2, equal_nan=True)
        assert np.array_equal(good_b.result(timeout=120), want_b, equal_nan=True)
        stats = srv.latency_stats()
        st = srv.stats()
    assert stats["failed"] == 1 and st["arena_used"] == st["reserved"]


def test_bad_request_fails_alone(cohorts):
    from jamma_amd.serve import Request, Server

    rc = cohorts["a"][0]
    srv = Server(slice_snps=64)
    srv.add_cohort(rc)
    with srv:
        bad = srv.submit(Request("a", np.zeros(rc.n + 1), 0, 100))
        good = srv.submit(Request("a", np.random.default_rng(2).normal(size=rc.n), 0, 100))
        with pytest.raises(ValueError):
            bad.result(timeout=60)
        assert good.result(timeout=120).shape == (100, 12)


def test_wide_spectrum_cohort_many_covariates():
    """A cohort whose spectrum compresses to T in (384, 512] (the kernel with 8 terms
    per lane and 4 moment rows in registers) with requests of 0-10 covariates, so that
    one batch holds problems of different c and rows read from memory: every result
    is the same alone and in the mix, and matches the oracle within 1e-4 in log10 p."""
    from jamma_amd.device import round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.serve import Request, ResidentCohort, Server

    rng = np.random.default_rng(77)
    n, p = 1500, 500
    d = np.sort(rng.uniform(0.01, 5.0, n))
    U, _ = np.linalg.qr(rng.normal(size=(n, n)))
    G = balding_nichols(n, p, miss=0.01, seed=77)
    ld = round_up(n, 64)
    g8 = np.zeros((p, ld), np.int8)
    g8[:, :n] = np.where(np.isnan(G), -1, G).astype(np.int8).T
    counts = np.stack([(G == 0).sum(0), (G == 1).sum(0), (G == 2).sum(0), np.isnan(G).sum(0)], 1)
    mean32, _, _ = stats_from_counts(counts, np.float32)
    rc = ResidentCohort("wide", to_dev(g8, torch.int8), ld, to_dev(mean32, torch.float64), d, U)
    assert 384 < rc.cohort.n_terms <= 512, rc.cohort.n_terms
    reqs = []
    for i, ncov in enumerate((0, 10, 3, 7, 1, 5)):
        W = np.column_stack([np.ones(n), rng.normal(size=(n, ncov))]) if ncov else None
        y = U @ (rng.normal(size=n) * np.sqrt(0.5 * d + 0.5)) + 0.2 * np.nan_to_num(G[:, i])
        length = int(rng.integers(40, 160))
        start = int(rng.integers(0, p - length))
        reqs.append(Request("wide", y, start, start + length, covariates=W, mode=4))
    alone = []
    for r in reqs:
        srv = Server(slice_snps=48, max_batch_snps=192)
        srv.add_cohort(rc)
        with srv:
            alone.append(srv.submit(_clone(r)).result(timeout=120))
    srv = Server(slice_snps=48, max_batch_snps=192)
    srv.add_cohort(rc)
    with srv:
        srv.pause()
        futs = [srv.submit(_clone(r)) for r in reqs]
        srv.wait_prepared()
        srv.resume()
        mixed = [f.result(timeout=300) for f in futs]
        stats = srv.latency_stats()
    assert stats["max_concurrent"] >= 2, stats
    UT = U.T
    for r, a, m in zip(reqs, alone, mixed):
        assert np.array_equal(a, m, equal_nan=True)
        X = G[:, r.snp_start:r.snp_end]
        X = np.where(np.isnan(X), mean32[r.snp_start:r.snp_end][None, :], X)
        W = r.covariates if r.covariates is not None else np.ones((n, 1))
        UtW, Uty = UT @ W, UT @ r.phenotype
        lam0, ll0 = ol.null_model_mle(d, UtW, Uty, W.shape[1])
        ref = ol.assoc_chunk(4, W.shape[1], d, UtW, Uty, UT @ X, 1.0 / (lam0 * d + 1.0), ll0)
        for k, col in (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)):
            dl = np.abs(np.log10(m[:, col]) - np.log10(ref[k]))
            assert np.nanmax(dl) <= 1e-4, (W.shape[1], k, np.nanmax(dl))
