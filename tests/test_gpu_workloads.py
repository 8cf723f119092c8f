"""Slices of the large reference-shaped workloads (GPU), sized to finish in a test.

* scale form (5 Balding-Nichols subpopulations, Fst 0.02, 0.5 % missing) at
  n = 16,385 samples: above 16,384 the cohort takes the three-plane bf16 split of
  U^T, and the eigendecomposition is the library's own 64-bit path at a size that is
  not a multiple of its 128-column panels. The eigendecomposition is checked by its
  residual and orthogonality (rocBLAS FP64 GEMMs, independent of the library), the
  mixed path against the FP64 device path on 2,048 SNPs at the acceptance bar, and
  the FP64 path against the oracle on 128 SNPs downstream of the same U.
* LOCO multi-phenotype form: n = 2,000, 1,000 SNPs on 5 chromosomes (lengths
  proportional to 1/sqrt(index)), 8 phenotypes with genetic correlation 0.3, an
  intercept + 3 covariates, -lmm 4: run_lmm_loco_multi against the oracle's LOCO
  kinships, eigendecompositions and association per phenotype (loco.py:54-593).
"""

import numpy as np
import pytest
import torch

from oracle import kinship as ok
from oracle import lmm as ol
from oracle import plink as opl
from tests.test_gpu_configs import _device_bn
from tests.test_gpu_parity import ALL, _arrays, _assert_matches_oracle
from tests.util import balding_nichols, eig_residuals

pytestmark = pytest.mark.gpu


def test_scale_form_three_planes_n16385():
    from jamma_amd.device import round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.kinship.compute import KinshipAccumulator
    from jamma_amd.lmm.eigen import eigh_device
    from jamma_amd.lmm.engine import AssocProblem, Cohort

    n, p_kin, p_test, p_oracle = 16_385, 12_000, 2048, 128
    ld = round_up(n, 64)
    g = _device_bn(n, p_kin + p_test, 5, 0.02, 0.005, 50, ld)
    gg = g[:, :n]
    counts = torch.stack([(gg == 0).sum(1), (gg == 1).sum(1), (gg == 2).sum(1), (gg < 0).sum(1)],
                         1).cpu().numpy().astype(np.int64)
    del gg
    mu64, _, poly = stats_from_counts(counts, np.float64)
    mean32, _, _ = stats_from_counts(counts, np.float32)
    keep = np.where(poly[:p_kin])[0]
    acc = KinshipAccumulator(n)
    gk = g[:p_kin][torch.as_tensor(keep, device=g.device)].contiguous()
    for s in range(0, len(keep), 8192):
        e = min(s + 8192, len(keep))
        acc.add_tile(gk[s:e], ld, e - s, mu64[keep[s:e]], True)
    del gk
    Kd = acc.finalize(len(keep))
    del acc
    d, U = eigh_device(Kd)
    res, orth = eig_residuals(Kd, d, U)
    print(f"n = {n}: residual {res:.2e}, orthogonality {orth:.2e}", flush=True)
    assert res <= 1e-11 and orth <= 1e-12, (res, orth)
    del Kd
    rng = np.random.default_rng(50)
    gt = g[p_kin:].contiguous()
    del g
    Xh = gt[:, :n].cpu().numpy().astype(np.float64).T
    m32 = mean32[p_kin:].astype(np.float64)
    Xh = np.where(Xh < 0, m32[None, :], Xh)
    causal = rng.choice(p_test, 40, replace=False)
    gv = Xh[:, causal] @ rng.normal(size=40)
    gv = (gv - gv.mean()) / gv.std()
    W = np.column_stack([np.ones(n), rng.normal(size=(n, 3))])
    y = np.sqrt(0.5) * gv + np.sqrt(0.5) * rng.normal(size=n) + W[:, 1:] @ rng.normal(0, 0.3, 3)
    cohort = Cohort(d, U, precision="both")
    assert cohort.bf16_planes == 3
    prob = AssocProblem(cohort, y, W)
    means = to_dev(m32, torch.float64)
    out = {prec: prob.associate(gt, ld, means, p_test, 1, prec).cpu().numpy()
           for prec in ("f64", "mixed")}
    a, b = out["mixed"], out["f64"]
    dl = np.abs(np.log10(a[:, 2]) - np.log10(b[:, 2]))
    print(f"mixed vs FP64: max |dlog10 p_wald| = {np.nanmax(dl):.2e}", flush=True)
    assert np.nanmax(dl) <= 1e-4
    for thr in (0.05, 0.01, 1e-3, 5e-8):
        assert np.array_equal(a[:, 2] < thr, b[:, 2] < thr), thr
    assert np.array_equal(np.sign(a[:, 0]), np.sign(b[:, 0]))
    sample = np.sort(rng.choice(p_test, p_oracle, replace=False))
    UT = np.ascontiguousarray(U.t().contiguous().cpu().numpy())
    del U, cohort, prob
    UtW, Uty = UT @ W, UT @ y
    ref = ol.assoc_chunk(1, W.shape[1], d, UtW, Uty, UT @ Xh[:, sample])
    cols = {"beta": 0, "se": 1, "p_wald": 2, "l_remle": 3, "logl_H1": 4}
    got = {k: b[sample, c] for k, c in cols.items()}
    _assert_matches_oracle(got, ref, tuple(cols))


def _loco_slice(seed=5):
    n, p, nchr, n_ph = 2000, 1000, 5, 8
    w = 1.0 / np.sqrt(np.arange(1, nchr + 1))
    L = np.round(p * w / w.sum()).astype(int)
    L[-1] = p - L[:-1].sum()
    chrom = np.repeat([str(c + 1) for c in range(nchr)], L)
    G = balding_nichols(n, p, n_pop=3, fst=0.05, miss=0.01, seed=seed)
    rng = np.random.default_rng(seed)
    Gi = np.where(np.isnan(G), np.nanmean(G, axis=0), G)
    causal = rng.choice(p, 60, replace=False)
    C = np.full((n_ph, n_ph), 0.3) + 0.7 * np.eye(n_ph)  # genetic correlation 0.3
    B = rng.multivariate_normal(np.zeros(n_ph), C, size=60)
    gvals = Gi[:, causal] @ B
    gvals = (gvals - gvals.mean(0)) / gvals.std(0)
    W = np.column_stack([np.ones(n), rng.normal(size=(n, 3))])
    h2 = rng.uniform(0.2, 0.8, n_ph)
    Y = np.sqrt(h2) * gvals + np.sqrt(1 - h2) * rng.normal(size=(n, n_ph)) + \
        W[:, 1:] @ rng.normal(0, 0.3, (3, n_ph))
    Y[7, :] = np.nan  # one sample without phenotypes (the shared valid-sample set)
    return G, chrom, Y, W


def test_loco_multi_phenotype_slice(tmp_path):
    from jamma_amd.lmm.loco import run_lmm_loco_multi

    G, chrom, Y, W = _loco_slice()
    pre = tmp_path / "loco8"
    opl.write_plink(pre, G, phenotypes=Y[:, 0], chromosomes=chrom)
    res = run_lmm_loco_multi(pre, Y, covariates=W, lmm_mode=4, precision="f64")
    mixed = run_lmm_loco_multi(pre, Y, covariates=W, lmm_mode=4, precision="mixed")
    assert len(res) == Y.shape[1] == len(mixed)
    # oracle: LOCO kinships over every sample with the LMM thresholds, one
    # eigendecomposition per chromosome on the valid samples, shared by the phenotypes
    loco = ok.loco_kinship(G, chrom, 0.01, 0.05)
    valid = ~np.isnan(Y[:, 0])
    eig = {c: ol.eigendecompose(loco[c][np.ix_(valid, valid)]) for c in loco}
    for k in range(Y.shape[1]):
        parts = []
        for c in sorted(eig):
            sel = np.where(chrom == c)[0]
            r = ol.run_lmm(G[:, sel], Y[:, k], covariates=W, eigenvalues=eig[c][0],
                           eigenvectors=eig[c][1], lmm_mode=4, stat_dtype=np.float64)
            r["global_indices"] = sel[r["snp_indices"]]
            parts.append(r)
        order = np.concatenate([r["global_indices"] for r in parts])
        assert [x.rs for x in res[k]] == [f"rs{j:06d}" for j in order]
        cat = {key: np.concatenate([r[key] for r in parts]) for key in ALL}
        _assert_matches_oracle(_arrays(res[k], ALL), cat, ALL)
        a, b = _arrays(mixed[k], ("p_wald", "p_lrt", "p_score")), _arrays(res[k], ("p_wald", "p_lrt", "p_score"))
        for key in a:
            assert np.max(np.abs(np.log10(a[key]) - np.log10(b[key]))) <= 1e-4, (k, key)
