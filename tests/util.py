"""Shared helpers for the tests: fixture loaders and synthetic generators."""

from __future__ import annotations

import numpy as np

from tests.conftest import GOLDEN


def load_assoc(path):
    rows = [line.rstrip("\n").split("\t") for line in open(path)]
    head = rows[0]
    cols = {k: [r[i] for r in rows[1:]] for i, k in enumerate(head)}
    out = {}
    for k, v in cols.items():
        try:
            out[k] = np.array([float(x) for x in v])
        except ValueError:
            out[k] = np.array(v)
    return out


def synthetic_prefix():
    return GOLDEN / "gemma_synthetic" / "test"


def fam_phenotypes(prefix):
    vals = []
    for line in open(f"{prefix}.fam"):
        t = line.split()
        vals.append(np.nan if t[5] in ("-9", "NA") else float(t[5]))
    return np.array(vals)


def balding_nichols(n, p, n_pop=3, fst=0.05, miss=0.01, seed=0):
    """Genotypes (n, p) float64 with NaN: BN subpopulations, p ~ U(0.05, 0.5)."""
    rng = np.random.default_rng(seed)
    anc = rng.uniform(0.05, 0.5, p)
    pops = rng.integers(0, n_pop, n)
    G = np.empty((n, p))
    for k in range(n_pop):
        pk = rng.beta(anc * (1 - fst) / fst, (1 - anc) * (1 - fst) / fst)
        idx = pops == k
        G[idx] = rng.binomial(2, pk, size=(int(idx.sum()), p))
    if miss > 0:
        G[rng.random((n, p)) < miss] = np.nan
    return G


def phenotype(G, n_causal=20, h2=0.5, n_cov=3, seed=1):
    """Quantitative phenotype with causal SNPs, Gaussian covariates + intercept."""
    rng = np.random.default_rng(seed)
    n, p = G.shape
    Gi = np.where(np.isnan(G), np.nanmean(G, axis=0), G)
    causal = rng.choice(p, size=min(n_causal, p), replace=False)
    g = Gi[:, causal] @ rng.normal(size=len(causal))
    g = (g - g.mean()) / (g.std() + 1e-12)
    W = np.column_stack([np.ones(n), rng.normal(size=(n, n_cov))]) if n_cov else None
    y = np.sqrt(h2) * g + np.sqrt(1 - h2) * rng.normal(size=n)
    if W is not None:
        y = y + W[:, 1:] @ rng.normal(scale=0.3, size=n_cov)
    return y, W


def eig_residuals(Kd, d, U):
    """(max |K U - U D| / max |d|, max |U^T U - I|) with torch's FP64 GEMM."""
    import torch

    dd = torch.as_tensor(np.asarray(d), dtype=torch.float64, device=Kd.device)
    Uc = U.contiguous()
    R = Kd @ Uc
    R -= Uc * dd[None, :]
    res = float(R.abs().max()) / float(dd.abs().max())
    del R
    O = Uc.t().contiguous() @ Uc
    O.diagonal().sub_(1.0)
    orth = float(O.abs().max())
    return res, orth
