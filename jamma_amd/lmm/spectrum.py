"""Compressed eigen-spectrum plan for the per-SNP lambda search.

Every likelihood evaluation of the reference (likelihood_jax.py:148-339) is a
set of weighted sums over the n eigen-indices,

    Pab0_ab(lambda) = sum_i v_a[i] v_b[i] / (1 + lambda d_i),
    logdet_h(lambda) = sum_i log|1 + lambda d_i|,

and the reference performs ~149 of them per SNP (50-point grid + 20 golden
steps + 3, for REML and MLE each, plus Wald and Score). Sorted eigenvalues are
grouped into clusters [i0, i1) of relative half-width r around a centre d_c.
With delta_i = d_i - d_c, h_c = 1/(1+lambda d_c) and z_c = lambda h_c:

    1/(1 + lambda d_i)     = h_c * sum_m (-z_c delta_i)^m
    log(1 + lambda d_i)    = log(1 + lambda d_c) + sum_{m>=1} (-1)^(m+1) (z_c delta_i)^m / m

and |z_c delta_i| <= |delta_i| / d_c <= r for every lambda >= 0, so M_c terms
give a relative truncation error below r^M_c / (1 - r) for the whole lambda
range. The per-cluster moments  sum_{i in c} v_a v_b delta_i^m  are computed
once per SNP; afterwards an evaluation costs sum_c M_c instead of n. Exact
zeros form one cluster with M = 1 (delta = 0), negative eigenvalues and
clusters whose M_c would reach their size are kept as singletons (exact).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

MAX_TERMS_PER_CLUSTER = 16


@dataclass
class SpectrumPlan:
    n: int
    d: np.ndarray            # (n,) eigenvalues, ascending, thresholded
    cl_start: np.ndarray     # (C,) int32
    cl_size: np.ndarray      # (C,) int32
    cl_center: np.ndarray    # (C,) float64
    cl_terms: np.ndarray     # (C,) int32  M_c
    cl_toff: np.ndarray      # (C,) int32  offset of (c, m=0) in the term list
    n_terms: int             # T = sum M_c
    term_cluster: np.ndarray  # (T,) int32
    term_m: np.ndarray       # (T,) int32
    idx_cluster: np.ndarray  # (n,) int32  cluster of eigen-index i
    idx_delta: np.ndarray    # (n,) float64  d_i - d_c
    radius: float
    # pieces of at most seg_max eigen-indices (balanced work for the moments)
    seg_max: int = 0
    seg_start: np.ndarray = None   # (S,) int32
    seg_len: np.ndarray = None     # (S,) int32
    seg_cluster: np.ndarray = None  # (S,) int32
    seg_slot: np.ndarray = None    # (S,) int32  first partial slot of the piece
    cl_seg0: np.ndarray = None     # (C,) int32  first piece of cluster c
    cl_nseg: np.ndarray = None     # (C,) int32
    n_slots: int = 0


def segment_plan(pl: "SpectrumPlan", seg_max: int | None = None) -> "SpectrumPlan":
    """Cut every cluster into near-equal pieces of at most seg_max eigen-indices
    (default max(64, min(512, ceil(n / 20)))); fills the seg_* fields in place."""
    n = pl.n
    if seg_max is None:
        seg_max = max(64, min(512, -(-n // 20)))
    starts, lens, clus, seg0, nseg = [], [], [], [], []
    for c in range(len(pl.cl_size)):
        size, st = int(pl.cl_size[c]), int(pl.cl_start[c])
        k = -(-size // seg_max)
        seg0.append(len(starts))
        nseg.append(k)
        base, extra = divmod(size, k)
        pos = st
        for q in range(k):
            ln = base + (1 if q < extra else 0)
            starts.append(pos)
            lens.append(ln)
            clus.append(c)
            pos += ln
    clus = np.array(clus, dtype=np.int32)
    m = pl.cl_terms[clus].astype(np.int64)
    slot = np.zeros(len(clus), dtype=np.int64)
    slot[1:] = np.cumsum(m)[:-1]
    pl.seg_max = int(seg_max)
    pl.seg_start = np.array(starts, dtype=np.int32)
    pl.seg_len = np.array(lens, dtype=np.int32)
    pl.seg_cluster = clus
    pl.seg_slot = slot.astype(np.int32)
    pl.cl_seg0 = np.array(seg0, dtype=np.int32)
    pl.cl_nseg = np.array(nseg, dtype=np.int32)
    pl.n_slots = int(m.sum())
    return pl


def _terms_for(rc: float, eps: float) -> int:
    if rc <= 0.0:
        return 1
    m = math.ceil(math.log(eps * (1.0 - rc)) / math.log(rc))
    return max(1, m)


def _greedy(d: np.ndarray, r: float, eps: float):
    n = d.shape[0]
    clusters = []  # (start, size, center, M)
    i = 0
    ratio = (1.0 + r) / (1.0 - r)
    while i < n:
        di = d[i]
        if di < 0.0:
            clusters.append((i, 1, di, 1))
            i += 1
            continue
        if di == 0.0:
            j = i
            while j + 1 < n and d[j + 1] == 0.0:
                j += 1
            clusters.append((i, j - i + 1, 0.0, 1))
            i = j + 1
            continue
        j = i
        lim = di * ratio
        while j + 1 < n and d[j + 1] <= lim:
            j += 1
        size = j - i + 1
        if size == 1:
            clusters.append((i, 1, di, 1))
        else:
            lo, hi = d[i], d[j]
            center = 0.5 * (lo + hi)
            rc = (hi - lo) / (hi + lo)
            m = _terms_for(rc, eps)
            if m > MAX_TERMS_PER_CLUSTER or m >= size:
                for k in range(i, j + 1):
                    clusters.append((k, 1, d[k], 1))
            else:
                clusters.append((i, size, center, m))
        i = j + 1
    return clusters


def plan_spectrum(d: np.ndarray, eps: float = 1e-16,
                  radii=(0.01, 0.02, 0.05, 0.1, 0.15, 0.2, 0.3)) -> SpectrumPlan:
    """Choose the clustering (over a few target radii) with the fewest terms."""
    d = np.ascontiguousarray(d, dtype=np.float64)
    if d.ndim != 1 or d.shape[0] == 0:
        raise ValueError("eigenvalues must be a non-empty vector")
    if np.any(np.diff(d) < 0):
        raise ValueError("eigenvalues must be sorted ascending")
    best = None
    for r in radii:
        cl = _greedy(d, r, eps)
        t = sum(c[3] for c in cl)
        if best is None or t < best[0]:
            best = (t, r, cl)
    t, r, cl = best
    start = np.array([c[0] for c in cl], dtype=np.int32)
    size = np.array([c[1] for c in cl], dtype=np.int32)
    center = np.array([c[2] for c in cl], dtype=np.float64)
    mterms = np.array([c[3] for c in cl], dtype=np.int32)
    toff = np.zeros(len(cl), dtype=np.int32)
    toff[1:] = np.cumsum(mterms)[:-1]
    term_cluster = np.repeat(np.arange(len(cl), dtype=np.int32), mterms)
    term_m = (np.arange(t, dtype=np.int32) - np.repeat(toff, mterms)).astype(np.int32)
    idx_cluster = np.repeat(np.arange(len(cl), dtype=np.int32), size)
    idx_delta = d - center[idx_cluster]
    pl = SpectrumPlan(n=d.shape[0], d=d, cl_start=start, cl_size=size, cl_center=center,
                      cl_terms=mterms, cl_toff=toff, n_terms=int(t),
                      term_cluster=term_cluster, term_m=term_m, idx_cluster=idx_cluster,
                      idx_delta=idx_delta, radius=r)
    return segment_plan(pl)
