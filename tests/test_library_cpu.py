"""CPU tests of the product: the C ABI loads and exports every declared symbol,
the host-compiled special functions match SciPy, and the host logic (spectrum
plan, SNP statistics from counts, GEMMA writers) is right. No GPU calls."""

from __future__ import annotations

import re

import numpy as np
import pytest
from scipy import special, stats

from jamma_amd import _lib
from jamma_amd.io.plink import stats_from_counts
from jamma_amd.lmm.io import format_line
from jamma_amd.lmm.spectrum import MAX_TERMS_PER_CLUSTER, plan_spectrum
from jamma_amd.lmm.stats import AssocResult
from oracle import lmm as ol
from tests.conftest import ROOT
from tests.util import GOLDEN


def test_library_exports_every_declared_symbol():
    header = (ROOT / "include" / "jamma_hip.h").read_text()
    declared = set(re.findall(r"^(?:int|double|int64_t)\s+(jm_\w+)\s*\(", header, flags=re.M))
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(_lib.lib, name), name
    assert declared == set(_lib.EXPORTED)
    assert _lib.lib.jm_version() == 101


def test_error_reporting_without_gpu_work():
    # argument validation happens before any device call
    rc = _lib.lib.jm_lmm_assoc(None, 4, None, 0, 0, None, None)
    assert rc == -1
    assert "null problem" in _lib.last_error()


@pytest.mark.parametrize("df", [3, 10, 97, 1408, 9995, 49995, 1_000_000])
def test_host_betainc_matches_scipy(df):
    a = df / 2.0
    for F in np.concatenate([np.logspace(-8, 4.5, 80), [0.5, 1.0, 3.84, 10.0]]):
        z = df / (df + F)
        ref = special.betainc(a, 0.5, z)
        got = _lib.lib.jm_host_betainc(a, 0.5, z)
        if ref > 1e-300:
            assert abs(got - ref) <= 1e-9 * ref, (df, F, got, ref)
    assert _lib.lib.jm_host_betainc(a, 0.5, 0.0) == 0.0
    assert _lib.lib.jm_host_betainc(a, 0.5, 1.0) == 1.0


def test_host_chi2_matches_scipy():
    for x in np.concatenate([[0.0], np.logspace(-10, 3.2, 120)]):
        ref = stats.chi2.sf(x, 1)
        got = _lib.lib.jm_host_chi2_sf1(x)
        if ref > 1e-300:
            assert abs(got - ref) <= 1e-12 * ref


def test_host_pvalue_guards_match_the_oracle():
    """The device's p-value helpers (compiled for the host) follow the reference's
    guards as the oracle restates them (likelihood_jax.py:231-251, 668-673): F <= 0 ->
    p = 1, a NaN statistic -> NaN (jnp.maximum / jnp.clip propagate it), a negative LRT
    statistic -> p = 1, NaN log-likelihood -> NaN."""
    df = 97.0
    for f in (-3.0, 0.0, 1e-12, 0.3, 4.0, 250.0, np.inf, np.nan):
        ref = float(ol._f_pvalue(df, np.array([f]), np.array([True]))[0])
        got = _lib.lib.jm_host_f_pvalue(df, f)
        if np.isnan(ref):
            assert np.isnan(got), f
        else:
            assert abs(got - ref) <= 1e-9 * max(ref, 1e-300), (f, got, ref)
    for h1, h0 in ((-10.0, -12.0), (-12.0, -10.0), (-5.0, -5.0), (np.nan, -3.0), (-3.0, np.nan)):
        ref = float(ol.lrt_pvalue(np.array([h1]), h0)[0])
        got = _lib.lib.jm_host_lrt_pvalue(h1, h0)
        if np.isnan(ref):
            assert np.isnan(got), (h1, h0)
        else:
            assert abs(got - ref) <= 1e-12 * ref, (h1, h0, got, ref)


# ------------------------------------------------------------ spectrum plan
def _kinship_spectrum(n=400, p=4000, seed=0):
    from oracle import kinship as ok
    from tests.util import balding_nichols

    G = balding_nichols(n, p, seed=seed)
    d, _ = ol.eigendecompose(ok.centered_kinship(G))
    return d


def test_spectrum_plan_invariants():
    d = _kinship_spectrum()
    pl = plan_spectrum(d)
    assert pl.n_terms <= len(d)
    assert np.all(pl.cl_terms >= 1) and np.all(pl.cl_terms <= MAX_TERMS_PER_CLUSTER)
    assert pl.cl_size.sum() == len(d)
    assert np.all(np.diff(pl.cl_start) > 0)
    for c in range(len(pl.cl_size)):
        s, e = pl.cl_start[c], pl.cl_start[c] + pl.cl_size[c]
        dc = pl.cl_center[c]
        if dc > 0:
            assert np.max(np.abs(d[s:e] - dc)) <= pl.radius * dc * (1 + 1e-12)
        else:
            assert np.all(d[s:e] == dc) and pl.cl_terms[c] == 1
    # exact zero eigenvalue (centred kinship) forms a single-term cluster
    assert (d == 0).sum() >= 1


def test_spectrum_compressed_sums_match_direct():
    """The compressed evaluation the lambda kernel uses reproduces the direct
    n-term sums to ~1e-14 for lambda across [1e-5, 1e5] (pure NumPy check of
    the plan; the device kernel is checked against the oracle on the GPU)."""
    d = _kinship_spectrum(seed=1)
    pl = plan_spectrum(d)
    rng = np.random.default_rng(0)
    v = rng.normal(size=len(d))
    mom = np.zeros(pl.n_terms)
    L = np.zeros(pl.n_terms)
    for t in range(pl.n_terms):
        c, m = pl.term_cluster[t], pl.term_m[t]
        s, e = pl.cl_start[c], pl.cl_start[c] + pl.cl_size[c]
        mom[t] = np.sum(v[s:e] * pl.idx_delta[s:e] ** m)
        L[t] = np.sum(pl.idx_delta[s:e] ** m)
    dc = pl.cl_center[pl.term_cluster]
    m = pl.term_m
    for lam in np.exp(np.linspace(np.log(1e-5), np.log(1e5), 23)):
        h = 1.0 / (1.0 + lam * dc)
        z = lam * h
        kap = h * (-z) ** m
        ell = np.where(m == 0, np.log(np.abs(1 + lam * dc)), (-1.0) ** (m + 1) * z ** m / np.maximum(m, 1))
        direct = np.sum(v / (1 + lam * d))
        assert abs(np.sum(kap * mom) - direct) <= 1e-12 * np.sum(np.abs(v / (1 + lam * d)))
        ld = np.sum(np.log(np.abs(1 + lam * d)))
        assert abs(np.sum(ell * L) - ld) <= 1e-12 * max(1.0, abs(ld))


def test_spectrum_plan_rejects_unsorted():
    with pytest.raises(ValueError):
        plan_spectrum(np.array([1.0, 0.5]))


# ------------------------------------------------------------ SNP statistics
def test_stats_from_counts_match_numpy_float32():
    rng = np.random.default_rng(9)
    G = rng.integers(0, 3, size=(513, 200)).astype(np.float32)
    G[rng.random(G.shape) < 0.03] = np.nan
    G[:, 5] = 1.0  # monomorphic
    G[:, 6] = np.nan  # all missing
    G[:, 7] = np.where(rng.random(513) < 0.5, np.nan, 2.0)  # monomorphic with missing
    counts = np.stack([(G == 0).sum(0), (G == 1).sum(0), (G == 2).sum(0), np.isnan(G).sum(0)], 1)
    mean, miss, poly = stats_from_counts(counts, np.float32)
    with np.errstate(invalid="ignore"), pytest.warns(RuntimeWarning):
        ref_mean = np.nan_to_num(np.nanmean(G, axis=0), nan=0.0).astype(np.float64)
        ref_var = np.nan_to_num(np.nanvar(G, axis=0), nan=0.0)
    assert np.array_equal(mean, ref_mean)
    assert np.array_equal(miss, np.isnan(G).sum(0))
    assert np.array_equal(poly, ref_var > 0)


# ------------------------------------------------------------ writers
def test_writer_reproduces_gemma_lines():
    path = GOLDEN / "gemma_score" / "gemma_score.assoc.txt"
    lines = open(path).read().splitlines()
    head = lines[0].split("\t")
    for line in lines[1:6]:
        f = dict(zip(head, line.split("\t")))
        r = AssocResult(chr=f["chr"], rs=f["rs"], ps=int(f["ps"]), n_miss=int(f["n_miss"]),
                        allele1=f["allele1"], allele0=f["allele0"], af=float(f["af"]),
                        beta=float(f["beta"]), se=float(f["se"]), p_score=float(f["p_score"]))
        assert format_line(r, "score") == line


def test_writer_all_columns_order():
    r = AssocResult("1", "rs1", 100, 0, "A", "G", 0.25, 1.5, 0.25, logl_H1=-10.0, l_remle=2.0,
                    p_wald=1e-3, p_score=2e-3, l_mle=3.0, p_lrt=4e-3)
    cols = format_line(r, "all").split("\t")
    assert cols[6] == "0.250"
    assert cols[7:] == [f"{v:.6e}" for v in (1.5, 0.25, -10.0, 2.0, 3.0, 1e-3, 4e-3, 2e-3)]


@pytest.mark.parametrize("n_cvt", [1, 2, 4, 7])
def test_pair_tables_reproduce_pab_row0(n_cvt):
    """The moment layout the device kernels use (engine.pair_tables) rebuilds the
    oracle's Pab row 0 for every packed pair at any lambda (NumPy emulation)."""
    from jamma_amd.lmm.engine import pair_tables

    rng = np.random.default_rng(n_cvt)
    n = 150
    d = np.sort(rng.uniform(0.0, 4.0, n))
    d[:2] = 0.0
    pl = plan_spectrum(d)
    UtW = rng.normal(size=(n, n_cvt))
    Uty = rng.normal(size=n)
    Utx = rng.normal(size=(n, 3))

    def moments(z, partners):  # rows: partners..., z*z
        out = np.zeros((len(partners) + 1, pl.n_terms))
        for t in range(pl.n_terms):
            c, m = pl.term_cluster[t], pl.term_m[t]
            s, e = pl.cl_start[c], pl.cl_start[c] + pl.cl_size[c]
            w = pl.idx_delta[s:e] ** m
            for b, v in enumerate(partners):
                out[b, t] = np.sum(z[s:e] * v[s:e] * w)
            out[-1, t] = np.sum(z[s:e] ** 2 * w)
        return out

    partners = [UtW[:, k] for k in range(n_cvt)] + [Uty]
    Zs = partners + [np.ones(n)]
    smom = np.stack([moments(z, partners) for z in Zs])  # (c+2, c+2, T)
    pair_row, shared_src = pair_tables(n_cvt)
    dc = pl.cl_center[pl.term_cluster]
    for lam in (1e-5, 0.3, 7.0, 1e5):
        h = 1 / (1 + lam * dc)
        kap = h * (-(lam * h)) ** pl.term_m
        ref = ol.pab_batch(n_cvt, 1.0 / (lam * d + 1.0), UtW, Uty, Utx)[:, 0, :]
        for s in range(3):
            xm = moments(Utx[:, s], partners)
            for j in range(len(pair_row)):
                row = xm[pair_row[j]] if pair_row[j] >= 0 else smom[shared_src[j][0], shared_src[j][1]]
                got = np.sum(kap * row)
                assert abs(got - ref[s, j]) <= 1e-11 * max(1.0, abs(ref[s, j])), (lam, s, j)


@pytest.mark.parametrize("seg_max", [None, 7, 64])
def test_segment_plan_pieces_fold_to_cluster_moments(seg_max):
    """Cluster pieces cover each cluster exactly, in order, with at most seg_max
    indices; per-piece partial sums folded in piece order equal the direct
    cluster sums (NumPy emulation of k_mom_seg + k_mom_fold)."""
    from jamma_amd.lmm.spectrum import segment_plan

    d = _kinship_spectrum(seed=2)
    pl = plan_spectrum(d)
    if seg_max is not None:
        pl = segment_plan(pl, seg_max)
    assert pl.seg_len.min() >= 1 and pl.seg_len.max() <= pl.seg_max
    for c in range(len(pl.cl_size)):
        s0, k = pl.cl_seg0[c], pl.cl_nseg[c]
        assert np.all(pl.seg_cluster[s0:s0 + k] == c)
        assert pl.seg_start[s0] == pl.cl_start[c]
        assert np.sum(pl.seg_len[s0:s0 + k]) == pl.cl_size[c]
        assert np.all(np.diff(pl.seg_start[s0:s0 + k]) == pl.seg_len[s0:s0 + k - 1])
    assert pl.n_slots == int(np.sum(pl.cl_terms[pl.seg_cluster]))
    rng = np.random.default_rng(3)
    z = rng.normal(size=len(d))
    work = np.zeros(pl.n_slots)
    for q in range(len(pl.seg_start)):
        c = pl.seg_cluster[q]
        sl = slice(pl.seg_start[q], pl.seg_start[q] + pl.seg_len[q])
        for m in range(pl.cl_terms[c]):
            work[pl.seg_slot[q] + m] = np.sum(z[sl] * pl.idx_delta[sl] ** m)
    for t in range(pl.n_terms):
        c, m = pl.term_cluster[t], pl.term_m[t]
        folded = sum(work[pl.seg_slot[pl.cl_seg0[c] + p] + m] for p in range(pl.cl_nseg[c]))
        sl = slice(pl.cl_start[c], pl.cl_start[c] + pl.cl_size[c])
        direct = np.sum(z[sl] * pl.idx_delta[sl] ** m)
        assert abs(folded - direct) <= 1e-12 * max(1.0, np.sum(np.abs(z[sl] * pl.idx_delta[sl] ** m)))


def test_compact_wy_factor_reproduces_reflector_product():
    """T from the V^T V products (LAPACK dlarft, forward columnwise) gives
    H_0 H_1 ... H_{k-1} = I - V T V^T, including a trivial reflector (tau = 0)."""
    from jamma_amd.lmm.eigen import _compact_wy_t

    rng = np.random.default_rng(2)
    L, k = 40, 9
    V = np.tril(rng.normal(size=(L, k)), -1) + np.eye(L, k)  # unit lower trapezoidal
    tau = 2.0 / np.sum(V * V, axis=0)  # orthogonal reflectors
    tau[4] = 0.0
    H = np.eye(L)
    for i in range(k):
        H = H @ (np.eye(L) - tau[i] * np.outer(V[:, i], V[:, i]))
    T = _compact_wy_t(tau, V.T @ V)
    assert np.allclose(np.triu(T), T)
    assert np.max(np.abs(np.eye(L) - V @ T @ V.T - H)) < 1e-13


def test_library_eigen_limit_is_the_32_bit_bound():
    from jamma_amd.lmm import eigen as eg

    assert eg._LIB_MAX_N ** 2 <= 2**31 - 1 < (eg._LIB_MAX_N + 1) ** 2
    assert eg._LARGE_LEAF <= eg._LIB_MAX_N


@pytest.mark.parametrize("own_min", [8192, 10**9])
def test_eigh_device_never_routes_beyond_32_bits_to_rocsolver(monkeypatch, own_min):
    """eigh_device sends every n > 46,340 to the 64-bit path, even if the switch-over
    size were raised past it, and never opens rocSOLVER for it. A meta tensor stands in
    for the matrix (shape only), and the solvers are replaced by stubs, so nothing
    touches a GPU."""
    import torch

    from jamma_amd.lmm import eigen as eg

    monkeypatch.setattr(eg, "_OWN_MIN_N", own_min)
    assert eg._select_solver(eg._LIB_MAX_N + 1) is eg._eigh_large
    assert eg._select_solver(50_000) is eg._eigh_large
    assert eg._select_solver(7) is eg._dsyevd
    assert eg._select_solver(eg._LIB_MAX_N) is (eg._dsyevd if own_min > eg._LIB_MAX_N
                                               else eg._eigh_large)

    def no_solver(*a, **k):
        raise AssertionError("rocSOLVER opened for a matrix beyond its 32-bit limit")

    calls = []

    def large_stub(Kd, progress=None):
        calls.append(tuple(Kd.shape))
        return torch.arange(Kd.shape[0], dtype=torch.float64) + 1.0, None

    monkeypatch.setattr(eg, "_solver", no_solver)
    monkeypatch.setattr(eg, "_dsyevd", no_solver)
    monkeypatch.setattr(eg, "_eigh_large", large_stub)
    monkeypatch.setattr(eg, "require_gpu", lambda: torch.device("meta"))
    monkeypatch.setattr(eg, "to_dev", lambda a, dtype=None, device=None: a)
    K = torch.empty((50_000, 50_000), dtype=torch.float64, device="meta")
    d, U = eg.eigh_device(K)
    assert calls == [(50_000, 50_000)] and d.shape == (50_000,) and U is None
    # the 32-bit routine itself refuses such a size before touching the device
    monkeypatch.undo()
    with pytest.raises(ValueError, match="32-bit"):
        eg._dsyevd(K)


def _merge_case(kind, seed):
    rng = np.random.default_rng(seed)
    if kind == "coincident":  # every pole twice (mirror-image halves): Givens deflation
        d1 = np.sort(rng.normal(size=40))
        dc = np.concatenate([d1, d1])
    elif kind == "tiny-weights":
        dc = np.concatenate([np.sort(rng.normal(size=30)), np.sort(rng.normal(size=30))])
    else:
        dc = np.concatenate([np.sort(rng.normal(size=25)), np.sort(rng.normal(size=35))])
    z = rng.normal(size=len(dc))
    if kind == "tiny-weights":
        z[rng.choice(len(z), 12, replace=False)] = 1e-19
    z *= np.sqrt(2.0) / np.linalg.norm(z)  # torn weights have |z| = sqrt 2
    return dc, z, 0.7


@pytest.mark.parametrize("kind", ["generic", "coincident", "tiny-weights"])
def test_merge_deflation_keeps_the_spectrum(kind):
    """dlaed2's deflation: the deflated values plus the spectrum of the reduced secular
    problem are the spectrum of D + rho z z^T, and in the rotated basis the deflated
    directions decouple (the convention jm_apply_rotations applies to the merge matrix)."""
    from jamma_amd.lmm.eigen import _deflate

    dc, z, rho = _merge_case(kind, 3)
    Mfull = np.diag(dc) + rho * np.outer(z, z)
    order, ds, zs, nd, dval, drow, rots, rho2 = _deflate(dc, z, rho)
    assert len(nd) + len(dval) == len(dc)
    if kind == "coincident":
        assert len(rots) > 0
    if kind == "tiny-weights":
        assert len(dval) >= 12
    dl = np.array([ds[k] for k in nd])
    w = np.array([zs[k] for k in nd])
    assert np.all(np.diff(dl) > 0)  # strictly ascending poles for the secular solver
    red = np.linalg.eigvalsh(np.diag(dl) + rho2 * np.outer(w, w)) if len(nd) else np.zeros(0)
    got = np.sort(np.concatenate([red, dval]))
    ref = np.linalg.eigvalsh(Mfull)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    # rotated basis: Q = G_1 G_2 ... (columns p, q of each G mixed by (c, s))
    Q = np.eye(len(dc))
    for p, q, c, s in rots:
        qp, qq = Q[:, p].copy(), Q[:, q].copy()
        Q[:, p], Q[:, q] = c * qp + s * qq, -s * qp + c * qq
    R = Q.T @ Mfull @ Q
    for v, r in zip(dval, drow):
        off = np.abs(R[r]).copy()
        off[r] = 0.0
        assert abs(R[r, r] - v) <= 1e-13 * np.max(np.abs(ref))
        assert np.max(off) <= 1e-13 * np.max(np.abs(ref))


def test_noalloc_moments_rejects_more_than_its_scratch():
    """jm_spectrum_moments_noalloc checks the partials' size against its fixed scratch
    before any device call: a plan with fake (never dereferenced) array pointers and
    one column too many returns JM_ERR_ARG with a message, on a machine without a GPU."""
    import ctypes

    pl = _lib.SpectrumPlanC()
    pl.n, pl.n_terms, pl.n_clusters, pl.n_seg, pl.n_slots = 1000, 200, 50, 60, 250
    for f in ("cl_terms", "cl_toff", "cl_seg0", "cl_nseg", "seg_start", "seg_len",
              "seg_cluster", "seg_slot", "delta", "term_cluster", "term_m"):
        setattr(pl, f, 8)
    cap = int(_lib.lib.jm_moments_scratch_bytes())
    per_col = int(_lib.lib.jm_moments_work_bytes(ctypes.byref(pl), 1, 5))
    assert per_col == 6 * 250 * 8
    too_many = cap // per_col + 1
    rc = _lib.lib.jm_spectrum_moments_noalloc(8, 0, too_many, too_many, 8, 1000, 5,
                                              ctypes.byref(pl), 8, 200, None)
    assert rc == -1 and "scratch" in _lib.last_error()
    # zero columns is a no-op that succeeds without touching a device
    assert _lib.lib.jm_spectrum_moments_noalloc(8, 0, 1, 0, 8, 1000, 5, ctypes.byref(pl), 8,
                                                200, None) == 0
