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


def test_hwe_pvalues_closed_form():
    """compute_hwe_pvalues (core/snp_filter.py:94-150) on crafted genotype counts:
    counts in exact HWE give chi2 = 0 -> p = 1; (50, 0, 50) has p = 1/2 and expected
    (25, 50, 25), so chi2 = 25 + 50 + 25 = 100; (81, 18, 1) is exact HWE at p = 0.9;
    (60, 20, 20): p = 0.7, expected (49, 42, 9), chi2 = 121/49 + 484/42 + 121/9;
    monomorphic and empty SNPs give NaN chi2 -> p = 1 by the reference's convention."""
    from jamma_amd.lmm.runner_streaming import hwe_pvalues

    aa = np.array([25, 50, 81, 60, 100, 0, 0])
    ab = np.array([50, 0, 18, 20, 0, 0, 0])
    bb = np.array([25, 50, 1, 20, 0, 0, 40])
    chi = np.array([0.0, 100.0, 0.0, 121 / 49 + 484 / 42 + 121 / 9, 0.0, 0.0, 0.0])
    got = hwe_pvalues(aa, ab, bb)
    want = stats.chi2.sf(chi, 1)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-300)
    assert got[0] == 1.0 and got[4] == 1.0 and got[5] == 1.0 and got[6] == 1.0
    assert abs(got[1] - special.erfc(np.sqrt(50.0))) <= 1e-12 * got[1]
    # the filter keeps p >= threshold (runner_streaming.py:298-305)
    assert list(got >= 1e-3) == [True, False, True, False, True, True, True]


# ------------------------------------------------------------ native writer
FIXTURE_FILES = {
    "wald": "gemma_synthetic/gemma_assoc.assoc.txt",
    "lrt": "gemma_synthetic/gemma_lrt.assoc.txt",
    "score": "gemma_score/gemma_score.assoc.txt",
    "all": "gemma_all_tests/gemma_all_covar.assoc.txt",
}


@pytest.mark.parametrize("kind", list(FIXTURE_FILES))
def test_native_rows_reproduce_gemma_files(kind):
    """jm_format_assoc_rows re-creates every data line of the GEMMA fixture files byte
    for byte from their parsed fields (af as .3f, statistics as .6e), in the files'
    own column sets (GEMMA 0.98's -lmm 1 / -lmm 4 files carry no logl_H1 column)."""
    from jamma_amd.lmm.io import format_rows, snp_text_parts

    lines = (GOLDEN / FIXTURE_FILES[kind]).read_text().splitlines()
    head = lines[0].split("\t")
    rows = [dict(zip(head, ln.split("\t"))) for ln in lines[1:]]
    infos = [{"chr": r["chr"], "rs": r["rs"], "pos": int(r["ps"]), "a1": r["allele1"],
              "a0": r["allele0"]} for r in rows]
    stat_cols = head[7:]
    stats = np.array([[float(r[c]) for c in stat_cols] for r in rows])
    data = format_rows(snp_text_parts(infos), [int(r["n_miss"]) for r in rows],
                       [float(r["af"]) for r in rows], stats, n_threads=3)
    assert data.decode().splitlines() == lines[1:]


def test_native_rows_match_python_formatting_on_hard_values():
    """Values where a formatter could differ from Python's format(): ties at the 7th
    significant digit, subnormals, huge and tiny exponents, +-0, +-inf, NaN of either
    sign, af rounding at .0005; checked against the reference's f-strings."""
    from jamma_amd.lmm.io import format_rows, snp_text_parts

    rng = np.random.default_rng(4)
    special_vals = [0.0, -0.0, np.inf, -np.inf, np.nan, -np.nan, 5e-324, -2.5e-308,
                    1.0000005, 1.0000015, 2.5e-7, 9.9999995e10, 1e300, -1e-300,
                    0.1, 1 / 3, 123456.5, 0.00012345675]
    rand = (rng.standard_normal(3000) * 10.0 ** rng.integers(-300, 300, 3000)).tolist()
    vals = np.array(special_vals + rand)
    nst = 5
    k = len(vals) // nst
    stats = vals[: k * nst].reshape(k, nst)
    af = np.concatenate([[0.0005, 0.0015, 0.0025, 0.1235, 0.9995, 1.0, np.nan],
                         rng.uniform(0, 1, k - 7)])
    n_miss = rng.integers(0, 10**6, k)
    infos = [{"chr": str(1 + j % 22), "rs": f"rs{j}", "pos": j * 7, "a1": "A", "a0": "T"}
             for j in range(k)]
    got = format_rows(snp_text_parts(infos), n_miss, af, stats, n_threads=7).decode()
    want = "".join("\t".join([i["chr"], i["rs"], str(i["pos"]), str(int(m)), "A", "T",
                              f"{a:.3f}"] + [f"{v:.6e}" for v in row]) + "\n"
                   for i, m, a, row in zip(infos, n_miss, af, stats))
    assert got == want


def test_native_g10_matches_reference_writer():
    """%.10g rows (kinship / eigen files) equal the reference's f"{v:.10g}" joins."""
    from jamma_amd.lmm.io import format_matrix_g10

    rng = np.random.default_rng(5)
    A = rng.standard_normal((37, 23)) * 10.0 ** rng.integers(-12, 12, (37, 23))
    A[0, :6] = [0.0, -0.0, np.nan, 1e-5, 123456789012.0, 2.0 / 3.0]
    want = "".join("\t".join(f"{v:.10g}" for v in row) + "\n" for row in A)
    assert format_matrix_g10(A, n_threads=4).decode() == want
    d = rng.standard_normal(11)
    assert format_matrix_g10(d).decode() == "".join(f"{v:.10g}\n" for v in d)


def test_incremental_writer_tiles_keep_order(tmp_path):
    """write_rows queues tiles on the background thread; mixing them with per-result
    write() keeps submission order, and the file equals the per-result writer's."""
    from jamma_amd.lmm.io import IncrementalAssocWriter, snp_text_parts
    from jamma_amd.lmm.runner_streaming import _make_result

    rng = np.random.default_rng(6)
    n = 1000
    out = rng.standard_normal((n, _lib.JM_NOUT))
    infos = [{"chr": "2", "rs": f"rs{j}", "pos": 10 * j, "a1": "C", "a0": "G"} for j in range(n)]
    af = rng.uniform(0, 1, n)
    miss = rng.integers(0, 50, n)
    for mode, kind in ((1, "wald"), (2, "lrt"), (3, "score"), (4, "all")):
        a, b = tmp_path / f"a{mode}.txt", tmp_path / f"b{mode}.txt"
        with IncrementalAssocWriter(a, kind) as w:
            for t0 in range(0, 600, 150):
                sl = slice(t0, t0 + 150)
                w.write_rows(snp_text_parts(infos[sl]), miss[sl], af[sl], out[sl])
            for j in range(600, n):
                w.write(_make_result(mode, infos[j], af[j], miss[j], out[j]))
        with IncrementalAssocWriter(b, kind) as w:
            for j in range(n):
                w.write(_make_result(mode, infos[j], af[j], miss[j], out[j]))
        assert a.read_bytes() == b.read_bytes()


# ------------------------------------------------------------ kinship / eigen files
def test_kinship_file_parse_matches_loadtxt(tmp_path):
    """The native parser reads the GEMMA kinship fixture bit-identically to
    np.loadtxt, and write_kinship_matrix reproduces the reference's %.10g text."""
    from jamma_amd.kinship.io import read_kinship_matrix, write_kinship_matrix

    path = GOLDEN / "gemma_synthetic" / "gemma_kinship.cXX.txt"
    K = read_kinship_matrix(path, n_samples=100)
    assert np.array_equal(K, np.loadtxt(path))
    out = tmp_path / "k.cXX.txt"
    write_kinship_matrix(K, out)
    assert out.read_text() == "".join("\t".join(f"{v:.10g}" for v in row) + "\n" for row in K)
    with pytest.raises(ValueError, match="does not match expected"):
        read_kinship_matrix(path, n_samples=99)


def test_eigen_files_round_trip_and_validation(tmp_path):
    """write_eigen_files / read_eigen_files (lmm/eigen_io.py:42-164): GEMMA text at 10
    significant digits round-trips to 5e-10 relative; the reference's shape checks
    raise with its messages."""
    from jamma_amd.lmm.eigen_io import read_eigen_files, write_eigen_files

    rng = np.random.default_rng(8)
    n = 57
    d = np.sort(rng.uniform(0, 3, n))
    U, _ = np.linalg.qr(rng.normal(size=(n, n)))
    dp, up = write_eigen_files(d, U, tmp_path / "e", "cohort")
    assert dp.name == "cohort.eigenD.txt" and up.name == "cohort.eigenU.txt"
    assert dp.read_text() == "".join(f"{v:.10g}\n" for v in d)
    d2, U2 = read_eigen_files(dp, up, n_samples=n)
    assert np.array_equal(d2, np.loadtxt(dp)) and np.array_equal(U2, np.loadtxt(up))
    np.testing.assert_allclose(d2, d, rtol=5e-10, atol=0)
    np.testing.assert_allclose(U2, U, rtol=5e-10, atol=0)
    with pytest.raises(ValueError, match="expected n_samples=58"):
        read_eigen_files(dp, up, n_samples=58)
    bad_u = tmp_path / "rect.txt"
    bad_u.write_text("1\t2\t3\n4\t5\t6\n")
    with pytest.raises(ValueError, match="must be square"):
        read_eigen_files(dp, bad_u)
    small_u = tmp_path / "small.txt"
    small_u.write_text("1\t0\n0\t1\n")
    with pytest.raises(ValueError, match="does not match eigenvector"):
        read_eigen_files(dp, small_u)
    ragged = tmp_path / "ragged.txt"
    ragged.write_text("1\t2\n3\n")
    with pytest.raises(ValueError, match="different lengths"):
        read_eigen_files(dp, ragged)


def test_pipeline_flag_validation(tmp_path):
    """PipelineRunner.validate_inputs (pipeline.py:142-238) without a GPU: -hwe with
    -loco, hwe out of range, -k with -loco, -d without -u."""
    from jamma_amd.pipeline import PipelineConfig, PipelineRunner

    pre = GOLDEN / "gemma_synthetic" / "test"

    def check(match, **kw):
        with pytest.raises(ValueError, match=match):
            PipelineRunner(PipelineConfig(bfile=pre, **kw)).validate_inputs()

    check("not yet supported with -loco", loco=True, hwe_threshold=1e-6)
    check("must be >= 0", hwe_threshold=-0.1)
    check(r"must be in \[0, 1\]", hwe_threshold=1.5)
    check("mutually exclusive", loco=True, kinship_file=GOLDEN / "gemma_synthetic" /
          "gemma_kinship.cXX.txt")
    check("must be provided together", eigenvalue_file=tmp_path / "x.eigenD.txt")
    PipelineRunner(PipelineConfig(bfile=pre, hwe_threshold=1e-6)).validate_inputs()


def test_guard_rows_selects_p_values_near_thresholds():
    """Rows whose p-value (of the mode's tests) lies within the band, in log10 units,
    of 0.05, 0.01, 1e-3 or 5e-8 are the ones the mixed path recomputes in FP64; NaN and
    p = 0 never qualify."""
    from jamma_amd.lmm.engine import guard_rows

    out = np.full((8, _lib.JM_NOUT), 0.5)
    out[1, 2] = 0.05 * 10 ** 1.5e-4     # Wald just above 0.05, inside the band
    out[2, 7] = 0.01 * 10 ** -3e-4      # LRT outside the band
    out[3, 8] = 5e-8                    # Score on the threshold
    out[4, 2] = np.nan
    out[5, 2] = 0.0
    out[6, 7] = 1e-3 * 10 ** -1e-4      # LRT inside
    out[7, 2] = 1e-3 * 10 ** 5e-5       # Wald inside, but mode 2 does not test Wald
    assert list(guard_rows(out, 4)) == [1, 3, 6, 7]
    assert list(guard_rows(out, 1)) == [1, 7]
    assert list(guard_rows(out, 2)) == [6]
    assert list(guard_rows(out, 3)) == [3]
    assert list(guard_rows(out, 4, band=1e-6)) == [3]


def test_jamma_import_alias():
    """`jamma.*` imports of a reference caller resolve to this package's modules (the
    same module objects), with runner_jax mapped to the module holding the in-memory
    runner; modules outside the hot path raise ImportError."""
    import importlib

    import jamma
    import jamma_amd.kinship.compute
    import jamma_amd.lmm.runner_streaming

    assert importlib.import_module("jamma.kinship.compute") is jamma_amd.kinship.compute
    assert importlib.import_module("jamma.lmm.runner_jax") is jamma_amd.lmm.runner_streaming
    from jamma.kinship import compute_centered_kinship
    from jamma.lmm import AssocResult as A1
    from jamma.lmm.stats import AssocResult as A2

    assert A1 is A2 and compute_centered_kinship is jamma_amd.kinship.compute.compute_centered_kinship
    assert jamma.gwas is jamma_amd.gwas
    with pytest.raises(ImportError):
        importlib.import_module("jamma.cli")
