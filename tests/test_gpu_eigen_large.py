"""The large-cohort eigendecomposition (64-bit offsets, n beyond 46,340) at small n.

The path is forced at small sizes with small panels and tear thresholds, so that
every stage runs several times: blocked Householder panels with trailing updates,
several levels of divide and conquer, both kinds of deflation (small weights and
Givens rotations of coincident poles) and the compact-WY back-transformation.
Checks against LAPACK (numpy): the spectrum, the residual K U - U diag(d), and
U^T U = I. A structural mistake in any stage (a wrong reflector, a missed
rotation, a transposed merge) breaks the residual or the orthogonality by many
orders of magnitude.
"""

import numpy as np
import pytest
import torch

from tests.util import balding_nichols

pytestmark = pytest.mark.gpu


def _check(K, d, U, tol=1e-11):
    K = np.asarray(K)
    d = d.cpu().numpy() if isinstance(d, torch.Tensor) else np.asarray(d)
    U = U.cpu().numpy()
    n = K.shape[0]
    scale = max(1.0, float(np.max(np.abs(np.linalg.eigvalsh(K)))))
    ref = np.linalg.eigvalsh(K)
    assert np.all(np.diff(d) >= 0)
    assert np.max(np.abs(d - ref)) <= tol * scale, np.max(np.abs(d - ref))
    assert np.max(np.abs(K @ U - U * d[None, :])) <= tol * scale
    assert np.max(np.abs(U.T @ U - np.eye(n))) <= 1e-12 * max(1.0, n / 100)


def _kinship(n, p, seed):
    G = balding_nichols(n, p, seed=seed)
    mu = np.nanmean(G, axis=0)
    X = np.where(np.isnan(G), mu, G) - mu
    return X @ X.T / p


@pytest.mark.parametrize("n,nb,leaf", [(2, 4, 1), (37, 8, 8), (300, 32, 50), (1031, 64, 200),
                                       (1500, 128, 400)])
def test_large_path_matches_lapack(n, nb, leaf):
    from jamma_amd.lmm.eigen import _eigh_large

    K = _kinship(n, 2 * n + 50, seed=n)
    d, U = _eigh_large(torch.as_tensor(K, device="cuda"), nb=nb, leaf=leaf)
    _check(K, d, U)


def test_large_path_random_symmetric():
    from jamma_amd.lmm.eigen import _eigh_large

    rng = np.random.default_rng(3)
    X = rng.normal(size=(517, 517))
    K = (X + X.T) / 2
    d, U = _eigh_large(torch.as_tensor(K, device="cuda"), nb=24, leaf=70)
    _check(K, d, U)


def test_large_path_repeated_and_zero_eigenvalues():
    """Three identical blocks (every eigenvalue threefold) and a rank-20 kinship of
    300 samples (280 zero eigenvalues): the deflation branches."""
    from jamma_amd.lmm.eigen import _eigh_large

    rng = np.random.default_rng(4)
    B = rng.normal(size=(100, 100))
    K1 = np.kron(np.eye(3), B @ B.T / 100)
    d, U = _eigh_large(torch.as_tensor(K1, device="cuda"), nb=16, leaf=40)
    _check(K1, d, U)
    X = rng.normal(size=(300, 20))
    K2 = X @ X.T / 20
    d, U = _eigh_large(torch.as_tensor(K2, device="cuda"), nb=32, leaf=60)
    _check(K2, d, U)


def test_tridiagonal_merge_with_coincident_poles():
    """T2 = J T1 J (mirror image): the two halves have the same spectrum, so the
    merge deflates by Givens rotations; Wilkinson's W+ has close pairs."""
    from jamma_amd.lmm.eigen import _tri_eig

    rng = np.random.default_rng(5)
    m = 120
    d1, e1 = rng.normal(size=m), rng.uniform(0.1, 1.0, m - 1)
    cases = [(np.concatenate([d1, d1[::-1]]), np.concatenate([e1, [0.5], e1[::-1]]), 120),
             (np.abs(np.arange(-100, 101)).astype(float), np.ones(200), 30)]
    for d, e, leaf in cases:
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w, Q = _tri_eig(d, e, torch.device("cuda"), leaf)
        Qh = Q.cpu().numpy().T  # rows of Q are the eigenvectors
        n = len(d)
        assert np.max(np.abs(w - np.linalg.eigvalsh(T))) <= 1e-12 * np.max(np.abs(w))
        assert np.max(np.abs(T @ Qh - Qh * w[None, :])) <= 1e-12 * np.max(np.abs(w))
        assert np.max(np.abs(Qh.T @ Qh - np.eye(n))) <= 1e-12


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_transposes(ta, tb):
    from jamma_amd import _lib
    from jamma_amd.device import stream_ptr

    rng = np.random.default_rng(10 * ta + tb)
    m, n, k = 203, 77, 45
    A = rng.normal(size=(k, m) if ta else (m, k))
    B = rng.normal(size=(n, k) if tb else (k, n))
    C0 = rng.normal(size=(m, n))
    ref = -0.5 * ((A.T if ta else A) @ (B.T if tb else B)) + 2.0 * C0
    # column-major device buffers: a (r x c) matrix is the row-major tensor of its transpose
    Ad = torch.as_tensor(np.ascontiguousarray(A.T), device="cuda")
    Bd = torch.as_tensor(np.ascontiguousarray(B.T), device="cuda")
    Cd = torch.as_tensor(np.ascontiguousarray(C0.T), device="cuda")
    _lib.call("jm_dgemm", ta, tb, m, n, k, -0.5, Ad.data_ptr(), A.shape[0], Bd.data_ptr(),
              B.shape[0], 2.0, Cd.data_ptr(), m, 0, stream_ptr())
    got = Cd.cpu().numpy().T
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_dgemm_lower_only_leaves_upper_tiles():
    from jamma_amd import _lib
    from jamma_amd.device import stream_ptr

    rng = np.random.default_rng(11)
    n, k = 300, 16
    A, B = rng.normal(size=(n, k)), rng.normal(size=(n, k))
    Cd = torch.zeros((n, n), dtype=torch.float64, device="cuda")
    Ad = torch.as_tensor(A.T.copy(), device="cuda")  # kept alive across the call
    Bd = torch.as_tensor(B.T.copy(), device="cuda")
    _lib.call("jm_dgemm", 0, 1, n, n, k, 1.0, Ad.data_ptr(), n, Bd.data_ptr(), n, 0.0,
              Cd.data_ptr(), n, 1, stream_ptr())
    got = Cd.cpu().numpy().T  # column-major C
    ref = A @ B.T
    low = np.tril(np.ones((n, n), bool))
    assert np.max(np.abs(got[low] - ref[low])) <= 1e-12 * np.max(np.abs(ref))
    # tiles entirely above the diagonal (i < j for every element) are untouched
    i_tile, j_tile = np.arange(n)[:, None] // 64, np.arange(n)[None, :] // 64
    untouched = (i_tile + 1) * 64 <= j_tile * 64
    assert untouched.any() and np.all(got[untouched] == 0.0)


def test_eigh_device_routes_large_cohorts(monkeypatch):
    """eigh_device sends large cohorts through the 64-bit path; the thresholding and
    warnings are the same."""
    import jamma_amd.lmm.eigen as eg

    K = _kinship(260, 120, seed=7)  # rank-deficient: p < n
    monkeypatch.setattr(eg, "_OWN_MIN_N", 100)
    monkeypatch.setattr(eg, "_LARGE_LEAF", 64)
    with pytest.warns(UserWarning, match="close to zero"):
        d, U = eg.eigh_device(K)
    ref = np.linalg.eigvalsh(K)
    assert int(np.sum(d == 0.0)) == int(np.sum(np.abs(ref) < 1e-10)) >= 260 - 120
    Un = U.cpu().numpy()
    assert np.max(np.abs(K @ Un - Un * d[None, :])) <= 1e-11 * np.max(d)


@pytest.mark.parametrize("N,lda", [(1, 1), (63, 70), (64, 64), (65, 100), (1000, 1003), (4097, 4100)])
def test_symv_lower_reads_the_lower_triangle(N, lda):
    """y = A x from the lower triangle only: the strictly upper part holds garbage."""
    from jamma_amd import _lib
    from jamma_amd.device import stream_ptr

    rng = np.random.default_rng(N)
    S = rng.normal(size=(N, N))
    S = S + S.T
    stored = np.full((lda, N), np.nan)  # column-major N x N with leading dimension lda
    stored[:N, :] = np.where(np.tril(np.ones((N, N), bool)), S, 1e300)
    x = rng.normal(size=N)
    Ad = torch.as_tensor(np.ascontiguousarray(stored.T), device="cuda")
    xd = torch.as_tensor(x, device="cuda")
    yd = torch.full((N,), np.nan, dtype=torch.float64, device="cuda")
    nw = int(_lib.lib.jm_symv_work_doubles(N))
    wk = torch.full((nw,), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("jm_symv_lower", N, Ad.data_ptr(), lda, xd.data_ptr(), yd.data_ptr(), wk.data_ptr(),
              nw, stream_ptr())
    ref = S @ x
    assert np.max(np.abs(yd.cpu().numpy() - ref)) <= 1e-12 * np.max(np.abs(ref)) * np.sqrt(N)


@pytest.mark.parametrize("splits", [1, 3, 16])
def test_dgemm_splitk_matches(splits):
    from jamma_amd import _lib
    from jamma_amd.device import stream_ptr

    rng = np.random.default_rng(splits)
    m, n, k = 64, 300, 2000
    A = rng.normal(size=(k, m))  # op(A) = A^T
    B = rng.normal(size=(k, n))
    C0 = rng.normal(size=(m, n))
    ref = 1.5 * (A.T @ B) - C0
    Ad = torch.as_tensor(np.ascontiguousarray(A.T), device="cuda")
    Bd = torch.as_tensor(np.ascontiguousarray(B.T), device="cuda")
    Cd = torch.as_tensor(np.ascontiguousarray(C0.T), device="cuda")
    wk = torch.full((splits * m * n,), np.nan, dtype=torch.float64, device="cuda")
    _lib.call("jm_dgemm_splitk", 1, 0, m, n, k, 1.5, Ad.data_ptr(), k, Bd.data_ptr(), k, -1.0,
              Cd.data_ptr(), m, splits, wk.data_ptr(), wk.numel(), stream_ptr())
    got = Cd.cpu().numpy().T
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)) * np.sqrt(k)


def test_dgemm2_two_pairs_in_one_pass():
    from jamma_amd import _lib
    from jamma_amd.device import stream_ptr

    rng = np.random.default_rng(12)
    n, k = 333, 64
    A, W = rng.normal(size=(n, k)), rng.normal(size=(n, k))
    C0 = rng.normal(size=(n, n))
    ref = C0 - (A @ W.T + W @ A.T)
    Ad = torch.as_tensor(A.T.copy(), device="cuda")
    Wd = torch.as_tensor(W.T.copy(), device="cuda")
    Cd = torch.as_tensor(C0.T.copy(), device="cuda")
    _lib.call("jm_dgemm2", 0, 1, n, n, k, -1.0, Ad.data_ptr(), Wd.data_ptr(), Wd.data_ptr(),
              Ad.data_ptr(), n, n, 1.0, Cd.data_ptr(), n, 0, stream_ptr())
    got = Cd.cpu().numpy().T
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("n,p,leaf", [(8192, 4096, None), (12_289, 24_000, None),
                                      (12_289, 24_000, 5000)])
def test_production_sizes_residual_and_orthogonality(n, p, leaf):
    """eigh_device where it selects the 64-bit path (n >= 8,192), at its production
    panel width: kinship-like matrices built on the device (rank-deficient at
    n = 8,192 with p = n / 2; full rank and not a multiple of the 128-column panel at
    n = 12,289, also with the tridiagonal problem torn into 5,000-row blocks so the
    divide-and-conquer merges run at this size). max |K U - U D| <= 1e-11 max |d| and
    max |U^T U - I| <= 1e-12 (torch FP64 GEMMs, independent of the library)."""
    from jamma_amd.lmm import eigen
    from tests.util import eig_residuals

    g = torch.Generator(device="cuda").manual_seed(n + p)
    af = torch.rand(p, device="cuda", dtype=torch.float64, generator=g) * 0.45 + 0.05
    X = (torch.rand((n, p), device="cuda", dtype=torch.float64, generator=g) < af).double() + \
        (torch.rand((n, p), device="cuda", dtype=torch.float64, generator=g) < af).double()
    X -= X.mean(0, keepdim=True)
    K = (X @ X.t()) / p
    del X
    K = 0.5 * (K + K.t())
    if leaf is None:
        d, U = eigen.eigh_device(K)
    else:
        dd, U = eigen._eigh_large(K, leaf=leaf)
        d = dd.cpu().numpy()
    res, orth = eig_residuals(K, d, U)
    print(f"n = {n}, p = {p}, leaf = {leaf}: residual {res:.2e}, orthogonality {orth:.2e}",
          flush=True)
    assert np.all(np.diff(d) >= 0)
    assert res <= 1e-11 and orth <= 1e-12, (res, orth)
