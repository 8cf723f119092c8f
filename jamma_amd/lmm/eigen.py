"""Eigendecomposition of the kinship on the device.

Mirrors eigendecompose_kinship (src/jamma/lmm/eigen.py:28-128): ascending
eigenvalues, |d| < threshold set to 0, warnings for negative / many zero
eigenvalues. The reference calls host LAPACK dsyevd through numpy.linalg.eigh;
here rocSOLVER's FP64 routines run on the resident device matrix, in the three
steps of a divide-and-conquer solver: Householder reduction to tridiagonal form
(dsytrd), the orthogonal factor of that reduction (dorgtr), and divide and conquer
on the tridiagonal matrix applied to that factor (dstedc). They are called directly
rather than through torch.linalg.eigh: the hipSOLVER layer torch goes through sizes
its workspace in 32-bit integers and fails for cohorts of ~30k samples, while
rocSOLVER manages its own workspace through the rocBLAS handle. Taking the steps one
by one lets a caller report progress between them.

Those routines take 32-bit sizes and have no 64-bit variant, so they could serve
only n <= 46,340 (n * n <= 2^31 - 1); the reference reaches larger cohorts with
ILP64 LAPACK (src/jamma/lmm/eigen.py:9,39). They are used below 8,192 samples.
From 8,192 on (where it is also faster) the eigendecomposition is _eigh_large,
which keeps every offset 64-bit:
  1. blocked Householder tridiagonalisation (LAPACK dsytrd/dlatrd, lower), every
     step a kernel of this library: the lower-triangle symv (jm_symv_lower, each
     stored element read once per column step), the reflector (jm_tridiag_larfg),
     the panel's matrix-vector products (jm_tridiag_panel_update / _step), and the
     FP64 MFMA GEMM (jm_dgemm) for the rank-2k trailing updates;
  2. Cuppen divide and conquer on the tridiagonal matrix: halves are torn off
     until each fits rocSOLVER's dstedc, and pairs are merged through the
     rank-one update D + rho z z^T (deflation as LAPACK dlaed2; secular roots,
     Gu-Eisenstat weights and eigenvectors on the device, jm_secular_*), then
     Q = blockdiag(Q1, Q2) M with jm_dgemm;
  3. back-transformation by the reflectors in compact WY blocks (I - V T V^T),
     T from the V^T v products the reduction already forms (LAPACK dlarft).
"""

from __future__ import annotations

import ctypes
import os
import warnings

import numpy as np
import torch

from jamma_amd import _lib
from jamma_amd.device import ptr, require_gpu, stream_ptr, to_dev

_RB_EVECT_ORIGINAL = 211     # rocblas_evect_original
_RB_EVECT_TRIDIAGONAL = 212  # rocblas_evect_tridiagonal
_RB_FILL_LOWER = 122         # rocblas_fill_lower
_LIB_MAX_N = 46_340          # largest n with n * n <= 2^31 - 1
# From this size on the 64-bit path is also the faster one (one MI355X, kinship-like
# matrices: n = 8,192 0.65 s vs rocSOLVER 0.93 s; 20,000 4.8 s vs 7.6 s; 30,000 14.3 s
# vs 22.9 s; eigenvalues equal to 1e-15, residuals and orthogonality alike).
_OWN_MIN_N = 8192
_LARGE_PANEL = 128           # reflectors per panel / per compact-WY block
_LARGE_LEAF = 30_000         # largest tridiagonal block handed to rocSOLVER dstedc
_SOLVER = {}


def _solver(device: torch.device):
    """(rocsolver lib, rocblas handle) for a device, created once per process."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    if idx in _SOLVER:
        return _SOLVER[idx]
    libdir = os.path.join(os.path.dirname(torch.__file__), "lib")
    # torch's own rocBLAS/rocSOLVER: already mapped in-process, same HIP runtime
    rb = ctypes.CDLL(os.path.join(libdir, "librocblas.so"), mode=ctypes.RTLD_GLOBAL)
    rs = ctypes.CDLL(os.path.join(libdir, "librocsolver.so"), mode=ctypes.RTLD_GLOBAL)
    rb.rocblas_create_handle.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rb.rocblas_create_handle.restype = ctypes.c_int
    rb.rocblas_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    rb.rocblas_set_stream.restype = ctypes.c_int
    vp, ci = ctypes.c_void_p, ctypes.c_int
    rs.rocsolver_dsytrd.argtypes = [vp, ci, ci, vp, ci, vp, vp, vp]
    rs.rocsolver_dorgtr.argtypes = [vp, ci, ci, vp, ci, vp]
    rs.rocsolver_dstedc.argtypes = [vp, ci, ci, vp, vp, vp, ci, vp]
    for f in (rs.rocsolver_dsytrd, rs.rocsolver_dorgtr, rs.rocsolver_dstedc):
        f.restype = ci
    h = ctypes.c_void_p()
    with torch.cuda.device(idx):
        st = rb.rocblas_create_handle(ctypes.byref(h))
    if st != 0:
        raise RuntimeError(f"rocblas_create_handle failed with status {st}")
    _SOLVER[idx] = (rb, rs, h)
    return _SOLVER[idx]


def _dsyevd(Kd: torch.Tensor, progress=None):
    """Eigen-pairs of a symmetric device float64 matrix: (d ascending, U columns).

    `progress(step)` (optional) is called after each of the three solver steps, once
    the step's work has finished on the device.
    """
    n = Kd.shape[0]
    if n > _LIB_MAX_N:
        raise ValueError(f"rocSOLVER's 32-bit routines cannot take n = {n}")
    rb, rs, h = _solver(Kd.device)
    A = Kd.clone(memory_format=torch.contiguous_format)  # overwritten with eigenvectors
    D = torch.empty(n, dtype=torch.float64, device=Kd.device)
    E = torch.empty(max(n, 1), dtype=torch.float64, device=Kd.device)
    tau = torch.empty(max(n, 1), dtype=torch.float64, device=Kd.device)
    info = torch.zeros(1, dtype=torch.int32, device=Kd.device)

    def done(name, st):
        if st != 0:
            raise RuntimeError(f"rocsolver_{name} failed with status {st} (n={n})")
        if progress is not None:
            torch.cuda.current_stream().synchronize()
            progress(name)

    with torch.cuda.device(Kd.device):
        stream = torch.cuda.current_stream()
        st = rb.rocblas_set_stream(h, ctypes.c_void_p(stream.cuda_stream))
        if st != 0:
            raise RuntimeError(f"rocblas_set_stream failed with status {st}")
        # the row-major symmetric matrix is its own column-major transpose
        done("dsytrd", rs.rocsolver_dsytrd(h, _RB_FILL_LOWER, n, A.data_ptr(), n, D.data_ptr(),
                                           E.data_ptr(), tau.data_ptr()))
        done("dorgtr", rs.rocsolver_dorgtr(h, _RB_FILL_LOWER, n, A.data_ptr(), n,
                                           tau.data_ptr()))
        done("dstedc", rs.rocsolver_dstedc(h, _RB_EVECT_ORIGINAL, n, D.data_ptr(), E.data_ptr(),
                                           A.data_ptr(), n, info.data_ptr()))
        code = int(info.item())  # synchronises the stream
    if code != 0:
        raise RuntimeError(f"rocsolver_dstedc did not converge: info={code} (n={n})")
    del E, tau
    # column-major result: A's row j (row-major view) is eigenvector j, so U = A^T
    return D, A.t()


def _ok(name, st):
    if st != 0:
        raise RuntimeError(f"{name} failed with status {st}")


def _tridiagonalize(A: torch.Tensor, nb: int, progress=None):
    """Householder reduction of the symmetric device matrix A (n x n, float64; as a
    column-major matrix it is the same) to tridiagonal form, LAPACK dsytrd/dlatrd
    with UPLO = 'L' and panels of nb columns.

    A is overwritten: column c (row c of the tensor) holds reflector c's vector
    below the diagonal, with its unit entry at row c + 1 stored as 1.
    Returns (d, e, tau) on the device and, per panel, the nb x nb matrix whose
    strictly upper part is V^T V of the panel's reflectors (the dlatrd products
    W(0:i, i) = V(:, 0:i)^T v_i), for the compact-WY factors.
    """
    n = A.shape[0]
    dev = A.device
    f64 = torch.float64
    e = torch.zeros(max(n - 1, 1), dtype=f64, device=dev)
    tau = torch.zeros(max(n - 1, 1), dtype=f64, device=dev)
    Wt = torch.zeros((nb, n), dtype=f64, device=dev)  # column-major W (n x nb, ld n)
    nwork = int(_lib.lib.jm_symv_work_doubles(n))
    symv_work = torch.empty(max(nwork, 1), dtype=f64, device=dev)
    npw = int(_lib.lib.jm_tridiag_panel_work_doubles(n))
    panel_work = torch.empty(npw, dtype=f64, device=dev)
    base, wb, ld = A.data_ptr(), Wt.data_ptr(), n
    st = stream_ptr()

    def pa(r, c):
        return base + 8 * (c * ld + r)

    def pw(r, c):
        return wb + 8 * (c * ld + r)

    # three launches per column (jm_tridiag_column); JM_TRIDIAG_UNFUSED=1 takes the
    # seven-launch sequence of the separate entry points (for comparisons)
    fused = os.environ.get("JM_TRIDIAG_UNFUSED", "0") != "1"
    if fused:
        ncol = int(_lib.lib.jm_tridiag_column_work_doubles(n))
        col_work = torch.empty(max(ncol, 1), dtype=f64, device=dev)
    gram = []
    next_report = 8192
    for p0 in range(0, n, nb):
        Np = n - p0
        nbp = min(nb, Np)
        if fused:
            tp, ep = tau.data_ptr() + 8 * p0, e.data_ptr() + 8 * p0
            for i in range(nbp):
                _lib.call("jm_tridiag_column", Np, i, pa(p0, p0), ld, pw(0, 0), ld, tp, ep,
                          col_work.data_ptr(), ncol, st)
            il = nbp - 1
            if il < Np - 1:  # the panel's last reflector: its w still needs dlatrd's finish
                _lib.call("jm_tridiag_finish", Np - il - 1, tau.data_ptr() + 8 * (p0 + il),
                          pw(il + 1, il), pa(p0 + il + 1, p0 + il), st)
        for i in range(0 if fused else nbp):
            gc = p0 + i
            if i > 0:  # apply the panel's earlier reflectors to column i
                _lib.call("jm_tridiag_panel_update", Np - i, i, pa(gc, p0), ld, pw(i, 0), ld,
                          pa(gc, gc), st)
            if i == Np - 1:
                continue
            L = Np - i - 1
            v = pa(gc + 1, gc)
            _lib.call("jm_tridiag_larfg", L, v, pa(gc + 2, gc) if L > 1 else None,
                      tau.data_ptr() + 8 * gc, e.data_ptr() + 8 * gc, st)
            w = pw(i + 1, i)
            _lib.call("jm_symv_lower", L, pa(gc + 1, gc + 1), ld, v, w, symv_work.data_ptr(),
                      nwork, st)
            if i > 0:  # w -= V (W^T v) + W (V^T v); V^T v kept in W(0:i, i)
                _lib.call("jm_tridiag_panel_step", L, i, pa(gc + 1, p0), ld, pw(i + 1, 0), ld,
                          v, w, pw(0, i), panel_work.data_ptr(), npw, st)
            _lib.call("jm_tridiag_finish", L, tau.data_ptr() + 8 * gc, w, v, st)
        gram.append(Wt[:nbp, :nbp].t().clone())  # W(0:nbp, 0:nbp), row-major
        if Np > nbp:  # trailing update A22 -= V W^T + W V^T (lower triangle)
            N2 = Np - nbp
            _lib.call("jm_dgemm2", 0, 1, N2, N2, nbp, -1.0, pa(p0 + nbp, p0), pw(nbp, 0),
                      pw(nbp, 0), pa(p0 + nbp, p0), ld, ld, 1.0, pa(p0 + nbp, p0 + nbp), ld, 1,
                      st)
        if progress is not None and p0 + nbp >= next_report and Np > nbp:
            torch.cuda.current_stream().synchronize()
            progress(f"tridiagonal reduction: {p0 + nbp} / {n} columns")
            next_report += 8192
    d = A.diagonal().clone()
    return d, e[: max(n - 1, 0)], tau, gram


def _stedc(d: np.ndarray, e: np.ndarray, device):
    """rocSOLVER dstedc on a tridiagonal block (n * n <= 2^31 - 1): (w ascending host,
    Q device (n, n) whose row j is eigenvector j)."""
    n = len(d)
    rb, rs, h = _solver(device)
    D = torch.as_tensor(np.ascontiguousarray(d), dtype=torch.float64).to(device)
    E = torch.zeros(max(n, 1), dtype=torch.float64, device=device)
    if n > 1:
        E[: n - 1] = torch.as_tensor(np.ascontiguousarray(e), dtype=torch.float64).to(device)
    C = torch.empty((n, n), dtype=torch.float64, device=device)
    info = torch.zeros(1, dtype=torch.int32, device=device)
    _ok("rocblas_set_stream", rb.rocblas_set_stream(h, ctypes.c_void_p(stream_ptr())))
    _ok("rocsolver_dstedc", rs.rocsolver_dstedc(h, _RB_EVECT_TRIDIAGONAL, n, D.data_ptr(),
                                                E.data_ptr(), C.data_ptr(), n, info.data_ptr()))
    code = int(info.item())
    if code != 0:
        raise RuntimeError(f"rocsolver_dstedc did not converge: info={code} (n={n})")
    return D.cpu().numpy(), C


def _deflate(dc: np.ndarray, z: np.ndarray, rho: float):
    """Deflation of diag(dc) + rho z z^T as LAPACK dlaed2 (host, O(n)).

    z is scaled to unit norm and rho doubled (|z| = sqrt 2 as torn); a pole deflates
    when rho |z_k| <= tol, or, when it is within tol of the previous candidate pole
    (|t c s| <= tol), by a Givens rotation (c, s) of the two basis vectors that zeroes
    its weight; tol = 8 eps max(max|d|, max|z|). Returns (order, ds, zs, nd, defl_val,
    defl_row, rots, rho): the sort order of dc, the updated sorted poles and weights,
    the sorted positions that stay in the secular equation, the deflated eigenvalues
    and their basis rows, and the rotations (row p, row q, c, s) in the order applied.
    """
    z = z / np.sqrt(2.0)
    rho = 2.0 * rho
    order = np.argsort(dc, kind="stable")
    ds = dc[order].astype(np.float64).tolist()
    zs = z[order].astype(np.float64).tolist()
    n = len(ds)
    tol = 8.0 * (np.finfo(np.float64).eps / 2) * max(float(np.max(np.abs(dc))),
                                                      float(np.max(np.abs(z))))
    nd, defl_val, defl_row, rots = [], [], [], []
    if rho * float(np.max(np.abs(z))) <= tol:
        return order, ds, zs, nd, list(ds), order.tolist(), rots, rho
    pj = -1
    for k in range(n):
        if rho * abs(zs[k]) <= tol:
            defl_val.append(ds[k])
            defl_row.append(int(order[k]))
            continue
        if pj < 0:
            pj = k
            continue
        s_, c_ = zs[pj], zs[k]
        tau = float(np.hypot(c_, s_))
        t = ds[k] - ds[pj]
        c_, s_ = c_ / tau, -s_ / tau
        if abs(t * c_ * s_) <= tol:
            zs[k], zs[pj] = tau, 0.0
            rots.append((int(order[pj]), int(order[k]), c_, s_))
            dp = ds[pj] * c_ * c_ + ds[k] * s_ * s_
            ds[k] = ds[pj] * s_ * s_ + ds[k] * c_ * c_
            defl_val.append(dp)
            defl_row.append(int(order[pj]))
        else:
            nd.append(pj)
        pj = k
    if pj >= 0:
        nd.append(pj)
    return order, ds, zs, nd, defl_val, defl_row, rots, rho


def _rank_one_merge(dc: np.ndarray, z: np.ndarray, rho: float, device):
    """Eigen-pairs of diag(dc) + rho z z^T (rho >= 0, |z| = sqrt 2 as torn): (w ascending
    host, M device (n, n) row-major with M[a, c] = component a of eigenvector c).
    Deflation on the host (_deflate), the secular equation and the eigenvectors of the
    rest on the device, then the deflation rotations applied to M's rows."""
    n = len(dc)
    f64 = torch.float64
    order, ds, zs, nd, defl_val, defl_row, rots, rho = _deflate(dc, z, rho)
    M = torch.zeros((n, n), dtype=f64, device=device)
    K = len(nd)
    st = stream_ptr()
    lam = np.zeros(0)
    if K:
        dl = torch.as_tensor([ds[k] for k in nd], dtype=f64).to(device)
        wh = np.array([zs[k] for k in nd])
        w = torch.as_tensor(wh, dtype=f64).to(device)
        w2 = w * w
        work = torch.empty((5, K), dtype=f64, device=device)  # dorg, tau, lam, zhat, inv_norm
        _lib.call("jm_secular_solve", K, ptr(dl), ptr(w), ptr(w2), rho, float(np.sum(wh * wh)),
                  ptr(work[0]), ptr(work[1]), ptr(work[2]), ptr(work[3]), ptr(work[4]), st)
        lam = work[2].cpu().numpy()
    vals = np.concatenate([lam, np.asarray(defl_val, dtype=np.float64)])
    perm = np.argsort(vals, kind="stable")
    col = np.empty(n, dtype=np.int64)
    col[perm] = np.arange(n)
    rows_nd = torch.as_tensor(order[nd].astype(np.int64) if K else np.zeros(0, np.int64)).to(device)
    cols_nd = torch.as_tensor(col[:K]).to(device)
    rows_d = torch.as_tensor(np.asarray(defl_row, dtype=np.int64)).to(device)
    cols_d = torch.as_tensor(col[K:]).to(device)
    if K:
        _lib.call("jm_secular_vectors", K, ptr(dl), ptr(work[3]), ptr(work[0]), ptr(work[1]),
                  ptr(work[4]), ptr(rows_nd), ptr(cols_nd), n - K, ptr(rows_d), ptr(cols_d),
                  ptr(M), n, st)
    elif n:
        _lib.call("jm_secular_vectors", 0, None, None, None, None, None, None, None, n,
                  ptr(rows_d), ptr(cols_d), ptr(M), n, st)
    if rots:
        pq = torch.as_tensor(np.array([(p, q) for p, q, _, _ in rots], dtype=np.int64)).to(device)
        cs = torch.as_tensor(np.array([(c, s) for _, _, c, s in rots], dtype=np.float64)).to(device)
        _lib.call("jm_apply_rotations", len(rots), ptr(pq), ptr(cs), n, ptr(M), n, st)
    torch.cuda.current_stream().synchronize()  # host arrays above stay alive until here
    return vals[perm], M


def _tri_eig(d: np.ndarray, e: np.ndarray, device, leaf: int):
    """Eigen-pairs of the symmetric tridiagonal (d, e): (w ascending host, Q device (n, n)
    whose row j is eigenvector j). Cuppen tear at the middle until blocks fit dstedc."""
    n = len(d)
    if n <= leaf:
        return _stedc(d, e, device)
    m = n // 2
    beta = float(e[m - 1])
    rho, sgn = abs(beta), (1.0 if beta >= 0 else -1.0)
    d1, d2 = d[:m].copy(), d[m:].copy()
    d1[-1] -= rho
    d2[0] -= rho
    w1, Q1 = _tri_eig(d1, e[: m - 1], device, leaf)
    w2, Q2 = _tri_eig(d2, e[m:], device, leaf)
    # z = Q^T (e_{m-1} + sgn e_m): the last row of Q1 and the first row of Q2
    z = np.concatenate([Q1[:, m - 1].cpu().numpy(), sgn * Q2[:, 0].cpu().numpy()])
    w, M = _rank_one_merge(np.concatenate([w1, w2]), z, rho, device)
    Q = torch.empty((n, n), dtype=torch.float64, device=device)  # column-major Q
    st = stream_ptr()
    # Q(0:m, :) = Q1 M(0:m, :) and Q(m:, :) = Q2 M(m:, :); M row-major = M^T column-major
    _lib.call("jm_dgemm", 0, 1, m, n, m, 1.0, ptr(Q1), m, ptr(M), n, 0.0, ptr(Q), n, 0, st)
    _lib.call("jm_dgemm", 0, 1, n - m, n, n - m, 1.0, ptr(Q2), n - m, M.data_ptr() + 8 * m * n,
              n, 0.0, Q.data_ptr() + 8 * m, n, 0, st)
    torch.cuda.current_stream().synchronize()
    del Q1, Q2, M
    return w, Q


def _compact_wy_t(tau: np.ndarray, G: np.ndarray) -> np.ndarray:
    """Upper-triangular T with H_0 ... H_{k-1} = I - V T V^T (LAPACK dlarft, forward,
    columnwise): T[i, i] = tau_i, T[:i, i] = -tau_i T[:i, :i] (V^T v_i)[:i], where the
    strictly upper part of G holds V^T V."""
    k = len(tau)
    T = np.zeros((k, k))
    for i in range(k):
        T[i, i] = tau[i]
        if i:
            T[:i, i] = -tau[i] * (T[:i, :i] @ G[:i, i])
    return T


def _back_transform(A: torch.Tensor, tau: torch.Tensor, gram, Q: torch.Tensor, nb: int):
    """Q (column-major, n x n) <- H_0 H_1 ... H_{n-2} Q, the reflectors of
    _tridiagonalize applied in blocks of nb from the last block to the first."""
    n = A.shape[0]
    dev = A.device
    st = stream_ptr()
    tau_h = tau.cpu().numpy()
    blocks = []
    for b, p0 in enumerate(range(0, n, nb)):
        r = min(nb, n - 1 - p0)  # reflectors of this panel (the last column has none)
        if r > 0:
            T = _compact_wy_t(tau_h[p0:p0 + r], gram[b].cpu().numpy()[:r, :r])
            blocks.append((p0, r, T))
    Y = torch.empty((nb, n), dtype=torch.float64, device=dev)   # column-major (r x n, ld r)
    Y2 = torch.empty((nb, n), dtype=torch.float64, device=dev)
    max_split = 16
    work = torch.empty(max_split * nb * n, dtype=torch.float64, device=dev)
    tiles = -(-n // 64)  # output tiles of V^T Q (one tile row)
    for p0, r, T in reversed(blocks):
        L = n - 1 - p0
        Vt = A[p0:p0 + r, p0 + 1:].clone()  # column-major V (L x r, ld L)
        Vt[:, :r] = torch.triu(Vt[:, :r], 1) + torch.eye(r, dtype=torch.float64, device=dev)
        Td = torch.as_tensor(np.ascontiguousarray(T.T)).to(dev)  # column-major T
        QR = Q.data_ptr() + 8 * (p0 + 1)
        # V^T Q_R has one row of output tiles and a long K: split K to fill the device
        splits = int(min(max_split, max(1, -(-4096 // tiles)), max(1, L // 1024)))
        _lib.call("jm_dgemm_splitk", 1, 0, r, n, L, 1.0, ptr(Vt), L, QR, n, 0.0, ptr(Y), r,
                  splits, ptr(work), work.numel(), st)
        _lib.call("jm_dgemm", 0, 0, r, n, r, 1.0, ptr(Td), r, ptr(Y), r, 0.0, ptr(Y2), r, 0, st)
        _lib.call("jm_dgemm", 0, 0, L, n, r, -1.0, ptr(Vt), L, ptr(Y2), r, 1.0, QR, n, 0, st)
        del Vt, Td


def _eigh_large(Kd: torch.Tensor, progress=None, nb: int | None = None,
                leaf: int | None = None):
    """(d ascending device, U columns) of a symmetric device float64 matrix with 64-bit
    offsets throughout (for n beyond the library solvers' 32-bit limit). nb: panel
    width; leaf: largest tridiagonal block solved by dstedc."""
    nb = _LARGE_PANEL if nb is None else nb
    leaf = _LARGE_LEAF if leaf is None else leaf
    if nb < 1 or leaf < 1 or leaf > _LIB_MAX_N:
        raise ValueError("panel width and leaf size must be in [1, 46340]")
    A = Kd.clone(memory_format=torch.contiguous_format)
    d, e, tau, gram = _tridiagonalize(A, nb, progress)
    torch.cuda.current_stream().synchronize()
    if progress is not None:
        progress("tridiagonal reduction")
    w, Q = _tri_eig(d.cpu().numpy(), e.cpu().numpy(), A.device, leaf)
    if progress is not None:
        progress("tridiagonal divide and conquer")
    _back_transform(A, tau, gram, Q, nb)
    torch.cuda.current_stream().synchronize()
    if progress is not None:
        progress("back-transformation")
    del A
    # Q column-major: tensor row j is eigenvector j, so U = Q^T as a tensor
    return torch.as_tensor(w, dtype=torch.float64).to(Kd.device), Q.t()


def _select_solver(n: int):
    """The solver eigh_device uses for an n x n matrix: rocSOLVER below _OWN_MIN_N, the
    64-bit path from there on. rocSOLVER's 32-bit routines are never chosen above
    _LIB_MAX_N, whatever _OWN_MIN_N is."""
    return _dsyevd if n < min(_OWN_MIN_N, _LIB_MAX_N + 1) else _eigh_large


def eigh_device(K, threshold: float = 1e-10, progress=None):
    """-> (d host float64 ascending thresholded, U device float64 (n x n)).

    U[:, j] is the eigenvector of d[j]; U is a transposed view whose transpose
    (U^T, eigenvectors as rows) is contiguous. `progress(step)`, if given, is called
    after each solver step.
    """
    Kd = to_dev(K, torch.float64, require_gpu())
    if Kd.ndim != 2 or Kd.shape[0] != Kd.shape[1]:
        raise ValueError(f"Kinship matrix must be square, got shape {tuple(Kd.shape)}")
    if Kd.shape[0] == 0:
        raise ValueError("Kinship matrix is empty")
    d, U = _select_solver(Kd.shape[0])(Kd, progress)
    del Kd
    d = d.cpu().numpy()
    n_neg = int(np.sum(d < -threshold))
    if n_neg:
        warnings.warn(f"Kinship matrix has {n_neg} negative eigenvalue(s). Matrix may not be "
                      "positive semi-definite.", stacklevel=2)
    d = np.where(np.abs(d) < threshold, 0.0, d)
    n_zero = int(np.sum(d == 0.0))
    if n_zero > 1:
        warnings.warn(f"Kinship matrix has {n_zero} eigenvalues close to zero. Matrix may be "
                      "rank-deficient.", stacklevel=2)
    return d, U


def eigendecompose_kinship(K, threshold: float = 1e-10, *, check_memory: bool = True):
    """Drop-in for the reference API: returns host (eigenvalues, eigenvectors)."""
    del check_memory
    d, U = eigh_device(K, threshold)
    return d, U.cpu().numpy()
