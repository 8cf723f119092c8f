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
by one lets a caller report progress between them (the reduction alone takes about
a minute at n = 50,000).
"""

from __future__ import annotations

import ctypes
import os
import warnings

import numpy as np
import torch

from jamma_amd.device import require_gpu, to_dev

_RB_EVECT_ORIGINAL = 211  # rocblas_evect_original
_RB_FILL_LOWER = 122      # rocblas_fill_lower
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
    d, U = _dsyevd(Kd, progress)
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
