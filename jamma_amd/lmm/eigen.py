"""Eigendecomposition of the kinship on the device.

Mirrors eigendecompose_kinship (src/jamma/lmm/eigen.py:28-128): ascending
eigenvalues, |d| < threshold set to 0, warnings for negative / many zero
eigenvalues. The reference calls host LAPACK dsyevd through numpy.linalg.eigh;
here the FP64 symmetric eigensolver of the ROCm stack (rocSOLVER syevd, reached
through torch.linalg.eigh) runs on the resident device matrix.
"""

from __future__ import annotations

import warnings

import numpy as np
import torch

from jamma_amd.device import require_gpu, to_dev


def eigh_device(K, threshold: float = 1e-10):
    """-> (d host float64 ascending thresholded, U device float64 (n x n))."""
    Kd = to_dev(K, torch.float64, require_gpu())
    if Kd.ndim != 2 or Kd.shape[0] != Kd.shape[1]:
        raise ValueError(f"Kinship matrix must be square, got shape {tuple(Kd.shape)}")
    d, U = torch.linalg.eigh(Kd)
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
