"""GEMMA .eigenD.txt / .eigenU.txt I/O (mirrors src/jamma/lmm/eigen_io.py:42-164)."""

from __future__ import annotations

from pathlib import Path

import numpy as np


def read_eigen_files(d_path: Path, u_path: Path, n_samples: int | None = None):
    d = np.loadtxt(d_path, dtype=np.float64).reshape(-1)
    U = np.loadtxt(u_path, dtype=np.float64)
    if U.ndim != 2 or U.shape[0] != U.shape[1] or U.shape[0] != d.shape[0]:
        raise ValueError(f"eigen files disagree: {d.shape} eigenvalues, eigenvectors {U.shape}")
    if n_samples is not None and d.shape[0] != n_samples:
        raise ValueError(f"eigen files have {d.shape[0]} samples, expected {n_samples}")
    return d, U


def write_eigen_files(eigenvalues, eigenvectors, output_dir: Path, prefix: str = "result"):
    output_dir = Path(output_dir)
    output_dir.mkdir(parents=True, exist_ok=True)
    d_path = output_dir / f"{prefix}.eigenD.txt"
    u_path = output_dir / f"{prefix}.eigenU.txt"
    with open(d_path, "w") as f:
        for v in np.asarray(eigenvalues):
            f.write(f"{v:.10g}\n")
    with open(u_path, "w") as f:
        for row in np.asarray(eigenvectors):
            f.write("\t".join(f"{v:.10g}" for v in row) + "\n")
    return d_path, u_path
