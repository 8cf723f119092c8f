"""GEMMA .eigenD.txt / .eigenU.txt I/O (mirrors src/jamma/lmm/eigen_io.py:18-164).

Format (GEMMA param.cpp WriteVector / WriteMatrix, as the reference writes it): one
eigenvalue per line, eigenvector rows tab separated, 10 significant digits (%.10g),
no header. Writing goes through the library's multi-threaded formatter (an n = 10,000
eigenvector file is 10^8 numbers); reading validates the shapes with the reference's
messages.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from jamma_amd.lmm.io import format_matrix_g10, parse_matrix_text


def read_eigenvalues(path: Path) -> np.ndarray:
    """One value per line (a single row of values is accepted too, as np.loadtxt)."""
    return parse_matrix_text(path).reshape(-1)


def read_eigenvectors(path: Path) -> np.ndarray:
    return parse_matrix_text(path)


def read_eigen_files(eigenD_path: Path, eigenU_path: Path, n_samples: int | None = None):
    """Both files with the reference's consistency checks (eigen_io.py:42-97)."""
    eigenvalues = read_eigenvalues(eigenD_path)
    eigenvectors = read_eigenvectors(eigenU_path)
    n_eval = eigenvalues.shape[0]
    if eigenvectors.ndim != 2:
        raise ValueError(f"Eigenvector file must contain a 2D matrix, got {eigenvectors.ndim}D "
                         f"array from {eigenU_path}")
    n_rows, n_cols = eigenvectors.shape
    if n_rows != n_cols:
        raise ValueError(f"Eigenvector matrix must be square, got shape ({n_rows}, {n_cols}) from "
                         f"{eigenU_path}")
    if n_eval != n_rows:
        raise ValueError(f"Eigenvalue count ({n_eval}) does not match eigenvector dimensions "
                         f"({n_rows} x {n_cols}). Files may be mismatched: {eigenD_path}, "
                         f"{eigenU_path}")
    if n_samples is not None and n_eval != n_samples:
        raise ValueError(f"Eigenvalue count ({n_eval}) does not match expected n_samples="
                         f"{n_samples}. Eigen files may be from a different dataset.")
    return eigenvalues, eigenvectors


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):  # a device tensor (e.g. eigh_device's U)
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def write_eigenvalues(eigenvalues, path: Path) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(format_matrix_g10(_host(eigenvalues).reshape(-1)))


def write_eigenvectors(eigenvectors, path: Path) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_bytes(format_matrix_g10(_host(eigenvectors)))


def write_eigen_files(eigenvalues, eigenvectors, output_dir: Path, prefix: str = "result"):
    """{prefix}.eigenD.txt and {prefix}.eigenU.txt in output_dir (eigen_io.py:140-164)."""
    output_dir = Path(output_dir)
    d_path = output_dir / f"{prefix}.eigenD.txt"
    u_path = output_dir / f"{prefix}.eigenU.txt"
    write_eigenvalues(eigenvalues, d_path)
    write_eigenvectors(eigenvectors, u_path)
    return d_path, u_path
