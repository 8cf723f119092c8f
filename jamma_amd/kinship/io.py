"""Kinship matrix I/O in GEMMA .cXX.txt format (src/jamma/kinship/io.py:9-108)."""

from __future__ import annotations

from collections.abc import Iterator
from pathlib import Path

import numpy as np


def read_kinship_matrix(path: Path, n_samples: int | None = None) -> np.ndarray:
    from jamma_amd.lmm.io import parse_matrix_text

    K = parse_matrix_text(path)
    if K.ndim != 2 or K.shape[0] != K.shape[1]:
        raise ValueError(f"Kinship matrix must be square, got shape {K.shape}")
    if n_samples is not None and K.shape[0] != n_samples:
        raise ValueError(f"Kinship matrix dimension {K.shape[0]} does not match expected "
                         f"n_samples={n_samples}")
    if not np.allclose(K, K.T, rtol=1e-10):
        raise ValueError("Kinship matrix is not symmetric")
    return K


def write_kinship_matrix(K: np.ndarray, path: Path) -> None:
    """10 significant digits (%.10g), tab separated, no header (the library's
    multi-threaded formatter; the same bytes as the reference's f-strings)."""
    from jamma_amd.lmm.io import format_matrix_g10

    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    if hasattr(K, "detach"):
        K = K.detach().cpu().numpy()
    path.write_bytes(format_matrix_g10(np.asarray(K, dtype=np.float64)))


def write_loco_kinship_matrices(loco_kinships: Iterator[tuple[str, np.ndarray]], output_dir: Path,
                                prefix: str = "result") -> list[Path]:
    paths = []
    for chr_name, K in loco_kinships:
        p = Path(output_dir) / f"{prefix}.loco.cXX.chr{chr_name}.txt"
        write_kinship_matrix(K, p)
        paths.append(p)
    return paths
