"""GEMMA covariate file reader (mirrors src/jamma/io/covariate.py:19-95)."""

from __future__ import annotations

from pathlib import Path

import numpy as np


def read_covariate_file(path: Path) -> tuple[np.ndarray, np.ndarray]:
    """Whitespace-delimited, no header, "NA" = missing -> (covariates, indicator)."""
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if s:
                rows.append(s.split())
    if not rows:
        raise ValueError(f"Covariate file is empty: {path}")
    n_cvt = len(rows[0])
    for i, r in enumerate(rows):
        if len(r) != n_cvt:
            raise ValueError(f"Covariate file row {i + 1} has {len(r)} columns but expected "
                             f"{n_cvt} (based on first row)")
    cov = np.zeros((len(rows), n_cvt), dtype=np.float64)
    ind = np.ones(len(rows), dtype=np.int32)
    for i, r in enumerate(rows):
        for j, v in enumerate(r):
            if v == "NA":
                cov[i, j] = np.nan
                ind[i] = 0
            else:
                try:
                    cov[i, j] = float(v)
                except ValueError as e:
                    raise ValueError(f"Covariate file row {i + 1}, column {j + 1}: cannot parse "
                                     f"'{v}' as numeric (use 'NA' for missing)") from e
    return cov, ind
