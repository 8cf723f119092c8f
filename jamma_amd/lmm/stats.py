"""Per-SNP association result record (mirrors src/jamma/lmm/stats.py:39-64)."""

from __future__ import annotations

from dataclasses import dataclass


@dataclass
class AssocResult:
    """Association test result for a single SNP (GEMMA output fields).

    Wald (-lmm 1): logl_H1, l_remle, p_wald; LRT (-lmm 2): l_mle, p_lrt;
    Score (-lmm 3): p_score; all (-lmm 4): every field.
    """

    chr: str
    rs: str
    ps: int
    n_miss: int
    allele1: str
    allele0: str
    af: float
    beta: float
    se: float
    logl_H1: float | None = None
    l_remle: float | None = None
    p_wald: float | None = None
    p_score: float | None = None
    l_mle: float | None = None
    p_lrt: float | None = None
