"""GEMMA .assoc.txt writers (byte-compatible with src/jamma/lmm/io.py:12-262)."""

from __future__ import annotations

from pathlib import Path

from jamma_amd.lmm.stats import AssocResult

HEADER_WALD = "chr\trs\tps\tn_miss\tallele1\tallele0\taf\tbeta\tse\tlogl_H1\tl_remle\tp_wald"
HEADER_SCORE = "chr\trs\tps\tn_miss\tallele1\tallele0\taf\tbeta\tse\tp_score"
HEADER_LRT = "chr\trs\tps\tn_miss\tallele1\tallele0\taf\tl_mle\tp_lrt"
HEADER_ALL = ("chr\trs\tps\tn_miss\tallele1\tallele0\taf\t"
              "beta\tse\tlogl_H1\tl_remle\tl_mle\tp_wald\tp_lrt\tp_score")

_HEADERS = {"wald": HEADER_WALD, "score": HEADER_SCORE, "lrt": HEADER_LRT, "all": HEADER_ALL}
_STAT_FIELDS = {
    "wald": ("beta", "se", "logl_H1", "l_remle", "p_wald"),
    "score": ("beta", "se", "p_score"),
    "lrt": ("l_mle", "p_lrt"),
    "all": ("beta", "se", "logl_H1", "l_remle", "l_mle", "p_wald", "p_lrt", "p_score"),
}


def _prefix(r: AssocResult) -> list[str]:
    return [r.chr, r.rs, str(r.ps), str(r.n_miss), r.allele1, r.allele0, f"{r.af:.3f}"]


def format_line(result: AssocResult, test_type: str = "wald") -> str:
    """af as .3f, statistics as .6e, tab separated (lmm/io.py:12-135)."""
    return "\t".join(_prefix(result) + [f"{getattr(result, k):.6e}" for k in _STAT_FIELDS[test_type]])


def format_assoc_line(result: AssocResult) -> str:
    return format_line(result, "wald")


def format_assoc_line_score(result: AssocResult) -> str:
    return format_line(result, "score")


def format_assoc_line_lrt(result: AssocResult) -> str:
    return format_line(result, "lrt")


def format_assoc_line_all(result: AssocResult) -> str:
    return format_line(result, "all")


def write_assoc_results(results: list[AssocResult], path: Path) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "w") as f:
        f.write(HEADER_WALD + "\n")
        for r in results:
            f.write(format_line(r, "wald") + "\n")


class IncrementalAssocWriter:
    """Context manager writing results as produced (lmm/io.py:172-262)."""

    def __init__(self, path: Path, test_type: str = "wald"):
        self.path = Path(path)
        self.test_type = test_type if test_type in _HEADERS else "wald"
        self._file = None
        self._count = 0

    def __enter__(self) -> "IncrementalAssocWriter":
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self._file = open(self.path, "w")
        self._file.write(_HEADERS[self.test_type] + "\n")
        return self

    def write(self, result: AssocResult) -> None:
        if self._file is None:
            raise RuntimeError("Writer not opened. Use as context manager.")
        self._file.write(format_line(result, self.test_type) + "\n")
        self._count += 1

    def write_batch(self, results: list[AssocResult]) -> None:
        for r in results:
            self.write(r)

    def __exit__(self, exc_type, exc_val, exc_tb) -> None:
        if self._file:
            self._file.close()
            self._file = None

    @property
    def count(self) -> int:
        return self._count
