"""GEMMA .assoc.txt writers (byte-compatible with src/jamma/lmm/io.py:12-262).

Two paths produce the same bytes: ``format_line`` per AssocResult (the reference's
per-result API), and ``format_rows`` / ``IncrementalAssocWriter.write_rows`` for
whole SNP tiles straight from the device's result rows: the library's native
formatter (jm_format_assoc_rows, several threads) fills one buffer per tile, and
the writer formats and writes tiles on a background thread, in order, so the text
output overlaps the next tile's device work.
"""

from __future__ import annotations

import ctypes
import os
import threading
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from jamma_amd import _lib
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


# device result columns (jm_lmm_assoc's JM_OUT_*) written per test type, in file order
# (_yield_chunk_results, results.py:222-286: -lmm 3 takes the Score beta / se)
OUT_COLUMNS = {
    "wald": (0, 1, 4, 3, 2),
    "score": (9, 10, 8),
    "lrt": (5, 7),
    "all": (0, 1, 4, 3, 5, 2, 7, 8),
}
FORMAT_THREADS = max(1, min(16, os.cpu_count() or 1))


def snp_text_parts(infos) -> tuple[bytes, np.ndarray, bytes, np.ndarray]:
    """The fixed text of each SNP's rows: "chr\trs\tps\t" and "allele1\tallele0\t"
    (from dicts with chr, rs, pos/ps, a1/allele1, a0/allele0), as two byte buffers
    with int64 offsets (len(infos) + 1 each)."""
    pre = [f"{i['chr']}\t{i['rs']}\t{i.get('pos', i.get('ps', 0))}\t".encode() for i in infos]
    mid = [f"{i.get('a1', i.get('allele1', ''))}\t{i.get('a0', i.get('allele0', ''))}\t".encode()
           for i in infos]

    def pack(parts):
        off = np.zeros(len(parts) + 1, dtype=np.int64)
        np.cumsum([len(x) for x in parts], out=off[1:])
        return b"".join(parts), off

    pb, po = pack(pre)
    mb, mo = pack(mid)
    return pb, po, mb, mo


def format_rows(parts, n_miss, af, stats, n_threads: int = FORMAT_THREADS) -> bytes:
    """GEMMA rows for the SNPs of `parts` (snp_text_parts) with their n_miss, af and
    (rows x n_stats) statistics, formatted by the native formatter."""
    pb, po, mb, mo = parts
    n = len(po) - 1
    n_miss = np.ascontiguousarray(n_miss, dtype=np.int64)
    af = np.ascontiguousarray(af, dtype=np.float64)
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    if n_miss.shape != (n,) or af.shape != (n,) or stats.ndim != 2 or stats.shape[0] != n:
        raise ValueError("format_rows: per-row arrays disagree with the SNP text parts")
    if n == 0:
        return b""
    cap = len(pb) + len(mb) + n * (24 + 8 + 16 * stats.shape[1])
    buf = ctypes.create_string_buffer(cap)
    po = np.ascontiguousarray(po, dtype=np.int64)
    mo = np.ascontiguousarray(mo, dtype=np.int64)
    got = _lib.lib.jm_format_assoc_rows(n, pb, po.ctypes.data, n_miss.ctypes.data, mb,
                                        mo.ctypes.data, af.ctypes.data, stats.ctypes.data,
                                        stats.shape[1], stats.shape[1], buf, cap, n_threads)
    if got < 0:
        raise RuntimeError(f"jm_format_assoc_rows failed ({got})")
    return buf.raw[:got]


def format_matrix_g10(a, n_threads: int = FORMAT_THREADS) -> bytes:
    """Tab-separated %.10g rows (GEMMA .cXX.txt / .eigenU.txt; a column vector gives
    .eigenD.txt), formatted by the native formatter."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    rows, cols = a.shape
    out = []
    step = max(1, (64 << 20) // max(1, 18 * cols))  # ~64 MB of text per call
    for r0 in range(0, rows, step):
        blk = a[r0:r0 + step]
        cap = blk.shape[0] * (cols * 18 + 1)
        buf = ctypes.create_string_buffer(max(cap, 1))
        got = _lib.lib.jm_format_g10_rows(blk.shape[0], cols, blk.ctypes.data, cols, buf, cap,
                                          n_threads)
        if got < 0:
            raise RuntimeError(f"jm_format_g10_rows failed ({got})")
        out.append(buf.raw[:got])
    return b"".join(out)


def parse_matrix_text(path, n_threads: int = FORMAT_THREADS) -> np.ndarray:
    """A whitespace-separated numeric text file (one row per line) as a float64
    matrix, parsed by the library's multi-threaded parser (np.loadtxt semantics for
    these files: blank lines skipped, every row the same length)."""
    data = Path(path).read_bytes()
    # upper bound on the values: one per separator run
    cap = data.count(b"\t") + data.count(b" ") + data.count(b"\n") + 1
    out = np.empty(cap, dtype=np.float64)
    rows, cols = ctypes.c_int64(0), ctypes.c_int64(0)
    got = _lib.lib.jm_parse_text_rows(data, len(data), out.ctypes.data, cap, ctypes.byref(rows),
                                      ctypes.byref(cols), n_threads)
    if got < 0:
        why = {-3: "rows of different lengths", -4: "a value that is not a number"}.get(got, got)
        raise ValueError(f"{path}: cannot parse as a numeric matrix ({why})")
    return out[: rows.value * cols.value].reshape(rows.value, cols.value)


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
    """Context manager writing results as produced (lmm/io.py:172-262).

    ``write`` / ``write_batch`` take AssocResults (the reference's API);
    ``write_rows`` takes a whole tile of device result rows and returns at once: the
    tile is formatted natively and written on the writer's background thread, tiles
    in submission order. ``__exit__`` (and ``flush``) wait for every pending tile and
    re-raise the first error of the background thread."""

    def __init__(self, path: Path, test_type: str = "wald"):
        self.path = Path(path)
        self.test_type = test_type if test_type in _HEADERS else "wald"
        self._file = None
        self._count = 0
        self._pool = None
        self._pending = []
        self._lock = threading.Lock()

    def __enter__(self) -> "IncrementalAssocWriter":
        self.path.parent.mkdir(parents=True, exist_ok=True)
        self._file = open(self.path, "wb")
        self._file.write((_HEADERS[self.test_type] + "\n").encode())
        return self

    def _check_open(self):
        if self._file is None:
            raise RuntimeError("Writer not opened. Use as context manager.")

    def write(self, result: AssocResult) -> None:
        self._check_open()
        self.flush()  # keep the order with tiles submitted before
        self._file.write((format_line(result, self.test_type) + "\n").encode())
        self._count += 1

    def write_batch(self, results: list[AssocResult]) -> None:
        for r in results:
            self.write(r)

    def write_rows(self, parts, n_miss, af, out_rows) -> None:
        """Queue one tile: SNP text parts (snp_text_parts), n_miss, af and the device
        result rows (n x JM_NOUT); the columns of this writer's test type are written."""
        self._check_open()
        stats = np.ascontiguousarray(np.asarray(out_rows)[:, OUT_COLUMNS[self.test_type]])
        n_miss = np.array(n_miss, dtype=np.int64)
        af = np.array(af, dtype=np.float64)
        if self._pool is None:
            self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="jm-writer")

        def job():
            data = format_rows(parts, n_miss, af, stats)
            with self._lock:
                self._file.write(data)

        self._pending.append(self._pool.submit(job))
        self._count += stats.shape[0]
        # bound the queue: at most a few tiles of text in flight
        while len(self._pending) > 4:
            self._pending.pop(0).result()

    def flush(self) -> None:
        while self._pending:
            self._pending.pop(0).result()

    def __exit__(self, exc_type, exc_val, exc_tb) -> None:
        try:
            if exc_type is None:
                self.flush()
            else:
                for f in self._pending:
                    f.cancel()
                self._pending = []
        finally:
            if self._pool is not None:
                self._pool.shutdown(wait=True)
                self._pool = None
            if self._file:
                self._file.close()
                self._file = None

    @property
    def count(self) -> int:
        return self._count
