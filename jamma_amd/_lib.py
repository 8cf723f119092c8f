"""ctypes binding of libjamma_hip.so (the C ABI declared in include/jamma_hip.h).

The library is required: importing this module raises if it is missing, and
every call raises RuntimeError with the library's own message on failure.
There is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(os.environ.get("JM_LIBRARY", Path(__file__).resolve().parent / "lib" / "libjamma_hip.so"))

JM_NOUT = 12
OUT = {
    "beta": 0, "se": 1, "p_wald": 2, "l_remle": 3, "logl_H1": 4, "l_mle": 5, "logl_mle": 6,
    "p_lrt": 7, "p_score": 8, "beta_score": 9, "se_score": 10, "evals": 11,
}


class MomCol(ctypes.Structure):
    """Mirror of jm_mom_col."""

    _fields_ = [("V", c_void_p), ("ldv", c_int64), ("nv", c_int32), ("reserved", c_int32),
                ("out", c_void_p), ("ldm", c_int64)]


class LmmProblem(ctypes.Structure):
    """Mirror of jm_lmm_problem."""

    _fields_ = [
        ("n", c_int32), ("n_cvt", c_int32), ("n_terms", c_int32), ("n_clusters", c_int32),
        ("n_grid", c_int32), ("n_iter", c_int32),
        ("l_min", c_double), ("l_max", c_double),
        ("term_cluster", c_void_p), ("term_m", c_void_p), ("cl_center", c_void_p),
        ("logdet_mom", c_void_p), ("shared_mom", c_void_p), ("pair_row", c_void_p),
        ("grid_loglam", c_void_p),
        ("grid_kappa", c_void_p), ("grid_shared", c_void_p), ("grid_logdet", c_void_p),
        ("null_kappa", c_void_p), ("null_shared", c_void_p),
        ("lambda_null", c_double), ("logl_null", c_double),
        ("cl_weight", c_void_p), ("iab_shared", c_void_p),
    ]


class SpectrumPlanC(ctypes.Structure):
    """Mirror of jm_spectrum_plan."""

    _fields_ = [
        ("n", c_int32), ("n_terms", c_int32), ("n_clusters", c_int32), ("n_seg", c_int32),
        ("n_slots", c_int32), ("reserved", c_int32),
        ("cl_terms", c_void_p), ("cl_toff", c_void_p), ("cl_seg0", c_void_p), ("cl_nseg", c_void_p),
        ("seg_start", c_void_p), ("seg_len", c_void_p), ("seg_cluster", c_void_p),
        ("seg_slot", c_void_p), ("delta", c_void_p), ("term_cluster", c_void_p),
        ("term_m", c_void_p),
    ]


class ServeConfig(ctypes.Structure):
    """Mirror of jm_serve_config."""

    _fields_ = [("slice_snps", c_int64), ("max_batch", c_int64), ("max_reqs", c_int32),
                ("max_inflight", c_int32), ("max_skips", c_int32), ("max_cvt", c_int32),
                ("flags", c_int32), ("reserved", c_int32)]


class ServeCohortDesc(ctypes.Structure):
    """Mirror of jm_serve_cohort_desc."""

    _fields_ = [("n", c_int32), ("precision", c_int32), ("n_snps", c_int64), ("g", c_void_p),
                ("ld_g", c_int64), ("means", c_void_p), ("UT", c_void_p), ("ld_u", c_int64),
                ("planes", c_void_p), ("nplanes", c_int32), ("n_terms", c_int32),
                ("n_clusters", c_int32), ("max_batch", c_int32)]


class ServeDone(ctypes.Structure):
    """Mirror of jm_serve_done."""

    _fields_ = [("id", c_int64), ("status", c_int32), ("reserved", c_int32),
                ("message", ctypes.c_char * 240)]


class ServeStats(ctypes.Structure):
    """Mirror of jm_serve_stats_t."""

    _fields_ = [("queued", c_int64), ("active", c_int64), ("steps", c_int64), ("snps", c_int64),
                ("arena_used", c_int64), ("arena_bytes", c_int64), ("max_active", c_int64),
                ("reserved", c_int64)]


JM_SERVE_F64, JM_SERVE_MIXED, JM_SERVE_STUB, JM_SERVE_MANUAL = 0, 1, 1, 2

_SIGS = {
    "jm_version": ([], c_int),
    "jm_last_error": ([ctypes.c_char_p, c_size_t], c_int),
    "jm_decode_bed": ([c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_int64,
                       c_void_p, c_void_p], c_int),
    "jm_transpose_i8": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64, c_void_p,
                         c_int64, c_void_p], c_int),
    "jm_center_columns_f64": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p], c_int),
    "jm_impute_rows_f64": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p], c_int),
    "jm_kinship_syrk_i8": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p], c_int),
    "jm_pack_gram_2bit": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_int64, c_int64,
                           c_void_p], c_int),
    "jm_kinship_syrk_2bit": ([c_void_p, c_int64, c_int64, c_int64, c_void_p, c_void_p], c_int),
    "jm_kinship_miss_ws_len": ([c_int64, c_int64, c_int64], c_int64),
    "jm_kinship_impute_terms": ([c_void_p, c_int64, c_void_p, c_int64, c_int64, c_int64, c_void_p,
                                 c_void_p, c_void_p, c_int, c_void_p, c_int64, c_int64, c_void_p],
                                c_int),
    "jm_kinship_finalize": ([c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                             c_double, c_int64, c_double, c_void_p, c_void_p], c_int),
    "jm_rotate_f64": ([c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64,
                       c_void_p, c_int64, c_void_p], c_int),
    "jm_gemm_nt_f64": ([c_void_p, c_int64, c_void_p, c_int64, c_int64, c_int64, c_int64, c_void_p,
                        c_int64, c_void_p], c_int),
    "jm_packed_ut_bytes": ([c_int64, c_int], c_int64),
    "jm_packed_geno_bytes": ([c_int64, c_int64], c_int64),
    "jm_pack_ut_bf16x3": ([c_void_p, c_int64, c_int64, c_int, c_void_p, c_void_p], c_int),
    "jm_pack_genotypes": ([c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                           c_void_p, c_void_p], c_int),
    "jm_rotate_bf16x3": ([c_void_p, c_int, c_void_p, c_void_p, c_int64, c_int64, c_void_p,
                          c_int64, c_void_p], c_int),
    "jm_moments_work_bytes": ([c_void_p, c_int64, c_int], c_int64),
    "jm_spectrum_moments": ([c_void_p, c_int, c_int64, c_int64, c_void_p, c_int64, c_int, c_void_p,
                             c_void_p, c_void_p, c_int64, c_void_p], c_int),
    "jm_moments_scratch_bytes": ([], c_int64),
    "jm_spectrum_moments_noalloc": ([c_void_p, c_int, c_int64, c_int64, c_void_p, c_int64, c_int,
                                     c_void_p, c_void_p, c_int64, c_void_p], c_int),
    "jm_spectrum_moments_multi": ([c_void_p, c_int, c_int64, c_int64, c_void_p, c_int, c_void_p,
                                   c_void_p, c_void_p], c_int),
    "jm_lmm_assoc_multi": ([c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_int, c_void_p,
                            c_void_p, c_int64, c_int64, c_void_p, c_void_p], c_int),
    "jm_lmm_null": ([POINTER(LmmProblem), c_void_p, c_void_p], c_int),
    "jm_lmm_prepare": ([POINTER(LmmProblem), c_void_p], c_int),
    "jm_lmm_assoc": ([POINTER(LmmProblem), c_int, c_void_p, c_int64, c_int64, c_void_p, c_void_p],
                     c_int),
    "jm_dgemm": ([c_int, c_int, c_int64, c_int64, c_int64, c_double, c_void_p, c_int64, c_void_p,
                  c_int64, c_double, c_void_p, c_int64, c_int, c_void_p], c_int),
    "jm_symv_work_doubles": ([c_int64], c_int64),
    "jm_symv_lower": ([c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64,
                       c_void_p], c_int),
    "jm_dgemm2": ([c_int, c_int, c_int64, c_int64, c_int64, c_double, c_void_p, c_void_p,
                   c_void_p, c_void_p, c_int64, c_int64, c_double, c_void_p, c_int64, c_int,
                   c_void_p], c_int),
    "jm_dgemm_splitk": ([c_int, c_int, c_int64, c_int64, c_int64, c_double, c_void_p, c_int64,
                         c_void_p, c_int64, c_double, c_void_p, c_int64, c_int, c_void_p, c_int64,
                         c_void_p], c_int),
    "jm_tridiag_larfg": ([c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "jm_tridiag_column_work_doubles": ([c_int64], c_int64),
    "jm_tridiag_column": ([c_int64, c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                           c_void_p, c_int64, c_void_p], c_int),
    "jm_tridiag_panel_update": ([c_int64, c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                                 c_void_p], c_int),
    "jm_tridiag_panel_work_doubles": ([c_int64], c_int64),
    "jm_tridiag_panel_step": ([c_int64, c_int, c_void_p, c_int64, c_void_p, c_int64, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_int64, c_void_p], c_int),
    "jm_tridiag_finish": ([c_int64, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "jm_secular_solve": ([c_int64, c_void_p, c_void_p, c_void_p, c_double, c_double, c_void_p,
                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "jm_secular_vectors": ([c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                            c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p],
                           c_int),
    "jm_apply_rotations": ([c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_int64, c_void_p],
                           c_int),
    "jm_format_assoc_rows": ([c_int64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_int, c_int64, c_void_p, c_int64, c_int],
                             c_int64),
    "jm_format_g10_rows": ([c_int64, c_int64, c_void_p, c_int64, c_void_p, c_int64, c_int],
                           c_int64),
    "jm_parse_text_rows": ([c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_int],
                           c_int64),
    "jm_host_betainc": ([c_double, c_double, c_double], c_double),
    "jm_host_chi2_sf1": ([c_double], c_double),
    "jm_host_f_pvalue": ([c_double, c_double], c_double),
    "jm_host_lrt_pvalue": ([c_double, c_double], c_double),
    "jm_serve_cohort_ws_bytes": ([POINTER(ServeConfig), POINTER(ServeCohortDesc), c_void_p],
                                 c_int64),
    "jm_serve_request_ws_bytes": ([POINTER(ServeConfig), c_int, c_int], c_int64),
    "jm_serve_create": ([POINTER(ServeConfig), c_void_p, c_int64, POINTER(c_void_p)], c_int),
    "jm_serve_add_cohort": ([c_void_p, POINTER(ServeCohortDesc), c_void_p, POINTER(c_int)],
                            c_int),
    "jm_serve_submit": ([c_void_p, c_int, POINTER(LmmProblem), c_void_p, c_int64, c_int, c_int64,
                         c_int64, c_void_p, POINTER(c_int64)], c_int),
    "jm_serve_poll": ([c_void_p, POINTER(ServeDone), c_int, POINTER(c_int)], c_int),
    "jm_serve_wait": ([c_void_p, c_int], c_int),
    "jm_serve_stats": ([c_void_p, POINTER(ServeStats)], c_int),
    "jm_serve_run_once": ([c_void_p], c_int64),
    "jm_serve_pause": ([c_void_p, c_int], c_int),
    "jm_serve_cohort_stream": ([c_void_p, c_int, POINTER(c_void_p)], c_int),
    "jm_serve_inject_step_error": ([c_void_p, c_int], c_int),
    "jm_serve_destroy": ([c_void_p], c_int),
}

EXPORTED = tuple(_SIGS)


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"libjamma_hip.so not found at {LIB_PATH}; build it with "
            "`python jamma_amd/csrc/build.py` (hipcc, gfx950)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib.jm_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed (rc={rc}): {last_error()}")


def call(name: str, *args) -> None:
    """Call an int-returning entry point and raise on a non-zero status."""
    check(getattr(lib, name)(*args), name)
