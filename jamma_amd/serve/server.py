"""Continuous-batching association server over resident cohorts.

A cohort (kinship eigendecomposition + int8 genotypes + imputation means) stays
resident in HBM. A request is one phenotype (+ covariates) against a contiguous SNP
range of a cohort. Scheduling runs in the library's native runtime (`jm_serve_*`,
csrc/serve.cpp), which owns a scheduler thread and one HBM arena given to it here:

* this module prepares each request on worker threads (rotated covariates and
  phenotype, their spectrum moments, the device null model and grid tables, on the
  cohort's stream, in order with its steps) and hands the prepared problem to the
  runtime with a pinned host result block of its own;
* the runtime admits requests in submission order against the arena (first fit, with
  aging: a request that later ones overtook `max_skips` times stops further
  overtaking), and every step of a cohort takes the next slice of each active
  request and launches ONE genotype pack + rotation, ONE spectrum-moments launch and
  ONE lambda-search launch over the whole batch on the cohort's stream, with the
  step's descriptors in pinned slots and every device buffer carved from the arena
  (no allocation on the serving path); steps are queued ahead of the device up to
  `max_inflight` and completions come back through events;
* a poller thread here waits on the runtime's completion records and resolves the
  requests' futures; a step that fails fails that cohort's active requests alone.
Per-SNP results do not depend on the batch: a SNP's column of the rotation, its
moments and its search are computed by the same instructions whatever else is in the
step.

The reference has no serving layer; its per-request work is
run_lmm_association_streaming (src/jamma/lmm/runner_streaming.py:94-628) with a
precomputed eigendecomposition (prepare.py:77-111).
"""

from __future__ import annotations

import ctypes
import itertools
import threading
import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from jamma_amd import _lib
from jamma_amd.device import ptr, require_gpu
from jamma_amd.lmm.engine import AssocProblem, Cohort


@dataclass
class Request:
    cohort: str
    phenotype: np.ndarray
    snp_start: int
    snp_end: int
    covariates: np.ndarray | None = None
    mode: int = 4
    id: int = -1
    t_submit: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    future: Future = field(default_factory=Future, repr=False)

    @property
    def n_snps(self) -> int:
        return self.snp_end - self.snp_start


class ResidentCohort:
    """Eigendecomposition + genotypes of one cohort, resident in HBM."""

    def __init__(self, name: str, genotypes_i8: torch.Tensor, ld: int, means: torch.Tensor,
                 eigenvalues, eigenvectors, precision: str = "mixed"):
        self.name = name
        self.g = genotypes_i8  # (p x ld) int8 SNP-major, -1 missing
        self.ld = ld
        self.p = genotypes_i8.shape[0]
        self.means = means  # (p,) float64 (float32-derived imputation means)
        self.precision = precision
        self.cohort = Cohort(eigenvalues, eigenvectors,
                             precision="both" if precision == "mixed" else "f64")
        self.n = self.cohort.n

    def resident_bytes(self) -> int:
        return self.cohort.hbm_bytes() + self.g.numel() + self.means.numel() * 8

    def desc(self, max_batch: int = 0) -> _lib.ServeCohortDesc:
        """The runtime's description of this cohort (device arrays stay ours); max_batch
        caps the SNPs of one step (0: the server's cap)."""
        co = self.cohort
        mixed = self.precision == "mixed"
        return _lib.ServeCohortDesc(
            n=self.n, precision=_lib.JM_SERVE_MIXED if mixed else _lib.JM_SERVE_F64,
            n_snps=self.p, g=ptr(self.g), ld_g=self.ld, means=ptr(self.means), UT=ptr(co.UT),
            ld_u=co.ld_u, planes=ptr(co.split) if mixed else None,
            nplanes=co.bf16_planes if mixed else 0, n_terms=co.n_terms,
            n_clusters=int(co.plan_c.n_clusters), max_batch=int(max_batch))


def serve_config(slice_snps: int = 2048, max_batch_snps: int = 16384, max_inflight: int = 2,
                 max_skips: int = 8, max_reqs: int = 64, max_cvt: int = 12) -> _lib.ServeConfig:
    return _lib.ServeConfig(slice_snps=slice_snps, max_batch=max(max_batch_snps, slice_snps),
                            max_reqs=max_reqs, max_inflight=max_inflight, max_skips=max_skips,
                            max_cvt=max_cvt, flags=0)


def cohort_step_snps(n: int, cfg: _lib.ServeConfig, step_work: float) -> int:
    """SNPs per step of a cohort of n samples: at most step_work / n^2 (the rotation
    of a step costs 2 n^2 FLOP per SNP), whole slices, between one slice and the
    server's max_batch. Short steps of a large cohort bound how long other work queued
    behind them on the device (another cohort's step, a request's setup) waits."""
    sl = int(cfg.slice_snps)
    cap = int(step_work // (float(n) * n)) // sl * sl
    return max(sl, min(int(cfg.max_batch), cap))


def cohort_workspace_bytes(rc: ResidentCohort, cfg: _lib.ServeConfig, max_batch: int = 0) -> int:
    """Arena bytes of a cohort's step workspace under `cfg` (steps of at most
    max_batch SNPs; 0: cfg.max_batch)."""
    return int(_lib.lib.jm_serve_cohort_ws_bytes(ctypes.byref(cfg),
                                                 ctypes.byref(rc.desc(max_batch)),
                                                 ctypes.byref(rc.cohort.plan_c)))


def request_workspace_bytes(rc: ResidentCohort, cfg: _lib.ServeConfig, n_cvt: int) -> int:
    """Arena bytes an admitted request holds (its moments slice)."""
    return int(_lib.lib.jm_serve_request_ws_bytes(ctypes.byref(cfg), rc.cohort.n_terms,
                                                  n_cvt + 1))


class Server:
    """Python front of the native serving runtime (one per device)."""

    def __init__(self, hbm_budget_bytes: int | None = None, slice_snps: int = 2048,
                 max_batch_snps: int = 16384, max_inflight: int = 2, setup_workers: int = 2,
                 max_skips: int = 8, max_reqs: int = 64, max_cvt: int = 12,
                 step_work: float | None = None):
        """step_work: n^2 x SNPs of one step of a cohort (default: max_batch_snps SNPs
        at n = 10,000); see cohort_step_snps."""
        self.device = require_gpu()
        self.cfg = serve_config(slice_snps, max_batch_snps, max_inflight, max_skips, max_reqs,
                                max_cvt)
        self.step_work = float(step_work) if step_work is not None else \
            float(self.cfg.max_batch) * 1e8
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget_bytes = int(free * 0.5)
        self.budget = max(int(hbm_budget_bytes), 256)
        # the runtime's HBM: every device buffer it uses is carved from this block
        self.arena = torch.empty((self.budget,), dtype=torch.uint8, device=self.device)
        self._h = ctypes.c_void_p()
        _lib.call("jm_serve_create", ctypes.byref(self.cfg), ptr(self.arena), self.budget,
                  ctypes.byref(self._h))
        self.cohorts: dict[str, ResidentCohort] = {}
        self._cid: dict[str, int] = {}
        self.step_snps: dict[str, int] = {}  # SNPs per step of each cohort
        self._streams: dict[str, torch.cuda.ExternalStream] = {}
        self._reserved = 0
        self._pending: dict[int, tuple] = {}  # runtime id -> (request, problem, out)
        self._setups = 0
        self._lock = threading.Lock()
        self._pool = ThreadPoolExecutor(max_workers=max(1, setup_workers),
                                        thread_name_prefix="jm-setup")
        self._ids = itertools.count()
        self._poller = None
        self._stop = False
        self.completed: list[Request] = []
        self.failed: list[Request] = []
        self._final_stats = None

    # -- cohorts --------------------------------------------------------------
    def add_cohort(self, rc: ResidentCohort) -> None:
        mb = cohort_step_snps(rc.n, self.cfg, self.step_work)
        need = cohort_workspace_bytes(rc, self.cfg, mb)
        if need > self.budget - self._reserved:
            raise MemoryError(f"the HBM budget cannot hold cohort {rc.name!r}'s step workspace "
                              f"({need} bytes)")
        cid = ctypes.c_int()
        _lib.call("jm_serve_add_cohort", self._h, ctypes.byref(rc.desc(mb)),
                  ctypes.byref(rc.cohort.plan_c), ctypes.byref(cid))
        self.step_snps[rc.name] = mb
        h = ctypes.c_void_p()
        _lib.call("jm_serve_cohort_stream", self._h, cid.value, ctypes.byref(h))
        # request setup runs on the cohort's own stream: no extra stream shares a
        # hardware queue with (and waits behind) another cohort's steps
        self._streams[rc.name] = torch.cuda.ExternalStream(h.value, device=rc.cohort.device)
        self._reserved += (need + 255) // 256 * 256
        self.cohorts[rc.name] = rc
        self._cid[rc.name] = cid.value

    # -- requests -------------------------------------------------------------
    def submit(self, req: Request) -> Future:
        rc = self.cohorts.get(req.cohort)
        if rc is None:
            raise KeyError(f"unknown cohort {req.cohort!r}")
        if not (0 <= req.snp_start < req.snp_end <= rc.p):
            raise ValueError(f"SNP range [{req.snp_start}, {req.snp_end}) outside [0, {rc.p})")
        if req.mode not in (1, 2, 3, 4):
            raise ValueError("mode must be 1..4")
        c = 1 if req.covariates is None else np.asarray(req.covariates).shape[1]
        if c > self.cfg.max_cvt:
            raise ValueError(f"{c} covariate columns exceed the server's max_cvt "
                             f"({self.cfg.max_cvt})")
        if request_workspace_bytes(rc, self.cfg, c) > self.budget - self._reserved:
            raise MemoryError("request can never fit in the HBM budget")
        req.id = next(self._ids)
        req.t_submit = time.perf_counter()
        with self._lock:
            self._setups += 1
        self._pool.submit(self._setup, req)
        return req.future

    def _setup(self, req: Request) -> None:
        """Worker thread: the request's device tables on the cohort's setup stream, then
        the hand-over to the runtime."""
        rc = self.cohorts[req.cohort]
        try:
            st = self._streams[req.cohort]
            with torch.cuda.device(rc.cohort.device), torch.cuda.stream(st):
                prob = AssocProblem(rc.cohort, req.phenotype, req.covariates)
                st.synchronize()  # tables complete before the steps read them
            # pinned host result block: each step copies its rows there on the cohort's
            # stream, so a completed request needs no further device-to-host copy
            out = torch.empty((req.n_snps, _lib.JM_NOUT), dtype=torch.float64, pin_memory=True)
            rid = ctypes.c_int64()
            with self._lock:  # registered before the runtime can complete it
                _lib.call("jm_serve_submit", self._h, self._cid[req.cohort],
                          ctypes.byref(prob.prob), ptr(prob.V), prob.V.shape[1], int(req.mode),
                          req.snp_start, req.snp_end, out.data_ptr(), ctypes.byref(rid))
                self._pending[rid.value] = (req, prob, out)
                req.t_admit = time.perf_counter()
        except Exception as e:  # noqa: BLE001 - reported to the request
            req.t_done = time.perf_counter()
            self.failed.append(req)
            req.future.set_exception(e)
        finally:
            with self._lock:
                self._setups -= 1

    def _resolve(self, rec) -> None:
        with self._lock:
            req, prob, out = self._pending.pop(rec.id)
        if rec.status == 0:
            res = out.numpy().copy()  # the step's event has completed: the rows are final
            req.t_done = time.perf_counter()
            self.completed.append(req)
            req.future.set_result(res)
        else:
            req.t_done = time.perf_counter()
            self.failed.append(req)
            req.future.set_exception(RuntimeError(
                f"serving step failed (rc={rec.status}): {rec.message.decode(errors='replace')}"))

    def _poll_loop(self) -> None:
        recs = (_lib.ServeDone * 64)()
        got = ctypes.c_int()
        while True:
            _lib.lib.jm_serve_wait(self._h, 20)  # blocks in native code, GIL released
            _lib.call("jm_serve_poll", self._h, recs, 64, ctypes.byref(got))
            for i in range(got.value):
                self._resolve(recs[i])
            with self._lock:
                idle = not self._pending and self._setups == 0
            if self._stop and idle and got.value == 0:
                return

    def start(self) -> "Server":
        self._stop = False
        self._poller = threading.Thread(target=self._poll_loop, name="jm-serve-poll", daemon=True)
        self._poller.start()
        return self

    def stop(self) -> None:
        """Waits for every submitted request, then shuts the runtime down."""
        self._pool.shutdown(wait=True)
        self._stop = True
        if self._poller:
            self._poller.join()
            self._poller = None
        if self._h:
            self._final_stats = self.stats()
            _lib.call("jm_serve_destroy", self._h)
            self._h = ctypes.c_void_p()
            self._streams.clear()  # the runtime's streams are gone
            self.arena = None  # back to the caching allocator

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._pool.shutdown(wait=True)
                _lib.lib.jm_serve_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:  # noqa: BLE001 - interpreter teardown
            pass

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    def pause(self) -> None:
        """No admission and no new step until resume(); steps already launched finish."""
        _lib.call("jm_serve_pause", self._h, 1)

    def resume(self) -> None:
        _lib.call("jm_serve_pause", self._h, 0)

    def wait_prepared(self, timeout: float = 300.0) -> None:
        """Blocks until every submitted request has been prepared and handed to the
        runtime (or failed in preparation)."""
        t_end = time.perf_counter() + timeout
        while True:
            with self._lock:
                if self._setups == 0:
                    return
            if time.perf_counter() > t_end:
                raise TimeoutError("request preparation did not finish")
            time.sleep(0.002)

    def inject_step_error(self, cohort: str) -> None:
        """Failure-handling hook: the cohort's next step fails before any launch."""
        _lib.call("jm_serve_inject_step_error", self._h, self._cid[cohort])

    # -- reporting -------------------------------------------------------------
    def stats(self) -> dict:
        if not self._h:
            return dict(self._final_stats or {})
        st = _lib.ServeStats()
        _lib.call("jm_serve_stats", self._h, ctypes.byref(st))
        return {k: int(getattr(st, k)) for k, _ in st._fields_}

    def latency_stats(self) -> dict:
        lat = np.array([r.t_done - r.t_submit for r in self.completed])
        if lat.size == 0:
            return {}
        st = self.stats() if self._h else self._final_stats
        snps = sum(r.n_snps for r in self.completed)
        span = max(r.t_done for r in self.completed) - min(r.t_submit for r in self.completed)
        return {"requests": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "snps": int(snps), "snps_per_s": snps / span if span > 0 else float("nan"),
                "batches": st["steps"], "max_concurrent": st["max_active"],
                "failed": len(self.failed)}
