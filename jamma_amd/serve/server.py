"""Continuous-batching association server over resident cohorts.

A cohort (kinship eigendecomposition + int8 genotypes + imputation means) stays
resident in HBM. A request is one phenotype (+ covariates) against a
contiguous SNP range of a cohort. Requests are admitted against an HBM budget
as memory frees up; every scheduler step takes the next slice of SNPs of every
active request of a cohort and runs them as ONE batch: one rotation launch
(gathered genotype rows), one spectrum-moments launch (per-column request
descriptors) and one lambda-search launch (per-SNP problem index). Requests
that arrive while others run join the next step, so their work overlaps on
the device; each cohort uses its own stream. Per-SNP results do not depend on
the batch composition (every output is computed by a fixed instruction
sequence of its own column / SNP).

The reference has no serving layer; its per-request work is
run_lmm_association_streaming (src/jamma/lmm/runner_streaming.py:94-628) with
a precomputed eigendecomposition (prepare.py:77-111).
"""

from __future__ import annotations

import ctypes
import itertools
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np
import torch

from jamma_amd import _lib
from jamma_amd.device import ptr, require_gpu, round_up, to_dev
from jamma_amd.lmm.engine import ctypes_ref, AssocProblem, Cohort


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
        self.stream = torch.cuda.Stream()  # default priority

    def resident_bytes(self) -> int:
        return self.cohort.hbm_bytes() + self.g.numel() + self.means.numel() * 8


MOMCOL_DTYPE = np.dtype([("V", np.uint64), ("ldv", np.int64), ("nv", np.int32),
                         ("reserved", np.int32), ("out", np.uint64), ("ldm", np.int64)])
assert MOMCOL_DTYPE.itemsize == ctypes.sizeof(_lib.MomCol)


class _Active:
    """Per-request device state while admitted."""

    def __init__(self, req: Request, rc: ResidentCohort, slice_snps: int):
        self.req = req
        self.rc = rc
        self.prob = AssocProblem(rc.cohort, req.phenotype, req.covariates)
        self.next = req.snp_start
        self.out = torch.empty((req.n_snps, _lib.JM_NOUT), dtype=torch.float64, device=rc.cohort.device)
        T = rc.cohort.n_terms
        self.rows = self.prob.n_cvt + 2
        self.xmom = torch.empty((slice_snps, self.rows, T), dtype=torch.float64,
                                device=rc.cohort.device)

    @staticmethod
    def cost_bytes(rc: ResidentCohort, req: Request, slice_snps: int) -> int:
        c = 1 if req.covariates is None else np.asarray(req.covariates).shape[1]
        T = rc.cohort.n_terms
        ni = (c + 3) * (c + 2) // 2
        tables = (T * 51 + 51 * ni + ni * T + 4 * T) * 8 + (c + 1) * rc.n * 8 * 2
        return tables + slice_snps * (c + 2) * T * 8 + req.n_snps * _lib.JM_NOUT * 8

    def done(self) -> bool:
        return self.next >= self.req.snp_end


class AdmissionQueue:
    """First-fit admission in arrival order against a byte budget (host logic).

    ``try_admit(cost, create)`` is called for queued requests in order; a request
    is admitted when ``used + cost <= budget`` and ``create`` succeeds. A request
    whose ``create`` raises fails alone (its error is returned) without holding
    budget; requests that do not fit stay queued in order.
    """

    def __init__(self, budget: int):
        self.budget = int(budget)
        self.used = 0
        self.queue: list = []

    def fits_ever(self, cost: int) -> bool:
        return cost <= self.budget

    def admit(self, cost_of, create):
        """-> (admitted [(item, obj, cost)], failed [(item, exc)])."""
        admitted, failed, keep = [], [], []
        for item in self.queue:
            cost = cost_of(item)
            if self.used + cost <= self.budget:
                try:
                    obj = create(item)
                except Exception as e:  # noqa: BLE001 - reported to the request
                    failed.append((item, e))
                    continue
                self.used += cost
                admitted.append((item, obj, cost))
            else:
                keep.append(item)
        self.queue = keep
        return admitted, failed

    def release(self, cost: int) -> None:
        self.used -= cost


class Server:
    """Scheduler thread owning the device work of every resident cohort."""

    def __init__(self, hbm_budget_bytes: int | None = None, slice_snps: int = 2048,
                 max_batch_snps: int = 16384):
        self.device = require_gpu()
        self.cohorts: dict[str, ResidentCohort] = {}
        self.slice_snps = slice_snps
        self.max_batch = max_batch_snps
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget_bytes = int(free * 0.5)
        self.adm = AdmissionQueue(hbm_budget_bytes)
        self._active: dict[str, list[_Active]] = {}
        self._cv = threading.Condition()
        self._ids = itertools.count()
        self._thread = None
        self._stop = False
        self.completed: list[Request] = []
        self.snps_done = 0
        self.batches = 0
        self.max_concurrent = 0

    # -- cohorts --------------------------------------------------------------
    def add_cohort(self, rc: ResidentCohort) -> None:
        self.cohorts[rc.name] = rc
        self._active[rc.name] = []

    # -- requests -------------------------------------------------------------
    def submit(self, req: Request) -> Future:
        rc = self.cohorts.get(req.cohort)
        if rc is None:
            raise KeyError(f"unknown cohort {req.cohort!r}")
        if not (0 <= req.snp_start < req.snp_end <= rc.p):
            raise ValueError(f"SNP range [{req.snp_start}, {req.snp_end}) outside [0, {rc.p})")
        if req.mode not in (1, 2, 3, 4):
            raise ValueError("mode must be 1..4")
        if not self.adm.fits_ever(_Active.cost_bytes(rc, req, self.slice_snps)):
            raise MemoryError("request can never fit in the HBM budget")
        req.id = next(self._ids)
        req.t_submit = time.perf_counter()
        with self._cv:
            self.adm.queue.append(req)
            self._cv.notify()
        return req.future

    def start(self) -> "Server":
        self._stop = False
        self._thread = threading.Thread(target=self._loop, name="jm-serve", daemon=True)
        self._thread.start()
        return self

    def stop(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread:
            self._thread.join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- scheduler --------------------------------------------------------------
    def _admit(self) -> None:
        """First-fit admission in arrival order against the HBM budget."""

        def create(req):
            rc = self.cohorts[req.cohort]
            with torch.cuda.stream(rc.stream):
                return _Active(req, rc, self.slice_snps)

        admitted, failed = self.adm.admit(
            lambda r: _Active.cost_bytes(self.cohorts[r.cohort], r, self.slice_snps), create)
        for req, exc in failed:
            req.future.set_exception(exc)
        for req, act, cost in admitted:
            act.cost = cost
            req.t_admit = time.perf_counter()
            self._active[req.cohort].append(act)

    def _step_cohort(self, name: str) -> list[_Active]:
        acts = self._active[name]
        if not acts:
            return []
        rc = self.cohorts[name]
        co = rc.cohort
        T = co.n_terms
        with torch.cuda.stream(rc.stream):
            cols, vptr, ldv, nv, xptr, snp_prob, dests = [], [], [], [], [], [], []
            budget = self.max_batch
            for k, a in enumerate(acts):
                if budget <= 0:
                    break
                take = min(self.slice_snps, a.req.snp_end - a.next, budget)
                if take <= 0:
                    continue
                budget -= take
                cols.append(np.arange(a.next, a.next + take, dtype=np.int32))
                stride = a.rows * T * 8
                xptr.append(a.xmom.data_ptr() + stride * np.arange(take, dtype=np.uint64))
                vptr.append(np.full(take, a.prob.V.data_ptr(), dtype=np.uint64))
                ldv.append(np.full(take, a.prob.V.shape[1], dtype=np.int64))
                nv.append(np.full(take, a.prob.n_cvt + 1, dtype=np.int32))
                snp_prob.append(np.full(take, k, dtype=np.int32))
                dests.append((a, a.next - a.req.snp_start, take))
                a.next += take
            if not cols:
                return []
            B = int(sum(len(c) for c in cols))
            col_idx = to_dev(np.concatenate(cols), torch.int32)
            Z, f32, ldz = co.rotate_genotypes(rc.g, rc.ld, rc.means, B, rc.precision, col_idx)
            desc = np.zeros(B, dtype=MOMCOL_DTYPE)
            desc["V"] = np.concatenate(vptr)
            desc["ldv"] = np.concatenate(ldv)
            desc["nv"] = np.concatenate(nv)
            xp = np.concatenate(xptr)
            desc["out"] = xp
            desc["ldm"] = T
            desc_arr = desc
            desc_dev = to_dev(np.frombuffer(desc_arr, dtype=np.uint8), torch.uint8)
            nv_max = max(a.prob.n_cvt for a in acts) + 1
            work = co.moments_work(B, nv_max)
            _lib.call("jm_spectrum_moments_multi", ptr(Z), int(f32), ldz, B, ptr(desc_dev),
                      nv_max, ctypes_ref(co.plan_c), ptr(work), rc.stream.cuda_stream)
            probs = (_lib.LmmProblem * len(acts))(*[a.prob.prob for a in acts])
            probs_dev = to_dev(np.frombuffer(probs, dtype=np.uint8), torch.uint8)
            modes = to_dev(np.array([a.req.mode for a in acts], dtype=np.int32), torch.int32)
            sp = to_dev(np.concatenate(snp_prob), torch.int32)
            sx = to_dev(xp.view(np.int64), torch.int64)
            out = torch.empty((B, _lib.JM_NOUT), dtype=torch.float64, device=co.device)
            max_cvt = max(a.prob.n_cvt for a in acts)
            _lib.call("jm_lmm_assoc_multi", ptr(probs_dev), ptr(modes), len(acts), max_cvt, T,
                      ptr(sp), ptr(sx), T, B, ptr(out), rc.stream.cuda_stream)
            o = 0
            for a, off, take in dests:
                a.out[off:off + take].copy_(out[o:o + take])
                o += take
            self.batches += 1
            self.snps_done += B
            finished = [a for a in acts if a.done()]
            if finished:
                rc.stream.synchronize()
            # keep the host arrays referenced by the launches alive until here
            del desc_arr, probs
        return finished

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not self._stop and not self.adm.queue and not any(self._active.values()):
                    self._cv.wait()
                if self._stop and not self.adm.queue and not any(self._active.values()):
                    return
                self._admit()
            self.max_concurrent = max(self.max_concurrent,
                                      sum(len(v) for v in self._active.values()))
            for name in list(self._active):
                finished = self._step_cohort(name)
                for a in finished:
                    res = a.out.cpu().numpy()
                    a.req.t_done = time.perf_counter()
                    self._active[name].remove(a)
                    with self._cv:
                        self.adm.release(a.cost)
                    self.completed.append(a.req)
                    a.req.future.set_result(res)

    # -- reporting -------------------------------------------------------------
    def latency_stats(self) -> dict:
        lat = np.array([r.t_done - r.t_submit for r in self.completed])
        if lat.size == 0:
            return {}
        snps = sum(r.n_snps for r in self.completed)
        span = max(r.t_done for r in self.completed) - min(r.t_submit for r in self.completed)
        return {"requests": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "snps": int(snps), "snps_per_s": snps / span if span > 0 else float("nan"),
                "batches": self.batches, "max_concurrent": self.max_concurrent}
