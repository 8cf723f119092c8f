"""Continuous-batching association server over resident cohorts.

A cohort (kinship eigendecomposition + int8 genotypes + imputation means) stays
resident in HBM. A request is one phenotype (+ covariates) against a contiguous SNP
range of a cohort. The scheduler thread only launches device work and never waits
for it:

* admission reserves HBM budget in arrival order (first fit, with aging: a request
  that later ones have overtaken ``max_skips`` times blocks further overtaking until
  it fits); the per-request setup (rotated covariates / phenotype, their spectrum
  moments, the device null model, the grid tables) then runs on a worker thread and
  a setup stream of the cohort, so the null model's host read-back never stalls the
  scheduler;
* every scheduler step takes the next slice of every active request of a cohort and
  launches, on that cohort's stream, one pass of the hot path over the whole batch:
  ONE genotype pack + rotation (a gather of the requests' SNP rows), ONE spectrum-
  moments launch (a descriptor per column: its request's partner rows and output
  block) and ONE lambda search + tests launch (per-SNP problem index; 16 lanes per SNP
  for spectra of up to 256 terms, each block reloading a problem's rows where the
  batch moves to the next request); the step's descriptors go up in one pinned,
  non-blocking copy, and the results are copied into each request's rows on the
  stream;
* at most ``max_inflight`` steps per cohort are queued ahead of the device; steps,
  result copies (to pinned memory) and completions are tracked with events that the
  loop polls, so requests that arrive while others run join the next step and their
  work overlaps on the device (one stream per cohort, all at default priority);
* an error in a step fails that cohort's active requests alone (their budget is
  released) and the loop goes on.
Per-SNP results do not depend on the batch: a SNP's column of the rotation, its
moments and its search are computed by the same instructions whatever else is in
the step.

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
from jamma_amd.lmm.engine import AssocProblem, Cohort, ctypes_ref


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
        self.stream = torch.cuda.Stream()        # steps (default priority)
        self.setup_stream = torch.cuda.Stream()  # per-request setup (default priority)

    def resident_bytes(self) -> int:
        return self.cohort.hbm_bytes() + self.g.numel() + self.means.numel() * 8


class _Active:
    """Per-request device state while admitted."""

    def __init__(self, req: Request, rc: ResidentCohort, prob: AssocProblem, slice_snps: int):
        self.req = req
        self.rc = rc
        self.prob = prob
        self.next = req.snp_start
        dev = rc.cohort.device
        T = rc.cohort.n_terms
        self.rows = prob.n_cvt + 2
        self.out = torch.empty((req.n_snps, _lib.JM_NOUT), dtype=torch.float64, device=dev)
        self.xmom = torch.empty((slice_snps, self.rows, T), dtype=torch.float64, device=dev)
        self.host = None       # pinned copy of out once the last slice is launched
        self.done_event = None
        self.cost = 0

    @staticmethod
    def cost_bytes(rc: ResidentCohort, req: Request, slice_snps: int) -> int:
        """HBM held by the request while admitted: problem tables, rotated covariates,
        its moments slice and result rows, and its share of a step's buffers (rotated
        columns, packed genotypes, moment partials) at full slice."""
        c = 1 if req.covariates is None else np.asarray(req.covariates).shape[1]
        T = rc.cohort.n_terms
        ni = (c + 3) * (c + 2) // 2
        n = rc.n
        tables = (T * 51 + 51 * ni + ni * T + 4 * T) * 8 + (c + 1) * n * 8 * 3
        own = slice_snps * (c + 2) * T * 8 + req.n_snps * _lib.JM_NOUT * 8
        z_col = n * (8 if rc.precision == "f64" else 4) + (n + 64)  # Z + packed genotypes
        work_col = (c + 2) * rc.cohort.plan.n_slots * 8
        return tables + own + slice_snps * (z_col + work_col)

    def remaining(self) -> int:
        return self.req.snp_end - self.next


class AdmissionQueue:
    """First-fit admission in arrival order against a byte budget, with aging (host
    logic, no device work).

    ``admit(cost_of)`` walks the queue in order and reserves budget for every item
    with ``used + cost <= budget``; an item that does not fit stays queued and counts
    one skip each time a later item is admitted past it. Once the oldest waiting item
    has ``max_skips`` skips, nothing behind it is admitted until it fits, so a large
    request cannot be starved by a stream of small ones."""

    def __init__(self, budget: int, max_skips: int = 8):
        self.budget = int(budget)
        self.used = 0
        self.max_skips = int(max_skips)
        self.queue: list = []
        self._skips: dict = {}

    def fits_ever(self, cost: int) -> bool:
        return cost <= self.budget

    def admit(self, cost_of):
        """-> [(item, cost)] admitted, budget reserved, in queue order."""
        admitted, keep = [], []
        blocked = False
        waiting = []  # items passed over so far in this walk
        for item in self.queue:
            cost = cost_of(item)
            if not blocked and self.used + cost <= self.budget:
                self.used += cost
                admitted.append((item, cost))
                for w in waiting:
                    self._skips[id(w)] = self._skips.get(id(w), 0) + 1
            else:
                keep.append(item)
                waiting.append(item)
                if self._skips.get(id(item), 0) >= self.max_skips:
                    blocked = True  # aged: no more overtaking
        for item, _ in admitted:
            self._skips.pop(id(item), None)
        self.queue = keep
        return admitted

    def skips(self, item) -> int:
        return self._skips.get(id(item), 0)

    def release(self, cost: int) -> None:
        self.used -= cost


_MOMCOL = np.dtype([("V", "<u8"), ("ldv", "<i8"), ("nv", "<i4"), ("res", "<i4"), ("out", "<u8"),
                    ("ldm", "<i8")])


def _step_descriptors(dests, co, T):
    """The device descriptors of one step, laid out in one pinned buffer: the gather
    index of every batch column (rotation), its moments descriptor (jm_mom_col: the
    request's partner rows, the column's block of the request's moments), its problem
    index and moments address (lambda search), then the problems and their modes.
    -> (pinned uint8 tensor, {name: (byte offset, bytes)})."""
    B = int(sum(t for _, _, t in dests))
    sizes = {"col_idx": 4 * B, "cols": _MOMCOL.itemsize * B, "snp_prob": 4 * B,
             "snp_xmom": 8 * B, "probs": ctypes.sizeof(_lib.LmmProblem) * len(dests),
             "modes": 4 * len(dests)}
    views, off = {}, 0
    for k, nb in sizes.items():
        views[k] = (off, nb)
        off += (nb + 15) // 16 * 16
    stage = torch.empty((max(off, 16),), dtype=torch.uint8, pin_memory=True)
    raw = stage.numpy()

    def arr(k, dt):
        o, nb = views[k]
        return raw[o:o + nb].view(dt)

    col_idx, cols = arr("col_idx", np.int32), arr("cols", _MOMCOL)
    snp_prob, snp_xmom = arr("snp_prob", np.int32), arr("snp_xmom", np.uint64)
    o = 0
    for k, (a, _, take) in enumerate(dests):
        rows = a.prob.n_cvt + 2
        blk = rows * T * 8
        r = np.arange(take)
        col_idx[o:o + take] = a.next - take + r
        addr = np.uint64(a.xmom.data_ptr()) + (r * blk).astype(np.uint64)
        c = cols[o:o + take]
        c["V"] = a.prob.V.data_ptr()
        c["ldv"] = a.prob.V.shape[1]
        c["nv"] = a.prob.n_cvt + 1
        c["res"] = 0
        c["out"] = addr
        c["ldm"] = T
        snp_prob[o:o + take] = k
        snp_xmom[o:o + take] = addr
        o += take
    probs = (_lib.LmmProblem * len(dests))(*[a.prob.prob for a, _, _ in dests])
    po, pn = views["probs"]
    raw[po:po + pn] = np.frombuffer(bytes(probs), dtype=np.uint8)
    arr("modes", np.int32)[:] = [int(a.req.mode) for a, _, _ in dests]
    return stage, views


class Server:
    """Scheduler thread owning the device work of every resident cohort."""

    def __init__(self, hbm_budget_bytes: int | None = None, slice_snps: int = 2048,
                 max_batch_snps: int = 16384, max_inflight: int = 2, setup_workers: int = 2,
                 max_skips: int = 8):
        self.device = require_gpu()
        self.cohorts: dict[str, ResidentCohort] = {}
        self.slice_snps = slice_snps
        self.max_batch = max_batch_snps
        self.max_inflight = max_inflight
        if hbm_budget_bytes is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget_bytes = int(free * 0.5)
        self.adm = AdmissionQueue(hbm_budget_bytes, max_skips)
        self._active: dict[str, list[_Active]] = {}
        self._inflight: dict[str, list[torch.cuda.Event]] = {}
        self._setup: list = []  # (req, cost, future of AssocProblem)
        self._pool = ThreadPoolExecutor(max_workers=max(1, setup_workers),
                                        thread_name_prefix="jm-setup")
        self._cv = threading.Condition()
        self._ids = itertools.count()
        self._thread = None
        self._stop = False
        self.completed: list[Request] = []
        self.failed: list[Request] = []
        self.snps_done = 0
        self.batches = 0
        self.max_concurrent = 0

    # -- cohorts --------------------------------------------------------------
    def add_cohort(self, rc: ResidentCohort) -> None:
        self.cohorts[rc.name] = rc
        self._active[rc.name] = []
        self._inflight[rc.name] = []

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
        self._pool.shutdown(wait=True)

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()

    # -- scheduler --------------------------------------------------------------
    def _create(self, req: Request) -> AssocProblem:
        """Worker thread: the request's device tables on the cohort's setup stream."""
        rc = self.cohorts[req.cohort]
        with torch.cuda.device(rc.cohort.device), torch.cuda.stream(rc.setup_stream):
            prob = AssocProblem(rc.cohort, req.phenotype, req.covariates)
            rc.setup_stream.synchronize()  # tables complete before the steps read them
        return prob

    def _fail(self, req: Request, exc: BaseException, cost: int) -> None:
        with self._cv:
            self.adm.release(cost)
        req.t_done = time.perf_counter()
        self.failed.append(req)
        if not req.future.done():
            req.future.set_exception(exc)

    def _admit(self) -> None:
        with self._cv:
            admitted = self.adm.admit(
                lambda r: _Active.cost_bytes(self.cohorts[r.cohort], r, self.slice_snps))
        for req, cost in admitted:
            req.t_admit = time.perf_counter()
            self._setup.append((req, cost, self._pool.submit(self._create, req)))
        keep = []
        for req, cost, fut in self._setup:
            if not fut.done():
                keep.append((req, cost, fut))
                continue
            exc = fut.exception()
            if exc is not None:
                self._fail(req, exc, cost)
                continue
            rc = self.cohorts[req.cohort]
            try:
                with torch.cuda.stream(rc.stream):
                    act = _Active(req, rc, fut.result(), self.slice_snps)
            except Exception as e:  # noqa: BLE001 - reported to the request
                self._fail(req, e, cost)
                continue
            act.cost = cost
            self._active[req.cohort].append(act)
        self._setup = keep

    def _launch_step(self, name: str) -> bool:
        """Launch the next step of a cohort if it has work and room in flight."""
        infl = self._inflight[name]
        while infl and infl[0].query():
            infl.pop(0)
        if len(infl) >= self.max_inflight:
            return False
        acts = [a for a in self._active[name] if a.remaining() > 0]
        if not acts:
            return False
        rc = self.cohorts[name]
        co = rc.cohort
        T = co.n_terms
        with torch.cuda.stream(rc.stream):
            dests = []
            budget = self.max_batch
            for a in acts:
                if budget <= 0:
                    break
                take = min(self.slice_snps, a.remaining(), budget)
                budget -= take
                dests.append((a, a.next - a.req.snp_start, take))
                a.next += take
            B = int(sum(t for _, _, t in dests))
            stage, views = _step_descriptors(dests, co, T)
            dev_buf = torch.empty((stage.numel(),), dtype=torch.uint8, device=co.device)
            dev_buf.copy_(stage, non_blocking=True)  # one upload, no host wait
            dv = {k: dev_buf[o:o + nb] for k, (o, nb) in views.items()}
            Z, f32, ldz = co.rotate_genotypes(rc.g, rc.ld, rc.means, B, rc.precision,
                                              dv["col_idx"].view(torch.int32))
            nv_max = max(a.prob.n_cvt for a, _, _ in dests) + 1
            work = co.moments_work(B, nv_max)
            st = rc.stream.cuda_stream
            _lib.call("jm_spectrum_moments_multi", Z.data_ptr(), int(f32), ldz, B,
                      ptr(dv["cols"]), nv_max, ctypes_ref(co.plan_c), ptr(work), st)
            sout = torch.empty((B, _lib.JM_NOUT), dtype=torch.float64, device=co.device)
            _lib.call("jm_lmm_assoc_multi", ptr(dv["probs"]), ptr(dv["modes"]), len(dests),
                      nv_max - 1, T, int(dests[0][0].prob.prob.n_clusters),
                      ptr(dv["snp_prob"]), ptr(dv["snp_xmom"]), T, B, ptr(sout), st)
            o = 0
            for a, off, take in dests:
                a.out[off:off + take].copy_(sout[o:o + take])
                o += take
                if a.remaining() == 0:
                    a.host = torch.empty(a.out.shape, dtype=torch.float64, pin_memory=True)
                    a.host.copy_(a.out, non_blocking=True)
                    a.done_event = torch.cuda.Event()
                    a.done_event.record(rc.stream)
            ev = torch.cuda.Event()
            ev.record(rc.stream)
            infl.append(ev)
            del Z, work, dev_buf, sout  # stream-ordered frees: reused only by later work
        self.batches += 1
        self.snps_done += B
        return True

    def _fail_cohort(self, name: str, exc: BaseException) -> None:
        """A step failed: its cohort's active requests fail alone; the device work
        already queued on the cohort's stream is drained before their buffers go."""
        rc = self.cohorts[name]
        try:
            rc.stream.synchronize()
        except Exception:  # noqa: BLE001 - the stream's own error is the one reported
            pass
        for a in self._active[name]:
            self._fail(a.req, exc, a.cost)
        self._active[name] = []
        self._inflight[name] = []

    def _poll_done(self) -> int:
        n_done = 0
        for name, acts in self._active.items():
            for a in list(acts):
                if a.done_event is None or not a.done_event.query():
                    continue
                res = a.host.numpy().copy()
                a.req.t_done = time.perf_counter()
                acts.remove(a)
                with self._cv:
                    self.adm.release(a.cost)
                self.completed.append(a.req)
                a.req.future.set_result(res)
                n_done += 1
        return n_done

    def _busy(self) -> bool:
        return bool(self.adm.queue or self._setup or any(self._active.values()))

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not self._stop and not self._busy():
                    self._cv.wait()
                if self._stop and not self._busy():
                    return
            self._admit()
            self.max_concurrent = max(self.max_concurrent,
                                      sum(len(v) for v in self._active.values()))
            launched = False
            for name in list(self._active):
                try:
                    launched |= self._launch_step(name)
                except Exception as e:  # noqa: BLE001 - isolate the cohort's requests
                    self._fail_cohort(name, e)
            done = self._poll_done()
            if not launched and not done:
                # everything is queued on the device (or still being set up): wait a
                # little for completions, setups or new requests
                with self._cv:
                    self._cv.wait(timeout=0.0005)

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
                "batches": self.batches, "max_concurrent": self.max_concurrent,
                "failed": len(self.failed)}
