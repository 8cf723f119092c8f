"""Device association engine: resident cohort state + per-phenotype problem.

A ``Cohort`` keeps, in HBM, what every association request against one kinship
shares: the eigenvalues and their compressed-spectrum plan, U^T in FP64 and
(for the mixed path) its split into bf16 planes (hi + lo by default, hi + mid + lo
on request). An ``AssocProblem`` adds one
phenotype and covariate set: the rotated covariates/phenotype, their spectrum
moments, the null model (device) and the grid tables. ``associate`` then runs
one SNP tile through rotation -> spectrum moments -> lambda search + tests.

Reference flow being replaced: src/jamma/lmm/runner_streaming.py:334-566
(eigendecomposition reuse, UtW/Uty, null model, per-chunk UtG rotation and
the JAX chain of likelihood_jax.py).
"""

from __future__ import annotations

import numpy as np
import torch

from jamma_amd import _lib
from jamma_amd.device import ptr, require_gpu, round_up, stream_ptr, to_dev
from jamma_amd.lmm.spectrum import SpectrumPlan, plan_spectrum

PRECISIONS = ("f64", "mixed")
MAX_CVT = 20


def get_ab_index(a: int, b: int, n_cvt: int) -> int:
    """GEMMA packed index (src/jamma/lmm/likelihood.py:25-43), 1-based a, b."""
    cols = n_cvt + 2
    a1, b1 = (a, b) if a <= b else (b, a)
    return (2 * cols - a1 + 2) * (a1 - 1) // 2 + b1 - a1


def pair_tables(n_cvt: int) -> tuple[np.ndarray, np.ndarray]:
    """Where each packed GEMMA pair (a, b) gets its moments.

    pair_row[j] >= 0: the pair involves the genotype (column n_cvt+1); its moments
    are row pair_row[j] of the per-SNP block, whose rows are the genotype's
    partners W_1..W_c (rows 0..c-1), y (row c) and the genotype itself (row c+1).
    pair_row[j] = -1: a covariate/phenotype pair; shared_src[j] = (column, row)
    of the shared moment block computed from Z = [UtW | Uty | 1] with partners
    [UtW | Uty] (row c+1 = the column with itself).
    """
    c = n_cvt
    ni = (c + 3) * (c + 2) // 2
    pair_row = np.full(ni, -1, dtype=np.int32)
    shared_src = np.zeros((ni, 2), dtype=np.int64)
    xc, yc = c + 1, c + 2
    for a in range(1, c + 3):
        for b in range(a, c + 3):
            j = get_ab_index(a, b, c)
            if a == xc or b == xc:
                other = b if a == xc else a
                pair_row[j] = (c + 1) if other == xc else (c if other == yc else other - 1)
            else:
                col = a - 1 if a <= c else c
                shared_src[j] = (col, c + 1) if a == b else (col, (b - 1) if b <= c else c)
    return pair_row, shared_src


TWO_PLANE_MAX_N = 16384  # largest cohort rotated with the two-plane split by default

# Significance thresholds whose agreement the mixed path guarantees, and the band (in
# log10 p, twice the 1e-4 accuracy bar of the mixed path) around each of them inside
# which a mixed-precision p-value is recomputed on the FP64 path: a p-value can only
# land on the other side of a threshold than its FP64 value if it lies within its
# error of the threshold, so after the recomputation every reported p-value agrees
# with the FP64 path on every threshold.
GUARD_THRESHOLDS = (0.05, 0.01, 1e-3, 5e-8)
GUARD_BAND = 2e-4
_P_COLS = {1: (2,), 2: (7,), 3: (8,), 4: (2, 7, 8)}


def guard_rows(out, mode: int, band: float | None = None):
    """Rows of a result block (device tensor or host array, n x JM_NOUT) with a p-value
    of the mode within `band` (log10 units; default GUARD_BAND) of a guarded threshold."""
    band = GUARD_BAND if band is None else band
    cols = list(_P_COLS[mode])
    logt = np.log10(np.array(GUARD_THRESHOLDS))
    if isinstance(out, torch.Tensor):
        lp = torch.log10(out[:, cols])
        t = torch.as_tensor(logt, dtype=lp.dtype, device=lp.device)
        near = (lp[:, :, None] - t).abs() <= band
        return torch.nonzero(near.any(dim=2).any(dim=1)).flatten()
    with np.errstate(divide="ignore", invalid="ignore"):
        lp = np.log10(np.asarray(out)[:, cols])
    near = np.abs(lp[:, :, None] - logt) <= band
    return np.nonzero(near.any(axis=2).any(axis=1))[0]


class Cohort:
    """Eigendecomposition of one kinship, resident on the device."""

    def __init__(self, eigenvalues, eigenvectors, precision: str = "mixed", device=None,
                 bf16_planes: int | None = None):
        """bf16_planes: 2 = U^T as hi + lo (16 significant bits; the rotation issues 3
        bf16 products per term: hi X1 + lo X1 + hi X2), 3 = hi + mid + lo (24 bits,
        4 products). None picks 2 up to TWO_PLANE_MAX_N samples and 3 above: the
        relative error of a rotated value is about 2^-17 with two planes, and
        |delta log10 p| grows with -log10 p, which reaches ~60 in cohorts of 50,000
        (there two planes measured 1.5e-4)."""
        if precision not in PRECISIONS + ("both",):
            raise ValueError(f"precision must be one of {PRECISIONS + ('both',)}")
        d = eigenvalues.detach().cpu().numpy() if isinstance(eigenvalues, torch.Tensor) \
            else np.asarray(eigenvalues, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        self.n = n = d.shape[0]
        if bf16_planes is None:
            bf16_planes = 2 if n <= TWO_PLANE_MAX_N else 3
        if bf16_planes not in (2, 3):
            raise ValueError("bf16_planes must be 2 or 3")
        self.bf16_planes = bf16_planes
        self.device = device or require_gpu()
        self.d = d
        self.plan: SpectrumPlan = plan_spectrum(d)
        U = to_dev(eigenvectors, torch.float64, self.device)
        if tuple(U.shape) != (n, n):
            raise ValueError(f"eigenvectors must be ({n}, {n}), got {tuple(U.shape)}")
        self.UT = U.t().contiguous()  # UT[i][k] = U[k][i]
        del U
        self.ld_u = n
        self.precision = precision
        self.split = None
        if precision in ("mixed", "both"):
            nbytes = int(_lib.lib.jm_packed_ut_bytes(n, bf16_planes))
            planes = torch.empty((nbytes // 2,), dtype=torch.int16, device=self.device)
            _lib.call("jm_pack_ut_bf16x3", ptr(self.UT), self.ld_u, n, bf16_planes, ptr(planes),
                      stream_ptr())
            self.split = planes  # packed bf16 planes of U^T
        p = self.plan
        dev = self.device
        self.cl_start = to_dev(p.cl_start, torch.int32, dev)
        self.cl_size = to_dev(p.cl_size, torch.int32, dev)
        self.cl_terms = to_dev(p.cl_terms, torch.int32, dev)
        self.cl_toff = to_dev(p.cl_toff, torch.int32, dev)
        self.cl_center = to_dev(p.cl_center, torch.float64, dev)
        self.cl_weight = to_dev(p.cl_size.astype(np.float64), torch.float64, dev)
        self.delta = to_dev(p.idx_delta, torch.float64, dev)
        self.term_cluster = to_dev(p.term_cluster, torch.int32, dev)
        self.term_m = to_dev(p.term_m, torch.int32, dev)
        # device plan for the moments kernels (cluster pieces + partial slots)
        self._plan_dev = [to_dev(a, torch.int32, dev) for a in
                          (p.cl_seg0, p.cl_nseg, p.seg_start, p.seg_len, p.seg_cluster, p.seg_slot)]
        pc = _lib.SpectrumPlanC()
        pc.n, pc.n_terms, pc.n_clusters = n, p.n_terms, len(p.cl_size)
        pc.n_seg, pc.n_slots = len(p.seg_start), p.n_slots
        pc.cl_terms, pc.cl_toff = ptr(self.cl_terms), ptr(self.cl_toff)
        (pc.cl_seg0, pc.cl_nseg, pc.seg_start, pc.seg_len, pc.seg_cluster,
         pc.seg_slot) = [ptr(a) for a in self._plan_dev]
        pc.delta = ptr(self.delta)
        pc.term_cluster, pc.term_m = ptr(self.term_cluster), ptr(self.term_m)
        self.plan_c = pc

    @property
    def n_terms(self) -> int:
        return self.plan.n_terms

    def hbm_bytes(self) -> int:
        b = self.UT.numel() * 8
        if self.split is not None:
            b += self.split.numel() * 2
        return b

    # -- rotation -----------------------------------------------------------
    def rotate_dense(self, rows: torch.Tensor) -> torch.Tensor:
        """UT @ rows^T for rows (m x n) float64 -> (n x m) float64."""
        rows = rows.contiguous()
        m = rows.shape[0]
        out = torch.empty((self.n, m), dtype=torch.float64, device=self.device)
        _lib.call("jm_gemm_nt_f64", ptr(self.UT), self.ld_u, ptr(rows), rows.shape[1], self.n, m,
                  self.n, ptr(out), m, stream_ptr())
        return out

    def rotate_genotypes(self, g: torch.Tensor, ld_g: int, means: torch.Tensor, n_snps: int,
                         precision: str, col_idx: torch.Tensor | None = None, timer=None
                         ) -> tuple[torch.Tensor, bool, int]:
        """Z = U^T X for a genotype tile (SNP-major int8 rows, -1 missing -> means).
        With col_idx, column s uses genotype row / mean col_idx[s]. timer: optional pair
        of torch.cuda.Event recorded around the rotation kernel alone."""
        ldz = max(1, n_snps)
        if precision == "f64":
            Z = torch.empty((self.n, ldz), dtype=torch.float64, device=self.device)
            if timer:
                timer[0].record()
            _lib.call("jm_rotate_f64", ptr(self.UT), self.ld_u, ptr(g), ld_g, ptr(means),
                      ptr(col_idx), self.n, n_snps, ptr(Z), ldz, stream_ptr())
            if timer:
                timer[1].record()
            return Z, False, ldz
        if precision == "mixed":
            if self.split is None:
                raise ValueError("cohort was built without the bf16 split (precision='f64')")
            Z = torch.empty((self.n, ldz), dtype=torch.float32, device=self.device)
            if n_snps:
                gbytes = int(_lib.lib.jm_packed_geno_bytes(self.n, n_snps))
                gp = torch.empty((gbytes,), dtype=torch.uint8, device=self.device)
                mm = torch.empty((n_snps,), dtype=torch.int32, device=self.device)
                st = stream_ptr()
                _lib.call("jm_pack_genotypes", ptr(g), ld_g, ptr(col_idx), ptr(means), self.n,
                          n_snps, ptr(gp), ptr(mm), st)
                if timer:
                    timer[0].record()
                _lib.call("jm_rotate_bf16x3", ptr(self.split), self.bf16_planes, ptr(gp), ptr(mm),
                          self.n, n_snps, ptr(Z), ldz, st)
                if timer:
                    timer[1].record()
            return Z, True, ldz
        raise ValueError(f"unknown precision {precision!r}")

    def moments(self, Z: torch.Tensor, z_is_f32: bool, ldz: int, n_cols: int, V: torch.Tensor,
                nv: int, out: torch.Tensor | None = None) -> torch.Tensor:
        """Spectrum moments (n_cols x (nv+1) x T) of the columns of Z."""
        T = self.n_terms
        if out is None:
            out = torch.empty((n_cols, nv + 1, T), dtype=torch.float64, device=self.device)
        if n_cols:
            work = self.moments_work(n_cols, nv)
            _lib.call("jm_spectrum_moments", ptr(Z), int(z_is_f32), ldz, n_cols, ptr(V),
                      V.shape[1], nv, ctypes_ref(self.plan_c), ptr(work), ptr(out), T,
                      stream_ptr())
        return out

    def moments_work(self, n_cols: int, nv: int) -> torch.Tensor:
        """Workspace of the moments kernels for n_cols columns and nv partner rows."""
        nbytes = int(_lib.lib.jm_moments_work_bytes(ctypes_ref(self.plan_c), n_cols, nv))
        return torch.empty((max(nbytes, 8) // 8,), dtype=torch.float64, device=self.device)


class AssocProblem:
    """One phenotype (+ covariates) against a resident cohort."""

    def __init__(self, cohort: Cohort, phenotype, covariates=None, l_min: float = 1e-5,
                 l_max: float = 1e5, n_grid: int = 50, n_refine: int = 10):
        self.cohort = cohort
        dev = cohort.device
        n = cohort.n
        y = np.asarray(phenotype.detach().cpu() if isinstance(phenotype, torch.Tensor) else phenotype,
                       dtype=np.float64)
        if y.shape != (n,):
            raise ValueError(f"phenotype must have {n} entries")
        W = np.ones((n, 1)) if covariates is None else np.asarray(covariates, dtype=np.float64)
        if W.ndim != 2 or W.shape[0] != n:
            raise ValueError("covariates must be (n_samples, n_cvt)")
        c = W.shape[1]
        if n <= c + 1:
            raise ValueError(f"Over-parameterized model: {n} samples with {c} covariates leaves "
                             f"df={n - c - 1} (need at least {c + 2} samples)")
        if c > MAX_CVT:
            raise ValueError(f"at most {MAX_CVT} covariate columns are supported, got {c}")
        self.n_cvt = c
        self.n = n
        T = cohort.n_terms
        ni = (c + 3) * (c + 2) // 2
        self.n_index = ni
        # rotated covariates + phenotype: R = U^T [W | y]   (n x (c+1))
        rows = to_dev(np.vstack([W.T, y[None, :]]), torch.float64, dev)
        R = cohort.rotate_dense(rows)
        self.V = R.t().contiguous()  # partner rows: UtW_1..UtW_c, Uty   ((c+1) x n)
        Zs = torch.cat([R, torch.ones((n, 1), dtype=torch.float64, device=dev)], dim=1).contiguous()
        smom = cohort.moments(Zs, False, c + 2, c + 2, self.V, c + 1)  # ((c+2), (c+2), T)
        pair_row, shared_src = pair_tables(c)
        idx_col = torch.as_tensor(shared_src[:, 0], device=dev)
        idx_row = torch.as_tensor(shared_src[:, 1], device=dev)
        shared = smom[idx_col, idx_row].contiguous()
        shared[torch.as_tensor(pair_row >= 0, device=dev)] = 0.0
        self.shared_mom = shared                       # (ni, T)
        self.logdet_mom = smom[c + 1, c + 1].contiguous()  # (T,)
        self.pair_row = to_dev(pair_row, torch.int32, dev)
        loglam = np.linspace(np.log(l_min), np.log(l_max), n_grid)
        self.grid_loglam = to_dev(loglam, torch.float64, dev)
        self.grid_kappa = torch.empty((T, n_grid), dtype=torch.float64, device=dev)
        self.grid_shared = torch.empty((n_grid, ni), dtype=torch.float64, device=dev)
        self.grid_logdet = torch.empty((n_grid,), dtype=torch.float64, device=dev)
        self.null_kappa = torch.empty((T,), dtype=torch.float64, device=dev)
        self.null_shared = torch.empty((ni,), dtype=torch.float64, device=dev)
        self.iab_shared = torch.empty((ni,), dtype=torch.float64, device=dev)
        pr = _lib.LmmProblem()
        pr.n, pr.n_cvt, pr.n_terms = n, c, T
        pr.n_clusters = len(cohort.plan.cl_size)
        pr.n_grid, pr.n_iter = n_grid, max(n_refine, 20)
        pr.l_min, pr.l_max = l_min, l_max
        pr.term_cluster, pr.term_m = ptr(cohort.term_cluster), ptr(cohort.term_m)
        pr.cl_center, pr.cl_weight = ptr(cohort.cl_center), ptr(cohort.cl_weight)
        pr.logdet_mom, pr.shared_mom = ptr(self.logdet_mom), ptr(self.shared_mom)
        pr.pair_row, pr.grid_loglam = ptr(self.pair_row), ptr(self.grid_loglam)
        pr.grid_kappa, pr.grid_shared = ptr(self.grid_kappa), ptr(self.grid_shared)
        pr.grid_logdet = ptr(self.grid_logdet)
        pr.null_kappa, pr.null_shared = ptr(self.null_kappa), ptr(self.null_shared)
        pr.iab_shared = ptr(self.iab_shared)
        self.prob = pr
        # null model MLE on the device (likelihood.py:571-607)
        nullout = torch.zeros(4, dtype=torch.float64, device=dev)
        _lib.call("jm_lmm_null", ctypes_ref(pr), ptr(nullout), stream_ptr())
        lam0, logl0 = nullout[:2].cpu().tolist()
        pr.lambda_null, pr.logl_null = lam0, logl0
        self.lambda_null, self.logl_H0 = lam0, logl0
        _lib.call("jm_lmm_prepare", ctypes_ref(pr), stream_ptr())

    def assoc_moments(self, xmom: torch.Tensor, n_snps: int, mode: int,
                      out: torch.Tensor | None = None) -> torch.Tensor:
        """lambda search + tests from per-SNP moments (n_snps x (c+2) x T)."""
        if out is None:
            out = torch.empty((max(n_snps, 1), _lib.JM_NOUT), dtype=torch.float64,
                              device=self.cohort.device)
        if n_snps:
            _lib.call("jm_lmm_assoc", ctypes_ref(self.prob), int(mode), ptr(xmom),
                      self.cohort.n_terms, n_snps, ptr(out), stream_ptr())
        return out

    def retest_f64(self, g: torch.Tensor, ld_g: int, means: torch.Tensor, rows, mode: int,
                   out: torch.Tensor) -> int:
        """Recompute SNP rows `rows` of g (with their means, indexed alike) on the FP64
        path and write them into the same rows of `out`. Returns the count."""
        if len(rows) == 0:
            return 0
        idx = torch.as_tensor(rows, device=self.cohort.device)
        col_idx = idx.to(torch.int32).contiguous()
        co = self.cohort
        k = int(idx.numel())
        Z, f32, ldz = co.rotate_genotypes(g, ld_g, means, k, "f64", col_idx)
        xm = co.moments(Z, f32, ldz, k, self.V, self.n_cvt + 1)
        o = self.assoc_moments(xm, k, mode)
        out[idx.to(torch.int64)] = o[:k]
        return k

    def associate_dense(self, X: torch.Tensor, means: torch.Tensor, n_snps: int,
                        mode: int) -> torch.Tensor:
        """One tile of dosage genotypes: X SNP-major float64 rows (n_snps x n, NaN =
        missing, imputed in place with `means`), FP64 rotation (jm_gemm_nt_f64), then
        moments and the lambda search + tests as for hard calls."""
        co = self.cohort
        n = co.n
        if tuple(X.shape) != (n_snps, n) or X.dtype != torch.float64 or not X.is_contiguous():
            raise ValueError("X must be contiguous float64 (n_snps x n)")
        st = stream_ptr()
        _lib.call("jm_impute_rows_f64", ptr(X), n_snps, n, n, ptr(means), st)
        ldz = max(1, n_snps)
        Z = torch.empty((n, ldz), dtype=torch.float64, device=co.device)
        _lib.call("jm_gemm_nt_f64", ptr(co.UT), co.ld_u, ptr(X), n, n, n_snps, n, ptr(Z), ldz, st)
        xmom = co.moments(Z, False, ldz, n_snps, self.V, self.n_cvt + 1)
        del Z
        return self.assoc_moments(xmom, n_snps, mode)

    def associate(self, g: torch.Tensor, ld_g: int, means: torch.Tensor, n_snps: int, mode: int,
                  precision: str = "mixed") -> torch.Tensor:
        """One SNP tile: rotation -> moments -> lambda search + tests (device out)."""
        co = self.cohort
        Z, f32, ldz = co.rotate_genotypes(g, ld_g, means, n_snps, precision)
        xmom = co.moments(Z, f32, ldz, n_snps, self.V, self.n_cvt + 1)
        del Z
        return self.assoc_moments(xmom, n_snps, mode)


def assoc_multi(problems, modes, snp_prob, snp_xmom, out: torch.Tensor, stream=None,
                width_cvt: int = 0) -> None:
    """jm_lmm_assoc_multi: SNP s is tested under problems[snp_prob[s]] (AssocProblems of
    one cohort) in mode modes[snp_prob[s]], from the moments block at device address
    snp_xmom[s] ((n_cvt+2) x T float64 rows); its results go to row s of out.
    width_cvt (0: the widest problem's n_cvt) picks the kernel variant: launches with
    one width_cvt give a SNP the same bits whatever else is in the batch."""
    if not problems:
        raise ValueError("no problems")
    co = problems[0].cohort
    if any(p.cohort is not co for p in problems):
        raise ValueError("all problems of one launch must share the cohort")
    snp_prob = np.ascontiguousarray(snp_prob, dtype=np.int32)
    snp_xmom = np.ascontiguousarray(snp_xmom, dtype=np.uint64)
    B = snp_prob.shape[0]
    if snp_xmom.shape != (B,) or out.shape[0] < B or out.shape[1] != _lib.JM_NOUT:
        raise ValueError("snp_prob / snp_xmom / out shapes disagree")
    if B and (snp_prob.min() < 0 or snp_prob.max() >= len(problems)):
        raise ValueError("snp_prob out of range")
    if B == 0:
        return
    T = co.n_terms
    probs = (_lib.LmmProblem * len(problems))(*[p.prob for p in problems])
    dev = co.device
    probs_dev = to_dev(np.frombuffer(probs, dtype=np.uint8), torch.uint8, dev)
    modes_dev = to_dev(np.asarray(modes, dtype=np.int32), torch.int32, dev)
    sp = to_dev(snp_prob, torch.int32, dev)
    sx = to_dev(snp_xmom.view(np.int64), torch.int64, dev)
    max_cvt = max(p.n_cvt for p in problems)
    _lib.call("jm_lmm_assoc_multi", ptr(probs_dev), ptr(modes_dev), len(problems), max_cvt,
              int(width_cvt), T, int(problems[0].prob.n_clusters), ptr(sp), ptr(sx), T, B,
              ptr(out), stream_ptr(stream))
    if stream is not None:
        # the launch reads these device copies on `stream`: the caching allocator must not
        # hand their memory out again before that stream has passed the launch
        for t in (probs_dev, modes_dev, sp, sx):
            t.record_stream(stream)


def ctypes_ref(pr):
    import ctypes
    return ctypes.byref(pr)
