#!/usr/bin/env python
"""Benchmark of the LMM association hot path on one MI355X.

Workload ("precision form"): n = 10,000 samples x 100,000 SNPs,
3 Balding-Nichols subpopulations (Fst = 0.05), 1% missing calls, h2 = 0.5 with
200 causal SNPs, 3 Gaussian covariates + intercept, all tests (-lmm 4).
Synthetic data is generated on the device. Untimed setup: genotype generation,
integer kinship GEMM (timed separately), eigendecomposition, null model.

A step = association of every filtered SNP with the inputs resident in HBM
(int8 genotypes, means, U^T and its bf16 split): rotation -> spectrum moments
-> lambda search + Wald/LRT/Score, one tile of TILE SNPs at a time.
`value` is the mixed-precision path's SNPs/s; the FP64 path is timed the same
way and reported beside it with the parity of the two.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SNPs/s at n=10k, mixed vs FP64 GPU path at |Δ-log10p|≤1e-4; req p50/p99"
PEAK = {"bf16": 2.5e15, "int8": 5.0e15, "fp4": 10.1e15, "fp64": 78.6e12, "hbm": 8.0e12}


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def gen_cohort(n: int, p: int, n_pop: int, fst: float, miss: float, seed: int, ld: int):
    """Balding-Nichols genotypes on the device: SNP-major int8 (p x ld), -1 = missing."""
    import torch

    dev = torch.device("cuda")
    torch.manual_seed(seed)
    anc = torch.rand(p, device=dev, dtype=torch.float64) * 0.45 + 0.05
    pops = torch.randint(0, n_pop, (n,), device=dev)
    a = anc * (1 - fst) / fst
    b = (1 - anc) * (1 - fst) / fst
    pk = torch.distributions.Beta(a, b).sample((n_pop,)).to(torch.float32)  # (n_pop, p)
    g = torch.zeros((p, ld), dtype=torch.int8, device=dev)
    blk = 8192
    for s in range(0, p, blk):
        e = min(s + blk, p)
        pr = pk[:, s:e].t()[:, pops]  # (blk, n) allele freq per SNP/sample
        x = (torch.rand(pr.shape, device=dev) < pr).to(torch.int8) + \
            (torch.rand(pr.shape, device=dev) < pr).to(torch.int8)
        x[torch.rand(pr.shape, device=dev) < miss] = -1
        g[s:e, :n] = x
    return g


def cpu_model() -> str:
    """Host CPU model name (first 'model name' line of /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def counts_of(g, n):
    import torch

    gg = g[:, :n]
    c = torch.stack([(gg == 0).sum(1), (gg == 1).sum(1), (gg == 2).sum(1), (gg < 0).sum(1)], 1)
    return c.cpu().numpy().astype("int64")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=10_000)
    ap.add_argument("--p", type=int, default=100_000)
    ap.add_argument("--ncov", type=int, default=3)
    ap.add_argument("--mode", type=int, default=4)
    ap.add_argument("--tile", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=2048,
                    help="SNPs of the CPU baseline sample (about 15 s of host work at n=10k)")
    ap.add_argument("--planes", type=int, default=None, choices=[2, 3],
                    help="bf16 planes of U^T on the mixed path (2: 3 MFMA products, 3: 4; "
                         "default: the cohort's choice by sample size)")
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 path (profiling runs)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-guard", action="store_true",
                    help="skip the FP64 recomputation of mixed p-values near the thresholds")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving session that adds request p50/p99 to the line")
    ap.add_argument("--workload", default="precision",
                    choices=["precision", "small", "scale", "loco", "pipeline", "serving", "copy"],
                    help="precision: the headline config; scale: n=50,000 x 200,000 SNPs, 5 "
                         "subpopulations (Fst 0.02), 0.5%% missing, Wald test in 10,000-SNP "
                         "tiles; loco: n=20,000 x 300,000 SNPs on 19 chromosomes, 8 "
                         "correlated phenotypes, leave-one-chromosome-out kinships, -lmm 4; "
                         "pipeline: the precision form written as a PLINK fileset and run "
                         "through gwas() (.bed on disk -> .assoc.txt); "
                         "serving: resident cohorts + Poisson request stream (p50/p99); "
                         "small: the plain FP64 method at n=1,940 x 12,226 SNPs (-lmm 1) with "
                         "the oracle timed on the whole job")
    ap.add_argument("--pops", type=int, default=3, help="Balding-Nichols subpopulations")
    ap.add_argument("--fst", type=float, default=0.05)
    ap.add_argument("--miss", type=float, default=0.01, help="missing-call rate")
    ap.add_argument("--kin-tile", type=int, default=16384,
                    help="SNPs per genotype tile streamed into the kinship accumulator")
    ap.add_argument("--cohorts", default="2000,10000,30000", help="serving cohort sizes")
    ap.add_argument("--requests", type=int, default=120)
    ap.add_argument("--rate", type=float, default=8.0, help="serving arrivals per second")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (collectives and "
              "sharding are design-only); run with --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "serving":
        return serving(args)
    if args.workload == "copy":
        return copy_calibration()
    if args.workload == "loco":
        return loco(args)
    if args.workload == "pipeline":
        return pipeline(args)
    if args.workload == "small":
        return small(args)
    if args.workload == "scale":
        args.n, args.p, args.pops, args.fst, args.miss = 50_000, 200_000, 5, 0.02, 0.005
        args.tile, args.kin_tile, args.mode = 10_000, 10_000, 1
        args.no_serving, args.no_cpu, args.steps, args.warmup = True, True, 1, 1

    import numpy as np
    import torch

    from jamma_amd.device import require_gpu, round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.kinship.compute import KinshipAccumulator
    from jamma_amd.lmm.eigen import eigh_device
    from jamma_amd.lmm.engine import GUARD_BAND, AssocProblem, Cohort, guard_rows

    dev = require_gpu()
    n, p = args.n, args.p
    ld = round_up(n, 64)
    t0 = time.perf_counter()
    g = gen_cohort(n, p, args.pops, args.fst, args.miss, args.seed, ld)
    torch.cuda.synchronize()
    log(f"generated {n} x {p} genotypes in {time.perf_counter() - t0:.1f}s")

    counts = counts_of(g, n)
    mean32, miss, poly = stats_from_counts(counts, np.float32)
    af = mean32 / 2
    maf = np.minimum(af, 1 - af)
    keep = (maf >= 0.01) & (miss / n <= 0.05) & poly
    idx = np.where(keep)[0]
    gf = g[torch.as_tensor(idx, device=dev)].contiguous() if len(idx) < p else g
    del g
    pf = gf.shape[0]
    means = to_dev(mean32[idx], torch.float64)
    mu64, _, _ = stats_from_counts(counts[idx], np.float64)
    miss_f = counts[idx, 3]
    log(f"{pf} SNPs pass the filters")

    # ---- kinship: exact integer GEMM on the FP4 MFMA (kinship-filter = var > 0 over all
    # samples); tiles of 16384 SNPs are staged 2-bit packed by the accumulator, the GEMM runs
    # as one launch over all staged SNPs, timed with HIP events on its stream
    acc = KinshipAccumulator(n)
    acc.syrk_timer = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ktile = args.kin_tile
    for s in range(0, pf, ktile):
        e = min(s + ktile, pf)
        acc.add_tile(gf[s:e], ld, e - s, mu64[s:e], True, int(miss_f[s:e].sum()))
    acc.flush()
    torch.cuda.synchronize()
    syrk_ms = sum(a.elapsed_time(b) for a, b in acc.syrk_timer)
    syrk_launches = len(acc.syrk_timer)
    t1 = time.perf_counter()
    i8mfma = syrk_int8_mfma(acc) if syrk_launches == 1 else None
    t_check = time.perf_counter() - t1  # not part of the kinship stage
    K = acc.finalize(pf)
    torch.cuda.synchronize()
    kin_s = time.perf_counter() - t0 - t_check
    del acc
    log(f"kinship {kin_s:.2f}s (GEMM {syrk_ms:.2f} ms in {syrk_launches} launch(es) over "
        f"{pf} SNPs, input tiles of {ktile})")

    t0 = time.perf_counter()
    d, U = eigh_device(K, progress=lambda step: log(
        f"eigendecomposition: {step} done at {time.perf_counter() - t0:.1f}s"))
    torch.cuda.synchronize()
    eig_s = time.perf_counter() - t0
    del K
    log(f"eigendecomposition {eig_s:.1f}s")

    # ---- phenotype: 200 causal SNPs, h2 = 0.5, 3 covariates + intercept ----
    rng = np.random.default_rng(args.seed)
    causal = np.sort(rng.choice(pf, 200, replace=False))
    gc = gf[torch.as_tensor(causal, device=dev)][:, :n].to(torch.float64)
    gc = torch.where(gc < 0, means[torch.as_tensor(causal, device=dev)][:, None], gc)
    gv = (torch.as_tensor(rng.normal(size=200), device=dev)[None, :] @ gc).squeeze(0)
    gv = ((gv - gv.mean()) / gv.std()).cpu().numpy()
    W = np.column_stack([np.ones(n), rng.normal(size=(n, args.ncov))])
    y = np.sqrt(0.5) * gv + np.sqrt(0.5) * rng.normal(size=n) + W[:, 1:] @ rng.normal(0, 0.3, args.ncov)

    t0 = time.perf_counter()
    cohort = Cohort(d, U, precision="both", bf16_planes=args.planes)
    args.planes = cohort.bf16_planes
    prob = AssocProblem(cohort, y, W)
    torch.cuda.synchronize()
    log(f"cohort + problem setup {time.perf_counter() - t0:.2f}s: T={cohort.n_terms} terms "
        f"({len(cohort.plan.cl_size)} clusters, r={cohort.plan.radius}), "
        f"lambda0={prob.lambda_null:.4g}")

    tiles = [(s, min(s + args.tile, pf)) for s in range(0, pf, args.tile)]
    out_all = {}
    retested = []

    def run_step(precision, timers=None):
        outs = []
        for s, e in tiles:
            evs = None
            if timers is not None:
                evs = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
            co = cohort
            if evs:
                evs[0].record()
            Z, f32, ldz = co.rotate_genotypes(gf[s:e], ld, means[s:e], e - s, precision,
                                              timer=(evs[4], evs[5]) if evs else None)
            if evs:
                evs[1].record()
            xm = co.moments(Z, f32, ldz, e - s, prob.V, prob.n_cvt + 1)
            if evs:
                evs[2].record()
            o = prob.assoc_moments(xm, e - s, args.mode)
            if evs:
                evs[3].record()
                timers.append(evs)
            outs.append(o[: e - s])
        if precision == "mixed" and not args.no_guard:
            # p-values within the guard band of a significance threshold: FP64 path
            full = torch.cat(outs)
            rows = guard_rows(full, args.mode)
            retested.append(prob.retest_f64(gf, ld, means, rows, args.mode, full))
            outs = [full]
        return outs

    results = {}
    for precision in (["mixed"] if args.no_f64 else ["mixed", "f64"]):
        # at the scale workload one FP64 pass takes ~30 s: it runs once, without warmup
        for _ in range(0 if (precision == "f64" and args.workload == "scale") else args.warmup):
            run_step(precision)
        torch.cuda.synchronize()
        timers = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            outs = run_step(precision, timers if k == 0 else None)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        kern = np.array([[e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]), e[2].elapsed_time(e[3]),
                          e[4].elapsed_time(e[5])] for e in timers])
        out_all[precision] = torch.cat(outs).cpu().numpy()
        results[precision] = {"s_per_step": dt, "snps_per_s": pf / dt,
                              "kernel_ms_per_tile": kern.mean(axis=0).tolist(),
                              "kernel_ms_total": kern.sum(axis=0).tolist()}
        log(f"{precision}: {dt * 1e3:.1f} ms/step = {pf / dt:,.0f} SNPs/s; per-tile ms "
            f"rotate(+pack)/moments/assoc/rotation kernel = {np.round(kern.mean(axis=0), 3).tolist()}")

    # ---- parity of the two device paths ----
    parity = {}
    if "f64" in out_all:
        a, b = out_all["mixed"], out_all["f64"]
        tests = {1: (("p_wald", 2),), 2: (("p_lrt", 7),), 3: (("p_score", 8),)}.get(
            args.mode, (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)))
        for name, col in tests:
            with np.errstate(divide="ignore", invalid="ignore"):  # p = 0 (underflow) in both
                dl = np.abs(np.log10(a[:, col]) - np.log10(b[:, col]))
            flips = sum(int(np.sum((a[:, col] < t) != (b[:, col] < t)))
                        for t in (0.05, 0.01, 1e-3, 5e-8))
            parity[name] = {"max_abs_dlog10p": float(np.nanmax(dl)),
                            "significance_agreement": flips == 0, "significance_flips": flips,
                            "p_zero_in_both": int(np.sum((a[:, col] == 0) & (b[:, col] == 0)))}
        parity["beta_sign_agreement"] = bool(np.array_equal(np.sign(a[:, 0]), np.sign(b[:, 0])))
        parity["n_sig_5e-8_wald"] = int(np.sum(b[:, 2] < 5e-8))
        parity["guard_band_log10p"] = None if args.no_guard else GUARD_BAND
        parity["retested_f64_per_step"] = retested[-1] if retested else 0

    # ---- CPU baseline: the oracle on a bounded sample of the same workload ----
    cpu = None
    if not args.no_cpu and args.cpu_sample > 0:
        from oracle import lmm as ol

        sample = np.sort(rng.choice(pf, args.cpu_sample, replace=False))
        Uh = U.cpu().numpy()
        UT = np.ascontiguousarray(Uh.T)
        Xs = gf[torch.as_tensor(sample, device=dev)][:, :n].cpu().numpy().astype(np.float64).T
        ms = mean32[idx][sample]
        Xs = np.where(Xs < 0, ms[None, :], Xs)
        UtW, Uty = UT @ W, UT @ y
        lam0, ll0 = ol.null_model_mle(d, UtW, Uty, W.shape[1])
        Hi0 = 1.0 / (lam0 * d + 1.0)
        t0 = time.perf_counter()
        parts = []
        for c0 in range(0, len(sample), 256):  # bounded memory per oracle call
            Utx = UT @ Xs[:, c0:c0 + 256]
            parts.append(ol.assoc_chunk(4, W.shape[1], d, UtW, Uty, Utx, Hi0, ll0))
        ref = {k: np.concatenate([r[k] for r in parts]) for k in parts[0]}
        cpu_s = time.perf_counter() - t0
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": args.cpu_sample / cpu_s, "unit": "SNPs/s", "cores": cores,
               "cpu_model": cpu_model(), "kind": "port",
               "sample": f"{args.cpu_sample} SNPs of the same workload (rotation + REML/MLE grid+"
                         f"golden + Wald/LRT/Score, numpy FP64, {cores} host threads), "
                         f"{cpu_s:.1f}s"}
        for prec in out_all:
            o = out_all[prec][sample]
            dl = max(float(np.nanmax(np.abs(np.log10(o[:, c]) - np.log10(ref[k]))))
                     for k, c in (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)))
            parity[f"{prec}_vs_oracle_max_abs_dlog10p"] = dl
        log(f"cpu baseline {cpu['value']:.1f} SNPs/s on {cores} threads")

    mixed = results["mixed"]
    rot_ms = mixed["kernel_ms_per_tile"][3]  # the rotation kernel alone
    tile_snps = pf / len(tiles)
    achieved = 2.0 * n * n * tile_snps / (rot_ms * 1e-3)
    # the rotation kernel jm_rotate_bf16x3 launches for this plane count (JM_ROTATE_UNPIPED=1
    # selects the two-stage form for both)
    unpiped = os.environ.get("JM_ROTATE_UNPIPED", "0") == "1"
    rot_kernel = ("k_rotate_bf16x3_packed" if unpiped else
                  "k_rotate_bf16x3_pipe4" if args.planes == 2 else "k_rotate_bf16x3_pipe3")
    traffic, traffic_status = hbm_traffic(rot_kernel, n, args.tile)
    vendor = vendor_bf16_gemm_tflops(n, int(tile_snps), n)
    issue_tflops = (args.planes + 1) * achieved / 1e12
    log(f"rotation: {achieved / 1e12:.0f} TFLOP/s algorithmic, {issue_tflops:.0f} issued; "
        f"hipBLASLt bf16 GEMM at ({n} x {int(tile_snps)} x {n}): {vendor:.0f} TFLOP/s")
    line = {
        "metric": METRIC,
        "value": mixed["snps_per_s"],
        "unit": "SNPs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["s_per_step"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 rotation (fp32 acc) + f64 lambda/tests",
        "data": "synthetic (Balding-Nichols genotypes generated on device)",
        "config": {"workload": f"{'scale' if args.workload == 'scale' else 'precision form'}: "
                               f"n={n} x {p} SNPs ({pf} pass filters), {args.pops} BN "
                               f"subpops Fst={args.fst:g}, {args.miss * 100:g}% missing, "
                               f"h2=0.5 / 200 causal, "
                               f"{args.ncov} covariates + intercept, -lmm {args.mode}",
                   "n": n, "p": p, "n_cvt": args.ncov + 1, "tile_snps": args.tile},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["bf16"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["bf16"], "traffic": traffic,
                     "kernel": rot_kernel,
                     "algorithmic": f"2 n^2 FLOP per SNP (U^T x); the split with {args.planes} "
                                    f"planes of U^T issues {args.planes + 1} bf16 MFMA products "
                                    "per algorithmic FLOP",
                     "mfma_issue_frac": (args.planes + 1) * achieved / PEAK["bf16"],
                     "traffic_source": traffic_status,
                     "vendor_bf16_gemm_tflops": vendor,
                     "issued_over_vendor_gemm": issue_tflops / vendor,
                     "launch_ms": rot_ms, "snps_per_launch": tile_snps},
        "cpu_baseline": cpu,
        "fp64_path": results.get("f64"),
        "mixed_path": mixed,
        "mixed_over_fp64": (mixed["snps_per_s"] / results["f64"]["snps_per_s"]) if "f64" in results
        else None,
        "parity": parity,
        "kinship": kinship_record(n, pf, kin_s, syrk_ms, syrk_launches, i8mfma),
        "eigendecomposition_s": eig_s,
        "spectrum_terms": cohort.n_terms,
    }
    if not args.no_serving:
        # the metric's "req p50/p99": the serving configuration, after this workload's buffers go
        del cohort, prob, gf, means, U, outs, out_all
        torch.cuda.empty_cache()
        stats, wall, setup, sizes, sp = serving_session(args)
        line["serving"] = {**stats, "wall_s": wall, "cohorts": setup,
                           "workload": serving_workload(args, sizes, sp)}
    print(json.dumps(line), flush=True)
    return 0


def syrk_int8_mfma(acc) -> dict | None:
    """Outside the kinship's timing: the same GEMM over the staged 2-bit operand once more
    with the int8 MFMA (v_mfma_i32_16x16x64_i8) instead of the FP4 one, into a scratch S;
    its time and whether its S equals the FP4 kernel's bit for bit."""
    import torch

    from jamma_amd import _lib
    from jamma_amd.device import ptr, stream_ptr

    if not acc.packed or os.environ.get("JM_KINSHIP_SYRK_I8MFMA", "0") == "1":
        return None
    n, cols = acc.n, acc.last_flush_cols
    S2 = torch.zeros_like(acc._S)
    st = stream_ptr()
    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
    os.environ["JM_KINSHIP_SYRK_I8MFMA"] = "1"
    try:
        ev[0].record()
        _lib.call("jm_kinship_syrk_2bit", ptr(acc._stage), acc._stage_ld(), n, cols, ptr(S2), st)
        ev[1].record()
    finally:
        del os.environ["JM_KINSHIP_SYRK_I8MFMA"]
    torch.cuda.synchronize()
    low = torch.tril(torch.ones((n, n), dtype=torch.bool, device=S2.device))
    same = bool(torch.equal(S2[low], acc._S[low]))
    return {"ms": ev[0].elapsed_time(ev[1]), "S_identical_to_fp4": same}


def kinship_record(n: int, pf: int, kin_s: float, syrk_ms: float, launches: int,
                   i8form: dict | None = None) -> dict:
    """The kinship stage and its GEMM S = X0 X0^T: n(n+1)p ops (the symmetric rank-p update)
    against the peak of the MFMA it runs on, and against the int8 peak."""
    int8_operand = os.environ.get("JM_KINSHIP_INT8", "0") == "1"
    i8mfma = int8_operand or os.environ.get("JM_KINSHIP_SYRK_I8MFMA", "0") == "1"
    ops_s = float(n) * (n + 1) * pf / (syrk_ms * 1e-3)
    mfma = "int8" if i8mfma else "fp4"
    return {"seconds": kin_s, "gemm_ms": syrk_ms, "gemm_launches": launches,
            "gemm_operand": "int8" if int8_operand else "2-bit packed genotypes (0/1/2, missing 0)",
            "gemm_mfma": "v_mfma_i32_16x16x64_i8 (int32 sums)" if i8mfma else
            "v_mfma_scale_f32_16x16x128_f8f6f4, e2m1 operands (codes widened in registers), "
            "exact integer sums in fp32 (< 2^24 per launch)",
            "gemm_ops": "n(n+1)p (symmetric rank-p update: the ops the SYRK needs)",
            "gemm_pops": ops_s / 1e15,
            "gemm_frac_of_mfma_peak": ops_s / PEAK[mfma], "mfma_peak": f"{mfma} {PEAK[mfma] / 1e15:g} POP/s",
            "gemm_frac_of_int8_peak": ops_s / PEAK["int8"],
            "gemm_frac_of_int8_peak_2n2p": 2.0 * n * n * pf / (syrk_ms * 1e-3) / PEAK["int8"],
            "int8_mfma_form": None if i8form is None else {
                **i8form, "frac_of_int8_peak": float(n) * (n + 1) * pf / (i8form["ms"] * 1e-3)
                / PEAK["int8"]}}


KERNEL_SOURCES = {"k_rotate_bf16x3_pipe4": "rotate.hip", "k_rotate_bf16x3_pipe3": "rotate.hip",
                  "k_rotate_bf16x3_packed": "rotate.hip",
                  "k_gemm_f64_geno": "rotate.hip",
                  "k_syrk_f4_main": "kinship.hip", "k_syrk_f4_tail": "kinship.hip",
                  "k_syrk_p2_main": "kinship.hip", "k_syrk_p2_tail": "kinship.hip",
                  "k_syrk_i8_main": "kinship.hip", "k_syrk_i8_tail": "kinship.hip",
                  "k_impute_C2": "kinship.hip",
                  "k_mom_seg_f32": "moments.hip", "k_mom_fold": "moments.hip",
                  "k_lmm_grid": "lmm.hip", "k_lmm_assoc16": "lmm.hip"}


def source_hash(kernel: str) -> str | None:
    """First 12 hex digits of the SHA-256 of the source file holding `kernel`."""
    import hashlib

    name = KERNEL_SOURCES.get(kernel)
    if name is None:
        return None
    path = Path(__file__).resolve().parent / "jamma_amd" / "csrc" / name
    return hashlib.sha256(path.read_bytes()).hexdigest()[:12]


def hbm_traffic(kernel: str, n: int, tile: int):
    """HBM bytes per launch of `kernel` at (n, tile) from the committed rocprofv3 counter
    summary (profiles/hbm_traffic.txt: FETCH_SIZE x fetch scale + WRITE_SIZE x write scale,
    the scales measured on a copy kernel of known size), valid only while the kernel's
    source file has the hash recorded beside the line. -> (bytes or None, status)."""
    path = Path(__file__).resolve().parent / "profiles" / "hbm_traffic.txt"
    if not path.exists():
        return None, "not measured"
    want = source_hash(kernel)
    for line in path.read_text().splitlines():
        f = line.split()
        if len(f) >= 5 and not line.startswith("#") and f[0] == kernel and \
                int(f[1]) == n and int(f[2]) == tile:
            if len(f) < 6 or f[5] != want:
                return None, f"stale (measured on source {f[5] if len(f) > 5 else '?'}, now {want})"
            return float(f[3]) + float(f[4]), f"measured on source {want}"
    return None, "not measured at this shape"


def vendor_bf16_gemm_tflops(m: int, n: int, k: int) -> float:
    """hipBLASLt bf16 GEMM (torch.matmul) TFLOP/s at the rotation's shape on this device:
    the practical MFMA ceiling the rotation kernel is compared with."""
    import torch

    A = torch.randn(m, k, device="cuda", dtype=torch.bfloat16)
    B = torch.randn(k, n, device="cuda", dtype=torch.bfloat16)
    for _ in range(3):
        C = A @ B
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        C = A @ B
    e1.record()
    torch.cuda.synchronize()
    del A, B, C
    return 2.0 * m * n * k / (e0.elapsed_time(e1) / 10 * 1e-3) / 1e12


def loco(args) -> int:
    """Leave-one-chromosome-out, several phenotypes: n = 20,000 samples x 300,000 SNPs on
    19 chromosomes (length of chromosome c proportional to c^-0.5), 3 Balding-Nichols
    subpopulations (Fst 0.05, 1% missing), 8 phenotypes (h2 = 0.5, 200 causal SNPs,
    genetic correlation 0.3), intercept + 3 Gaussian covariates, -lmm 4.

    Mirrors run_lmm_loco_multi on device-resident genotypes: per-chromosome int8 partial
    Grams, K_c = (S_full - S_c) / (p - p_c), one eigendecomposition per chromosome shared by
    the 8 phenotypes, one rotation per SNP tile shared by the 8 phenotypes. The timed job
    (mixed precision) is kinship + 19 eigendecompositions + association; every SNP x
    phenotype is then recomputed on the FP64 device path (untimed) for parity."""
    import numpy as np
    import torch

    from jamma_amd.device import require_gpu, round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.kinship.compute import KinshipAccumulator
    from jamma_amd.lmm.eigen import eigh_device
    from jamma_amd.lmm.engine import GUARD_BAND, AssocProblem, Cohort, guard_rows

    dev = require_gpu()
    n, p, n_chr, n_ph, tile = 20_000, 300_000, 19, 8, args.tile
    ld = round_up(n, 64)
    w = np.arange(1, n_chr + 1) ** -0.5
    sizes = np.floor(p * w / w.sum()).astype(int)
    sizes[0] += p - sizes.sum()
    chrom = np.repeat(np.arange(n_chr), sizes)
    g = gen_cohort(n, p, 3, 0.05, 0.01, args.seed, ld)
    counts = counts_of(g, n)
    mean32, miss, poly = stats_from_counts(counts, np.float32)
    maf = np.minimum(mean32 / 2, 1 - mean32 / 2)
    keep = (maf >= 0.01) & (miss / n <= 0.05) & poly
    mu64, _, _ = stats_from_counts(counts, np.float64)
    log(f"loco: generated {n} x {p} genotypes, {int(keep.sum())} pass the filters, "
        f"chromosome sizes {sizes.tolist()}")
    rng = np.random.default_rng(args.seed)
    idx_all = np.where(keep)[0]
    causal = np.sort(rng.choice(idx_all, 200, replace=False))
    gc = g[torch.as_tensor(causal, device=dev)][:, :n].to(torch.float64)
    mc = to_dev(mean32[causal].astype(np.float64), torch.float64)
    gc = torch.where(gc < 0, mc[:, None], gc).cpu().numpy()
    b0 = rng.normal(size=200)
    G = np.empty((n, n_ph))
    for k in range(n_ph):  # genetic correlation 0.3 between every pair of traits
        gv = (np.sqrt(0.3) * b0 + np.sqrt(0.7) * rng.normal(size=200)) @ gc
        G[:, k] = (gv - gv.mean()) / gv.std()
    W = np.column_stack([np.ones(n), rng.normal(size=(n, 3))])
    Y = np.sqrt(0.5) * G + np.sqrt(0.5) * rng.normal(size=(n, n_ph)) + \
        (W[:, 1:] @ rng.normal(0, 0.3, (3, n_ph)))
    means = to_dev(mean32.astype(np.float64), torch.float64)

    def accumulate(sel):
        acc = KinshipAccumulator(n, max_snps=len(sel))
        for s in range(0, len(sel), 16384):
            ss = torch.as_tensor(sel[s:s + 16384], device=dev)
            acc.add_tile(g[ss].contiguous(), ld, len(ss), mu64[sel[s:s + 16384]], True,
                         int(counts[sel[s:s + 16384], 3].sum()))
        acc.release()
        return acc

    torch.cuda.synchronize()
    base_mem = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    t_all = time.perf_counter()
    # the full int8 Gram + FP64 corrections once; each chromosome's partial Gram is
    # recomputed from its own SNPs when its LOCO kinship is formed, then freed
    full = accumulate(idx_all)
    torch.cuda.synchronize()
    t_kin = time.perf_counter() - t_all
    # peaks of the kinship phases alone (the counter is reset before each, so an
    # eigendecomposition's workspace does not count), and of the whole run
    kin_peak = torch.cuda.max_memory_allocated() - base_mem
    run_peak = torch.cuda.max_memory_allocated()
    log(f"loco: full Gram in {t_kin:.2f}s")
    t_eig = t_rot = t_assoc = 0.0
    retested = 0
    out = {"mixed": [], "f64": []}
    for c in range(n_chr):
        t0 = time.perf_counter()
        run_peak = max(run_peak, torch.cuda.max_memory_allocated())
        torch.cuda.reset_peak_memory_stats()
        sel_c = idx_all[chrom[idx_all] == c]
        acc_c = accumulate(sel_c)
        K = full.finalize(full.p - acc_c.p, sub=acc_c)
        torch.cuda.synchronize()
        kin_peak = max(kin_peak, torch.cuda.max_memory_allocated() - base_mem)
        del acc_c
        t_kin += time.perf_counter() - t0
        t0 = time.perf_counter()
        d, U = eigh_device(K)
        del K
        torch.cuda.synchronize()
        t_eig += time.perf_counter() - t0
        t0 = time.perf_counter()
        cohort = Cohort(d, U, precision="both")
        del U
        probs = [AssocProblem(cohort, Y[:, k], W) for k in range(n_ph)]
        sel = idx_all[chrom[idx_all] == c]
        res = {"mixed": [], "f64": []}
        for prec in ("mixed", "f64"):
            if prec == "f64":
                torch.cuda.synchronize()
                t_assoc += time.perf_counter() - t0
            for s in range(0, len(sel), tile):
                ss = torch.as_tensor(sel[s:s + tile], device=dev)
                gt = g[ss].contiguous()
                Z, f32, ldz = cohort.rotate_genotypes(gt, ld, means[ss], len(ss), prec)
                outs = []
                for pr in probs:
                    o = pr.assoc_moments(cohort.moments(Z, f32, ldz, len(ss), pr.V,
                                                        pr.n_cvt + 1), len(ss), 4)[: len(ss)]
                    if prec == "mixed":  # guard band: p near a threshold -> FP64 path
                        retested += pr.retest_f64(gt, ld, means[ss], guard_rows(o, 4), 4, o)
                    outs.append(o)
                res[prec].append(torch.stack(outs, 1).cpu())
                del Z, gt
        for prec in res:
            out[prec].append(torch.cat(res[prec]))
        log(f"loco: chromosome {c + 1}: {len(sel)} SNPs x {n_ph} phenotypes, T={cohort.n_terms}, "
            f"{cohort.bf16_planes} bf16 planes")
        del cohort, probs
    total = time.perf_counter() - t_all
    a = torch.cat(out["mixed"]).numpy()  # (SNPs, phenotypes, outputs)
    b = torch.cat(out["f64"]).numpy()
    tests = a.shape[0] * n_ph
    parity = {}
    for name, col in (("p_wald", 2), ("p_lrt", 7), ("p_score", 8)):
        with np.errstate(divide="ignore", invalid="ignore"):
            dl = np.abs(np.log10(a[:, :, col]) - np.log10(b[:, :, col]))
        flips = sum(int(np.sum((a[:, :, col] < t) != (b[:, :, col] < t)))
                    for t in (0.05, 0.01, 1e-3, 5e-8))
        parity[name] = {"max_abs_dlog10p": float(np.nanmax(dl)), "significance_flips": flips}
    parity["beta_sign_agreement"] = bool(np.array_equal(np.sign(a[:, :, 0]), np.sign(b[:, :, 0])))
    parity["n_sig_5e-8_wald"] = int(np.sum(b[:, :, 2] < 5e-8))
    parity["guard_band_log10p"] = GUARD_BAND
    parity["retested_f64"] = int(retested)
    timed = t_kin + t_eig + t_assoc
    line = {
        "metric": "LOCO SNP x phenotype tests/s (kinships + eigendecompositions + association)",
        "value": tests / timed, "unit": "tests/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "integer kinship GEMM (fp4 MFMA, exact), f64 eigen, bf16x3 rotation + f64 lambda/tests",
        "data": "synthetic (Balding-Nichols genotypes generated on device)",
        "config": {"workload": f"loco: n={n} x {p} SNPs ({a.shape[0]} pass filters) on {n_chr} "
                               f"chromosomes (length ~ c^-0.5), {n_ph} phenotypes (genetic corr "
                               f"0.3, h2=0.5), intercept + 3 covariates, -lmm 4",
                   "n": n, "p": p, "chromosomes": n_chr, "phenotypes": n_ph, "tile_snps": tile},
        "seconds": {"kinship": t_kin, "eigendecompositions": t_eig, "association_mixed": t_assoc,
                    "timed_total": timed, "wall_including_f64_parity": total},
        "kinship_peak_hbm_gb": kin_peak / 1e9,
        "peak_hbm_gb": (max(run_peak, torch.cuda.max_memory_allocated()) - base_mem) / 1e9,
        "parity_mixed_vs_f64": parity,
    }
    print(json.dumps(line), flush=True)
    return 0


def pipeline(args) -> int:
    """The precision-form cohort (n = 10,000 x 100,000 SNPs, 3 Balding-Nichols
    subpopulations, Fst 0.05, 1% missing, h2 = 0.5 / 200 causal, 3 covariates +
    intercept) written as a PLINK fileset, then the reference-signature gwas() run end to
    end (-lmm 4): .bed read from disk and copied to the device, decode, pass-1 statistics,
    integer kinship GEMM, host kinship matrix (the API's contract), eigendecomposition, rotation,
    lambda search and tests, .assoc.txt written. Mixed precision, then FP64."""
    import tempfile

    import numpy as np
    import torch

    from jamma_amd import gwas
    from jamma_amd.device import require_gpu, round_up

    dev = require_gpu()
    n, p = 10_000, 100_000
    ld = round_up(n, 64)
    g = gen_cohort(n, p, 3, 0.05, 0.01, args.seed, ld)
    rng = np.random.default_rng(args.seed)
    causal = np.sort(rng.choice(p, 200, replace=False))
    gc = g[torch.as_tensor(causal, device=dev)][:, :n].to(torch.float64)
    cnt = (gc >= 0).sum(1).clamp(min=1)
    mu = torch.where(gc >= 0, gc, torch.zeros_like(gc)).sum(1) / cnt
    gc = torch.where(gc < 0, mu[:, None], gc)
    gv = (torch.as_tensor(rng.normal(size=200), device=dev)[None, :] @ gc).squeeze(0).cpu().numpy()
    gv = (gv - gv.mean()) / gv.std()
    W = np.column_stack([np.ones(n), rng.normal(size=(n, 3))])
    y = np.sqrt(0.5) * gv + np.sqrt(0.5) * rng.normal(size=n) + W[:, 1:] @ rng.normal(0, 0.3, 3)
    tmp = Path(tempfile.mkdtemp(prefix="jm_pipeline_"))
    prefix = tmp / "cohort"
    t0 = time.perf_counter()
    bps = (n + 3) // 4
    with open(f"{prefix}.bed", "wb") as f:  # PLINK SNP-major: A1 count 2/1/0 -> 00/10/11, missing 01
        f.write(bytes([0x6C, 0x1B, 0x01]))
        for s0 in range(0, p, 8192):
            gg = g[s0:s0 + 8192, :n]
            code = torch.full((gg.shape[0], bps * 4), 3, dtype=torch.uint8, device=dev)
            code[:, :n] = torch.where(gg < 0, 1, torch.where(gg == 2, 0, torch.where(gg == 1, 2, 3))
                                      ).to(torch.uint8)
            c4 = code.view(gg.shape[0], bps, 4)
            packed = c4[:, :, 0] | (c4[:, :, 1] << 2) | (c4[:, :, 2] << 4) | (c4[:, :, 3] << 6)
            f.write(packed.cpu().numpy().tobytes())
    with open(f"{prefix}.bim", "w") as f:
        f.writelines(f"1\trs{j:06d}\t0\t{1000 * (j + 1)}\tA\tG\n" for j in range(p))
    with open(f"{prefix}.fam", "w") as f:
        f.writelines(f"F{i}\tI{i}\t0\t0\t0\t{float(y[i])!r}\n" for i in range(n))
    np.savetxt(tmp / "covariates.txt", W, fmt="%.17g")
    del g, gc
    torch.cuda.empty_cache()
    log(f"pipeline: wrote {n} x {p} PLINK fileset ({(3 + bps * p) / 1e6:.0f} MB .bed) in "
        f"{time.perf_counter() - t0:.1f}s")
    runs = {}
    for prec in ("mixed", "f64"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = gwas(prefix, covariate_file=tmp / "covariates.txt", lmm_mode=4, output_dir=tmp,
                 output_prefix=f"result_{prec}", show_progress=False, precision=prec)
        wall = time.perf_counter() - t0
        runs[prec] = {"wall_s": wall, "snps_per_s": r.n_snps_tested / wall,
                      "n_snps_tested": r.n_snps_tested, **r.timing}
        log(f"pipeline {prec}: {wall:.1f}s end to end ({r.n_snps_tested / wall:,.0f} SNPs/s), "
            f"kinship {r.timing['kinship_s']:.1f}s, eigendecomposition + association "
            f"{r.timing['lmm_s']:.1f}s")
    line = {
        "metric": "end-to-end SNPs/s through gwas() (.bed on disk -> .assoc.txt)",
        "value": runs["mixed"]["snps_per_s"], "unit": "SNPs/s", "n_gpus": 1,
        "higher_is_better": True,
        "data": "synthetic (Balding-Nichols genotypes written as a PLINK fileset)",
        "config": {"workload": f"pipeline: precision form n={n} x {p} SNPs through gwas(), "
                               f"3 covariates + intercept, -lmm 4", "n": n, "p": p},
        "mixed": runs["mixed"], "f64": runs["f64"],
    }
    print(json.dumps(line), flush=True)
    return 0


def small(args) -> int:
    """The plain method at 1,940 samples x 12,226 SNPs, FP64: genotypes ~ Binomial(2, p_j)
    with p_j ~ U(0.05, 0.5), 0.5% of calls missing (mean-imputed); phenotype = 50 causal
    SNP effects beta ~ N(0, 0.05) + N(0, 1); intercept only; centred kinship, its
    eigendecomposition, then the per-SNP REML Wald test (-lmm 1).

    Device: integer kinship GEMM + 64-bit eigendecomposition + FP64 rotation and lambda search
    over every SNP. CPU: the oracle on the whole job (NumPy centred kinship, LAPACK
    eigh, the same association), on the host cores, with its p-values as the check."""
    import numpy as np
    import torch

    from jamma_amd.device import require_gpu, round_up, to_dev
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.kinship.compute import KinshipAccumulator
    from jamma_amd.lmm.eigen import eigh_device
    from jamma_amd.lmm.engine import AssocProblem, Cohort
    from oracle import kinship as ok
    from oracle import lmm as ol

    dev = require_gpu()
    n, p = 1940, 12226
    ld = round_up(n, 64)
    g = gen_cohort(n, p, 1, 1e-6, 0.005, args.seed, ld)  # one population: Binomial(2, p_j)
    counts = counts_of(g, n)
    mean32, miss, poly = stats_from_counts(counts, np.float32)
    af = mean32 / 2
    keep = (np.minimum(af, 1 - af) >= 0.01) & (miss / n <= 0.05) & poly
    idx = np.where(keep)[0]
    gf = g[torch.as_tensor(idx, device=dev)].contiguous()
    pf = gf.shape[0]
    means = to_dev(mean32[idx], torch.float64)
    mu64, _, _ = stats_from_counts(counts[idx], np.float64)
    rng = np.random.default_rng(args.seed)
    causal = np.sort(rng.choice(pf, 50, replace=False))
    Gh = gf[:, :n].cpu().numpy().astype(np.float64).T  # (n, pf), -1 = missing
    Gh[Gh < 0] = np.nan
    Gc = np.where(np.isnan(Gh[:, causal]), mu64[causal][None, :], Gh[:, causal])
    y = Gc @ rng.normal(0.0, 0.05, 50) + rng.normal(size=n)
    W = np.ones((n, 1))

    def device_job():
        acc = KinshipAccumulator(n)
        for s in range(0, pf, 16384):
            e = min(s + 16384, pf)
            acc.add_tile(gf[s:e], ld, e - s, mu64[s:e], True, int(counts[idx[s:e], 3].sum()))
        K = acc.finalize(pf)
        d, U = eigh_device(K)
        co = Cohort(d, U, precision="f64")
        prob = AssocProblem(co, y, W)
        torch.cuda.synchronize()
        t_a = time.perf_counter()
        outs = [prob.associate(gf[s:min(s + 8192, pf)], ld, means[s:min(s + 8192, pf)],
                               min(s + 8192, pf) - s, 1, "f64")[: min(s + 8192, pf) - s]
                for s in range(0, pf, 8192)]
        torch.cuda.synchronize()
        return torch.cat(outs).cpu().numpy(), time.perf_counter() - t_a

    for _ in range(max(1, args.warmup)):
        device_job()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    walls, assoc = [], []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        out, ta = device_job()
        walls.append(time.perf_counter() - t1)
        assoc.append(ta)
    wall = float(np.mean(walls))
    log(f"small: device job {wall * 1e3:.1f} ms ({pf} SNPs), association {np.mean(assoc) * 1e3:.1f} ms")

    # the oracle on the whole job (kinship -> eigh -> association), host threads
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    K_ref = ok.centered_kinship(Gh)
    t_kin = time.perf_counter() - t0
    d_ref, U_ref = ol.eigendecompose(K_ref)
    t_eig = time.perf_counter() - t0 - t_kin
    ref = ol.run_lmm(Gh, y, eigenvalues=d_ref, eigenvectors=U_ref, maf_threshold=0.0,
                     miss_threshold=1.0, lmm_mode=1, stat_dtype=np.float32)
    cpu_s = time.perf_counter() - t0
    t_assoc = cpu_s - t_kin - t_eig
    with np.errstate(divide="ignore"):
        dl = float(np.nanmax(np.abs(np.log10(out[:, 2]) - np.log10(ref["p_wald"]))))
    log(f"small: oracle {cpu_s:.1f}s on {cores} threads (kinship {t_kin:.1f}s, eigh {t_eig:.1f}s, "
        f"association {t_assoc:.1f}s); max |dlog10 p_wald| {dl:.2e}")
    line = {
        "metric": "whole-job SNPs/s, plain FP64 method (-gk 1, eigendecomposition, -lmm 1)",
        "value": pf / wall, "unit": "SNPs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic (Binomial(2, p_j) genotypes, 0.5% missing)",
        "config": {"workload": f"small: n={n} x {p} SNPs ({pf} pass the filters), intercept "
                               f"only, centred kinship + eigendecomposition + REML Wald",
                   "n": n, "p": p},
        "association_snps_per_s": pf / float(np.mean(assoc)),
        "parity": {"wald_vs_oracle_max_abs_dlog10p": dl, "n_snps": pf},
        "cpu_baseline": {"value": pf / cpu_s, "unit": "SNPs/s", "cores": cores,
                         "cpu_model": cpu_model(), "kind": "port",
                         "sample": f"the whole job ({pf} SNPs): NumPy centred kinship "
                                   f"{t_kin:.1f}s, LAPACK eigh {t_eig:.1f}s, association "
                                   f"{t_assoc:.1f}s",
                         "association_snps_per_s": pf / t_assoc},
    }
    print(json.dumps(line), flush=True)
    return 0


def copy_calibration() -> int:
    """A plain elementwise copy of known size (1 GiB read + 1 GiB written), used to
    scale the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 into HBM bytes."""
    import torch

    x = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel": "torch.neg float32", "bytes_read": x.numel() * 4,
                      "bytes_written": y.numel() * 4, "launches": 3}), flush=True)
    return 0


def serving_workload(args, sizes, p) -> str:
    return (f"serving: resident cohorts n={sizes} x {p} SNPs, {args.requests} requests, Poisson "
            f"{args.rate}/s, ranges LogUniform(1k,50k), 0-10 covariates, h2~U(0.2,0.8), -lmm 4")


def serving(args) -> int:
    """Serving config alone: resident cohorts, Poisson arrivals of (phenotype x SNP range)
    requests, HBM-budget admission, continuous batching; p50/p99 latency + SNPs/s."""
    stats, wall, setup, sizes, p = serving_session(args)
    line = {"metric": METRIC, "value": stats["snps_per_s"], "unit": "SNPs/s", "n_gpus": 1,
            "steps": args.requests, "warmup": 0, "ms_per_step": wall * 1e3 / args.requests,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16x3 rotation (fp32 acc) + f64 lambda/tests",
            "data": "synthetic (Balding-Nichols genotypes generated on device)",
            "config": {"workload": serving_workload(args, sizes, p)},
            "serving": {**stats, "wall_s": wall, "cohorts": setup}}
    print(json.dumps(line), flush=True)
    return 0


def serving_session(args):
    """Build the resident cohorts, replay the Poisson request stream, return
    (latency stats, wall seconds, per-cohort setup, cohort sizes, SNPs per cohort)."""
    import numpy as np
    import torch

    from jamma_amd.device import require_gpu, round_up
    from jamma_amd.io.plink import stats_from_counts
    from jamma_amd.kinship.compute import KinshipAccumulator
    from jamma_amd.lmm.eigen import eigh_device
    from jamma_amd.serve import Request, ResidentCohort, Server

    dev = require_gpu()
    sizes = [int(x) for x in args.cohorts.split(",")]
    p = 60_000
    rng = np.random.default_rng(args.seed)
    srv = Server(slice_snps=4096, max_batch_snps=16384)
    setup = {}
    for ci, n in enumerate(sizes):
        t0 = time.perf_counter()
        ld = round_up(n, 64)
        g = gen_cohort(n, p, 3, 0.05, 0.01, args.seed + ci, ld)
        counts = counts_of(g, n)
        mean32, _, _ = stats_from_counts(counts, np.float32)
        mu64, _, poly = stats_from_counts(counts, np.float64)
        acc = KinshipAccumulator(n)
        for s in range(0, p, 8192):
            e = min(s + 8192, p)
            acc.add_tile(g[s:e], ld, e - s, mu64[s:e], True, int(counts[s:e, 3].sum()))
        K = acc.finalize(p)
        del acc
        t1 = time.perf_counter()
        d, U = eigh_device(K)
        del K
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rc = ResidentCohort(f"n{n}", g, ld, torch.as_tensor(mean32, device=dev), d, U, "mixed")
        del U
        srv.add_cohort(rc)
        setup[f"n{n}"] = {"kinship_s": t1 - t0, "eigh_s": t2 - t1, "terms": rc.cohort.n_terms}
        log(f"cohort n={n}: kinship {t1 - t0:.1f}s eigh {t2 - t1:.1f}s T={rc.cohort.n_terms}")
    names = list(srv.cohorts)
    reqs = []
    for i in range(args.requests):
        name = names[int(rng.integers(0, len(names)))]
        rc = srv.cohorts[name]
        n = rc.n
        h2 = rng.uniform(0.2, 0.8)
        ncov = int(rng.integers(0, 11))
        W = np.column_stack([np.ones(n), rng.normal(size=(n, ncov))]) if ncov else None
        # genetic phenotype: 50 causal SNPs of the cohort (mean-imputed genotypes on the
        # device), standardised genetic value scaled to h2, plus noise and covariate effects
        causal = torch.as_tensor(np.sort(rng.choice(rc.p, 50, replace=False)), device=dev)
        gx = rc.g[causal, :n].to(torch.float64)
        gx = torch.where(gx < 0, rc.means[causal][:, None], gx)
        gv = (torch.as_tensor(rng.normal(size=50), device=dev) @ gx).cpu().numpy()
        gv = (gv - gv.mean()) / (gv.std() + 1e-12)
        y = np.sqrt(h2) * gv + np.sqrt(1 - h2) * rng.normal(size=n)
        if ncov:
            y = y + W[:, 1:] @ rng.normal(0, 0.3, ncov)
        length = int(np.exp(rng.uniform(np.log(1000), np.log(50000))))
        start = int(rng.integers(0, rc.p - length))
        reqs.append(Request(name, y, start, start + length, covariates=W, mode=4))
    gaps = rng.exponential(1.0 / args.rate, size=len(reqs))
    log(f"serving {len(reqs)} requests at {args.rate}/s")
    with srv:
        futs = []
        t0 = time.perf_counter()
        for r, gap in zip(reqs, gaps):
            time.sleep(gap)
            futs.append(srv.submit(r))
        for f in futs:
            f.result(timeout=1200)
        wall = time.perf_counter() - t0
    stats = srv.latency_stats()
    log(f"serving: p50 {stats['p50_s'] * 1e3:.1f} ms, p99 {stats['p99_s'] * 1e3:.1f} ms, "
        f"{stats['snps_per_s']:,.0f} SNPs/s")
    # saturation: the same requests all submitted at once (a fresh server on the same
    # resident cohorts), SNPs/s over the time to drain them
    sat = Server(slice_snps=4096, max_batch_snps=16384)
    for rc in srv.cohorts.values():
        sat.add_cohort(rc)
    with sat:
        t0 = time.perf_counter()
        futs = [sat.submit(Request(r.cohort, r.phenotype, r.snp_start, r.snp_end, r.covariates,
                                   r.mode)) for r in reqs]
        for f in futs:
            f.result(timeout=1200)
        sat_wall = time.perf_counter() - t0
    total = sum(r.n_snps for r in reqs)
    stats["saturation_snps_per_s"] = total / sat_wall
    stats["saturation_wall_s"] = sat_wall
    stats["offered_snps_per_s"] = total / wall
    log(f"serving saturation: {total / sat_wall:,.0f} SNPs/s ({len(reqs)} requests at once, "
        f"{sat_wall:.2f} s)")
    return stats, wall, setup, sizes, p


if __name__ == "__main__":
    sys.exit(main())
