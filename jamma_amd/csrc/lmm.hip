// Per-SNP REML / MLE variance-ratio search fused with the Wald, LRT and Score
// statistics, one wavefront per SNP, on the compressed eigen-spectrum.
//
// Reference semantics (paths relative to the reference repository root):
//   calc_pab_jax                 src/jamma/lmm/likelihood_jax.py:148-186
//   mle_log_likelihood_jax       likelihood_jax.py:189-228
//   _reml_with_precomputed_iab   likelihood_jax.py:280-339
//   _golden_section_refine       likelihood_jax.py:419-488
//   golden_section_optimize_lambda(_mle)  likelihood_jax.py:491-562, 830-883
//   calc_wald_stats_jax          likelihood_jax.py:601-675
//   calc_score_stats_jax         likelihood_jax.py:706-773
//   calc_lrt_pvalue_jax          likelihood_jax.py:231-251
//   null model (host golden)     src/jamma/lmm/likelihood.py:349-392, 395-474
//
// A likelihood evaluation at lambda needs row 0 of Pab, i.e. the weighted
// sums  sum_i v_a v_b / (1 + lambda d_i)  and  logdet_h = sum_i log|1+lambda d_i|.
// On the compressed spectrum each is a short dot product over T terms:
//   Pab0_ab = sum_t kappa_t(lambda) * mom_ab[t],  kappa_t = h_c (-z_c)^m,
//   logdet  = sum_t ell_t(lambda) * L[t],
// with h_c = 1/(1+lambda d_c), z_c = lambda h_c (cluster c, power m of term t).
//
// Wave layout of an evaluation: lanes over terms compute kappa (to LDS) and
// the logdet partials; lanes over packed pairs form the dot products; the
// Pab recursion then runs on an LDS column. REML and MLE golden searches run
// in lockstep in the two half-waves.

#include "jm_common.h"

#include <cmath>

namespace jm {

// ---------------------------------------------------------------------------
// Special functions (host + device)
// ---------------------------------------------------------------------------

// Continued fraction for I_x(a, b) (modified Lentz; Numerical Recipes 6.4 /
// DLMF 8.17.22).
__host__ __device__ static double betacf(double a, double b, double x) {
  const double fpmin = 1e-300;
  const double eps = 1e-16;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0;
  double d = 1.0 - qab * x / qap;
  if (fabs(d) < fpmin) d = fpmin;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m <= 5000; ++m) {
    const double m2 = 2.0 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (fabs(d) < fpmin) d = fpmin;
    c = 1.0 + aa / c;
    if (fabs(c) < fpmin) c = fpmin;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (fabs(del - 1.0) < eps) break;
  }
  return h;
}

// Stirling remainder of log Gamma: lgamma(x) - [(x-1/2) log x - x + log(2 pi)/2]
// (asymptotic series, x >= 10: truncation < 1e-17 relative).
__host__ __device__ static double stirling_corr(double x) {
  const double r = 1.0 / x, r2 = r * r;
  return r * (1.0 / 12.0 - r2 * (1.0 / 360.0 - r2 * (1.0 / 1260.0 - r2 * (1.0 / 1680.0))));
}

// lgamma(a + b) - lgamma(a), without the cancellation of two large lgamma
// values when a is large (the F-test has a = df / 2 up to ~1e5).
__host__ __device__ static double lgamma_diff(double a, double b) {
  if (a < 10.0) return lgamma(a + b) - lgamma(a);
  const double ab = a + b;
  return (a - 0.5) * log1p(b / a) + b * log(ab) - b + (stirling_corr(ab) - stirling_corr(a));
}

// Regularised incomplete beta I_x(a, b), a, b > 0.
__host__ __device__ double betainc_reg(double a, double b, double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : NAN;
  if (!(x < 1.0)) return (x == 1.0) ? 1.0 : NAN;
  const double lg = (a >= b) ? lgamma_diff(a, b) - lgamma(b) : lgamma_diff(b, a) - lgamma(a);
  const double lbt = lg + a * log(x) + b * log1p(-x);
  if (x < (a + 1.0) / (a + b + 2.0)) return exp(lbt) * betacf(a, b, x) / a;
  return 1.0 - exp(lbt) * betacf(b, a, 1.0 - x) / b;
}

// chi-square(1) survival: P(X > x) = erfc(sqrt(x / 2)).
__host__ __device__ double chi2_sf1(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 1.0 : NAN;
  return erfc(sqrt(0.5 * x));
}

// F(1, df) upper tail through betainc, exactly as likelihood_jax.py:668-673.
__host__ __device__ static double f_pvalue(double df, double f_stat) {
  double z = df / (df + fmax(f_stat, 1e-10));
  z = fmin(fmax(z, 0.0), 1.0);
  double p = betainc_reg(0.5 * df, 0.5, z);
  if (f_stat <= 0.0) p = 1.0;
  return p;
}

// GEMMA packed upper-triangle index (likelihood.py:25-43), 1-based a, b.
__host__ __device__ __forceinline__ int ab_index(int a, int b, int nc) {
  const int cols = nc + 2;
  if (a > b) {
    const int t = a;
    a = b;
    b = t;
  }
  return (2 * cols - a + 2) * (a - 1) / 2 + b - a;
}

// ---------------------------------------------------------------------------
// Pab recursion on an LDS column (likelihood_jax.py:176-184)
// ---------------------------------------------------------------------------
struct PabSummary {
  double hiw;            // sum over levels 0..nc of [log Pab_ii]+ - [log Iab_ii]+
  double pxx, pxy, pyy;  // level nc
  double pyy1;           // level nc + 1
};

// P[idx * stride] holds row 0 on entry; rows are updated in place (entries
// (a,b) with a,b > p only read level p-1 values of column p, which level p
// does not modify). iab_ld (may be null): log Iab_ii per level (0 when
// Iab_ii <= 0), subtracted in the reference's interleaved order
// (likelihood_jax.py:322-327). out_ld (may be null) receives this column's
// level logs (used to form iab_ld).
__device__ static void pab_recursion(double* P, int stride, int nc, PabSummary& o,
                                     const double* iab_ld, double* out_ld) {
  double hiw = 0.0;
  {
    const double v = P[ab_index(1, 1, nc) * stride];
    const double lv = (v > 0.0) ? log(v) : 0.0;
    if (out_ld) out_ld[0] = lv;
    hiw = hiw + lv;
    if (iab_ld) hiw = hiw - iab_ld[0];
  }
  for (int p = 1; p <= nc + 1; ++p) {
    const double ps_ww = P[ab_index(p, p, nc) * stride];
    const double inv = (ps_ww != 0.0) ? 1.0 / ps_ww : 0.0;
    for (int a = p + 1; a <= nc + 2; ++a) {
      const double ps_aw = P[ab_index(a, p, nc) * stride];
      for (int b = a; b <= nc + 2; ++b) {
        const int iab = ab_index(a, b, nc);
        const double ps_bw = P[ab_index(b, p, nc) * stride];
        P[iab * stride] = P[iab * stride] - ps_aw * ps_bw * inv;
      }
    }
    if (p == nc) {
      o.pxx = P[ab_index(nc + 1, nc + 1, nc) * stride];
      o.pxy = P[ab_index(nc + 1, nc + 2, nc) * stride];
      o.pyy = P[ab_index(nc + 2, nc + 2, nc) * stride];
    }
    if (p <= nc) {
      const double v = P[ab_index(p + 1, p + 1, nc) * stride];
      const double lv = (v > 0.0) ? log(v) : 0.0;
      if (out_ld) out_ld[p] = lv;
      hiw = hiw + lv;
      if (iab_ld) hiw = hiw - iab_ld[p];
    }
  }
  o.pyy1 = P[ab_index(nc + 2, nc + 2, nc) * stride];
  o.hiw = hiw;
}

__device__ __forceinline__ double clamp_pyy(double v) {
  return (v >= 0.0 && v < 1e-8) ? 1e-8 : v;
}

// REML value, likelihood_jax.py:303-339.
__device__ __forceinline__ double reml_value(const PabSummary& o, int nc, int n,
                                             double logdet_h) {
  const double df = (double)(n - nc - 1);
  const double pyy = clamp_pyy(o.pyy1);
  const double c = 0.5 * df * (log(df) - log(2.0 * M_PI) - 1.0);
  return c - 0.5 * logdet_h - 0.5 * o.hiw - 0.5 * df * log(pyy);
}

// MLE value, likelihood_jax.py:209-228 (level nc+1 P_yy).
__device__ __forceinline__ double mle_value(double pyy_level, int n, double logdet_h) {
  const double nn = (double)n;
  const double pyy = clamp_pyy(pyy_level);
  const double c = 0.5 * nn * (log(nn) - log(2.0 * M_PI) - 1.0);
  return c - 0.5 * logdet_h - 0.5 * nn * log(pyy);
}

// beta / se with GEMMA's safe_sqrt (likelihood_jax.py:645-661).
__device__ __forceinline__ void beta_se(double df, double pxx, double pxy, double pxyy,
                                        double& beta, double& se, double& tau) {
  tau = df / pxyy;
  if (pxx > 0.0) {
    beta = pxy / pxx;
    double var = 1.0 / (tau * pxx);
    if (fabs(var) < 0.001) var = fabs(var);
    se = sqrt(var);
  } else {
    beta = NAN;
    se = NAN;
  }
}

// kappa_t and ell_t at lambda for term (center dc, power m).
__device__ __forceinline__ void term_coeffs(double lam, double dc, int m, double& kap,
                                            double& ell) {
  const double v = lam * dc + 1.0;
  const double h = 1.0 / v;
  const double z = lam * h;
  double zm = 1.0;
  for (int k = 0; k < m; ++k) zm *= z;
  kap = (m & 1) ? -h * zm : h * zm;
  if (m == 0)
    ell = log(fabs(v));
  else
    ell = ((m & 1) ? zm : -zm) / (double)m;
}

// ---------------------------------------------------------------------------
// Shared evaluation context
// ---------------------------------------------------------------------------
struct WaveCtx {
  const jm_lmm_problem* pr;  // in kernel-argument memory (by value copy below)
  int nc, ni, T, lane;
  double* kap;  // LDS: 2 x T
  double* P;    // LDS: ni x 64
};

// Evaluate row-0 Pab for up to two lambdas (half-wave h uses lam[h]) and
// write it to column h of P. Returns this lane's half logdet_h.
// xrow: per-SNP moments base (n_rows x ldm) or nullptr for the null model
// (genotype pairs are zero).
__device__ static double eval_two(const jm_lmm_problem& pr, int nc, int ni, int T, int lane,
                                  double* kap, double* P, double lam0, double lam1,
                                  const double* xrow, int64_t ldm, const double* kap_table) {
  // (1) kappa for both lambdas, logdet partials
  double ld0 = 0.0, ld1 = 0.0;
  if (kap_table) {
    for (int t = lane; t < T; t += 64) {
      kap[t] = kap_table[t];
      kap[T + t] = kap_table[t];
    }
  } else {
    for (int t = lane; t < T; t += 64) {
      const double dc = pr.cl_center[pr.term_cluster[t]];
      const int m = pr.term_m[t];
      const double L = pr.logdet_mom[t];
      double k0, e0, k1, e1;
      term_coeffs(lam0, dc, m, k0, e0);
      term_coeffs(lam1, dc, m, k1, e1);
      kap[t] = k0;
      kap[T + t] = k1;
      ld0 += e0 * L;
      ld1 += e1 * L;
    }
  }
  ld0 = wave_sum(ld0);
  ld1 = wave_sum(ld1);
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  // (2) dot products: half h, pair slot j
  const int h = lane >> 5;
  const double* kh = kap + h * T;
  for (int j = lane & 31; j < ni; j += 32) {
    const int r = pr.pair_row[j];
    const double* row;
    if (r >= 0) {
      row = xrow ? xrow + (int64_t)r * ldm : nullptr;
    } else {
      row = pr.shared_mom + (int64_t)j * T;
    }
    double acc = 0.0;
    if (row) {
      for (int t = 0; t < T; ++t) acc += kh[t] * row[t];
    }
    P[j * 64 + h] = acc;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
  return h ? ld1 : ld0;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Grid / null tables: grid_kappa[t][g], grid_logdet[g], grid_shared[g][pair],
// null_kappa[t], null_shared[pair].
__global__ void k_lmm_prepare(jm_lmm_problem pr) {
  const int T = pr.n_terms;
  const int G = pr.n_grid;
  const int nc = pr.n_cvt;
  const int ni = (nc + 3) * (nc + 2) / 2;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  // kappa table over (t, g) and null kappa (g == G)
  for (int idx = tid; idx < T * (G + 1); idx += nthreads) {
    const int t = idx / (G + 1), g = idx % (G + 1);
    const double lam = (g < G) ? exp(pr.grid_loglam[g]) : pr.lambda_null;
    double kap, ell;
    term_coeffs(lam, pr.cl_center[pr.term_cluster[t]], pr.term_m[t], kap, ell);
    if (g < G)
      pr.grid_kappa[t * G + g] = kap;
    else
      pr.null_kappa[t] = kap;
  }
  // logdet per grid point + shared pair sums (one thread per (g, pair|logdet))
  for (int idx = tid; idx < (G + 1) * (ni + 1); idx += nthreads) {
    const int g = idx / (ni + 1), j = idx % (ni + 1);
    const double lam = (g < G) ? exp(pr.grid_loglam[g]) : pr.lambda_null;
    double acc = 0.0;
    if (j == ni) {
      for (int t = 0; t < T; ++t) {
        double kap, ell;
        term_coeffs(lam, pr.cl_center[pr.term_cluster[t]], pr.term_m[t], kap, ell);
        acc += ell * pr.logdet_mom[t];
      }
      if (g < G) pr.grid_logdet[g] = acc;
    } else {
      if (pr.pair_row[j] < 0) {
        const double* row = pr.shared_mom + (int64_t)j * T;
        for (int t = 0; t < T; ++t) {
          double kap, ell;
          term_coeffs(lam, pr.cl_center[pr.term_cluster[t]], pr.term_m[t], kap, ell);
          acc += kap * row[t];
        }
      }
      if (g < G)
        pr.grid_shared[g * ni + j] = acc;
      else
        pr.null_shared[j] = acc;
    }
  }
}

// Null model: golden_section_minimize (likelihood.py:395-474) over
// -mle_log_likelihood_null (likelihood.py:349-392), one wavefront.
__global__ void k_lmm_null(jm_lmm_problem pr, double* out) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x;
  const int nc = pr.n_cvt, T = pr.n_terms, n = pr.n;
  const int ni = (nc + 3) * (nc + 2) / 2;
  double* kap = smem;        // 2T
  double* P = smem + 2 * T;  // ni x 64
  auto f_null = [&](double lam) -> double {  // -logl (minimised)
    const double ldh = eval_two(pr, nc, ni, T, lane, kap, P, lam, lam, nullptr, 0, nullptr);
    PabSummary o;
    pab_recursion(P, 64, nc, o, nullptr, nullptr);
    __builtin_amdgcn_wave_barrier();
    // null MLE: P_yy at level nc (after the covariates)
    return -mle_value(o.pyy, n, ldh);
  };
  const int G = pr.n_grid;
  const double phi = 0.6180339887498949;
  const double lo = log(pr.l_min), hi = log(pr.l_max);
  const double step = (hi - lo) / (double)(G - 1);
  int best = 0;
  double best_val = f_null(exp(lo));
  for (int i = 1; i < G; ++i) {
    const double v = f_null(exp(lo + i * step));
    if (v < best_val) {
      best_val = v;
      best = i;
    }
  }
  double a = lo + (best > 0 ? best - 1 : 0) * step;
  double b = lo + (best + 1 < G ? best + 1 : G - 1) * step;
  double c = b - phi * (b - a);
  double d = a + phi * (b - a);
  double fc = f_null(exp(c));
  double fd = f_null(exp(d));
  for (int it = 0; it < pr.n_iter; ++it) {
    if (fc < fd) {
      b = d;
      d = c;
      fd = fc;
      c = b - phi * (b - a);
      fc = f_null(exp(c));
    } else {
      a = c;
      c = d;
      fc = fd;
      d = a + phi * (b - a);
      fd = f_null(exp(d));
    }
  }
  const double lam = exp((a + b) / 2.0);
  const double logl = -f_null(lam);
  if (lane == 0) {
    out[0] = lam;
    out[1] = logl;
  }
}

// argmax with NaN first and first-index ties (numpy / jnp.argmax semantics).
__device__ __forceinline__ bool better(double v, int i, double bv, int bi) {
  const bool vn = isnan(v), bn = isnan(bv);
  if (vn || bn) {
    if (vn && bn) return i < bi;
    return vn;
  }
  if (v > bv) return true;
  if (v < bv) return false;
  return i < bi;
}

__device__ __forceinline__ void wave_argmax(double& v, int& idx) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const double ov = shfl_xor_d(v, m);
    const int oi = __shfl_xor(idx, m, 64);
    if (better(ov, oi, v, idx)) {
      v = ov;
      idx = oi;
    }
  }
}

constexpr int kMaxLevels = 24;  // n_cvt <= 20 -> levels 0..n_cvt

// Main association kernel: one wavefront per SNP, grid-stride over SNPs.
// Per-wave LDS: kappa (2T) | P (ni_max x 64) | iab level logs (kMaxLevels).
// MULTI: SNP s belongs to problem snp_prob[s] (its own phenotype, covariates,
// null model and mode) and its moments start at snp_xmom[s]; all problems of
// one launch share the cohort (same T). Otherwise every SNP uses pr_one.
template <bool MULTI>
__global__ void __launch_bounds__(256) k_lmm_assoc(jm_lmm_problem pr_one, int mode_one,
                                                   const double* xmom, int64_t ldm, int64_t n_snps,
                                                   double* out, const jm_lmm_problem* probs,
                                                   const int32_t* modes, const int32_t* snp_prob,
                                                   const double* const* snp_xmom, int T_all,
                                                   int ni_max) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int per_wave = 2 * T_all + ni_max * 64 + kMaxLevels;
  double* kap = smem + (size_t)wib * per_wave;
  double* P = kap + 2 * T_all;
  double* iab_ld = P + ni_max * 64;
  const double phi = 0.6180339887498949;
  const int h = lane >> 5;

  for (int64_t s = (int64_t)blockIdx.x * wpb + wib; s < n_snps; s += (int64_t)gridDim.x * wpb) {
    const int pidx = MULTI ? snp_prob[s] : 0;
    const jm_lmm_problem pr = MULTI ? probs[pidx] : pr_one;
    const int mode = MULTI ? modes[pidx] : mode_one;
    const int nc = pr.n_cvt, T = pr.n_terms, n = pr.n, G = pr.n_grid;
    const int ni = (nc + 3) * (nc + 2) / 2;
    const bool do_reml = (mode == 1 || mode == 4);
    const bool do_mle = (mode == 2 || mode == 4);
    const bool do_score = (mode == 3 || mode == 4);
    const double df = (double)(n - nc - 1);
    const int nrows = nc + 2;  // per-SNP moment rows: W_1..W_c, y, xx
    const double* xm = MULTI ? snp_xmom[s] : xmom + s * (int64_t)nrows * ldm;
    int evals = 0;
    // ---- Iab (lambda = 0): row 0 = sum of the m = 0 terms ----
    {
      for (int j = lane; j < ni; j += 64) {
        const int r = pr.pair_row[j];
        const double* row = (r >= 0) ? xm + (int64_t)r * ldm : pr.shared_mom + (int64_t)j * T;
        double acc = 0.0;
        for (int t = 0; t < T; ++t)
          if (pr.term_m[t] == 0) acc += row[t];
        P[j * 64] = acc;
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      PabSummary oi;
      pab_recursion(P, 64, nc, oi, nullptr, iab_ld);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }

    // ---- grid stage: lane g evaluates REML and MLE at lambda_g ----
    double best_r = -INFINITY, best_m = -INFINITY;
    int bi_r = 0x7fffffff, bi_m = 0x7fffffff;
    if (do_reml || do_mle) {
      for (int g0 = 0; g0 < G; g0 += 64) {
        const int g = g0 + lane;
        if (g < G) {
          for (int j = 0; j < ni; ++j) {
            const int r = pr.pair_row[j];
            double acc;
            if (r >= 0) {
              const double* row = xm + (int64_t)r * ldm;
              acc = 0.0;
              for (int t = 0; t < T; ++t) acc += pr.grid_kappa[(int64_t)t * G + g] * row[t];
            } else {
              acc = pr.grid_shared[(int64_t)g * ni + j];
            }
            P[j * 64 + lane] = acc;
          }
          PabSummary o;
          pab_recursion(P + lane, 64, nc, o, iab_ld, nullptr);
          const double ldh = pr.grid_logdet[g];
          const double fr = reml_value(o, nc, n, ldh);
          const double fm = mle_value(o.pyy1, n, ldh);
          if (bi_r == 0x7fffffff || better(fr, g, best_r, bi_r)) {
            best_r = fr;
            bi_r = g;
          }
          if (bi_m == 0x7fffffff || better(fm, g, best_m, bi_m)) {
            best_m = fm;
            bi_m = g;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      evals += G;
      wave_argmax(best_r, bi_r);
      wave_argmax(best_m, bi_m);
    }

    // ---- golden-section refinement: REML in half 0, MLE in half 1 ----
    double my_lam = NAN, my_logl = NAN;
    PabSummary ofin;  // final Pab summary of this half (REML half -> Wald)
    ofin.pxx = ofin.pxy = ofin.pyy = ofin.pyy1 = ofin.hiw = NAN;
    if (do_reml || do_mle) {
      const int bi = h ? bi_m : bi_r;
      const int ilo = bi > 0 ? bi - 1 : 0;
      const int ihi = bi + 1 < G ? bi + 1 : G - 1;
      double a = pr.grid_loglam[ilo];
      double b = pr.grid_loglam[ihi];
      double c = b - phi * (b - a);
      double d = a + phi * (b - a);
      PabSummary o;
      auto eval = [&](double loglam_mine) -> double {
        const double lam_mine = exp(loglam_mine);
        const double lam_r = shfl_d(lam_mine, 0);
        const double lam_m = shfl_d(lam_mine, 32);
        const double ldh = eval_two(pr, nc, ni, T, lane, kap, P, lam_r, lam_m, xm, ldm, nullptr);
        pab_recursion(P + h, 64, nc, o, h ? nullptr : iab_ld, nullptr);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        return h ? mle_value(o.pyy1, n, ldh) : reml_value(o, nc, n, ldh);
      };
      double fc = eval(c);
      double fd = eval(d);
      for (int it = 0; it < pr.n_iter; ++it) {
        const bool keep_left = fc > fd;
        const double na = keep_left ? a : c;
        const double nb = keep_left ? d : b;
        const double ncp = nb - phi * (nb - na);
        const double ndp = na + phi * (nb - na);
        const double nl = eval(keep_left ? ncp : ndp);
        const double nfc = keep_left ? nl : fd;
        const double nfd = keep_left ? fc : nl;
        a = na;
        b = nb;
        c = ncp;
        d = ndp;
        fc = nfc;
        fd = nfd;
      }
      const double log_opt = (a + b) / 2.0;
      const double fin = eval(log_opt);
      ofin = o;
      my_lam = exp(log_opt);
      my_logl = fin;
      evals += 2 * (pr.n_iter + 3);
    }
    const double lam_mle = shfl_d(my_lam, 32);
    const double logl_mle = shfl_d(my_logl, 32);

    // ---- Score at the null lambda ----
    double sb = NAN, sse = NAN, ps = NAN;
    if (do_score) {
      eval_two(pr, nc, ni, T, lane, kap, P, pr.lambda_null, pr.lambda_null, xm, ldm,
               pr.null_kappa);
      PabSummary o;
      pab_recursion(P, 64, nc, o, nullptr, nullptr);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      __builtin_amdgcn_wave_barrier();
      const double pxyy = clamp_pyy(o.pyy1);
      double tau;
      beta_se(df, o.pxx, o.pxy, pxyy, sb, sse, tau);
      if (o.pxx > 0.0) {
        const double f = (double)n * (o.pxy * o.pxy) / (o.pyy * o.pxx);
        ps = f_pvalue(df, f);
      }
      evals += 1;
    }

    if (lane == 0) {
      double* os = out + s * JM_NOUT;
      os[JM_OUT_EVALS] = (double)evals;
      if (do_reml) {
        double beta, se, tau;
        const double pxyy = clamp_pyy(ofin.pyy1);
        beta_se(df, ofin.pxx, ofin.pxy, pxyy, beta, se, tau);
        double pw = NAN;
        if (ofin.pxx > 0.0) pw = f_pvalue(df, (ofin.pyy - pxyy) * tau);
        os[JM_OUT_BETA] = beta;
        os[JM_OUT_SE] = se;
        os[JM_OUT_P_WALD] = pw;
        os[JM_OUT_L_REMLE] = my_lam;
        os[JM_OUT_LOGL_H1] = my_logl;
      } else {
        os[JM_OUT_BETA] = os[JM_OUT_SE] = os[JM_OUT_P_WALD] = NAN;
        os[JM_OUT_L_REMLE] = os[JM_OUT_LOGL_H1] = NAN;
      }
      if (do_mle) {
        os[JM_OUT_L_MLE] = lam_mle;
        os[JM_OUT_LOGL_MLE] = logl_mle;
        os[JM_OUT_P_LRT] = chi2_sf1(fmax(2.0 * (logl_mle - pr.logl_null), 0.0));
      } else {
        os[JM_OUT_L_MLE] = os[JM_OUT_LOGL_MLE] = os[JM_OUT_P_LRT] = NAN;
      }
      os[JM_OUT_P_SCORE] = ps;
      os[JM_OUT_BETA_SCORE] = sb;
      os[JM_OUT_SE_SCORE] = sse;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace jm

using namespace jm;

static int check_problem(const jm_lmm_problem* pr) {
  JM_CHECK_ARG(pr != nullptr, "null problem");
  JM_CHECK_ARG(pr->n_cvt >= 1 && pr->n_cvt <= 20, "n_cvt out of range [1,20]: %d", pr->n_cvt);
  JM_CHECK_ARG(pr->n > pr->n_cvt + 1, "n=%d must exceed n_cvt+1", pr->n);
  JM_CHECK_ARG(pr->n_terms >= 1, "n_terms must be >= 1");
  JM_CHECK_ARG(pr->n_grid >= 2 && pr->n_grid <= 4096, "n_grid out of range");
  JM_CHECK_ARG(pr->n_iter >= 0 && pr->n_iter <= 1000, "n_iter out of range");
  JM_CHECK_ARG(pr->term_cluster && pr->term_m && pr->cl_center && pr->logdet_mom &&
                   pr->shared_mom && pr->pair_row && pr->grid_loglam,
               "missing problem array");
  return JM_OK;
}

static size_t lds_per_wave(const jm_lmm_problem* pr) {
  const int ni = (pr->n_cvt + 3) * (pr->n_cvt + 2) / 2;
  return (size_t)(2 * pr->n_terms + ni * 64 + kMaxLevels) * sizeof(double);
}

extern "C" int jm_lmm_prepare(const jm_lmm_problem* prob, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  JM_CHECK_ARG(prob->grid_kappa && prob->grid_shared && prob->grid_logdet && prob->null_kappa &&
                   prob->null_shared,
               "missing table array");
  hipLaunchKernelGGL(k_lmm_prepare, dim3(256), dim3(256), 0, (hipStream_t)stream, *prob);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_lmm_null(const jm_lmm_problem* prob, double* out, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  JM_CHECK_ARG(out != nullptr, "null out");
  const size_t lds = lds_per_wave(prob);
  JM_CHECK_ARG(lds <= 160 * 1024, "problem too large for the LDS budget (T=%d)", prob->n_terms);
  hipLaunchKernelGGL(k_lmm_null, dim3(1), dim3(64), lds, (hipStream_t)stream, *prob, out);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

static int launch_assoc(const jm_lmm_problem* one, int mode_one, const double* xmom, int64_t ldm,
                        int64_t n_snps, double* out, const jm_lmm_problem* probs,
                        const int32_t* modes, const int32_t* snp_prob,
                        const double* const* snp_xmom, int T, int max_cvt, void* stream) {
  const int ni_max = (max_cvt + 3) * (max_cvt + 2) / 2;
  const size_t per_wave = (size_t)(2 * T + ni_max * 64 + kMaxLevels) * sizeof(double);
  int wpb = (int)((64 * 1024) / per_wave);
  if (wpb > 4) wpb = 4;
  if (wpb < 1) {
    JM_CHECK_ARG(per_wave <= 160 * 1024, "problem too large for the LDS budget (T=%d)", T);
    wpb = 1;
  }
  int64_t blocks = (n_snps + wpb - 1) / wpb;
  const int64_t max_blocks = 256 * 16;
  if (blocks > max_blocks) blocks = max_blocks;
  jm_lmm_problem dummy;
  memset(&dummy, 0, sizeof(dummy));
  if (one)
    hipLaunchKernelGGL(k_lmm_assoc<false>, dim3((unsigned)blocks), dim3(64 * wpb), per_wave * wpb,
                       (hipStream_t)stream, *one, mode_one, xmom, ldm, n_snps, out, nullptr,
                       nullptr, nullptr, nullptr, T, ni_max);
  else
    hipLaunchKernelGGL(k_lmm_assoc<true>, dim3((unsigned)blocks), dim3(64 * wpb), per_wave * wpb,
                       (hipStream_t)stream, dummy, 0, nullptr, ldm, n_snps, out, probs, modes,
                       snp_prob, snp_xmom, T, ni_max);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_lmm_assoc(const jm_lmm_problem* prob, int mode, const double* xmom,
                            int64_t ldm, int64_t n_snps, double* out, void* stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  JM_CHECK_ARG(mode >= 1 && mode <= 4, "mode must be 1..4, got %d", mode);
  JM_CHECK_ARG(n_snps >= 0, "n_snps < 0");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(xmom && out, "null xmom/out");
  JM_CHECK_ARG(ldm >= prob->n_terms, "ldm < n_terms");
  JM_CHECK_ARG(prob->grid_kappa && prob->grid_shared && prob->grid_logdet && prob->null_kappa &&
                   prob->null_shared,
               "tables not prepared");
  return launch_assoc(prob, mode, xmom, ldm, n_snps, out, nullptr, nullptr, nullptr, nullptr,
                      prob->n_terms, prob->n_cvt, stream);
}

extern "C" int jm_lmm_assoc_multi(const jm_lmm_problem* probs, const int32_t* modes, int n_probs,
                                  int max_cvt, int n_terms, const int32_t* snp_prob,
                                  const double* const* snp_xmom, int64_t ldm, int64_t n_snps,
                                  double* out, void* stream) {
  JM_CHECK_ARG(n_snps >= 0 && n_probs >= 1, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(probs && modes && snp_prob && snp_xmom && out, "null pointer");
  JM_CHECK_ARG(max_cvt >= 1 && max_cvt <= 20, "max_cvt out of range [1,20]");
  JM_CHECK_ARG(n_terms >= 1 && ldm >= n_terms, "bad n_terms / ldm");
  return launch_assoc(nullptr, 0, nullptr, ldm, n_snps, out, probs, modes, snp_prob, snp_xmom,
                      n_terms, max_cvt, stream);
}

extern "C" double jm_host_betainc(double a, double b, double x) { return betainc_reg(a, b, x); }
extern "C" double jm_host_chi2_sf1(double x) { return chi2_sf1(x); }
