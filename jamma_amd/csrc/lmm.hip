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
#include <cstdlib>

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
__host__ __device__ __attribute__((noinline)) double betainc_reg(double a, double b, double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 0.0 : NAN;
  if (!(x < 1.0)) return (x == 1.0) ? 1.0 : NAN;
  const double lg = (a >= b) ? lgamma_diff(a, b) - lgamma(b) : lgamma_diff(b, a) - lgamma(a);
  const double lbt = lg + a * log(x) + b * log1p(-x);
  if (x < (a + 1.0) / (a + b + 2.0)) return exp(lbt) * betacf(a, b, x) / a;
  return 1.0 - exp(lbt) * betacf(b, a, 1.0 - x) / b;
}

// chi-square(1) survival: P(X > x) = erfc(sqrt(x / 2)).
__host__ __device__ __attribute__((noinline)) double chi2_sf1(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? 1.0 : NAN;
  return erfc(sqrt(0.5 * x));
}

// F(1, df) upper tail through betainc, exactly as likelihood_jax.py:668-673. A NaN
// statistic gives NaN, as jnp.maximum / jnp.clip propagate it (C's fmax / fmin would
// drop it and return p ~ 1).
__host__ __device__ static double f_pvalue(double df, double f_stat) {
  if (f_stat != f_stat) return NAN;  // NaN (host and device)
  double z = df / (df + fmax(f_stat, 1e-10));
  z = fmin(fmax(z, 0.0), 1.0);
  double p = betainc_reg(0.5 * df, 0.5, z);
  if (f_stat <= 0.0) p = 1.0;
  return p;
}

// LRT p-value, calc_lrt_pvalue_jax (likelihood_jax.py:231-251): chi2_1 tail of
// max(2 (logl_H1 - logl_H0), 0) with jnp.maximum's NaN propagation.
__host__ __device__ static double lrt_pvalue(double logl_h1, double logl_h0) {
  const double stat = 2.0 * (logl_h1 - logl_h0);
  return chi2_sf1(stat < 0.0 ? 0.0 : stat);
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
  // shared pair sums at lambda = 0 (Iab: kappa_t = [m_t == 0])
  for (int j = tid; j < ni; j += nthreads) {
    double acc = 0.0;
    if (pr.pair_row[j] < 0) {
      const double* row = pr.shared_mom + (int64_t)j * T;
      for (int t = 0; t < T; ++t)
        if (pr.term_m[t] == 0) acc += row[t];
    }
    pr.iab_shared[j] = acc;
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

// x (x + 1) (x + 2) / 6: the number of Pab recursion updates over the levels whose
// triangle sides are 1 .. x.
__host__ __device__ constexpr int tet(int x) { return x <= 0 ? 0 : x * (x + 1) * (x + 2) / 6; }

// Address-space views: constant (wave-uniform reads become scalar loads) and global
// (vector loads without the flat-address path).
typedef const double __attribute__((address_space(4))) cdbl;
typedef const int32_t __attribute__((address_space(4))) cint;
typedef const double __attribute__((address_space(1))) gdbl;

// KT (terms per lane in one block of terms) the launcher picks for T terms.
static int assoc_kt(int T) {
  int kt = (T + 63) / 64;
  if (kt > 8) kt = 8;
  if (kt == 5) kt = 6;
  if (kt == 7) kt = 8;
  return kt;
}

// k_lmm_assoc block-level LDS ahead of the shared rows (bytes): 64*KT terms x (2 doubles
// + 1 int), 32 clusters x 2 doubles and the recursion table of the widest problem.
static size_t assoc_lds_consts(int T, int max_cvt) {
  const int tb0 = 64 * assoc_kt(T);
  return (size_t)(2 * tb0 + 64 + tb0 / 2 + (tet(max_cvt + 1) + 1) / 2) * sizeof(double);
}

// k_lmm_assoc LDS per wave (doubles): two packed rows (ni + logdet slot each), the
// Iab level logs, and ni + 1 int slots.
__host__ __device__ constexpr int assoc_lds_per_wave(int ni_max) {
  return 2 * (ni_max + 1) + kMaxLevels + (ni_max + 2) / 2;
}
constexpr int kXR = 8;          // per-SNP moment rows held in registers

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ const double* uniform_ptr(const double* p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (const double*)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double bcast(double v, int src) {
  const uint64_t u = __double_as_longlong(v);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, src);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), src);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Value of the butterfly partner for lane bit BIT, without going through
// LDS: bits 0/1 quad_perm DPP (lane ^ 1, lane ^ 2), bit 2 row_half_mirror
// (i <-> 7 - i), bit 3 row_mirror (i <-> 15 - i), bits 4/5 the gfx950
// permlane16 / permlane32 swaps (lane ^ 16, lane ^ 32). Every partner differs
// from the lane in bit BIT and agrees on the higher bits, which is all a
// butterfly (and the index bookkeeping of the reduce-scatter) needs.
template <int BIT>
__device__ __forceinline__ uint32_t xpartner32(uint32_t v, int lane) {
  if constexpr (BIT == 0) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
  if constexpr (BIT == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
  if constexpr (BIT == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xF, 0xF, false);
  if constexpr (BIT == 3) return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xF, 0xF, false);
  if constexpr (BIT == 4) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? r[0] : r[1];
  }
  if constexpr (BIT == 5) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
  return v;
}

template <int BIT>
__device__ __forceinline__ double xpartner(double v, int lane) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = xpartner32<BIT>((uint32_t)u, lane);
  const uint32_t hi = xpartner32<BIT>((uint32_t)(u >> 32), lane);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Butterfly reduce-scatter of 8 per-lane values: on return every lane holds
// the full wave sum of value q = 4*b5 + 2*b4 + b3 (10 exchanges).
__device__ __forceinline__ double reduce8(double (&v)[8], int lane) {
  {
    const bool up = lane & 32;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double send = up ? v[i] : v[i + 4];
      const double keep = up ? v[i + 4] : v[i];
      v[i] = keep + xpartner<5>(send, lane);
    }
  }
  {
    const bool up = lane & 16;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double send = up ? v[i] : v[i + 2];
      const double keep = up ? v[i + 2] : v[i];
      v[i] = keep + xpartner<4>(send, lane);
    }
  }
  {
    const bool up = lane & 8;
    const double send = up ? v[0] : v[1];
    const double keep = up ? v[1] : v[0];
    v[0] = keep + xpartner<3>(send, lane);
  }
  v[0] = v[0] + xpartner<2>(v[0], lane);
  v[0] = v[0] + xpartner<1>(v[0], lane);
  v[0] = v[0] + xpartner<0>(v[0], lane);
  return v[0];
}

// Sum over the 32 lanes of a half-wave (both halves at once).
__device__ __forceinline__ double half_sum(double v, int lane) {
  v += xpartner<4>(v, lane);
  v += xpartner<3>(v, lane);
  v += xpartner<2>(v, lane);
  v += xpartner<1>(v, lane);
  v += xpartner<0>(v, lane);
  return v;
}

// x^m for 0 <= m < 16 by binary powering.
__device__ __forceinline__ double ipow16(double x, int m) {
  double r = 1.0, b = x;
  r = (m & 1) ? r * b : r;
  b = b * b;
  r = (m & 2) ? r * b : r;
  b = b * b;
  r = (m & 4) ? r * b : r;
  b = b * b;
  r = (m & 8) ? r * b : r;
  return r;
}

// Packed index of the pair (genotype, partner) for per-SNP moment row r
// (rows: W_1..W_c, y, x): W_b -> (b, c+1), y -> (c+1, c+2), x -> (c+1, c+1).
__device__ __forceinline__ int xpair_index(int r, int nc) {
  if (r < nc) return ab_index(r + 1, nc + 1, nc);
  if (r == nc) return ab_index(nc + 1, nc + 2, nc);
  return ab_index(nc + 1, nc + 1, nc);
}

// ---------------------------------------------------------------------------
// Grid stage of the REML and MLE searches
// ---------------------------------------------------------------------------
// _batch_grid_reml_with_iab / _batch_grid_mle (likelihood_jax.py:565-598,
// 801-827) and the argmax that brackets the golden section (:539-547, NaN
// first, first index on ties), one wavefront per SNP. Column k of a 64-column
// pass is lane k: column 0 is Iab (lambda = 0, kappa_t = [m_t == 0]), column
// k >= 1 grid point g = k - 1. A lane walks the T terms once per 8 moment rows:
// kappa at its grid point (a reciprocal per cluster, a multiply per term) times
// the SNP's moments (wave-uniform scalar loads). Row 0 of its Pab column is
// completed with the shared pairs (table) in LDS and the recursion runs per
// lane. The REML log-determinant term subtracts the Iab level logs in the
// reference's interleaved order once column 0 has them. Output: the argmax
// grid indices of REML and MLE, stored as doubles in the L_REMLE / L_MLE
// columns of the SNP's output row (k_lmm_assoc reads and replaces them).
template <bool MULTI>
__global__ void __launch_bounds__(256) k_lmm_grid(jm_lmm_problem pr_one, int mode_one,
                                                  const double* xmom, int64_t ldm, int64_t n_snps,
                                                  double* out, const jm_lmm_problem* probs,
                                                  const int32_t* modes, const int32_t* snp_prob,
                                                  const double* const* snp_xmom, int ni_max,
                                                  int ps) {
  // ps: LDS stride of the Pab columns = the lanes that hold one (G + 1 when the grid fits
  // one pass, else 64); the lanes past it compute nothing that is kept
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const bool col = lane < ps;
  double* P = smem + (size_t)wib * (ni_max * ps + kMaxLevels);  // column per lane
  double* iab = P + ni_max * ps;                                  // Iab level logs
  for (int64_t s = (int64_t)blockIdx.x * wpb + wib; s < n_snps; s += (int64_t)gridDim.x * wpb) {
    const int pidx = MULTI ? snp_prob[s] : 0;
    const int mode = MULTI ? modes[pidx] : mode_one;
    if (mode == 3) continue;  // Score only: no search
    const jm_lmm_problem& pr = MULTI ? probs[pidx] : pr_one;
    const int nc = pr.n_cvt, T = pr.n_terms, n = pr.n, G = pr.n_grid;
    const int ni = (nc + 3) * (nc + 2) / 2;
    const int nrows = nc + 2;
    // the SNP's moments and the term data are wave-uniform: constant address space,
    // so they come through scalar loads
    const cdbl* xm = (const cdbl*)uniform_ptr(MULTI ? snp_xmom[s]
                                                   : xmom + s * (int64_t)nrows * ldm);
    const cint* term_m = (const cint*)pr.term_m;
    const cint* term_cl = (const cint*)pr.term_cluster;
    const cdbl* cl_center = (const cdbl*)pr.cl_center;
    double best_r = -INFINITY, best_m = -INFINITY;
    int bi_r = 0x7fffffff, bi_m = 0x7fffffff;
    for (int k0 = 0; k0 <= G; k0 += 64) {
      const int k = k0 + lane;
      const bool act = k <= G;
      const bool is_iab = k == 0;
      const int g = act && !is_iab ? k - 1 : 0;
      // the lane's lambda: exactly the value the grid tables were built with
      const double lam_g = is_iab ? 0.0 : exp(pr.grid_loglam[g]);
      for (int r0 = 0; r0 < nrows; r0 += 8) {
        // rows past nrows repeat the last row (no branches in the term loop)
        int64_t roff[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) roff[r] = (int64_t)min(r0 + r, nrows - 1) * ldm;
        double acc[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc[r] = 0.0;
        // kappa of terms t .. t+3 at the lane's lambda, walking the terms in order: a
        // cluster's terms are contiguous with m = 0, 1, 2, ..., so h_c = 1/(1 + lambda d_c)
        // is formed once per cluster (a uniform branch) and (lambda h_c)^m by one multiply
        // per term, the operations term_coeffs uses for the prepared tables (lambda = 0
        // gives the Iab column's [m == 0], up to the sign of zeros)
        double ch = 1.0, cz = 0.0, czm = 1.0;
        auto kap4 = [&](int t0, double (&kv)[4]) {
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int t = t0 + u;
            double kt = 0.0;
            if (t < T) {
              const int m = term_m[t];
              if (m == 0) {
                ch = 1.0 / fma(lam_g, cl_center[term_cl[t]], 1.0);
                cz = lam_g * ch;
                czm = 1.0;
              } else {
                czm *= cz;
              }
              kt = (m & 1) ? -ch * czm : ch * czm;
            }
            kv[u] = kt;
          }
        };
        double kc[4], kn[4];
        kap4(0, kn);
        const int T4 = T & ~3;
        for (int t0 = 0; t0 < T4; t0 += 4) {  // contiguous moment reads: scalar x8 loads
#pragma unroll
          for (int u = 0; u < 4; ++u) kc[u] = kn[u];
          kap4(t0 + 4, kn);
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int r = 0; r < 8; ++r) acc[r] = fma(kc[u], xm[roff[r] + t0 + u], acc[r]);
        }
        for (int u = 0; u < T - T4; ++u)  // kn holds terms T4 .. T - 1
#pragma unroll
          for (int r = 0; r < 8; ++r) acc[r] = fma(kn[u], xm[roff[r] + T4 + u], acc[r]);
#pragma unroll
        for (int r = 0; r < 8; ++r)
          if (col && r0 + r < nrows) P[xpair_index(r0 + r, nc) * ps + lane] = acc[r];
      }
      if (col)
        for (int j = 0; j < ni; ++j)
          if (pr.pair_row[j] < 0)
            P[j * ps + lane] = is_iab ? pr.iab_shared[j] : pr.grid_shared[(int64_t)g * ni + j];
      wave_sync();
      PabSummary o;
      if (col) pab_recursion(P + lane, ps, nc, o, nullptr, nullptr);
      // the level pivots stay on the column's diagonal: (p+1, p+1) is final after level p
      if (k0 == 0 && lane == 0) {
        for (int p = 0; p <= nc; ++p) {
          const double v = P[ab_index(p + 1, p + 1, nc) * ps];
          iab[p] = (v > 0.0) ? log(v) : 0.0;
        }
      }
      wave_sync();
      if (act && !is_iab) {
        double hiw = 0.0;  // likelihood_jax.py:322-327 order
        for (int p = 0; p <= nc; ++p) {
          const double v = P[ab_index(p + 1, p + 1, nc) * ps + lane];
          hiw = hiw + ((v > 0.0) ? log(v) : 0.0);
          hiw = hiw - iab[p];
        }
        o.hiw = hiw;
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
      wave_sync();  // the next pass reuses the columns
    }
    wave_argmax(best_r, bi_r);
    wave_argmax(best_m, bi_m);
    if (lane == 0) {
      out[s * JM_NOUT + JM_OUT_L_REMLE] = (double)bi_r;
      out[s * JM_NOUT + JM_OUT_L_MLE] = (double)bi_m;
    }
  }
}

// Packed index of the k-th pair (a <= b) over the covariate / phenotype indices
// {1..c, c+2} (the pairs that do not involve the genotype), in packed order.
__device__ __forceinline__ int shared_pair_index(int k, int nc) {
  int a = 1;
  for (;; ++a) {
    if (a == nc + 1) continue;
    const int len = (nc + 2 - a) + 1 - (a <= nc ? 1 : 0);
    if (k < len) break;
    k -= len;
  }
  int b = a;
  for (;; ++b) {
    if (b == nc + 1) continue;
    if (k == 0) break;
    --k;
  }
  return ab_index(a, b, nc);
}

// Per-lane term data: lane l owns the KT contiguous terms l*KT .. l*KT+KT-1.
template <int KT>
struct Terms {
  double dc[KT], L[KT], inv_m[KT];
  int m[KT];
  bool act[KT];
};

// Main association kernel: one wavefront per SNP, grid-stride over SNPs.
// Per-wave LDS: P (ni_max x 64 doubles; the grid stage's per-lane columns,
// reused by the golden stage as two packed Pab rows) | iab logs | slot table.
// MULTI: SNP s belongs to problem snp_prob[s] (own phenotype, covariates,
// null model and mode) and its moments start at snp_xmom[s]; all problems of
// one launch share the cohort (same terms). Otherwise every SNP uses pr_one.
//
// An evaluation of REML at lambda_A and MLE at lambda_B (golden stage):
//   1. lanes over terms: kappa_t for both lambdas and the m >= 1 part of
//      logdet_h (lane l owns the KT contiguous terms l*KT .. of each block of
//      64*KT terms; T > 64*KT runs several blocks, kappa of blocks after the
//      first is recomputed per pair chunk);
//   2. dot products with the SNP's moment rows (XR of them in registers, the
//      rest and the shared rows from memory), reduced 8 values (4 pairs x 2
//      lambdas) at a time by reduce8 into two packed Pab rows in LDS;
//   3. the Pab recursion, half-wave A on the REML row and half-wave B on the
//      MLE row, parallel over the pairs of each level;
//   4. one round of logs (level pivots, P_yy, log|1 + lambda d_c| per cluster)
//      spread over the lanes, then a half-wave sum.
// Stages: 0 = Iab (lambda = 0, kappa_t = [m_t == 0]) followed by the grid;
// 1, 2 = first golden probes; then n_iter golden steps; the final evaluation
// at the bracket midpoints (Wald / LRT); the Score evaluation at lambda_null.
template <int KT, int XR, bool MB, bool MULTI>
__global__ void __launch_bounds__(256) k_lmm_assoc(jm_lmm_problem pr_one, int mode_one,
                                                   const double* xmom, int64_t ldm, int64_t n_snps,
                                                   double* out, const jm_lmm_problem* probs,
                                                   const int32_t* modes, const int32_t* snp_prob,
                                                   const double* const* snp_xmom, int ni_max,
                                                   int srow_n) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int per_wave = assoc_lds_per_wave(ni_max);
  // srow_n > 0 (single problem): the block keeps the srow_n shared pair rows
  // (covariate / phenotype moments) in LDS, in slot order, ahead of the waves
  // Cohort constants, staged once per block (all SNPs of a launch share the cohort):
  // for the first 64*KT terms the cluster centre, logdet weight / power and power, for
  // the first 32 clusters the centre and weight. Reading them from global memory in
  // every evaluation (a dependent chain term -> cluster -> centre) took most of its time.
  constexpr int TB0 = 64 * KT;
  double* c_tdc = smem;                        // TB0
  double* c_twl = c_tdc + TB0;                 // TB0
  double* c_cdc = c_twl + TB0;                 // 32
  double* c_cw = c_cdc + 32;                   // 32
  int* c_tm = (int*)(c_cw + 32);               // TB0 ints
  // single problem: per recursion level p the packed LDS offsets (ab | ap << 8 | bp << 16)
  // of the level's updates, level p starting at tet(nc + 1) - tet(nc + 2 - p)
  int* rtab = c_tm + TB0;                      // tet(ni_lv) ints (+1 pad)
  const int rtab_n = MULTI ? 0 : tet(pr_one.n_cvt + 1);
  double* srows = c_cw + 32 + TB0 / 2 + (rtab_n + 1) / 2;
  // shared rows: zero-padded to 64*KT terms (one block of terms) so that a lane reads its
  // KT values without bounds checks; full T when T > 64*KT
  const int sstride = MB ? pr_one.n_terms : TB0;
  const size_t srow_len = MULTI ? 0 : (size_t)srow_n * sstride;
  {
    const int64_t s0 = (int64_t)blockIdx.x * wpb;
    const jm_lmm_problem& p0 = MULTI ? probs[s0 < n_snps ? snp_prob[s0] : 0] : pr_one;
    for (int t = threadIdx.x; t < TB0; t += blockDim.x) {
      const int tc = t < p0.n_terms ? t : 0;
      const int m = t < p0.n_terms ? p0.term_m[tc] : 0;
      c_tm[t] = m;
      c_tdc[t] = p0.cl_center[p0.term_cluster[tc]];
      c_twl[t] = m > 0 ? p0.logdet_mom[tc] / (double)m : 0.0;
    }
    for (int c = threadIdx.x; c < 32; c += blockDim.x) {
      c_cdc[c] = c < p0.n_clusters ? p0.cl_center[c] : 0.0;
      c_cw[c] = c < p0.n_clusters ? p0.cl_weight[c] : 0.0;
    }
    __syncthreads();
  }
  if (!MULTI) {
    const int T1 = pr_one.n_terms, nc1 = pr_one.n_cvt, m = nc1 + 2;
    for (int k = 0; k < srow_n; ++k) {
      const int j = shared_pair_index(k, nc1);
      for (int t = threadIdx.x; t < sstride; t += blockDim.x)
        srows[(size_t)k * sstride + t] = t < T1 ? pr_one.shared_mom[(int64_t)j * T1 + t] : 0.0;
    }
    for (int e = threadIdx.x; e < rtab_n; e += blockDim.x) {
      int p = 1;
      while (tet(m - 1) - tet(m - p - 1) <= e) ++p;  // level of entry e
      int rem = e - (tet(m - 1) - tet(m - p)), ap = 0;
      const int Lw = m - p;
      while (rem >= Lw - ap) {
        rem -= Lw - ap;
        ++ap;
      }
      const int a = p + 1 + ap, b = a + rem;
      rtab[e] = ab_index(a, b, nc1) | (ab_index(a, p, nc1) << 8) | (ab_index(b, p, nc1) << 16);
    }
    __syncthreads();
  }
  double* P = srows + srow_len + (size_t)wib * per_wave;  // two packed rows
  double* PA = P;                              // row A: ni entries + logdet slot
  double* PB = P + ni_max + 1;                 // row B
  double* iab_ld = P + 2 * (ni_max + 1);      // kMaxLevels
  int* slot = (int*)(iab_ld + kMaxLevels);    // memory-phase slots (<= ni_max + 1 ints)
  const double phi = 0.6180339887498949;
  const int h = lane >> 5;  // half-wave: 0 -> REML / lambda A, 1 -> MLE / lambda B
  const int q = lane & 31;
  for (int64_t s = (int64_t)blockIdx.x * wpb + wib; s < n_snps; s += (int64_t)gridDim.x * wpb) {
    const int pidx = MULTI ? snp_prob[s] : 0;
    const jm_lmm_problem pr = MULTI ? probs[pidx] : pr_one;
    const int mode = MULTI ? modes[pidx] : mode_one;
    const int nc = pr.n_cvt, T = pr.n_terms, n = pr.n, G = pr.n_grid, C = pr.n_clusters;
    const int ni = (nc + 3) * (nc + 2) / 2;
    const bool do_reml = (mode == 1 || mode == 4);
    const bool do_mle = (mode == 2 || mode == 4);
    const bool do_score = (mode == 3 || mode == 4);
    const bool do_golden = do_reml || do_mle;
    const double df = (double)(n - nc - 1);
    const double nn = (double)n;
    const double cr = 0.5 * df * (log(df) - log(2.0 * M_PI) - 1.0);
    const double cm = 0.5 * nn * (log(nn) - log(2.0 * M_PI) - 1.0);
    const int nrows = nc + 2;  // per-SNP moment rows: W_1..W_c, y, xx
    const double* xm = uniform_ptr(MULTI ? snp_xmom[s] : xmom + s * (int64_t)nrows * ldm);

    // ---- the SNP's register rows (lane l owns terms l*KT .. l*KT+KT-1) ----
    double xr[XR][KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) {
      const int t = lane * KT + k;
#pragma unroll
      for (int r = 0; r < XR; ++r) xr[r][k] = (r < nrows && t < T) ? xm[(int64_t)r * ldm + t] : 0.0;
    }
    // memory-phase slots: genotype rows >= XR, then the shared pairs
    const int nxm = nrows > XR ? nrows - XR : 0;
    const int n_mem = nxm + (nc + 1) * (nc + 2) / 2;
    for (int u = lane; u < n_mem; u += 64) {
      int j;
      if (u < nxm) {
        j = xpair_index(XR + u, nc);
      } else {
        j = shared_pair_index(u - nxm, nc);
      }
      slot[u] = j;
    }
    wave_sync();

    double* os = out + s * JM_NOUT;
    // grid argmax indices left by k_lmm_grid, then columns a mode does not compute are NaN
    const int gi_r = do_golden ? (int)os[JM_OUT_L_REMLE] : 0;
    const int gi_m = do_golden ? (int)os[JM_OUT_L_MLE] : 0;
    wave_sync();
    if (lane < JM_NOUT) os[lane] = NAN;
    int evals = 0;
    int bi_r = 0, bi_m = 0;
    double aA = 0, bA = 0, cA = 0, dA = 0, fcA = 0, fdA = 0;
    double aB = 0, bB = 0, cB = 0, dB = 0, fcB = 0, fdB = 0;
    bool klA = false, klB = false;
    double lam_reml = NAN, lam_mle = NAN;
    const int st_final = 3 + pr.n_iter, st_score = 4 + pr.n_iter;

    for (int stage = 0;;) {
      // ---- the lambdas of this stage ----
      const bool lam0 = stage == 0;
      double lamA = 0.0, lamB = 0.0;
      // exponents of the two lambdas; half-wave h takes the exponential of its own, so a
      // stage waits for one exp instead of two in a row
      double xA = 0.0, xB = 0.0;
      bool take_exp = true;
      if (stage == 1) {
        xA = cA;
        xB = cB;
      } else if (stage == 2) {
        xA = dA;
        xB = dB;
      } else if (stage >= 3 && stage < st_final) {
        klA = fcA > fdA;
        klB = fcB > fdB;
        const double naA = klA ? aA : cA, nbA = klA ? dA : bA;
        const double naB = klB ? aB : cB, nbB = klB ? dB : bB;
        const double ncA = nbA - phi * (nbA - naA), ndA = naA + phi * (nbA - naA);
        const double ncB = nbB - phi * (nbB - naB), ndB = naB + phi * (nbB - naB);
        aA = naA; bA = nbA; cA = ncA; dA = ndA;
        aB = naB; bB = nbB; cB = ncB; dB = ndB;
        xA = klA ? cA : dA;
        xB = klB ? cB : dB;
      } else if (stage == st_final) {
        xA = (aA + bA) / 2.0;
        xB = (aB + bB) / 2.0;
      } else {
        take_exp = false;
        if (stage == st_score) lamA = lamB = pr.lambda_null;
      }
      if (take_exp) {
        const double e = exp(h ? xB : xA);
        lamA = bcast(e, 0);
        lamB = bcast(e, 32);
        if (stage == st_final) {
          lam_reml = lamA;
          lam_mle = lamB;
        }
      }

      // ---- 1. kappa and the m >= 1 logdet part, per block of 64*KT terms ----
      const int TB = 64 * KT;
      const int nblk = MB ? (T + TB - 1) / TB : 1;  // MB: T > 64 * KT
      auto kappa_block = [&](int tb, double (&kA)[KT], double (&kB)[KT], double& ldA,
                             double& ldB) {
        double pA = 1.0, pB = 1.0, hA = 0.0, hB = 0.0, nzA = 0.0, nzB = 0.0;
        ldA = ldB = 0.0;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
          const int t = tb * TB + lane * KT + k;
          const bool act = t < T;
          const int tc = act ? t : 0;
          const int mk = tb == 0 ? c_tm[lane * KT + k] : (act ? pr.term_m[tc] : 0);
          if (k == 0 || mk == 0) {
            const double dck =
                tb == 0 ? c_tdc[lane * KT + k] : pr.cl_center[pr.term_cluster[tc]];
            hA = 1.0 / fma(lamA, dck, 1.0);
            hB = 1.0 / fma(lamB, dck, 1.0);
            nzA = -lamA * hA;
            nzB = -lamB * hB;
          }
          if (k == 0) {
            pA = ipow16(nzA, mk);
            pB = ipow16(nzB, mk);
          } else {
            pA = mk == 0 ? 1.0 : pA * nzA;
            pB = mk == 0 ? 1.0 : pB * nzB;
          }
          if (lam0) {
            kA[k] = kB[k] = (act && mk == 0) ? 1.0 : 0.0;
          } else {
            kA[k] = act ? hA * pA : 0.0;
            kB[k] = act ? hB * pB : 0.0;
            if (mk > 0) {
              const double wl = tb == 0 ? c_twl[lane * KT + k] : pr.logdet_mom[tc] / (double)mk;
              ldA -= pA * wl;
              ldB -= pB * wl;
            }
          }
        }
      };
      double kA0[KT], kB0[KT], ldA0, ldB0;
      kappa_block(0, kA0, kB0, ldA0, ldB0);
      // ---- 2. dot products, reduced 4 pairs at a time into PA / PB ----
      const int qv = ((lane >> 5) & 1) * 4 + ((lane >> 4) & 1) * 2 + ((lane >> 3) & 1);
#pragma unroll
      for (int e0 = 0; e0 < XR; e0 += 4) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.0;
        for (int tb = 0; tb < nblk; ++tb) {
          double kA[KT], kB[KT], la, lb;
          if (tb == 0) {
#pragma unroll
            for (int k = 0; k < KT; ++k) {
              kA[k] = kA0[k];
              kB[k] = kB0[k];
            }
          } else {
            kappa_block(tb, kA, kB, la, lb);
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = e0 + e;
            if (r < XR) {
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const int t = tb * TB + lane * KT + k;
                const double x = (tb == 0) ? xr[r < XR ? r : 0][k]
                                           : ((r < nrows && t < T) ? xm[(int64_t)r * ldm + t] : 0.0);
                v[2 * e] = fma(kA[k], x, v[2 * e]);
                v[2 * e + 1] = fma(kB[k], x, v[2 * e + 1]);
              }
            }
          }
        }
        const double r = reduce8(v, lane);
        const int e = e0 + (qv >> 1);
        if ((lane & 7) == 0 && e < nrows && e < XR) ((qv & 1) ? PB : PA)[xpair_index(e, nc)] = r;
      }
      if (!MULTI && !MB && srow_len && nxm == 0) {
        // single problem, one block of terms, every shared row in LDS (zero-padded)
        for (int u0 = 0; u0 <= n_mem; u0 += 4) {
          double v[8];
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int u = u0 + e;
            v[2 * e] = 0.0;
            v[2 * e + 1] = 0.0;
            if (u < n_mem) {
              const double* row = srows + (size_t)u * TB0 + lane * KT;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const double x = row[k];
                v[2 * e] = fma(kA0[k], x, v[2 * e]);
                v[2 * e + 1] = fma(kB0[k], x, v[2 * e + 1]);
              }
            } else if (u == n_mem) {
              v[2 * e] = ldA0;
              v[2 * e + 1] = ldB0;
            }
          }
          const double r = reduce8(v, lane);
          const int u = u0 + (qv >> 1);
          if ((lane & 7) == 0 && u <= n_mem) {
            double* dst = (qv & 1) ? PB : PA;
            dst[u < n_mem ? slot[u] : ni] = r;  // slot ni: m >= 1 part of logdet_h
          }
        }
      } else
      for (int u0 = 0; u0 <= n_mem; u0 += 4) {
        double v[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = 0.0;
        for (int tb = 0; tb < nblk; ++tb) {
          double kA[KT], kB[KT], la, lb;
          if (tb == 0) {
#pragma unroll
            for (int k = 0; k < KT; ++k) {
              kA[k] = kA0[k];
              kB[k] = kB0[k];
            }
            la = ldA0;
            lb = ldB0;
          } else {
            kappa_block(tb, kA, kB, la, lb);
          }
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int u = u0 + e;
            if (u < n_mem && u >= nxm && srow_len) {  // shared row from LDS
              const double* row = srows + (size_t)(u - nxm) * sstride;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const int t = tb * TB + lane * KT + k;
                const double x = t < T ? row[t] : 0.0;
                v[2 * e] = fma(kA[k], x, v[2 * e]);
                v[2 * e + 1] = fma(kB[k], x, v[2 * e + 1]);
              }
            } else if (u < n_mem) {
              const double* row = (u < nxm) ? xm + (int64_t)(XR + u) * ldm
                                            : pr.shared_mom + (int64_t)slot[u] * T;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const int t = tb * TB + lane * KT + k;
                const double x = t < T ? row[t] : 0.0;
                v[2 * e] = fma(kA[k], x, v[2 * e]);
                v[2 * e + 1] = fma(kB[k], x, v[2 * e + 1]);
              }
            } else if (u == n_mem) {
              v[2 * e] += la;
              v[2 * e + 1] += lb;
            }
          }
        }
        const double r = reduce8(v, lane);
        const int u = u0 + (qv >> 1);
        if ((lane & 7) == 0 && u <= n_mem) {
          double* dst = (qv & 1) ? PB : PA;
          dst[u < n_mem ? slot[u] : ni] = r;  // slot ni: m >= 1 part of logdet_h
        }
      }
      wave_sync();
      // ---- 3. Pab recursion, half A on PA, half B on PB ----
      double* Ph = h ? PB : PA;
      double piv = 0.0, pyy_lvl = 0.0;
      for (int p = 1; p <= nc + 1; ++p) {
        const double d = Ph[ab_index(p, p, nc)];
        if (q == p - 1) piv = d;
        if (p == nc + 1) pyy_lvl = Ph[ab_index(nc + 2, nc + 2, nc)];
        const double inv = (d != 0.0) ? 1.0 / d : 0.0;
        const int Lw = nc + 2 - p;
        const int np = Lw * (Lw + 1) / 2;
        const int lv0 = MULTI ? 0 : tet(nc + 1) - tet(nc + 2 - p);  // this level's entries
        for (int idx = q; idx < np; idx += 32) {
          int iab, ia, ib;
          if (!MULTI) {
            const int e = rtab[lv0 + idx];
            iab = e & 255;
            ia = (e >> 8) & 255;
            ib = e >> 16;
          } else {
            int ap = 0, rem = idx;
            while (rem >= Lw - ap) {
              rem -= Lw - ap;
              ++ap;
            }
            const int a = p + 1 + ap, b = a + rem;
            iab = ab_index(a, b, nc);
            ia = ab_index(a, p, nc);
            ib = ab_index(b, p, nc);
          }
          const double pa = Ph[ia], pb = Ph[ib];
          Ph[iab] = Ph[iab] - pa * pb * inv;
        }
        wave_sync();
      }
      // ---- 4. logs: pivots (A), P_yy (A: lane c+1, B: lane 32), clusters ----
      const double pyy1 = Ph[ab_index(nc + 2, nc + 2, nc)];
      const double lam_h = h ? lamB : lamA;
      const int first = h ? 1 : nc + 2;  // first cluster lane of this half
      const int per = 32 - first;
      const int n_cl = lam0 ? 0 : C;
      double contrib = 0.0, lgp = 0.0;
      for (int c0 = 0;; c0 += per) {
        double arg = 1.0, w = 0.0;
        int kind = 0;  // 1: pivot, 2: P_yy, 3: cluster
        if (c0 == 0 && h == 0 && q <= nc) {
          kind = 1;
          arg = piv > 0.0 ? piv : 1.0;
        } else if (c0 == 0 && q == first - 1) {
          kind = 2;
          arg = clamp_pyy(pyy1);
        } else if (q >= first && c0 + q - first < n_cl) {
          const int c = c0 + q - first;
          kind = 3;
          arg = fabs(fma(lam_h, c < 32 ? c_cdc[c] : pr.cl_center[c], 1.0));
          w = c < 32 ? c_cw[c] : pr.cl_weight[c];
        }
        const double lg = log(arg);
        if (kind == 1) contrib += lam0 ? lg : lg - iab_ld[q];
        if (kind == 1 && lam0) iab_ld[q] = lg;
        if (kind == 2) lgp = lg;
        if (kind == 3) contrib += w * lg;
        if (c0 + per >= n_cl) break;
      }
      wave_sync();
      const double S = half_sum(contrib, lane);
      const double fA = cr - 0.5 * (bcast(S, 0) + PA[ni]) - 0.5 * df * bcast(lgp, nc + 1);
      const double fB = cm - 0.5 * (bcast(S, 32) + PB[ni]) - 0.5 * nn * bcast(lgp, 32);
      const double pyyA = bcast(pyy_lvl, 0);

      // ---- consume the stage ----
      if (stage == 0) {
        // Iab done; the grid stage ran in k_lmm_grid
        if (do_golden) {
          bi_r = gi_r;
          bi_m = gi_m;
          evals += G;
          const int iloA = bi_r > 0 ? bi_r - 1 : 0, ihiA = bi_r + 1 < G ? bi_r + 1 : G - 1;
          const int iloB = bi_m > 0 ? bi_m - 1 : 0, ihiB = bi_m + 1 < G ? bi_m + 1 : G - 1;
          aA = pr.grid_loglam[iloA];
          bA = pr.grid_loglam[ihiA];
          aB = pr.grid_loglam[iloB];
          bB = pr.grid_loglam[ihiB];
          cA = bA - phi * (bA - aA);
          dA = aA + phi * (bA - aA);
          cB = bB - phi * (bB - aB);
          dB = aB + phi * (bB - aB);
          stage = 1;
        } else {
          stage = st_score;
        }
      } else if (stage == 1) {
        fcA = fA;
        fcB = fB;
        stage = 2;
      } else if (stage == 2) {
        fdA = fA;
        fdB = fB;
        stage = 3;
      } else if (stage < st_final) {
        const double nfcA = klA ? fA : fdA, nfdA = klA ? fcA : fA;
        const double nfcB = klB ? fB : fdB, nfdB = klB ? fcB : fB;
        fcA = nfcA;
        fdA = nfdA;
        fcB = nfcB;
        fdB = nfdB;
        ++stage;
      } else if (stage == st_final) {
        // Wald at the REML optimum (likelihood_jax.py:601-675), LRT at the MLE optimum
        evals += 2 * (pr.n_iter + 3);
        if (lane == 0 && do_reml) {
          const double pxx = PA[ab_index(nc + 1, nc + 1, nc)];
          const double pxy = PA[ab_index(nc + 1, nc + 2, nc)];
          const double pxyy = clamp_pyy(PA[ab_index(nc + 2, nc + 2, nc)]);
          double beta, se, tau;
          beta_se(df, pxx, pxy, pxyy, beta, se, tau);
          os[JM_OUT_BETA] = beta;
          os[JM_OUT_SE] = se;
          os[JM_OUT_P_WALD] = (pxx > 0.0) ? f_pvalue(df, (pyyA - pxyy) * tau) : NAN;
          os[JM_OUT_L_REMLE] = lam_reml;
          os[JM_OUT_LOGL_H1] = fA;
        }
        if (lane == 0 && do_mle) {
          os[JM_OUT_L_MLE] = lam_mle;
          os[JM_OUT_LOGL_MLE] = fB;
          os[JM_OUT_P_LRT] = lrt_pvalue(fB, pr.logl_null);
        }
        if (!do_score) break;
        stage = st_score;
      } else {  // score at the null lambda (likelihood_jax.py:706-773)
        evals += 1;
        if (lane == 0) {
          const double pxx = PA[ab_index(nc + 1, nc + 1, nc)];
          const double pxy = PA[ab_index(nc + 1, nc + 2, nc)];
          const double pxyy = clamp_pyy(PA[ab_index(nc + 2, nc + 2, nc)]);
          double sb, sse, tau;
          beta_se(df, pxx, pxy, pxyy, sb, sse, tau);
          os[JM_OUT_BETA_SCORE] = sb;
          os[JM_OUT_SE_SCORE] = sse;
          os[JM_OUT_P_SCORE] = (pxx > 0.0) ? f_pvalue(df, nn * (pxy * pxy) / (pyyA * pxx)) : NAN;
        }
        break;
      }
      wave_sync();
    }

    if (lane == 0) os[JM_OUT_EVALS] = (double)evals;
    wave_sync();
  }
}

// ---------------------------------------------------------------------------
// Association search with 16 lanes per SNP (single problem, T <= 16 * KT)
// ---------------------------------------------------------------------------
// The same evaluations, stages and outputs as k_lmm_assoc, with a SNP's work
// spread over a 16-lane group (four SNPs per wavefront) instead of a whole
// wavefront: the serial parts of an evaluation (reductions, the Pab
// recursion, the logs) are then shared by four SNPs. Lane j of a group owns
// the KT contiguous terms j*KT .. j*KT+KT-1. An evaluation:
//   1. h_c = 1 / (1 + lambda d_c) for both lambdas, clusters spread over the
//      group's lanes, into the group's LDS table;
//   2. kappa_t = h_c (-lambda h_c)^m and the m >= 1 logdet partials per lane;
//   3. dot products with the SNP's moment rows (global memory), the logdet
//      partials and the shared pair rows (block LDS), 8 rows x 2 lambdas at a
//      time, reduce-scattered over the group into the two packed Pab rows;
//   4. the Pab recursion, lanes 0-7 on row A (REML), lanes 8-15 on row B (MLE);
//   5. the logs (pivots, P_yy, clusters) over the 8 lanes of each half.
// Every SNP of a launch runs the same stage sequence, so the four groups of a
// wavefront stay in lockstep; a group past n_snps repeats the last SNP and
// writes nothing.
constexpr int kGL = 16;  // lanes per SNP

// Reduce-scatter of 16 per-lane values over a 16-lane group: on return lane j
// holds the group sum of value j & 15 (partners from xpartner<3..0>).
__device__ __forceinline__ double reduce16(double (&v)[16], int lane) {
  {
    const bool up = lane & 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const double send = up ? v[i] : v[i + 8];
      const double keep = up ? v[i + 8] : v[i];
      v[i] = keep + xpartner<3>(send, lane);
    }
  }
  {
    const bool up = lane & 4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const double send = up ? v[i] : v[i + 4];
      const double keep = up ? v[i + 4] : v[i];
      v[i] = keep + xpartner<2>(send, lane);
    }
  }
  {
    const bool up = lane & 2;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double send = up ? v[i] : v[i + 2];
      const double keep = up ? v[i + 2] : v[i];
      v[i] = keep + xpartner<1>(send, lane);
    }
  }
  {
    const bool up = lane & 1;
    const double send = up ? v[0] : v[1];
    const double keep = up ? v[1] : v[0];
    v[0] = keep + xpartner<0>(send, lane);
  }
  return v[0];
}

// Sum over the 8 lanes of a half-group; every lane of the half gets the same bits.
__device__ __forceinline__ double oct_sum(double v, int lane) {
  v += xpartner<0>(v, lane);
  v += xpartner<1>(v, lane);
  v += xpartner<2>(v, lane);
  return v;
}

// LDS of k_lmm_assoc16 (bytes): block constants, the shared pair rows (none
// when SG: they are read from global memory) and 16 per-group areas (two packed
// rows with their logdet slot, Iab level logs and the h table of both lambdas),
// then the int tables.
__host__ __device__ constexpr int assoc16_group_doubles(int ni, int C) {
  return 2 * (ni + 1) + kMaxLevels + 2 * C;
}
static size_t assoc16_lds(int KT, int nc, int C, bool sg) {
  const int TB = kGL * KT;
  const int ni = (nc + 3) * (nc + 2) / 2;
  const int n_sh = (nc + 1) * (nc + 2) / 2;
  const int nrows = nc + 2;
  const int NR = nrows + 1 + n_sh;
  return (size_t)(TB + 2 * C + (sg ? 0 : (size_t)n_sh * TB) +
                  16 * (size_t)assoc16_group_doubles(ni, C)) *
             sizeof(double) +
         (size_t)(TB + tet(nc + 1) + NR) * sizeof(int);
}

// One block = 16 consecutive SNPs of the launch (4 per wavefront). MULTI: SNP s
// is tested under probs[snp_prob[s]] (problems of one cohort: the same spectrum
// plan, n and T) from the moments block at snp_xmom[s]; a block splits its 16 SNPs
// into runs of one problem and reloads the problem's LDS rows (shared pair rows,
// recursion table) when the problem changes, so a batch of requests runs as one
// launch. The LDS layout is that of the widest problem (max_cvt). SG: the shared
// pair rows are read from global memory (L2-resident: every SNP of a problem
// reads the same rows) instead of LDS, for layouts whose rows do not fit; the
// values and the order of the multiply-adds are the same, so are the results.
template <int KT, bool MULTI, bool SG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_lmm_assoc16(
    jm_lmm_problem pr_one, int mode_one, const double* __restrict__ xmom, int64_t ldm,
    int64_t n_snps, double* __restrict__ out, const jm_lmm_problem* __restrict__ probs,
    const int32_t* __restrict__ modes, const int32_t* __restrict__ snp_prob,
    const double* const* __restrict__ snp_xmom, int max_cvt) {
  extern __shared__ double smem[];
  constexpr int TB = kGL * KT;
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar loop control
  const int grp = lane >> 4;
  const int j = lane & 15;
  const int h = j >> 3;  // 0: row A (REML, lambda A), 1: row B (MLE, lambda B)
  const int q8 = j & 7;
  const jm_lmm_problem& p0 = MULTI ? probs[0] : pr_one;  // the cohort's plan, n and T
  const int T = p0.n_terms, n = p0.n, C = p0.n_clusters;
  const int ncm = MULTI ? max_cvt : pr_one.n_cvt;  // LDS layout of the widest problem
  const int ni_m = (ncm + 3) * (ncm + 2) / 2;
  const int n_sh_m = (ncm + 1) * (ncm + 2) / 2;
  const int per_grp = assoc16_group_doubles(ni_m, C);
  double* c_twl = smem;                       // TB: logdet weight / power (0: m == 0, t >= T)
  double* c_cdc = c_twl + TB;                 // C cluster centres
  double* c_cw = c_cdc + C;                   // C cluster weights
  double* srows = c_cw + C;                   // n_sh x TB shared pair rows, zero-padded
  double* groups = srows + (SG ? 0 : (size_t)n_sh_m * TB);  // 16 x per_grp
  int* c_tinf = (int*)(groups + 16 * per_grp);  // TB: m | cluster << 8
  int* rtab = c_tinf + TB;                    // tet(nc + 1) recursion updates
  int* vslot = rtab + tet(ncm + 1);           // NR: packed-row slot of each dot-product row
  for (int t = threadIdx.x; t < TB; t += blockDim.x) {
    const bool act = t < T;
    const int m = act ? p0.term_m[t] : 0;
    const int cl = act ? p0.term_cluster[t] : 0;
    c_tinf[t] = m | (cl << 8);
    c_twl[t] = (act && m > 0) ? p0.logdet_mom[t] / (double)m : 0.0;
  }
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    c_cdc[c] = p0.cl_center[c];
    c_cw[c] = p0.cl_weight[c];
  }
  // the problem's shared pair rows, recursion table and row slots
  auto load_problem = [&](const jm_lmm_problem& pr) {
    const int nc = pr.n_cvt;
    const int n_sh = (nc + 1) * (nc + 2) / 2;
    const int NR = nc + 2 + 1 + n_sh;
    if (!SG) {
      for (int k = 0; k < n_sh; ++k) {
        const int jj = shared_pair_index(k, nc);
        for (int t = threadIdx.x; t < TB; t += blockDim.x)
          srows[(size_t)k * TB + t] = t < T ? pr.shared_mom[(int64_t)jj * T + t] : 0.0;
      }
    }
    const int m = nc + 2;
    for (int e = threadIdx.x; e < tet(nc + 1); e += blockDim.x) {
      int p = 1;
      while (tet(m - 1) - tet(m - p - 1) <= e) ++p;  // level of entry e
      int rem = e - (tet(m - 1) - tet(m - p)), ap = 0;
      const int Lw = m - p;
      while (rem >= Lw - ap) {
        rem -= Lw - ap;
        ++ap;
      }
      const int a = p + 1 + ap, b = a + rem;
      rtab[e] = ab_index(a, b, nc) | (ab_index(a, p, nc) << 8) | (ab_index(b, p, nc) << 16);
    }
    for (int r = threadIdx.x; r < NR; r += blockDim.x)
      vslot[r] = r < nc + 2 ? xpair_index(r, nc)
                            : (r == nc + 2 ? (nc + 3) * (nc + 2) / 2
                                           : shared_pair_index(r - nc - 3, nc));
  };
  if (!MULTI) load_problem(pr_one);
  __syncthreads();

  int tinf[KT];
#pragma unroll
  for (int k = 0; k < KT; ++k) tinf[k] = c_tinf[j * KT + k];
  const double phi = 0.6180339887498949;
  const double nn = (double)n;
  const double cm = 0.5 * nn * (log(nn) - log(2.0 * M_PI) - 1.0);
  int loaded = MULTI ? -1 : 0;
  const int64_t n_units = (n_snps + 15) / 16;
  for (int64_t un = blockIdx.x; un < n_units; un += gridDim.x) {
   const int64_t u_lo = un * 16, u_hi = min(u_lo + 16, n_snps);
   for (int64_t sub = u_lo; sub < u_hi;) {
    int pidx = 0;
    int64_t s_end = u_hi;
    if constexpr (MULTI) {  // the run of SNPs of one problem starting at sub (block-uniform)
      pidx = snp_prob[sub];
      s_end = sub + 1;
      while (s_end < u_hi && snp_prob[s_end] == pidx) ++s_end;
      if (pidx != loaded) {
        __syncthreads();  // every wavefront is done with the previous problem's rows
        load_problem(probs[pidx]);
        __syncthreads();
        loaded = pidx;
      }
    }
    const jm_lmm_problem& pr = MULTI ? probs[pidx] : pr_one;
    const int mode = MULTI ? modes[pidx] : mode_one;
    const int nc = pr.n_cvt, G = pr.n_grid;
    const double* shm = pr.shared_mom;  // SG: the shared pair rows, (ni x T)
    const int ni = (nc + 3) * (nc + 2) / 2;
    const int nrows = nc + 2;  // per-SNP moment rows: W_1..W_c, y, xx
    const int n_sh = (nc + 1) * (nc + 2) / 2;
    const int NR = nrows + 1 + n_sh;  // dot-product rows: SNP rows, logdet, shared rows
    double* PA = groups + (size_t)(wib * 4 + grp) * per_grp;  // ni entries + logdet slot
    double* PB = PA + ni + 1;
    double* iab_ld = PB + ni + 1;  // kMaxLevels
    double* H = iab_ld + kMaxLevels;  // 2 C: (h_A, h_B) per cluster
    double* Ph = h ? PB : PA;
    const bool do_reml = (mode == 1 || mode == 4);
    const bool do_mle = (mode == 2 || mode == 4);
    const bool do_score = (mode == 3 || mode == 4);
    const bool do_golden = do_reml || do_mle;
    const double df = (double)(n - nc - 1);
    const double cr = 0.5 * df * (log(df) - log(2.0 * M_PI) - 1.0);
    const int st_final = 3 + pr.n_iter, st_score = 4 + pr.n_iter;
    const int64_t s_raw = u_lo + 4 * wib + grp;
    const bool active = s_raw >= sub && s_raw < s_end;
    const int64_t s = active ? s_raw : sub;  // an idle group repeats a SNP of the run, unwritten
    const double* xm = MULTI ? snp_xmom[s] : xmom + s * (int64_t)nrows * ldm;
    double* os = out + s * JM_NOUT;
    // grid argmax indices left by k_lmm_grid, then columns a mode does not compute are
    // NaN. An idle group stands in for a SNP that another wavefront may be writing: it
    // reads no index (0), and indices are kept inside the grid.
    const int gi_r = (do_golden && active) ? min(max((int)os[JM_OUT_L_REMLE], 0), G - 1) : 0;
    const int gi_m = (do_golden && active) ? min(max((int)os[JM_OUT_L_MLE], 0), G - 1) : 0;
    wave_sync();
    if (active && j < JM_NOUT) os[j] = NAN;
    int evals = 0;
    double aA = 0, bA = 0, cA = 0, dA = 0, fcA = 0, fdA = 0;
    double aB = 0, bB = 0, cB = 0, dB = 0, fcB = 0, fdB = 0;
    bool klA = false, klB = false;
    double lam_reml = NAN, lam_mle = NAN;
    for (int stage = 0;;) {
      // ---- the lambdas of this stage ----
      const bool lam0 = stage == 0;
      double lamA = 0.0, lamB = 0.0, xA = 0.0, xB = 0.0;
      bool take_exp = true;
      if (stage == 1) {
        xA = cA;
        xB = cB;
      } else if (stage == 2) {
        xA = dA;
        xB = dB;
      } else if (stage >= 3 && stage < st_final) {
        klA = fcA > fdA;
        klB = fcB > fdB;
        const double naA = klA ? aA : cA, nbA = klA ? dA : bA;
        const double naB = klB ? aB : cB, nbB = klB ? dB : bB;
        const double ncA = nbA - phi * (nbA - naA), ndA = naA + phi * (nbA - naA);
        const double ncB = nbB - phi * (nbB - naB), ndB = naB + phi * (nbB - naB);
        aA = naA; bA = nbA; cA = ncA; dA = ndA;
        aB = naB; bB = nbB; cB = ncB; dB = ndB;
        xA = klA ? cA : dA;
        xB = klB ? cB : dB;
      } else if (stage == st_final) {
        xA = (aA + bA) / 2.0;
        xB = (aB + bB) / 2.0;
      } else {
        take_exp = false;
        if (stage == st_score) lamA = lamB = pr.lambda_null;
      }
      if (take_exp) {
        // each half takes the exponential of its own lambda; lane 15 - j is in the other half
        const double e = exp(h ? xB : xA);
        const double o = xpartner<3>(e, lane);
        lamA = h ? o : e;
        lamB = h ? e : o;
        if (stage == st_final) {
          lam_reml = lamA;
          lam_mle = lamB;
        }
      }
      // ---- 1. h table ----
      for (int c = j; c < C; c += kGL) {
        const double dc = c_cdc[c];
        H[2 * c] = 1.0 / fma(lamA, dc, 1.0);
        H[2 * c + 1] = 1.0 / fma(lamB, dc, 1.0);
      }
      wave_sync();
      // ---- 2. kappa and the m >= 1 logdet partials ----
      double kA[KT], kB[KT];
      double ldA = 0.0, ldB = 0.0;
      {
        double pA = 1.0, pB = 1.0;
#pragma unroll
        for (int k = 0; k < KT; ++k) {
          const int t = j * KT + k;
          const int m = tinf[k] & 255, cl = tinf[k] >> 8;
          const double hA = H[2 * cl], hB = H[2 * cl + 1];
          const double nzA = -lamA * hA, nzB = -lamB * hB;
          if (k == 0) {
            pA = ipow16(nzA, m);
            pB = ipow16(nzB, m);
          } else {
            pA = m == 0 ? 1.0 : pA * nzA;
            pB = m == 0 ? 1.0 : pB * nzB;
          }
          const bool act = t < T;
          kA[k] = act ? hA * pA : 0.0;
          kB[k] = act ? hB * pB : 0.0;
          const double wl = c_twl[t];
          ldA -= pA * wl;
          ldB -= pB * wl;
        }
      }
      // ---- 3. dot products, 8 rows x 2 lambdas per reduce-scatter ----
      for (int b0 = 0; b0 < NR; b0 += 8) {
        double v[16];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int r = b0 + e;
          double sa = 0.0, sb = 0.0;
          if (r < nrows) {
            const double* row = xm + (int64_t)r * ldm + j * KT;
#pragma unroll
            for (int k = 0; k < KT; ++k) {
              const double x = (j * KT + k < T) ? row[k] : 0.0;
              sa = fma(kA[k], x, sa);
              sb = fma(kB[k], x, sb);
            }
          } else if (r == nrows) {
            sa = ldA;
            sb = ldB;
          } else if (r < NR) {
            if constexpr (SG) {  // packed pair index vslot[r] of shared pair r - nrows - 1
              const double* row = shm + (int64_t)vslot[r] * T + j * KT;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const double x = (j * KT + k < T) ? row[k] : 0.0;
                sa = fma(kA[k], x, sa);
                sb = fma(kB[k], x, sb);
              }
            } else {
              const double* row = srows + (size_t)(r - nrows - 1) * TB + j * KT;
#pragma unroll
              for (int k = 0; k < KT; ++k) {
                const double x = row[k];
                sa = fma(kA[k], x, sa);
                sb = fma(kB[k], x, sb);
              }
            }
          }
          v[2 * e] = sa;
          v[2 * e + 1] = sb;
        }
        const double r = reduce16(v, lane);
        const int row = b0 + (j >> 1);
        if (row < NR) ((j & 1) ? PB : PA)[vslot[row]] = r;
      }
      wave_sync();
      // ---- 4. Pab recursion: half A on PA, half B on PB ----
      double pyy_lvl = 0.0;
      for (int p = 1; p <= nc + 1; ++p) {
        const double d = Ph[ab_index(p, p, nc)];
        if (p == nc + 1) pyy_lvl = Ph[ab_index(nc + 2, nc + 2, nc)];
        const double inv = (d != 0.0) ? 1.0 / d : 0.0;
        const int Lw = nc + 2 - p;
        const int np = Lw * (Lw + 1) / 2;
        const int lv0 = tet(nc + 1) - tet(nc + 2 - p);
        for (int idx = q8; idx < np; idx += 8) {
          const int e = rtab[lv0 + idx];
          const int iab = e & 255, ia = (e >> 8) & 255, ib = e >> 16;
          const double pa = Ph[ia], pb = Ph[ib];
          Ph[iab] = Ph[iab] - pa * pb * inv;
        }
        wave_sync();
      }
      // ---- 5. logs: pivots (A), P_yy, clusters ----
      const double pyy1 = Ph[ab_index(nc + 2, nc + 2, nc)];
      const double lam_h = h ? lamB : lamA;
      const int n_piv = h ? 0 : nc + 1;
      const int n_cl = lam0 ? 0 : C;
      const int n_items = n_piv + 1 + n_cl;
      const double dfh = h ? nn : df;
      double u = 0.0;  // -0.5 x (this lane's share of S) - 0.5 dfh log P_yy
      for (int it = q8; it < n_items; it += 8) {
        if (it < n_piv) {
          const double piv = Ph[ab_index(it + 1, it + 1, nc)];
          const double lg = log(piv > 0.0 ? piv : 1.0);
          if (lam0) iab_ld[it] = lg;
          u -= 0.5 * (lam0 ? lg : lg - iab_ld[it]);
        } else if (it == n_piv) {
          u -= 0.5 * dfh * log(clamp_pyy(pyy1));
        } else {
          const int c = it - n_piv - 1;
          u -= 0.5 * (c_cw[c] * log(fabs(fma(lam_h, c_cdc[c], 1.0))));
        }
      }
      const double fh = (h ? cm : cr) - 0.5 * Ph[ni] + oct_sum(u, lane);
      const double fo = xpartner<3>(fh, lane);
      const double fA = h ? fo : fh, fB = h ? fh : fo;
      const double po = xpartner<3>(pyy_lvl, lane);
      const double pyyA = h ? po : pyy_lvl;
      const bool writer = active && j == 0;

      // ---- consume the stage ----
      if (stage == 0) {
        // Iab done; the grid stage ran in k_lmm_grid
        if (do_golden) {
          evals += G;
          const int iloA = gi_r > 0 ? gi_r - 1 : 0, ihiA = gi_r + 1 < G ? gi_r + 1 : G - 1;
          const int iloB = gi_m > 0 ? gi_m - 1 : 0, ihiB = gi_m + 1 < G ? gi_m + 1 : G - 1;
          aA = pr.grid_loglam[iloA];
          bA = pr.grid_loglam[ihiA];
          aB = pr.grid_loglam[iloB];
          bB = pr.grid_loglam[ihiB];
          cA = bA - phi * (bA - aA);
          dA = aA + phi * (bA - aA);
          cB = bB - phi * (bB - aB);
          dB = aB + phi * (bB - aB);
          stage = 1;
        } else {
          stage = st_score;
        }
      } else if (stage == 1) {
        fcA = fA;
        fcB = fB;
        stage = 2;
      } else if (stage == 2) {
        fdA = fA;
        fdB = fB;
        stage = 3;
      } else if (stage < st_final) {
        const double nfcA = klA ? fA : fdA, nfdA = klA ? fcA : fA;
        const double nfcB = klB ? fB : fdB, nfdB = klB ? fcB : fB;
        fcA = nfcA;
        fdA = nfdA;
        fcB = nfcB;
        fdB = nfdB;
        ++stage;
      } else if (stage == st_final) {
        // Wald at the REML optimum (likelihood_jax.py:601-675), LRT at the MLE optimum
        evals += 2 * (pr.n_iter + 3);
        if (writer && do_reml) {
          const double pxx = PA[ab_index(nc + 1, nc + 1, nc)];
          const double pxy = PA[ab_index(nc + 1, nc + 2, nc)];
          const double pxyy = clamp_pyy(PA[ab_index(nc + 2, nc + 2, nc)]);
          double beta, se, tau;
          beta_se(df, pxx, pxy, pxyy, beta, se, tau);
          os[JM_OUT_BETA] = beta;
          os[JM_OUT_SE] = se;
          os[JM_OUT_P_WALD] = (pxx > 0.0) ? f_pvalue(df, (pyyA - pxyy) * tau) : NAN;
          os[JM_OUT_L_REMLE] = lam_reml;
          os[JM_OUT_LOGL_H1] = fA;
        }
        if (writer && do_mle) {
          os[JM_OUT_L_MLE] = lam_mle;
          os[JM_OUT_LOGL_MLE] = fB;
          os[JM_OUT_P_LRT] = lrt_pvalue(fB, pr.logl_null);
        }
        if (!do_score) break;
        stage = st_score;
      } else {  // score at the null lambda (likelihood_jax.py:706-773)
        evals += 1;
        if (writer) {
          const double pxx = PA[ab_index(nc + 1, nc + 1, nc)];
          const double pxy = PA[ab_index(nc + 1, nc + 2, nc)];
          const double pxyy = clamp_pyy(PA[ab_index(nc + 2, nc + 2, nc)]);
          double sb, sse, tau;
          beta_se(df, pxx, pxy, pxyy, sb, sse, tau);
          os[JM_OUT_BETA_SCORE] = sb;
          os[JM_OUT_SE_SCORE] = sse;
          os[JM_OUT_P_SCORE] = (pxx > 0.0) ? f_pvalue(df, nn * (pxy * pxy) / (pyyA * pxx)) : NAN;
        }
        break;
      }
      wave_sync();
    }
    if (active && j == 0) os[JM_OUT_EVALS] = (double)evals;
    wave_sync();
    sub = s_end;
   }
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

template <int KT, int XR, bool MB>
static void launch_kt(const jm_lmm_problem* one, int mode_one, const double* xmom, int64_t ldm,
                      int64_t n_snps, double* out, const jm_lmm_problem* probs,
                      const int32_t* modes, const int32_t* snp_prob, const double* const* snp_xmom,
                      int ni_max, unsigned blocks, int wpb, size_t lds, int srow_n,
                      hipStream_t st) {
  if (one) {
    hipLaunchKernelGGL((k_lmm_assoc<KT, XR, MB, false>), dim3(blocks), dim3(64 * wpb), lds, st, *one,
                       mode_one, xmom, ldm, n_snps, out, nullptr, nullptr, nullptr, nullptr,
                       ni_max, srow_n);
  } else {
    jm_lmm_problem dummy;
    memset(&dummy, 0, sizeof(dummy));
    hipLaunchKernelGGL((k_lmm_assoc<KT, XR, MB, true>), dim3(blocks), dim3(64 * wpb), lds, st, dummy,
                       0, nullptr, ldm, n_snps, out, probs, modes, snp_prob, snp_xmom, ni_max, 0);
  }
}

static int launch_assoc(const jm_lmm_problem* one, int mode_one, const double* xmom, int64_t ldm,
                        int64_t n_snps, double* out, const jm_lmm_problem* probs,
                        const int32_t* modes, const int32_t* snp_prob,
                        const double* const* snp_xmom, int T, int max_cvt, int sel_cvt,
                        int n_clusters, void* stream) {
  // max_cvt sizes the LDS layouts (the widest problem of the launch); sel_cvt >=
  // max_cvt picks the lambda-search kernel (16 lanes or a wavefront per SNP, which
  // sum in different orders), so launches with one sel_cvt agree bit for bit
  JM_CHECK_ARG(T >= 1, "n_terms=%d < 1", T);
  if (sel_cvt < max_cvt) sel_cvt = max_cvt;
  const int ni_max = (max_cvt + 3) * (max_cvt + 2) / 2;
  hipStream_t st = (hipStream_t)stream;
  // grid stage (skips Score-only SNPs): one wave per SNP, a Pab column per lane in LDS
  {
    // Pab columns: G + 1 lanes hold one when a single-problem grid fits one pass (the
    // reference's 50 points: 51), which lets 4 waves per SIMD fit the LDS instead of 3
    const int ps = (one && one->n_grid + 1 <= 64) ? one->n_grid + 1 : 64;
    const size_t gw = (size_t)(ni_max * ps + kMaxLevels) * sizeof(double);
    int gwpb = (int)((64 * 1024) / gw);
    if (gwpb > 4) gwpb = 4;
    if (gwpb < 1) {
      JM_CHECK_ARG(gw <= 160 * 1024, "problem too large for the LDS budget (n_cvt=%d)", max_cvt);
      gwpb = 1;
    }
    int64_t gblocks = (n_snps + gwpb - 1) / gwpb;
    if (gblocks > 256 * 32) gblocks = 256 * 32;
    if (one) {
      if (mode_one != 3)
        hipLaunchKernelGGL(k_lmm_grid<false>, dim3((unsigned)gblocks), dim3(64 * gwpb), gw * gwpb,
                           st, *one, mode_one, xmom, ldm, n_snps, out, nullptr, nullptr, nullptr,
                           nullptr, ni_max, ps);
    } else {
      jm_lmm_problem dummy;
      memset(&dummy, 0, sizeof(dummy));
      hipLaunchKernelGGL(k_lmm_grid<true>, dim3((unsigned)gblocks), dim3(64 * gwpb), gw * gwpb, st,
                         dummy, 0, nullptr, ldm, n_snps, out, probs, modes, snp_prob, snp_xmom,
                         ni_max, ps);
    }
    JM_CHECK_LAUNCH();
  }
  // at most 256 terms: 16 lanes per SNP when the LDS fits (one problem, or a batch of
  // problems of one cohort; JM_LMM_WAVE_PER_SNP=1 keeps the wavefront-per-SNP kernel,
  // for comparisons)
  {
    const char* env = getenv("JM_LMM_WAVE_PER_SNP");
    const bool wave = env && env[0] == '1';
    const int kt16 = T <= 64 ? 4 : (T <= 128 ? 8 : (T <= 192 ? 12 : (T <= 256 ? 16 : 0)));
    // the shared pair rows go to LDS when they fit, else they are read from global
    // memory (same arithmetic): the kernel choice depends on T, the cluster count and
    // max_cvt only
    const bool sg = kt16 && assoc16_lds(kt16, max_cvt, n_clusters, false) > 64 * 1024;
    const size_t lds16 = kt16 ? assoc16_lds(kt16, max_cvt, n_clusters, sg) : 0;
    const bool fits = kt16 && assoc16_lds(kt16, sel_cvt, n_clusters, true) <= 64 * 1024;
    if (!wave && kt16 && n_clusters >= 1 && fits) {
      int64_t b16 = (n_snps + 15) / 16;
      if (b16 > 256 * 8) b16 = 256 * 8;
      jm_lmm_problem dummy;
      memset(&dummy, 0, sizeof(dummy));
#define JM_LAUNCH16_SG(KT_, SG_)                                                                \
  do {                                                                                          \
    if (one)                                                                                    \
      hipLaunchKernelGGL((k_lmm_assoc16<KT_, false, SG_>), dim3((unsigned)b16), dim3(256), lds16, \
                         st, *one, mode_one, xmom, ldm, n_snps, out, nullptr, nullptr, nullptr, \
                         nullptr, max_cvt);                                                     \
    else                                                                                        \
      hipLaunchKernelGGL((k_lmm_assoc16<KT_, true, SG_>), dim3((unsigned)b16), dim3(256), lds16,  \
                         st, dummy, 0, nullptr, ldm, n_snps, out, probs, modes, snp_prob,       \
                         snp_xmom, max_cvt);                                                    \
  } while (0)
#define JM_LAUNCH16(KT_)           \
  do {                             \
    if (sg)                        \
      JM_LAUNCH16_SG(KT_, true);   \
    else                           \
      JM_LAUNCH16_SG(KT_, false);  \
  } while (0)
      if (kt16 == 4)
        JM_LAUNCH16(4);
      else if (kt16 == 8)
        JM_LAUNCH16(8);
      else if (kt16 == 12)
        JM_LAUNCH16(12);
      else
        JM_LAUNCH16(16);
#undef JM_LAUNCH16
#undef JM_LAUNCH16_SG
      JM_CHECK_LAUNCH();
      return JM_OK;
    }
  }
  const size_t per_wave = (size_t)assoc_lds_per_wave(ni_max) * sizeof(double);
  // single problem: the shared pair rows go to LDS when they take <= 48 KB
  const int n_sh = (max_cvt + 1) * (max_cvt + 2) / 2;
  const int sstride = T > 512 ? T : 64 * assoc_kt(T);  // the kernel's padded row length
  const size_t sh_bytes = (size_t)n_sh * sstride * sizeof(double);
  const int srow_n = (one && sh_bytes <= 48 * 1024) ? n_sh : 0;
  const size_t sh = srow_n ? sh_bytes : 0;
  int wpb = (int)((80 * 1024 - sh) / per_wave);
  if (wpb > 4) wpb = 4;
  if (wpb < 1) {
    JM_CHECK_ARG(sh + per_wave <= 160 * 1024, "problem too large for the LDS budget (n_cvt=%d)",
                 max_cvt);
    wpb = 1;
  }
  int64_t blocks = (n_snps + wpb - 1) / wpb;
  const int64_t max_blocks = 256 * 16;
  if (blocks > max_blocks) blocks = max_blocks;
  const unsigned nb = (unsigned)blocks;
  const size_t lds = sh + per_wave * wpb + assoc_lds_consts(T, max_cvt);
  const int kt = (T + 63) / 64;  // > 8: blocks of 512 terms
  // moment rows held in registers: the c + 2 rows of the widest problem, up to 8
  const int xr = max_cvt + 2 <= 4 ? 4 : (max_cvt + 2 <= 6 ? 6 : 8);
#define JM_LAUNCH_KT(KT_, XR_, MB_)                                                            \
  launch_kt<KT_, XR_, MB_>(one, mode_one, xmom, ldm, n_snps, out, probs, modes, snp_prob,       \
                           snp_xmom, ni_max, nb, wpb, lds, srow_n, st)
#define JM_LAUNCH_XR(KT_)                  \
  do {                                     \
    if (xr == 4)                           \
      JM_LAUNCH_KT(KT_, 4, false);         \
    else if (xr == 6)                      \
      JM_LAUNCH_KT(KT_, 6, false);         \
    else                                   \
      JM_LAUNCH_KT(KT_, 8, false);         \
  } while (0)
  if (kt <= 1)
    JM_LAUNCH_XR(1);
  else if (kt <= 2)
    JM_LAUNCH_XR(2);
  else if (kt <= 3)
    JM_LAUNCH_XR(3);
  else if (kt <= 4)
    JM_LAUNCH_XR(4);
  else if (kt <= 6)
    JM_LAUNCH_KT(6, 4, false);
  else if (kt <= 8)
    JM_LAUNCH_KT(8, 4, false);
  else
    JM_LAUNCH_KT(8, 4, true);
#undef JM_LAUNCH_XR
#undef JM_LAUNCH_KT
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
                      prob->n_terms, prob->n_cvt, prob->n_cvt, prob->n_clusters, stream);
}

extern "C" int jm_lmm_assoc_multi(const jm_lmm_problem* probs, const int32_t* modes, int n_probs,
                                  int max_cvt, int width_cvt, int n_terms, int n_clusters,
                                  const int32_t* snp_prob, const double* const* snp_xmom,
                                  int64_t ldm, int64_t n_snps, double* out, void* stream) {
  JM_CHECK_ARG(n_snps >= 0 && n_probs >= 1, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(probs && modes && snp_prob && snp_xmom && out, "null pointer");
  JM_CHECK_ARG(max_cvt >= 1 && max_cvt <= 20, "max_cvt out of range [1,20]");
  JM_CHECK_ARG(width_cvt == 0 || (width_cvt >= max_cvt && width_cvt <= 20),
               "width_cvt must be 0 or in [max_cvt, 20]");
  JM_CHECK_ARG(n_terms >= 1 && ldm >= n_terms, "bad n_terms / ldm");
  JM_CHECK_ARG(n_clusters >= 1, "n_clusters < 1");
  return launch_assoc(nullptr, 0, nullptr, ldm, n_snps, out, probs, modes, snp_prob, snp_xmom,
                      n_terms, max_cvt, width_cvt ? width_cvt : max_cvt, n_clusters, stream);
}

extern "C" double jm_host_betainc(double a, double b, double x) { return betainc_reg(a, b, x); }
extern "C" double jm_host_chi2_sf1(double x) { return chi2_sf1(x); }
extern "C" double jm_host_f_pvalue(double df, double f_stat) { return f_pvalue(df, f_stat); }
extern "C" double jm_host_lrt_pvalue(double logl_h1, double logl_h0) {
  return lrt_pvalue(logl_h1, logl_h0);
}
