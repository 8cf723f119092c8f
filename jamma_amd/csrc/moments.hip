// Spectrum moments: the lambda-independent part of every likelihood
// evaluation, computed once per SNP from its rotated genotype column.
//
// For eigen-index i in cluster c (centre d_c, delta_i = d_i - d_c) and term
// t = toff_c + m:
//   mom[s][b][t]  = sum_{i in c} Z[i][s] * V[b][i] * delta_i^m     (b < nv)
//   mom[s][nv][t] = sum_{i in c} Z[i][s]^2 * delta_i^m
// These replace the per-SNP Uab materialisation of batch_compute_uab
// (src/jamma/lmm/likelihood_jax.py:342-393): a Pab row-0 entry at any lambda
// is then sum_t kappa_t(lambda) * mom[.][.][t] (see lmm.hip).
//
// Layout: Z is (n x ldz), SNP s contiguous within a row, so one thread per
// SNP reads coalesced rows while walking i; accumulators for the current
// cluster stay in registers and are flushed at the cluster end. Clusters are
// split into index-range groups across grid.y for parallelism; the summation
// order inside a cluster is fixed, so results are bitwise reproducible.

#include "jm_common.h"

namespace jm {

constexpr int kMomMaxTerms = 16;  // MAX_TERMS_PER_CLUSTER on the host side
constexpr int kMomGroup = 6;      // partner rows accumulated per pass
constexpr int kMomTerms = kMomMaxTerms;

template <typename ZT, bool MULTI>
__global__ void __launch_bounds__(64) k_moments(const ZT* __restrict__ Z, int64_t ldz, int64_t n,
                                                int64_t n_snps, const double* __restrict__ V_one,
                                                int64_t ldv_one, int nv_one,
                                                const jm_mom_col* __restrict__ cols, int nv_max,
                                                const int32_t* __restrict__ cl_start,
                                                const int32_t* __restrict__ cl_size,
                                                const int32_t* __restrict__ cl_terms,
                                                const int32_t* __restrict__ cl_toff,
                                                int n_clusters,
                                                const double* __restrict__ delta,
                                                double* __restrict__ mom_one, int64_t ldm_one) {
  const int64_t s = (int64_t)blockIdx.x * 64 + threadIdx.x;
  const bool active = s < n_snps;
  const int64_t sc = active ? s : 0;
  // this column's partner rows and output block
  const double* V = V_one;
  int64_t ldv = ldv_one, ldm = ldm_one;
  int nv = nv_one;
  double* ms;
  if (MULTI) {
    const jm_mom_col d = cols[sc];
    V = d.V;
    ldv = d.ldv;
    nv = d.nv;
    ldm = d.ldm;
    ms = d.out;
  } else {
    ms = mom_one + sc * (int64_t)(nv_one + 1) * ldm_one;
  }
  // cluster range of this group: clusters whose start lies in [i_lo, i_hi)
  const int64_t i_lo = n * blockIdx.y / gridDim.y;
  const int64_t i_hi = n * (blockIdx.y + 1) / gridDim.y;
  int c0 = 0, c1 = n_clusters;
  {  // first cluster with start >= i_lo
    int lo = 0, hi = n_clusters;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cl_start[mid] < i_lo) lo = mid + 1; else hi = mid;
    }
    c0 = lo;
    lo = c0;
    hi = n_clusters;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cl_start[mid] < i_hi) lo = mid + 1; else hi = mid;
    }
    c1 = lo;
  }
  const int nrows = nv + 1;
  const int nrows_max = (MULTI ? nv_max : nv_one) + 1;
  for (int g0 = 0; g0 < nrows_max; g0 += kMomGroup) {
    int ng = nrows - g0;
    ng = ng < 0 ? 0 : (ng > kMomGroup ? kMomGroup : ng);
    for (int c = c0; c < c1; ++c) {
      const int i0 = cl_start[c];
      const int sz = cl_size[c];
      const int M = cl_terms[c];
      double acc[kMomGroup][kMomTerms];
#pragma unroll
      for (int b = 0; b < kMomGroup; ++b)
#pragma unroll
        for (int m = 0; m < kMomTerms; ++m) acc[b][m] = 0.0;
      for (int i = i0; i < i0 + sz; ++i) {
        const double z = (double)Z[(int64_t)i * ldz + sc];
        const double dl = delta[i];
        double prod[kMomGroup];
#pragma unroll
        for (int b = 0; b < kMomGroup; ++b) {
          const int row = g0 + b;
          prod[b] = 0.0;
          if (b < ng) prod[b] = (row < nv) ? z * V[(int64_t)row * ldv + i] : z * z;
        }
        double pw = 1.0;
#pragma unroll
        for (int m = 0; m < kMomTerms; ++m) {
          if (m < M) {
#pragma unroll
            for (int b = 0; b < kMomGroup; ++b) acc[b][m] += prod[b] * pw;
            pw *= dl;
          }
        }
      }
      if (active) {
        const int toff = cl_toff[c];
#pragma unroll
        for (int b = 0; b < kMomGroup; ++b) {
          if (b < ng) {
            double* dst = ms + (int64_t)(g0 + b) * ldm + toff;
#pragma unroll
            for (int m = 0; m < kMomTerms; ++m)
              if (m < M) dst[m] = acc[b][m];
          }
        }
      }
    }
  }
}

}  // namespace jm

using namespace jm;

extern "C" int jm_spectrum_moments(const void* Z, int z_is_f32, int64_t ldz, int64_t n,
                                   int64_t n_snps, const double* V, int64_t ldv, int nv,
                                   const int32_t* cl_start, const int32_t* cl_size,
                                   const int32_t* cl_terms, const int32_t* cl_toff,
                                   int n_clusters, const double* delta, double* mom, int64_t ldm,
                                   void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0 && nv >= 0 && n_clusters >= 1, "bad sizes");
  JM_CHECK_ARG(ldz >= n_snps, "ldz < n_snps");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(Z && cl_start && cl_size && cl_terms && cl_toff && delta && mom, "null pointer");
  JM_CHECK_ARG(nv == 0 || (V && ldv >= n), "V missing or ldv < n");
  int groups = (int)((n + 255) / 256);
  if (groups > 64) groups = 64;
  if (groups < 1) groups = 1;
  const dim3 grid((unsigned)((n_snps + 63) / 64), (unsigned)groups);
  hipStream_t st = (hipStream_t)stream;
  if (z_is_f32)
    hipLaunchKernelGGL((k_moments<float, false>), grid, dim3(64), 0, st, (const float*)Z, ldz, n,
                       n_snps, V, ldv, nv, nullptr, 0, cl_start, cl_size, cl_terms, cl_toff,
                       n_clusters, delta, mom, ldm);
  else
    hipLaunchKernelGGL((k_moments<double, false>), grid, dim3(64), 0, st, (const double*)Z, ldz,
                       n, n_snps, V, ldv, nv, nullptr, 0, cl_start, cl_size, cl_terms, cl_toff,
                       n_clusters, delta, mom, ldm);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_spectrum_moments_multi(const void* Z, int z_is_f32, int64_t ldz, int64_t n,
                                         int64_t n_snps, const jm_mom_col* cols, int nv_max,
                                         const int32_t* cl_start, const int32_t* cl_size,
                                         const int32_t* cl_terms, const int32_t* cl_toff,
                                         int n_clusters, const double* delta, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0 && nv_max >= 0 && n_clusters >= 1, "bad sizes");
  JM_CHECK_ARG(ldz >= n_snps, "ldz < n_snps");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(Z && cols && cl_start && cl_size && cl_terms && cl_toff && delta, "null pointer");
  int groups = (int)((n + 255) / 256);
  if (groups > 64) groups = 64;
  if (groups < 1) groups = 1;
  const dim3 grid((unsigned)((n_snps + 63) / 64), (unsigned)groups);
  hipStream_t st = (hipStream_t)stream;
  if (z_is_f32)
    hipLaunchKernelGGL((k_moments<float, true>), grid, dim3(64), 0, st, (const float*)Z, ldz, n,
                       n_snps, nullptr, 0, 0, cols, nv_max, cl_start, cl_size, cl_terms, cl_toff,
                       n_clusters, delta, nullptr, 0);
  else
    hipLaunchKernelGGL((k_moments<double, true>), grid, dim3(64), 0, st, (const double*)Z, ldz, n,
                       n_snps, nullptr, 0, 0, cols, nv_max, cl_start, cl_size, cl_terms, cl_toff,
                       n_clusters, delta, nullptr, 0);
  JM_CHECK_LAUNCH();
  return JM_OK;
}
