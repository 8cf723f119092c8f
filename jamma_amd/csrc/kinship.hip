// Centred kinship (GEMMA -gk 1) on int8 MFMA with exact FP64 corrections.
//
// Reference: K = Xc Xc^T / p with Xc = impute_and_center(X) per batch
// (src/jamma/kinship/compute.py:46-57, 93-226, 502-690;
//  src/jamma/kinship/missing.py:22-61). With X0 = genotypes with missing calls
// set to 0, M the missing mask, mu_j the float64 nanmean and D = diag(mu):
//   Xc = X0 - (1 - M) D
//   Xc Xc^T = X0 X0^T - u 1^T - 1 u^T + s0 11^T + C + C^T
// where u_i = sum_j (observed ? g_ij mu_j : mu_j^2), s0 = sum_j mu_j^2 and
// C = Y D M^T with Y_ij = g_ij (observed) or mu_j / 2 (missing). X0 X0^T is an
// exact int32 GEMM on v_mfma_i32_16x16x64_i8; the corrections are FP64 and
// only C is sparse work (one FP64 FMA per (row, missing call) pair).

#include "jm_common.h"

namespace jm {

typedef int i4_t __attribute__((ext_vector_type(4)));

constexpr int K_BM = 128, K_BK = 128;
constexpr int K_LS = K_BK + 16;  // LDS row stride in bytes

// missing (-1) -> 0 on packed bytes
__device__ __forceinline__ uint32_t zero_neg_bytes(uint32_t w) {
  const uint32_t neg = (w >> 7) & 0x01010101u;
  return w & ~(neg * 0xFFu);
}

__device__ __forceinline__ void tri_block(int t, int& rb, int& cb) {
  int r = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  while (r * (r + 1) / 2 > t) --r;
  rb = r;
  cb = t - r * (r + 1) / 2;
}

// S += X0 X0^T for 128x128 tiles with row block >= column block.
__global__ void __launch_bounds__(256) k_syrk_i8(const int8_t* __restrict__ gT, int64_t ld,
                                                 int64_t n, int64_t K, int32_t* __restrict__ S) {
  __shared__ __attribute__((aligned(16))) uint8_t As[K_BM * K_LS];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[K_BM * K_LS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int rb, cb;
  tri_block(blockIdx.x, rb, cb);
  const int64_t i0 = (int64_t)rb * K_BM, j0 = (int64_t)cb * K_BM;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  i4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (i4_t){0, 0, 0, 0};

  for (int64_t k0 = 0; k0 < K; k0 += K_BK) {
    // 128 rows x 128 bytes per operand = 1024 x 16 B chunks; 4 per thread
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int idx = tid + 256 * u;
      const int r = idx >> 3, c = (idx & 7) * 16;
      uint4 wa = {0, 0, 0, 0}, wb = {0, 0, 0, 0};
      const int64_t ia = i0 + r, ib = j0 + r;
      if (k0 + c + 16 <= K) {
        if (ia < n) wa = *reinterpret_cast<const uint4*>(gT + ia * ld + k0 + c);
        if (ib < n) wb = *reinterpret_cast<const uint4*>(gT + ib * ld + k0 + c);
      } else if (k0 + c < K) {
        uint8_t ta[16], tb[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const bool in = (k0 + c + e) < K;
          ta[e] = (in && ia < n) ? (uint8_t)gT[ia * ld + k0 + c + e] : 0;
          tb[e] = (in && ib < n) ? (uint8_t)gT[ib * ld + k0 + c + e] : 0;
        }
        memcpy(&wa, ta, 16);
        memcpy(&wb, tb, 16);
      }
      wa.x = zero_neg_bytes(wa.x); wa.y = zero_neg_bytes(wa.y);
      wa.z = zero_neg_bytes(wa.z); wa.w = zero_neg_bytes(wa.w);
      wb.x = zero_neg_bytes(wb.x); wb.y = zero_neg_bytes(wb.y);
      wb.z = zero_neg_bytes(wb.z); wb.w = zero_neg_bytes(wb.w);
      *reinterpret_cast<uint4*>(&As[r * K_LS + c]) = wa;
      *reinterpret_cast<uint4*>(&Bs[r * K_LS + c]) = wb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < K_BK; kk += 64) {
      i4_t af[4], bf[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        af[t] = *reinterpret_cast<const i4_t*>(&As[(wr + t * 16 + lr) * K_LS + kk + 16 * lk]);
        bf[t] = *reinterpret_cast<const i4_t*>(&Bs[(wc + t * 16 + lr) * K_LS + kk + 16 * lk]);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + a * 16 + 4 * lk + r;
        const int64_t j = j0 + wc + b * 16 + lr;
        if (i < n && j < n) S[i * n + j] += acc[a][b][r];
      }
}

// u[i] += sum_j (g_ij >= 0 ? g_ij mu_j : mu_j^2), one block per sample row.
__global__ void __launch_bounds__(256) k_impute_u(const int8_t* __restrict__ gT, int64_t ld,
                                                  int64_t K, const double* __restrict__ mu,
                                                  double* __restrict__ u) {
  __shared__ double red[4];
  const int64_t i = blockIdx.x;
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < K; j += blockDim.x) {
    const int8_t gv = gT[i * ld + j];
    const double m = mu[j];
    acc += (gv >= 0) ? (double)gv * m : m * m;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) u[i] += ((red[0] + red[1]) + (red[2] + red[3]));
}

// C[k][i] += sum_{j: sample k missing at j} mu_j * Y_ij over one SNP tile.
// Block (k, i-chunk of 2048): the missing SNPs of row k are compacted into LDS
// (in SNP order), then each thread accumulates 8 consecutive samples i.
constexpr int C_CHUNK = 4096;
__global__ void __launch_bounds__(256) k_impute_C(const int8_t* __restrict__ g, int64_t ld_g,
                                                  const int8_t* __restrict__ gT, int64_t ld_gT,
                                                  int64_t n, int64_t K,
                                                  const double* __restrict__ mu,
                                                  double* __restrict__ C) {
  __shared__ int32_t list[C_CHUNK];
  __shared__ int32_t wsum[5];
  const int64_t k = blockIdx.x;
  const int64_t ibase = (int64_t)blockIdx.y * 2048 + threadIdx.x * 8;
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int64_t j0 = 0; j0 < K; j0 += C_CHUNK) {
    // each thread inspects 16 consecutive SNPs of row k
    const int64_t jt = j0 + threadIdx.x * 16;
    uint32_t mask = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e)
      if (jt + e < K && gT[k * ld_gT + jt + e] < 0) mask |= (1u << e);
    int cnt = __popc(mask);
    // exclusive scan of cnt across the block
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int v = __shfl_up(incl, off, 64);
      if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int wofs = 0;
    for (int q = 0; q < w; ++q) wofs += wsum[q];
    const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    int pos = wofs + incl - cnt;
    for (int e = 0; e < 16; ++e)
      if (mask & (1u << e)) list[pos++] = (int32_t)(jt + e);
    __syncthreads();
    for (int q = 0; q < total; ++q) {
      const int64_t j = list[q];
      const double m = mu[j];
      const double half = 0.5 * m;
      if (ibase < n) {
        const int8_t* gr = g + j * ld_g + ibase;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          if (ibase + e < n) {
            const int8_t gv = gr[e];
            acc[e] += m * ((gv >= 0) ? (double)gv : half);
          }
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (ibase + e < n) C[k * n + ibase + e] += acc[e];
}

__global__ void k_kin_finalize(const int32_t* __restrict__ S, const double* __restrict__ u,
                               const double* __restrict__ C, double s0,
                               const int32_t* __restrict__ Ss, const double* __restrict__ us,
                               const double* __restrict__ Cs, double s0s, int64_t n, double p,
                               double* __restrict__ K) {
  const int64_t total = n * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / n, k = idx % n;
    const int64_t src = ((i / K_BM) >= (k / K_BM)) ? idx : (k * n + i);
    int64_t sv = S[src];
    if (Ss) sv -= Ss[src];
    double ui = u[i], uk = u[k], c = 0.0, s = s0;
    if (us) {
      ui -= us[i];
      uk -= us[k];
      s -= s0s;
    }
    if (C) c = C[i * n + k] + C[k * n + i];
    if (Cs) c -= Cs[i * n + k] + Cs[k * n + i];
    K[idx] = ((double)sv - ui - uk + s + c) / p;
  }
}

}  // namespace jm

using namespace jm;

extern "C" int jm_kinship_syrk_i8(const int8_t* gT, int64_t ld_gT, int64_t n, int64_t n_snps,
                                  int32_t* S, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(gT && S, "null pointer");
  JM_CHECK_ARG(ld_gT >= n_snps && ld_gT % 16 == 0 && ((uintptr_t)gT % 16) == 0,
               "gT rows must be 16-byte aligned with ld_gT >= n_snps");
  const int64_t nb = (n + K_BM - 1) / K_BM;
  JM_CHECK_ARG(nb * (nb + 1) / 2 < (1LL << 31), "n too large");
  hipLaunchKernelGGL(k_syrk_i8, dim3((unsigned)(nb * (nb + 1) / 2)), dim3(256), 0,
                     (hipStream_t)stream, gT, ld_gT, n, n_snps, S);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_kinship_impute_terms(const int8_t* g, int64_t ld_g, const int8_t* gT,
                                       int64_t ld_gT, int64_t n, int64_t n_snps, const double* mu,
                                       double* u, double* C, int has_missing, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(gT && mu && u, "null pointer");
  JM_CHECK_ARG(ld_gT >= n_snps, "ld_gT < n_snps");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_impute_u, dim3((unsigned)n), dim3(256), 0, st, gT, ld_gT, n_snps, mu, u);
  JM_CHECK_LAUNCH();
  if (has_missing) {
    JM_CHECK_ARG(g && C && ld_g >= n, "missing-call correction needs g and C");
    const dim3 grid((unsigned)n, (unsigned)ceil_div(n, 2048));
    hipLaunchKernelGGL(k_impute_C, grid, dim3(256), 0, st, g, ld_g, gT, ld_gT, n, n_snps, mu, C);
    JM_CHECK_LAUNCH();
  }
  return JM_OK;
}

extern "C" int jm_kinship_finalize(const int32_t* S, const double* u, const double* C, double s0,
                                   const int32_t* S_sub, const double* u_sub, const double* C_sub,
                                   double s0_sub, int64_t n, double p_snps, double* K,
                                   void* stream) {
  JM_CHECK_ARG(n >= 1 && p_snps > 0, "bad sizes");
  JM_CHECK_ARG(S && u && K, "null pointer");
  JM_CHECK_ARG((S_sub == nullptr) == (u_sub == nullptr), "S_sub and u_sub go together");
  hipLaunchKernelGGL(k_kin_finalize, dim3(4096), dim3(256), 0, (hipStream_t)stream, S, u, C, s0,
                     S_sub, u_sub, C_sub, s0_sub, n, p_snps, K);
  JM_CHECK_LAUNCH();
  return JM_OK;
}
