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

// S += X0 X0^T on 256x256 lower tiles (row block >= column block), for X0^T
// with missing calls already 0 (jm_transpose_i8's out0). 4 waves, each a
// 128x128 tile of 8x8 MFMA tiles (int32 accumulators in AGPRs). Both operand
// panels (256 rows x 128 bytes each) are staged straight from global memory
// into double-buffered LDS with global_load_lds_dwordx4, one barrier per
// 128-deep K step; rows are XOR-swizzled in 16-byte chunks (chunk c of row r at
// c ^ ((r >> 1) & 7): with 128-byte rows the 4-bank group of a chunk is
// 8 (r & 1) + c, so this gives 16 distinct groups to the 16 lanes of a
// ds_read_b128 phase). Lower tiles are enumerated XCD-major: the blocks resident on
// one XCD sweep consecutive tiles of the same row panel, which then stays in
// that XCD's L2.
constexpr int S2_B = 256, S2_BK = 128;
constexpr int S2_PANEL = S2_B * S2_BK;  // 32 KB per operand
constexpr int S2_STAGE = 2 * S2_PANEL;  // 64 KB

typedef __attribute__((address_space(3))) void lds_void_t;

__global__ void __launch_bounds__(256) k_syrk_i8_v2(const int8_t* __restrict__ gT0, int64_t ld,
                                                    int64_t n, int64_t K,
                                                    int32_t* __restrict__ S) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * S2_STAGE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = (int)((n + S2_B - 1) / S2_B);
  const int total = nb * (nb + 1) / 2;
  int t;
  {  // XCD-major bijection of the lower tiles
    const int x = blockIdx.x % 8, j = blockIdx.x / 8;
    const int base = total / 8, rem = total % 8;
    t = x * base + (x < rem ? x : rem) + j;
  }
  int rb, cb;
  tri_block(t, rb, cb);
  const int64_t i0 = (int64_t)rb * S2_B, j0 = (int64_t)cb * S2_B;
  const int wr = (wave >> 1) * 128, wc = (wave & 1) * 128;
  const int lr = lane & 15, lk = lane >> 4;
  // DMA: per operand 32 instructions of 8 rows x 128 B; each wave issues 8 of each
  const int d_r = lane >> 3, d_cs = lane & 7;
  auto issue = [&](int stage, int64_t k0) {
    uint8_t* base = lds + stage * S2_STAGE;
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t r0 = op ? j0 : i0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = wave * 8 + u;  // 0..31
        const int row = 8 * q + d_r;
        int64_t gr = r0 + row;
        if (gr >= n) gr = n - 1;  // rows past the end load a valid row; never stored
        const int c = d_cs ^ ((row >> 1) & 7);
        __builtin_amdgcn_global_load_lds(gT0 + gr * ld + k0 + 16 * c,
                                         (lds_void_t*)(base + op * S2_PANEL + q * 1024), 16, 0, 0);
      }
    }
  };
  i4_t acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = (i4_t){0, 0, 0, 0};
  const int nk = (int)((K + S2_BK - 1) / S2_BK);
  issue(0, 0);
  for (int kt = 0; kt < nk; ++kt) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (kt + 1 < nk) issue((kt + 1) & 1, (int64_t)(kt + 1) * S2_BK);
    const uint8_t* base = lds + (kt & 1) * S2_STAGE;
#pragma unroll
    for (int kh = 0; kh < 2; ++kh) {  // two 64-deep MFMA steps per stage
      i4_t af[8], bf[8];
#pragma unroll
      for (int ti = 0; ti < 8; ++ti) {
        const int r = wr + 16 * ti + lr;
        const int c = (4 * kh + lk) ^ ((r >> 1) & 7);
        memcpy(&af[ti], base + r * S2_BK + 16 * c, 16);
      }
#pragma unroll
      for (int tj = 0; tj < 8; ++tj) {
        const int r = wc + 16 * tj + lr;
        const int c = (4 * kh + lk) ^ ((r >> 1) & 7);
        memcpy(&bf[tj], base + S2_PANEL + r * S2_BK + 16 * c, 16);
      }
#pragma unroll
      for (int ti = 0; ti < 8; ++ti)
#pragma unroll
        for (int tj = 0; tj < 8; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[ti], bf[tj], acc[ti][tj], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ti = 0; ti < 8; ++ti)
#pragma unroll
    for (int tj = 0; tj < 8; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t j = j0 + wc + tj * 16 + lr;
        if (i < n && j < n) S[i * n + j] += acc[ti][tj][r];
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
// (in SNP order), then each thread accumulates 8 consecutive samples i. Per
// SNP the contribution mu_j * Y_ij is one of four values indexed by the 2-bit
// code of g_ij (0, mu, 2 mu, missing -> mu^2 / 2), so the inner loop is a
// select and an add; four SNPs are in flight at a time.
constexpr int C_CHUNK = 4096;
__device__ __forceinline__ double code_val(uint32_t code, double v1, double v2, double vm) {
  const double lo = (code & 1u) ? v1 : 0.0;
  const double hi = (code & 1u) ? vm : v2;
  return (code & 2u) ? hi : lo;
}

__global__ void __launch_bounds__(256) k_impute_C(const int8_t* __restrict__ g, int64_t ld_g,
                                                  const int8_t* __restrict__ gT, int64_t ld_gT,
                                                  int64_t n, int64_t K,
                                                  const double* __restrict__ mu,
                                                  double* __restrict__ C) {
  __shared__ int32_t list[C_CHUNK];
  __shared__ int32_t wsum[5];
  const int64_t k = blockIdx.x;
  const int64_t ibase = (int64_t)blockIdx.y * 2048 + threadIdx.x * 8;
  const bool full = ibase + 8 <= n;  // 8 samples in range: one 8-byte load per SNP
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
    if (ibase < n) {
      int q = 0;
      for (; q + 4 <= total; q += 4) {
        uint64_t v[4];
        double m[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t j = list[q + u];
          m[u] = mu[j];
          const int8_t* gr = g + j * ld_g + ibase;
          if (full) {
            v[u] = *reinterpret_cast<const uint64_t*>(gr);
          } else {
            v[u] = 0;
            for (int e = 0; e < 8; ++e)
              if (ibase + e < n) v[u] |= (uint64_t)(uint8_t)gr[e] << (8 * e);
          }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double v1 = m[u], v2 = 2.0 * m[u], vm = m[u] * (0.5 * m[u]);
#pragma unroll
          for (int e = 0; e < 8; ++e)
            acc[e] += code_val((uint32_t)(v[u] >> (8 * e)) & 3u, v1, v2, vm);
        }
      }
      for (; q < total; ++q) {
        const int64_t j = list[q];
        const double mj = mu[j];
        const int8_t* gr = g + j * ld_g + ibase;
        const double v1 = mj, v2 = 2.0 * mj, vm = mj * (0.5 * mj);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (ibase + e < n) acc[e] += code_val((uint32_t)(uint8_t)gr[e] & 3u, v1, v2, vm);
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
    const int64_t src = (i >= k) ? idx : (k * n + i);  // lower triangle of S
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

extern "C" int jm_kinship_syrk_i8(const int8_t* gT0, int64_t ld_gT, int64_t n, int64_t n_snps,
                                  int32_t* S, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(gT0 && S, "null pointer");
  JM_CHECK_ARG(ld_gT >= n_snps && ld_gT % S2_BK == 0 && ((uintptr_t)gT0 % 16) == 0,
               "gT0 rows must be 16-byte aligned with ld_gT >= n_snps, a multiple of 128");
  const int64_t nb = (n + S2_B - 1) / S2_B;
  JM_CHECK_ARG(nb * (nb + 1) / 2 < (1LL << 31), "n too large");
  hipLaunchKernelGGL(k_syrk_i8_v2, dim3((unsigned)(nb * (nb + 1) / 2)), dim3(256), 0,
                     (hipStream_t)stream, gT0, ld_gT, n, n_snps, S);
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
    JM_CHECK_ARG(ld_g % 8 == 0 && ((uintptr_t)g % 8) == 0, "g rows must be 8-byte aligned");
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
