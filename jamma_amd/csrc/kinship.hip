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

// ---------------------------------------------------------------------------
// S += X0 X0^T on int8 MFMA
// ---------------------------------------------------------------------------
// 256x256 lower tiles (row panel >= column panel) of S, 4 waves of 128x128
// (8 x 8 v_mfma_i32_16x16x64_i8 tiles, 256 int32 accumulators per lane in
// AGPRs, one wave per SIMD). Both 256-row operand panels are copied 64 SNPs
// deep per stage (2 x 16 KB) straight into LDS (global_load_lds_dwordx4, SGPR
// base + per-lane offset), 4 stages, so the copies of the next 3 K steps are in
// flight while one is multiplied; one barrier per step. LDS rows are 64 B with
// 16-byte chunk c of row r at c ^ ((r >> 1) & 3): the 8 lanes of a
// ds_read_b128 quarter (rows r..r+7, same chunk) then cover all 32 banks.
//
// Work split: the tiles go in groups of 4 row panels, column-major inside a
// group, and each XCD takes a contiguous range of that order, so the 32 blocks
// resident on one XCD work on a 4 x 8 patch and share its panels in L2. Whole
// rounds of 256 tiles run as plain read-modify-write blocks; the tiles of a
// last, partial round are split along K into parts that add into S with
// integer atomics (exact, any order), so that round is as full as the others.
constexpr int SY_B = 256, SY_BK = 64, SY_NBUF = 4;
constexpr int SY_PANEL = SY_B * SY_BK;    // 16 KB per operand per stage
constexpr int SY_STAGE = 2 * SY_PANEL;    // 32 KB
constexpr int SY_CPO = SY_PANEL / 1024 / 4;  // 1 KB copies per operand per wave per stage (4)
constexpr int SY_CPW = 2 * SY_CPO;        // copies per wave per stage (8)

__device__ __forceinline__ void tri_block(int t, int& rb, int& cb) {
  int r = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  while (r * (r + 1) / 2 > t) --r;
  rb = r;
  cb = t - r * (r + 1) / 2;
}

// Tile t of the group order: groups of 4 row panels (tiles before group g:
// sum_{r < 4g} (r + 1) = 8g^2 + 2g), column-major inside a group (rows R0..R1
// for columns <= R0, rows cb..R1 for the columns of the group's own triangle).
__device__ __forceinline__ void group_tile(int t, int nb, int& rb, int& cb) {
  int g = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.125f);
  while (8 * (g + 1) * (g + 1) + 2 * (g + 1) <= t) ++g;
  while (8 * g * g + 2 * g > t) --g;
  const int R0 = 4 * g, R1 = min(4 * g + 3, nb - 1), h = R1 - R0 + 1;
  int u = t - (8 * g * g + 2 * g);
  if (u < (R0 + 1) * h) {
    cb = u / h;
    rb = R0 + u % h;
    return;
  }
  u -= (R0 + 1) * h;
  cb = R0 + 1;
  while (u >= R1 - cb + 1) {
    u -= R1 - cb + 1;
    ++cb;
  }
  rb = cb + u;
}

// Copy 16 bytes per lane from sbase + voff (SGPR base, 32-bit lane offset) to
// lds_base + 16 * lane; ordering as for lds_copy16.
__device__ __forceinline__ void lds_copy16_s(const void* sbase, uint32_t voff,
                                             const uint8_t* lds_base) {
  const uint32_t m0v = (uint32_t)(uintptr_t)(const lds_u8_t*)lds_base;
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1"
               :
               : "v"(voff), "s"(sbase), "s"(m0v)
               : "memory", "m0");
}

// ATOMIC = false: tile t = XCD-mapped block id < n_full, whole K, S += acc.
// ATOMIC = true: block u -> tile n_full + u / split, K part u % split, atomicAdd.
template <bool ATOMIC>
__global__ void __launch_bounds__(256) k_syrk_i8(const int8_t* __restrict__ gT0, int64_t ld,
                                                 int64_t n, int64_t K, int32_t* __restrict__ S,
                                                 int n_full, int split) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[SY_NBUF * SY_STAGE];
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nb = (int)((n + SY_B - 1) / SY_B);
  int t, part = 0;
  if (!ATOMIC) {
    const int x = blockIdx.x % 8, j = blockIdx.x / 8;
    const int base = n_full / 8, rem = n_full % 8;
    t = x * base + (x < rem ? x : rem) + j;
  } else {
    t = n_full + blockIdx.x / split;
    part = blockIdx.x % split;
  }
  int rb, cb;
  group_tile(t, nb, rb, cb);
  const int nk_all = (int)((K + SY_BK - 1) / SY_BK);
  const int kb0 = ATOMIC ? (int)((int64_t)nk_all * part / split) : 0;
  const int nk = (ATOMIC ? (int)((int64_t)nk_all * (part + 1) / split) : nk_all) - kb0;
  const int64_t i0 = (int64_t)rb * SY_B, j0 = (int64_t)cb * SY_B;
  const int wr = (wave >> 1) * 128, wc = (wave & 1) * 128;
  const int lr = lane & 15, lk = lane >> 4;
  // a copy moves 16 rows x 64 B: lane -> row d_r, slot d_s; the chunk in the slot is
  // d_s ^ ((row >> 1) & 3), which depends on the lane only (16-row copies)
  const int d_r = lane >> 2, d_s = lane & 3;
  const uint32_t voff = (uint32_t)(d_r * ld + 16 * (d_s ^ ((d_r >> 1) & 3)));
  auto issue = [&](uint8_t* stage, int64_t k0) {
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t r0 = op ? j0 : i0;
#pragma unroll
      for (int u = 0; u < SY_CPO; ++u) {
        const int q = wave * SY_CPO + u;  // rows 16q .. 16q+15 of the panel
        lds_copy16_s(gT0 + (r0 + 16 * q) * ld + k0, voff, stage + op * SY_PANEL + q * 1024);
      }
    }
  };
  i4_t acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = (i4_t){0, 0, 0, 0};
#pragma unroll
  for (int st = 0; st < SY_NBUF - 1; ++st)
    if (st < nk) issue(lds + st * SY_STAGE, (int64_t)(kb0 + st) * SY_BK);
  int cur = 0;  // LDS stage of K step kt
  for (int kt = 0; kt < nk; ++kt) {
    // this wave's copies of steps kt .. kt+NBUF-2 may be in flight (in order);
    // keep the later ones, wait for step kt
    if (kt + SY_NBUF - 2 < nk)
      wait_vmcnt<SY_CPW * (SY_NBUF - 2)>();
    else
      wait_vmcnt<0>();
    __syncthreads();  // every wave's copies of step kt landed; step kt-1's stage is free
    if (kt + SY_NBUF - 1 < nk) {
      const int nxt = cur == 0 ? SY_NBUF - 1 : cur - 1;
      issue(lds + nxt * SY_STAGE, (int64_t)(kb0 + kt + SY_NBUF - 1) * SY_BK);
    }
    const uint8_t* base = lds + cur * SY_STAGE;
    i4_t af[8], bf[8];
#pragma unroll
    for (int ti = 0; ti < 8; ++ti) {
      const int r = wr + 16 * ti + lr;
      memcpy(&af[ti], base + r * SY_BK + 16 * (lk ^ ((r >> 1) & 3)), 16);
    }
#pragma unroll
    for (int tj = 0; tj < 8; ++tj) {
      const int r = wc + 16 * tj + lr;
      memcpy(&bf[tj], base + SY_PANEL + r * SY_BK + 16 * (lk ^ ((r >> 1) & 3)), 16);
    }
#pragma unroll
    for (int ti = 0; ti < 8; ++ti)
#pragma unroll
      for (int tj = 0; tj < 8; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[ti], bf[tj], acc[ti][tj], 0, 0, 0);
    cur = cur + 1 == SY_NBUF ? 0 : cur + 1;
  }
#pragma unroll
  for (int ti = 0; ti < 8; ++ti) {
    // one 16-row group at a time: bounds the read-modify-write loads in flight
    asm volatile("" ::: "memory");
#pragma unroll
    for (int tj = 0; tj < 8; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t j = j0 + wc + tj * 16 + lr;
        if (i < n && j < n) {
          if (ATOMIC)
            atomicAdd(&S[i * n + j], acc[ti][tj][r]);
          else
            S[i * n + j] += acc[ti][tj][r];
        }
      }
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
  JM_CHECK_ARG(ld_gT >= n_snps && ld_gT % 128 == 0 && ((uintptr_t)gT0 % 16) == 0,
               "gT0 rows must be 16-byte aligned with ld_gT >= n_snps, a multiple of 128");
  JM_CHECK_ARG(16 * ld_gT < (1LL << 32), "ld_gT too large");
  const int64_t nb = (n + SY_B - 1) / SY_B;
  JM_CHECK_ARG(nb * (nb + 1) / 2 < (1LL << 28), "n too large");
  const int total = (int)(nb * (nb + 1) / 2);
  const int nk = (int)((n_snps + SY_BK - 1) / SY_BK);
  // whole rounds of 256 tiles run unsplit; the tiles of a partial last round are
  // split along K (at least 4 steps per part) to fill that round
  const int tail = total % 256;
  int split = tail ? 256 / tail : 1;
  if (split > nk / 4) split = nk / 4;
  if (split < 1) split = 1;
  const int n_full = split > 1 ? total - tail : total;
  hipStream_t st = (hipStream_t)stream;
  if (n_full > 0) {
    hipLaunchKernelGGL(k_syrk_i8<false>, dim3(n_full), dim3(256), 0, st, gT0, ld_gT, n, n_snps, S,
                       n_full, 1);
    JM_CHECK_LAUNCH();
  }
  if (total > n_full) {
    hipLaunchKernelGGL(k_syrk_i8<true>, dim3((total - n_full) * split), dim3(256), 0, st, gT0,
                       ld_gT, n, n_snps, S, n_full, split);
    JM_CHECK_LAUNCH();
  }
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
