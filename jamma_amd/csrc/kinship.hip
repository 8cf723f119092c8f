// Centred kinship (GEMMA -gk 1): exact integer GEMM on MFMA with exact FP64 corrections.
//
// Reference: K = Xc Xc^T / p with Xc = impute_and_center(X) per batch
// (src/jamma/kinship/compute.py:46-57, 93-226, 502-690;
//  src/jamma/kinship/missing.py:22-61). With X0 = genotypes with missing calls
// set to 0, M the missing mask, mu_j the float64 nanmean and D = diag(mu):
//   Xc = X0 - (1 - M) D
//   Xc Xc^T = X0 X0^T - u 1^T - 1 u^T + s0 11^T + C + C^T
// where u_i = sum_j (observed ? g_ij mu_j : mu_j^2), s0 = sum_j mu_j^2 and
// C = Y D M^T with Y_ij = g_ij (observed) or mu_j / 2 (missing). X0 X0^T is an
// exact integer GEMM (int32 S) on the FP4 MFMA from 2-bit packed codes (or on
// v_mfma_i32_16x16x64_i8); the corrections are FP64 and only C is sparse work.

#include "jm_common.h"

namespace jm {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// S += X0 X0^T on int8 MFMA
// ---------------------------------------------------------------------------
// 256x256 lower tiles (row panel >= column panel) of S, 8 waves of 128x64
// (8 x 4 v_mfma_i32_16x16x64_i8 tiles, 128 int32 accumulators per lane, two
// waves per SIMD: a single wave cannot issue these MFMAs back to back, two keep
// the unit fed). Both 256-row operand panels are copied 64 SNPs
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
// NW = 4: 2 x 2 waves of 128 x 128 (one wave per SIMD); NW = 8: 2 x 4 waves of
// 128 x 64 (two waves per SIMD, 128 accumulators each).
template <bool ATOMIC, int NW>
__global__ void __launch_bounds__(64 * NW) k_syrk_i8(const int8_t* __restrict__ gT0, int64_t ld,
                                                     int64_t n, int64_t K, int32_t* __restrict__ S,
                                                     int n_full, int split) {
  constexpr int WR = 8, WC = NW == 4 ? 8 : 4;  // 16-row / 16-column blocks per wave
  constexpr int NWC = 256 / (16 * WC);         // waves across the 256 columns
  constexpr int SY_CPO = 16 / NW;              // 1 KB copies per operand per wave per stage
  constexpr int SY_CPW = 2 * SY_CPO;           // copies per wave per stage
  static_assert((NW == 4 || NW == 8) && (NW / NWC) * WR * 16 == SY_B, "wave grid");
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
  const int wr = (wave / NWC) * 16 * WR, wc = (wave % NWC) * 16 * WC;
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
  i4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (i4_t){0, 0, 0, 0};
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
    i4_t af[WR], bf[WC];
#pragma unroll
    for (int ti = 0; ti < WR; ++ti) {
      const int r = wr + 16 * ti + lr;
      memcpy(&af[ti], base + r * SY_BK + 16 * (lk ^ ((r >> 1) & 3)), 16);
    }
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) {
      const int r = wc + 16 * tj + lr;
      memcpy(&bf[tj], base + SY_PANEL + r * SY_BK + 16 * (lk ^ ((r >> 1) & 3)), 16);
    }
#pragma unroll
    for (int ti = 0; ti < WR; ++ti)
#pragma unroll
      for (int tj = 0; tj < WC; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[ti], bf[tj], acc[ti][tj], 0, 0, 0);
    cur = cur + 1 == SY_NBUF ? 0 : cur + 1;
  }
#pragma unroll
  for (int ti = 0; ti < WR; ++ti) {
    // one 16-row group at a time: bounds the read-modify-write loads in flight
    asm volatile("" ::: "memory");
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
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

// ---------------------------------------------------------------------------
// S += X0 X0^T from 2-bit packed genotypes
// ---------------------------------------------------------------------------
// The int8 GEMM above moves 32 KB per 64 SNPs through L2 and LDS for every
// 256x256 tile and is bound by that stream (without its MFMAs it takes 90 % of
// its time). Hard-call genotypes with missing calls as 0 need two bits, so this
// form keeps the operand 2-bit packed in HBM, L2 and LDS (a quarter of the bytes)
// and widens it to int8 in registers right before the MFMAs.
//
// Packed operand gP (jm_pack_gram_2bit): sample-major rows of ldp bytes; SNP
// group b (SNPs 256b .. 256b + 255) of a row is 64 bytes, and its 16-SNP group
// Q (SNPs 256b + 16Q ..) is the 32-bit word at byte 64b + 16 (Q & 3) + 4 (Q >> 2)
// holding the code of SNP 256b + 16Q + e in bits 2e, 2e + 1 (0, 1, 2; missing 0).
// So the 16 bytes of a row at 16 lk hold the four words of groups lk, lk + 4,
// lk + 8, lk + 12: the SNPs that lane group lk contributes to each of the four
// 64-deep MFMA steps of the group (sub-step s = Q >> 2). A word widens to the
// four dwords (w >> 2k) & 0x03030303, k = 0..3: byte j of dword k is the code of
// SNP 16Q + 4j + k, a fixed permutation of the word's 16 K positions that the A
// and B fragments share, so the int32 products are those of the int8 GEMM.
//
// Kernel: the same 256x256 lower tiles, XCD patches, 8 waves of 128 x 64 and
// K-split tail as k_syrk_i8, 4 LDS stages of 2 x 16 KB, each stage 256 SNPs deep
// (one barrier per 256 SNPs); per stage a wave reads its 12 fragments once
// (ds_read_b128) and runs 4 sub-steps of 32 MFMAs, widening a fragment's word
// (7 VALU) before the MFMAs that use it.
//
// Measured at n = 10,000 over 99,952 SNPs (one launch): 3.38 ms against 4.45 ms for
// k_syrk_i8 (3.43 ms without the software-pipelined widening), bit-identical S.
// Tried: widening once per block into an int8 LDS stage from register-staged packed
// loads (4.25 ms: each lane's 16-byte loads touch 64 rows per instruction, and the
// int8 stage adds 32 KB of LDS writes per 64 SNPs), and a fixed MFMA/VALU
// interleave order (3.54 ms).
constexpr int SP_G = 256;  // SNPs per stage (64 bytes of a packed row)

__device__ __forceinline__ i4_t widen_codes(uint32_t w) {
  const uint32_t m = 0x03030303u;
  return (i4_t){(int)(w & m), (int)((w >> 2) & m), (int)((w >> 4) & m), (int)((w >> 6) & m)};
}

template <bool ATOMIC>
__global__ void __launch_bounds__(512) k_syrk_p2(const uint8_t* __restrict__ gP, int64_t ldp,
                                                 int64_t n, int64_t K, int32_t* __restrict__ S,
                                                 int n_full, int split) {
  constexpr int WR = 8, WC = 4, NWC = 4;
  constexpr int CPO = 2;        // 1 KB copies per operand per wave per stage
  constexpr int CPW = 2 * CPO;  // copies per wave per stage
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
  const int nk_all = (int)((K + SP_G - 1) / SP_G);
  const int kb0 = ATOMIC ? (int)((int64_t)nk_all * part / split) : 0;
  const int nk = (ATOMIC ? (int)((int64_t)nk_all * (part + 1) / split) : nk_all) - kb0;
  const int64_t i0 = (int64_t)rb * SY_B, j0 = (int64_t)cb * SY_B;
  const int wr = (wave / NWC) * 16 * WR, wc = (wave % NWC) * 16 * WC;
  const int lr = lane & 15, lk = lane >> 4;
  // a copy moves 16 rows x 64 B: lane -> row d_r, LDS slot d_s holding the row's
  // 16-byte chunk d_s ^ ((row >> 1) & 3)
  const int d_r = lane >> 2, d_s = lane & 3;
  const uint32_t voff = (uint32_t)(d_r * ldp + 16 * (d_s ^ ((d_r >> 1) & 3)));
  auto issue = [&](uint8_t* stage, int64_t kb) {
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t r0 = op ? j0 : i0;
#pragma unroll
      for (int u = 0; u < CPO; ++u) {
        const int q = wave * CPO + u;  // rows 16q .. 16q+15 of the panel
        lds_copy16_s(gP + (r0 + 16 * q) * ldp + kb * (SP_G / 4), voff,
                     stage + op * SY_PANEL + q * 1024);
      }
    }
  };
  i4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (i4_t){0, 0, 0, 0};
#pragma unroll
  for (int st = 0; st < SY_NBUF - 1; ++st)
    if (st < nk) issue(lds + st * SY_STAGE, kb0 + st);
  int cur = 0;  // LDS stage of K step kt
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + SY_NBUF - 2 < nk)
      wait_vmcnt<CPW * (SY_NBUF - 2)>();
    else
      wait_vmcnt<0>();
    __syncthreads();  // every wave's copies of step kt landed; step kt-1's stage is free
    if (kt + SY_NBUF - 1 < nk) {
      const int nxt = cur == 0 ? SY_NBUF - 1 : cur - 1;
      issue(lds + nxt * SY_STAGE, kb0 + kt + SY_NBUF - 1);
    }
    const uint8_t* base = lds + cur * SY_STAGE;
    u4_t ar[WR], br[WC];  // the four sub-steps' words of each fragment
#pragma unroll
    for (int ti = 0; ti < WR; ++ti) {
      const int r = wr + 16 * ti + lr;
      ar[ti] = *reinterpret_cast<const u4_t*>(base + r * 64 + 16 * (lk ^ ((r >> 1) & 3)));
    }
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) {
      const int r = wc + 16 * tj + lr;
      br[tj] = *reinterpret_cast<const u4_t*>(base + SY_PANEL + r * 64 + 16 * (lk ^ ((r >> 1) & 3)));
    }
    // software-pipelined widening: the next row block's A fragment and (early in a
    // sub-step) the next sub-step's B fragments are widened while MFMAs issue
    i4_t bq[2][WC];
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) bq[0][tj] = widen_codes(br[tj][0]);
    i4_t aq = widen_codes(ar[0][0]);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
#pragma unroll
      for (int ti = 0; ti < WR; ++ti) {
        i4_t an = aq;
        if (ti + 1 < WR)
          an = widen_codes(ar[ti + 1][s]);
        else if (s + 1 < 4)
          an = widen_codes(ar[0][s + 1]);
        if (ti == 1 && s + 1 < 4) {
#pragma unroll
          for (int tj = 0; tj < WC; ++tj) bq[(s + 1) & 1][tj] = widen_codes(br[tj][s + 1]);
        }
#pragma unroll
        for (int tj = 0; tj < WC; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_i32_16x16x64_i8(aq, bq[s & 1][tj], acc[ti][tj], 0, 0, 0);
        aq = an;
      }
    }
    cur = cur + 1 == SY_NBUF ? 0 : cur + 1;
  }
#pragma unroll
  for (int ti = 0; ti < WR; ++ti) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
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

// ---------------------------------------------------------------------------
// The same product on the FP4 MFMA
// ---------------------------------------------------------------------------
// v_mfma_scale_f32_16x16x128_f8f6f4 with both operands in e2m1 multiplies 16 x 128
// by 128 x 16 in the time v_mfma_i32_16x16x64_i8 takes for 16 x 64: twice the
// int8 rate. A 2-bit code c in the low two bits of a nibble is the e2m1 value c / 2
// (0, 0.5 subnormal, 1.0); a block scale of 2 (E8M0 128) on each operand makes the
// products c1 c2. Every product and partial sum is an integer below 2^24 while a
// launch covers at most 2^22 SNPs (|sum| <= 4 K), so the FP32 accumulation is exact
// in any order and S gets the same integers as from the int8 MFMA.
//
// Operand, LDS stages, tiles, XCD patches and K-split tail as k_syrk_p2. Lane group
// lk's 16 bytes of a row (words of 16-SNP groups lk, lk + 4, lk + 8, lk + 12) feed
// two 128-deep MFMAs per stage: words 0, 1 the first, words 2, 3 the second. A word
// widens to two dwords of 8 nibbles (w & 0x33333333 and (w >> 2) & 0x33333333: 3
// VALU per 16 codes); lane l of an MFMA operand holds K positions 32 (l >> 4) ..
// + 31 of its row in nibble order, so A and B see the same permutation of the
// stage's 256 SNPs.
constexpr int64_t SF4_KMAX = 1LL << 22;  // SNPs per launch for an exact FP32 sum

typedef int v8i_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v8i_t fp4_frag(uint32_t w0, uint32_t w1) {
  const uint32_t m = 0x33333333u;
  return (v8i_t){(int)(w0 & m), (int)((w0 >> 2) & m), (int)(w1 & m), (int)((w1 >> 2) & m),
                 0, 0, 0, 0};
}

template <bool ATOMIC>
__global__ void __launch_bounds__(512) k_syrk_f4(const uint8_t* __restrict__ gP, int64_t ldp,
                                                 int64_t n, int64_t K, int32_t* __restrict__ S,
                                                 int n_full, int split) {
  constexpr int WR = 8, WC = 4, NWC = 4;
  constexpr int CPO = 2;        // 1 KB copies per operand per wave per stage
  constexpr int CPW = 2 * CPO;  // copies per wave per stage
  constexpr int SC = 128;       // E8M0 block scale 2
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
  const int nk_all = (int)((K + SP_G - 1) / SP_G);
  const int kb0 = ATOMIC ? (int)((int64_t)nk_all * part / split) : 0;
  const int nk = (ATOMIC ? (int)((int64_t)nk_all * (part + 1) / split) : nk_all) - kb0;
  const int64_t i0 = (int64_t)rb * SY_B, j0 = (int64_t)cb * SY_B;
  const int wr = (wave / NWC) * 16 * WR, wc = (wave % NWC) * 16 * WC;
  const int lr = lane & 15, lk = lane >> 4;
  const int d_r = lane >> 2, d_s = lane & 3;
  const uint32_t voff = (uint32_t)(d_r * ldp + 16 * (d_s ^ ((d_r >> 1) & 3)));
  auto issue = [&](uint8_t* stage, int64_t kb) {
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t r0 = op ? j0 : i0;
#pragma unroll
      for (int u = 0; u < CPO; ++u) {
        const int q = wave * CPO + u;  // rows 16q .. 16q+15 of the panel
        lds_copy16_s(gP + (r0 + 16 * q) * ldp + kb * (SP_G / 4), voff,
                     stage + op * SY_PANEL + q * 1024);
      }
    }
  };
  f4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < SY_NBUF - 1; ++st)
    if (st < nk) issue(lds + st * SY_STAGE, kb0 + st);
  int cur = 0;  // LDS stage of K step kt
  for (int kt = 0; kt < nk; ++kt) {
    if (kt + SY_NBUF - 2 < nk)
      wait_vmcnt<CPW * (SY_NBUF - 2)>();
    else
      wait_vmcnt<0>();
    __syncthreads();  // every wave's copies of step kt landed; step kt-1's stage is free
    if (kt + SY_NBUF - 1 < nk) {
      const int nxt = cur == 0 ? SY_NBUF - 1 : cur - 1;
      issue(lds + nxt * SY_STAGE, kb0 + kt + SY_NBUF - 1);
    }
    const uint8_t* base = lds + cur * SY_STAGE;
    u4_t ar[WR], br[WC];
#pragma unroll
    for (int ti = 0; ti < WR; ++ti) {
      const int r = wr + 16 * ti + lr;
      ar[ti] = *reinterpret_cast<const u4_t*>(base + r * 64 + 16 * (lk ^ ((r >> 1) & 3)));
    }
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) {
      const int r = wc + 16 * tj + lr;
      br[tj] = *reinterpret_cast<const u4_t*>(base + SY_PANEL + r * 64 + 16 * (lk ^ ((r >> 1) & 3)));
    }
    // widening pipelined as in k_syrk_p2: the next row block's A fragment and the
    // second MFMA's B fragments are formed while MFMAs issue
    v8i_t bq[2][WC];
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) bq[0][tj] = fp4_frag(br[tj][0], br[tj][1]);
    v8i_t aq = fp4_frag(ar[0][0], ar[0][1]);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
#pragma unroll
      for (int ti = 0; ti < WR; ++ti) {
        v8i_t an = aq;
        if (ti + 1 < WR)
          an = fp4_frag(ar[ti + 1][2 * s], ar[ti + 1][2 * s + 1]);
        else if (s == 0)
          an = fp4_frag(ar[0][2], ar[0][3]);
        if (ti == 1 && s == 0) {
#pragma unroll
          for (int tj = 0; tj < WC; ++tj) bq[1][tj] = fp4_frag(br[tj][2], br[tj][3]);
        }
#pragma unroll
        for (int tj = 0; tj < WC; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
              aq, bq[s][tj], acc[ti][tj], 4, 4, 0, SC, 0, SC);
        aq = an;
      }
    }
    cur = cur + 1 == SY_NBUF ? 0 : cur + 1;
  }
#pragma unroll
  for (int ti = 0; ti < WR; ++ti) {
    asm volatile("" ::: "memory");
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t j = j0 + wc + tj * 16 + lr;
        if (i < n && j < n) {
          const int32_t v = (int32_t)acc[ti][tj][r];  // an exact integer
          if (ATOMIC)
            atomicAdd(&S[i * n + j], v);
          else
            S[i * n + j] += v;
        }
      }
  }
}

// Pack a SNP-major int8 tile (n_snps rows of ld_g, -1 = missing) into the 2-bit
// operand at SNP column `col0` (a multiple of 16) of gP. One block = 64 samples x
// 256 SNPs of the tile, staged through LDS (16-byte loads); each thread then
// builds 4 words (16 SNPs of one sample each). Words from col0 up to the end of
// the 256-SNP group holding the tile's last SNP are all written (codes past the
// tile are 0): the GEMM reads whole groups, which then hold zeros past the last
// SNP until a later tile fills them.
__global__ void __launch_bounds__(256) k_pack_gram_2bit(const int8_t* __restrict__ g, int64_t ld_g,
                                                        int64_t n_snps, int64_t n,
                                                        uint8_t* __restrict__ gP, int64_t ldp,
                                                        int64_t col0, int64_t n_words) {
  __shared__ uint32_t tile[256][17];  // SNP row x 64 samples (16 words of 4 code bytes)
  const int64_t s0 = (int64_t)blockIdx.x * 256, i0 = (int64_t)blockIdx.y * 64;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = t + 256 * q;
    const int row = idx >> 2, ch = (idx & 3) * 16;  // SNP row, first sample of 16
    const int64_t s = s0 + row, i = i0 + ch;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    if (s < n_snps && i < n) {
      const int8_t* src = g + s * ld_g + i;
      if (i + 16 <= n && ((uintptr_t)src % 16) == 0) {
        const uint4 v = *reinterpret_cast<const uint4*>(src);
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
      } else {
        for (int e = 0; e < 16 && i + e < n; ++e)
          w[e >> 2] |= (uint32_t)(uint8_t)src[e] << (8 * (e & 3));
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) tile[row][(ch >> 2) + e] = w[e];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = t + 256 * q;
    const int smp = idx & 63, grp = idx >> 6;  // sample, 16-SNP group of the block tile
    const int64_t i = i0 + smp;
    const int64_t wq = (s0 >> 4) + grp;  // word index relative to col0
    if (i >= n || wq >= n_words) continue;
    uint32_t word = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint32_t b = (tile[16 * grp + e][smp >> 2] >> (8 * (smp & 3))) & 0xFFu;
      const uint32_t code = (b & 0x80u) ? 0u : (b & 3u);  // missing -> 0
      word |= code << (2 * e);
    }
    const int64_t Q = (col0 >> 4) + wq;  // 16-SNP group of the row
    const int64_t off = (Q >> 4) * 64 + 16 * (Q & 3) + 4 * ((Q >> 2) & 3);
    *reinterpret_cast<uint32_t*>(gP + i * ldp + off) = word;
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

// C[k][i] += sum_{j: sample k missing at j} mu_j * Y_ij over one SNP tile, in four
// steps. The missing calls of every sample are first listed (k_miss_count,
// k_miss_scan, k_miss_fill: a CSR of SNP indices per sample, in SNP order, built once
// per tile), and the tile's genotypes are packed to 2-bit codes (k_pack_codes: the
// int8 value's low two bits, 0 / 1 / 2 / 3 = missing; 16 samples per 32-bit word).
// k_impute_C then gives each wavefront one sample k and 512 samples i (8 per lane);
// per listed SNP j the contribution is mu_j * x with x the code of g_ij converted to
// FP64 (0 / 1 / 2, and mu_j / 2 for a missing call), one fused multiply-add per
// element (the kernel is VALU-bound: a conversion, a compare, a 64-bit select and the
// FMA per element), one 4-byte load per two lanes, sixteen SNPs in flight, with the
// list and the means on scalar loads. The blocks are ordered so that each XCD sweeps
// the 512-sample chunks one after another: the codes of one chunk (128 B per SNP, 2 MB
// for 16k SNPs) stay in that XCD's L2 while all samples k read them. Each C element
// is summed by one lane in list order (deterministic).
constexpr int C_ICHUNK = 512;  // samples i per wavefront (8 per lane)
constexpr int C_WAVES = 4;     // samples k per block

// The factor multiplying mu_j in mu_j * Y_ij: the genotype 0 / 1 / 2 (code 0-2), or
// mu_j / 2 at a missing call (code 3, h = mu_j / 2): a conversion and one select.
__device__ __forceinline__ double code_x(uint32_t code, double h) {
  return code == 3u ? h : (double)code;
}

// cnt[k] = missing calls (-1) of sample k among the tile's K SNPs; gT rows are
// zero-padded to a multiple of 128 (16-byte loads stay inside the row).
__global__ void __launch_bounds__(256) k_miss_count(const int8_t* __restrict__ gT, int64_t ld,
                                                    int64_t n, int64_t K, int32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  int c = 0;
  for (int64_t j = (int64_t)lane * 16; j < K; j += 64 * 16) {
    const uint4 v = *reinterpret_cast<const uint4*>(gT + k * ld + j);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 16; ++e)
      c += (j + e < K && ((w[e >> 2] >> (8 * (e & 3))) & 0xFFu) == 0xFFu) ? 1 : 0;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m, 64);
  if (lane == 0) cnt[k] = c;
}

// In place: cnt[0..n) -> exclusive offsets off[0..n], off[n] = total (one block).
__global__ void __launch_bounds__(1024) k_miss_scan(int32_t* __restrict__ off, int64_t n) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t a = t * per, b = min(n, a + per);
  int64_t s = 0;
  for (int64_t i = a; i < b; ++i) s += off[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int q = 0; q < 1024; ++q) {
      const int64_t v = part[q];
      part[q] = run;
      run += v;
    }
    off[n] = (int32_t)run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t i = a; i < b; ++i) {
    const int32_t v = off[i];
    off[i] = (int32_t)run;
    run += v;
  }
}

// list[off[k] + q] = SNP index of the q-th missing call of sample k (SNP order);
// entries at or past cap are not written.
__global__ void __launch_bounds__(256) k_miss_fill(const int8_t* __restrict__ gT, int64_t ld,
                                                   int64_t n, int64_t K,
                                                   const int32_t* __restrict__ off,
                                                   int32_t* __restrict__ list, int64_t cap) {
  const int lane = threadIdx.x & 63;
  const int64_t k = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  int64_t pos = off[k];
  for (int64_t j0 = 0; j0 < K; j0 += 64 * 16) {
    const int64_t j = j0 + (int64_t)lane * 16;
    uint32_t mask = 0;
    if (j < K) {
      const uint4 v = *reinterpret_cast<const uint4*>(gT + k * ld + j);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 16; ++e)
        if (j + e < K && ((w[e >> 2] >> (8 * (e & 3))) & 0xFFu) == 0xFFu) mask |= 1u << e;
    }
    const int c = __popc(mask);
    int incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    int64_t p = pos + incl - c;
    for (int e = 0; e < 16; ++e)
      if (mask & (1u << e)) {
        if (p < cap) list[p] = (int32_t)(j + e);
        ++p;
      }
    pos += __shfl(incl, 63, 64);
  }
}

// gc[j][w] (ldc words per SNP row): the 2-bit codes of samples 16w .. 16w + 15 of
// SNP j (low bits first); samples past n are code 0.
__global__ void __launch_bounds__(256) k_pack_codes(const int8_t* __restrict__ g, int64_t ld_g,
                                                    int64_t n, int64_t K,
                                                    uint32_t* __restrict__ gc, int64_t ldc) {
  const int64_t words = (n + 15) / 16;
  const int64_t total = K * words;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / words, wd = e % words;
    const int64_t i0 = wd * 16;
    uint32_t code = 0;
    if (i0 + 16 <= n) {
      const uint4 v = *reinterpret_cast<const uint4*>(g + j * ld_g + i0);
      const uint32_t b[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int q = 0; q < 16; ++q) code |= ((b[q >> 2] >> (8 * (q & 3))) & 3u) << (2 * q);
    } else {
      for (int q = 0; i0 + q < n; ++q) code |= ((uint32_t)(uint8_t)g[j * ld_g + i0 + q] & 3u) << (2 * q);
    }
    gc[j * ldc + wd] = code;
  }
}

typedef const int32_t __attribute__((address_space(4))) c_i32;
typedef const double __attribute__((address_space(4))) c_f64;

__global__ void __launch_bounds__(256) k_impute_C(const uint32_t* __restrict__ gc, int64_t ldc,
                                                  int64_t n, const double* __restrict__ mu,
                                                  const int32_t* __restrict__ off,
                                                  const int32_t* __restrict__ list, int64_t cap,
                                                  double* __restrict__ C) {
  const int lane = threadIdx.x & 63;
  // wave-uniform values in scalar registers: the list, its offsets and the means then
  // come through scalar loads, and only the genotype codes are vector loads
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ni = (int)((n + C_ICHUNK - 1) / C_ICHUNK);
  const int nk = (int)((n + C_WAVES - 1) / C_WAVES);
  // XCD-contiguous ranges of the chunk-major order (blocks b, b + 8, ... share an XCD)
  const int total = ni * nk;
  const int x = blockIdx.x % 8, jb = blockIdx.x / 8;
  const int base = total / 8, rem = total % 8;
  const int t = x * base + (x < rem ? x : rem) + jb;
  const int ic = t / nk, kg = t % nk;
  const int64_t k = (int64_t)kg * C_WAVES + w;
  if (k >= n) return;
  const c_i32* soff = (const c_i32*)off;
  const c_i32* slist = (const c_i32*)list;
  const c_f64* smu = (const c_f64*)mu;
  const int64_t i0 = (int64_t)ic * C_ICHUNK + lane * 8;
  const bool live = i0 < n;
  const uint32_t* gl = gc + i0 / 16;  // word of this lane's 8 samples
  const int sh = (lane & 1) * 16;
  int64_t q0 = soff[k], q1 = soff[k + 1];
  if (q1 > cap) q1 = cap;
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  constexpr int U = 16;  // SNPs in flight
  int64_t q = q0;
  for (; q + U <= q1; q += U) {
    int32_t j[U];
    uint32_t v[U];
    double m[U];
#pragma unroll
    for (int u = 0; u < U; ++u) j[u] = slist[q + u];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      m[u] = smu[j[u]];
      v[u] = live ? gl[(int64_t)j[u] * ldc] >> sh : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double h = 0.5 * m[u];
#pragma unroll
      for (int e = 0; e < 8; ++e) acc[e] = fma(m[u], code_x((v[u] >> (2 * e)) & 3u, h), acc[e]);
    }
  }
  for (; q < q1; ++q) {
    const int64_t jj = slist[q];
    const double mj = smu[jj];
    const uint32_t vv = live ? gl[jj * ldc] >> sh : 0u;
    const double h = 0.5 * mj;
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = fma(mj, code_x((vv >> (2 * e)) & 3u, h), acc[e]);
  }
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (i0 + e < n) C[k * n + i0 + e] += acc[e];
}

// gc4[j][w] (ldc4 words per SNP row): the 2-bit codes of samples 8w .. 8w + 7 of SNP j,
// one per 4-bit slot (low slots first); samples past n are code 0.
__global__ void __launch_bounds__(256) k_pack_codes4(const int8_t* __restrict__ g, int64_t ld_g,
                                                     int64_t n, int64_t K,
                                                     uint32_t* __restrict__ gc4, int64_t ldc4) {
  const int64_t words = (n + 7) / 8;
  const int64_t total = K * words;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = e / words, wd = e % words;
    const int64_t i0 = wd * 8;
    uint64_t b = 0;
    if (i0 + 8 <= n) {
      b = *reinterpret_cast<const uint64_t*>(g + j * ld_g + i0);  // rows 8-byte aligned
    } else {
      for (int q = 0; i0 + q < n; ++q) b |= (uint64_t)(uint8_t)g[j * ld_g + i0 + q] << (8 * q);
    }
    uint32_t code = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) code |= (uint32_t)((b >> (8 * q)) & 3u) << (4 * q);
    gc4[j * ldc4 + wd] = code;
  }
}

// The same C[k][i] += sum_j mu_j Y_ij with the missing-call SNPs of sample k taken
// in pairs (j1, j2): the 16 possible values mu_j1 Y(c1) + mu_j2 Y(c2) of a pair (c =
// 2-bit code, Y(3) = mu / 2) are tabulated once per wavefront in LDS, and each lane
// then adds, per sample, the table entry that its two codes index. The codes come one
// per 4-bit slot (k_pack_codes4: a lane's 8 samples of a SNP are one word), so one OR
// of the pair's two words (w1 | w2 << 2) holds the 8 table indices, and a sample costs
// a bit-field extract, an address, an LDS read and one FP64 add per pair of SNPs,
// where the per-SNP form (k_impute_C) spends a conversion, a compare, a 64-bit select
// and an FMA per SNP. The table covers the missing code too, so the both-missing pairs
// need no separate pass. A last unpaired SNP pairs with a zero-weight partner. The 16
// entries of a table occupy 32 banks; lanes reading one entry share the access.
constexpr int CP_U = 8;  // pairs per table batch (16 SNPs of the list)

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__global__ void __launch_bounds__(256) k_impute_C2(const uint32_t* __restrict__ gc4, int64_t ldc4,
                                                   int64_t n, const double* __restrict__ mu,
                                                   const int32_t* __restrict__ off,
                                                   const int32_t* __restrict__ list, int64_t cap,
                                                   double* __restrict__ C) {
  __shared__ double tab[C_WAVES][CP_U * 16];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ni = (int)((n + C_ICHUNK - 1) / C_ICHUNK);
  const int nk = (int)((n + C_WAVES - 1) / C_WAVES);
  const int total = ni * nk;
  const int x = blockIdx.x % 8, jb = blockIdx.x / 8;
  const int base = total / 8, rem = total % 8;
  const int t = x * base + (x < rem ? x : rem) + jb;
  const int ic = t / nk, kg = t % nk;
  const int64_t k = (int64_t)kg * C_WAVES + w;
  if (k >= n) return;  // wave-uniform; the kernel has no block-level barrier
  const c_i32* soff = (const c_i32*)off;
  const c_i32* slist = (const c_i32*)list;
  const int64_t i0 = (int64_t)ic * C_ICHUNK + lane * 8;
  const bool live = i0 < n;
  const uint32_t* gl = gc4 + i0 / 8;  // the word of this lane's 8 samples
  int64_t q0 = soff[k], q1 = soff[k + 1];
  if (q1 > cap) q1 = cap;
  double* T = tab[w];
  double acc[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) acc[e] = 0.0;
  for (int64_t q = q0; q < q1; q += 2 * CP_U) {
    const int np = (int)min((int64_t)CP_U, (q1 - q + 1) / 2);  // pairs in this batch
    {  // tables: lane l fills entries 2l, 2l + 1 of pair l >> 3
      const int u = lane >> 3;
      if (u < np) {
        const int64_t qa = q + 2 * u, qb = qa + 1;
        const bool hb = qb < q1;
        const double ma = mu[list[qa]];
        const double mb = hb ? mu[list[qb]] : 0.0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int idx = (2 * lane + h) & 15;
          const int c1 = idx & 3, c2 = idx >> 2;
          const double y1 = c1 == 3 ? 0.5 * ma : (double)c1;
          const double y2 = c2 == 3 ? 0.5 * mb : (double)c2;
          T[u * 16 + idx] = fma(mb, y2, ma * y1);
        }
      }
    }
    // the pair index words of the batch (slot e of word u: the table index of sample e)
    uint32_t tw[CP_U];
#pragma unroll
    for (int u = 0; u < CP_U; ++u) {
      if (u < np) {
        const int64_t qa = q + 2 * u;
        const int64_t ja = slist[qa];
        const int64_t jb2 = qa + 1 < q1 ? (int64_t)slist[qa + 1] : ja;
        const uint32_t va = live ? gl[ja * ldc4] : 0u;
        const uint32_t vb = live ? gl[jb2 * ldc4] : 0u;
        tw[u] = va | (vb << 2);
      } else {
        tw[u] = 0u;
      }
    }
    wave_lds_sync();  // tables visible to every lane
#pragma unroll
    for (int u = 0; u < CP_U; ++u) {
      if (u < np) {
        const double* Tu = T + u * 16;
#pragma unroll
        for (int e = 0; e < 8; ++e) acc[e] += Tu[__builtin_amdgcn_ubfe(tw[u], 4 * e, 4)];
      }
    }
    wave_lds_sync();  // tables read before the next batch overwrites them
  }
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (i0 + e < n) C[k * n + i0 + e] += acc[e];
}

// K for the tile pair (ti, tk), ti >= tk, of 32 x 32 blocks: both K[i][k] (rows of
// block ti) and its mirror K[k][i] are written, and C[i][k] + C[k][i] is formed from
// two tiles staged in LDS, so every global access is a coalesced row segment (the
// element-wise form reads C and the upper triangle of S column-wise). S is read
// from its lower triangle only.
constexpr int KF_T = 32;
__global__ void __launch_bounds__(256) k_kin_finalize_tiled(
    const int32_t* __restrict__ S, const double* __restrict__ u, const double* __restrict__ C,
    double s0, const int32_t* __restrict__ Ss, const double* __restrict__ us,
    const double* __restrict__ Cs, double s0s, int64_t n, double p, double* __restrict__ K) {
  __shared__ double ct[KF_T][KF_T + 1];  // C[k][i] - Cs[k][i] of the mirror tile, as [k][i]
  __shared__ double kt[KF_T][KF_T + 1];  // K tile [i][k], read back transposed
  __shared__ int32_t st[KF_T][KF_T + 1]; // diagonal tile: S[i][k] for the mirror entries
  // linear block -> lower-triangle tile pair (ti, tk)
  const int64_t b = blockIdx.x;
  int64_t ti = (int64_t)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= b) ++ti;
  while (ti * (ti + 1) / 2 > b) --ti;
  const int64_t tk = b - ti * (ti + 1) / 2;
  const int64_t i0 = ti * KF_T, k0 = tk * KF_T;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 8 rows x 32 columns per pass
  // mirror tile: C[k][i] for k in block tk (rows), i in block ti (columns)
  for (int r = ty; r < KF_T; r += 8) {
    const int64_t k = k0 + r, i = i0 + tx;
    double c = 0.0;
    if (k < n && i < n) {
      if (C) c = C[k * n + i];
      if (Cs) c -= Cs[k * n + i];
    }
    ct[r][tx] = c;
    if (ti == tk) st[r][tx] = (k < n && i < n) ? S[k * n + i] - (Ss ? Ss[k * n + i] : 0) : 0;
  }
  __syncthreads();
  const double sv0 = s0 - (us ? s0s : 0.0);
  for (int r = ty; r < KF_T; r += 8) {
    const int64_t i = i0 + r, k = k0 + tx;
    double kv = 0.0;
    if (i < n && k < n) {
      int64_t sv;
      if (ti != tk || r >= tx) {  // lower triangle
        sv = S[i * n + k];
        if (Ss) sv -= Ss[i * n + k];
      } else {
        sv = st[tx][r];  // S[k][i], k > i: its lower-triangle entry
      }
      double ui = u[i], uk = u[k];
      if (us) {
        ui -= us[i];
        uk -= us[k];
      }
      double c = 0.0;
      if (C) c = C[i * n + k];
      if (Cs) c -= Cs[i * n + k];
      c += ct[tx][r];
      kv = ((double)sv - ui - uk + sv0 + c) / p;
      K[i * n + k] = kv;
    }
    kt[r][tx] = kv;
  }
  if (ti == tk) return;
  __syncthreads();
  for (int r = ty; r < KF_T; r += 8) {
    const int64_t k = k0 + r, i = i0 + tx;
    if (k < n && i < n) K[k * n + i] = kt[tx][r];
  }
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
  // two waves per SIMD (8 waves of 128 x 64): 4.7 ms at n = 10k over 99,952 SNPs, against
  // 5.7 ms for one wave per SIMD (4 waves of 128 x 128), which cannot issue the MFMAs
  // back to back
  if (n_full > 0) {
    hipLaunchKernelGGL((k_syrk_i8<false, 8>), dim3(n_full), dim3(512), 0, st, gT0, ld_gT, n,
                       n_snps, S, n_full, 1);
    JM_CHECK_LAUNCH();
  }
  if (total > n_full) {
    hipLaunchKernelGGL((k_syrk_i8<true, 8>), dim3((total - n_full) * split), dim3(512), 0, st,
                       gT0, ld_gT, n, n_snps, S, n_full, split);
    JM_CHECK_LAUNCH();
  }
  return JM_OK;
}

extern "C" int jm_pack_gram_2bit(const int8_t* g, int64_t ld_g, int64_t n_snps, int64_t n,
                                 uint8_t* gP, int64_t ldp, int64_t col0, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0 && ld_g >= n && col0 >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(g && gP && ((uintptr_t)gP % 4) == 0, "null or misaligned pointer");
  JM_CHECK_ARG(col0 % 16 == 0, "col0 must be a multiple of 16");
  JM_CHECK_ARG(ldp % 64 == 0, "ldp must be a multiple of 64 bytes");
  const int64_t end = (col0 + n_snps + SP_G - 1) / SP_G * SP_G;  // SNP columns written
  JM_CHECK_ARG(end <= 4 * ldp, "ldp too small: %lld bytes for %lld SNP columns", (long long)ldp,
               (long long)end);
  const int64_t n_words = (end - col0) / 16;
  const dim3 grid((unsigned)ceil_div(n_words, 16), (unsigned)ceil_div(n, 64));
  hipLaunchKernelGGL(k_pack_gram_2bit, grid, dim3(256), 0, (hipStream_t)stream, g, ld_g, n_snps, n,
                     gP, ldp, col0, n_words);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_kinship_syrk_2bit(const uint8_t* gP, int64_t ldp, int64_t n, int64_t n_snps,
                                    int32_t* S, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(gP && S, "null pointer");
  JM_CHECK_ARG(ldp % 64 == 0 && ((uintptr_t)gP % 16) == 0 &&
                   (n_snps + SP_G - 1) / SP_G * (SP_G / 4) <= ldp,
               "gP rows must be 16-byte aligned, ldp a multiple of 64 covering the SNP groups");
  JM_CHECK_ARG(16 * ldp < (1LL << 32), "ldp too large");
  const int64_t nb = (n + SY_B - 1) / SY_B;
  JM_CHECK_ARG(nb * (nb + 1) / 2 < (1LL << 28), "n too large");
  const int total = (int)(nb * (nb + 1) / 2);
  // FP4 MFMA (default) or the int8 MFMA on the same operand (JM_KINSHIP_SYRK_I8MFMA=1, read
  // per call, for comparisons); the FP4 form runs at most SF4_KMAX SNPs per launch (exact
  // FP32 sums)
  const char* env_i8 = getenv("JM_KINSHIP_SYRK_I8MFMA");
  const bool i8mfma = env_i8 && env_i8[0] == '1';
  const int64_t kmax = i8mfma ? n_snps : SF4_KMAX;
  hipStream_t st = (hipStream_t)stream;
  for (int64_t k0 = 0; k0 < n_snps; k0 += kmax) {
    const int64_t kn = n_snps - k0 < kmax ? n_snps - k0 : kmax;
    const uint8_t* g0 = gP + k0 / 4;  // k0 is a multiple of 256 SNPs: 64-byte aligned
    const int nk = (int)((kn + SP_G - 1) / SP_G);  // 256-SNP stages
    const int tail = total % 256;
    int split = tail ? 256 / tail : 1;
    if (split > nk / 4) split = nk / 4;
    if (split < 1) split = 1;
    const int n_full = split > 1 ? total - tail : total;
    if (n_full > 0) {
      if (i8mfma)
        hipLaunchKernelGGL(k_syrk_p2<false>, dim3(n_full), dim3(512), 0, st, g0, ldp, n, kn, S,
                           n_full, 1);
      else
        hipLaunchKernelGGL(k_syrk_f4<false>, dim3(n_full), dim3(512), 0, st, g0, ldp, n, kn, S,
                           n_full, 1);
      JM_CHECK_LAUNCH();
    }
    if (total > n_full) {
      if (i8mfma)
        hipLaunchKernelGGL(k_syrk_p2<true>, dim3((total - n_full) * split), dim3(512), 0, st, g0,
                           ldp, n, kn, S, n_full, split);
      else
        hipLaunchKernelGGL(k_syrk_f4<true>, dim3((total - n_full) * split), dim3(512), 0, st, g0,
                           ldp, n, kn, S, n_full, split);
      JM_CHECK_LAUNCH();
    }
  }
  return JM_OK;
}

// workspace (int32 units): offsets (n + 1) | lists (n_missing) | 2-bit codes (K rows of
// ceil(n / 16) words, rounded up to an even count so that rows stay 8-byte aligned)
static int64_t codes_ld(int64_t n) { return (((n + 15) / 16) + 1) / 2 * 2; }
static int64_t codes_ld4(int64_t n) { return (((n + 7) / 8) + 1) / 2 * 2; }

extern "C" int64_t jm_kinship_miss_ws_len(int64_t n, int64_t n_snps, int64_t n_missing) {
  if (n < 1 || n_snps < 0 || n_missing < 0) return 0;
  return n + 1 + (n_missing + 1) / 2 * 2 + n_snps * codes_ld4(n);
}

extern "C" int jm_kinship_impute_terms(const int8_t* g, int64_t ld_g, const int8_t* gT,
                                       int64_t ld_gT, int64_t n, int64_t n_snps, const double* mu,
                                       double* u, double* C, int has_missing, int32_t* miss_ws,
                                       int64_t miss_ws_len, int64_t n_missing, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(gT && mu && u, "null pointer");
  JM_CHECK_ARG(ld_gT >= n_snps, "ld_gT < n_snps");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_impute_u, dim3((unsigned)n), dim3(256), 0, st, gT, ld_gT, n_snps, mu, u);
  JM_CHECK_LAUNCH();
  if (has_missing) {
    JM_CHECK_ARG(g && C && ld_g >= n, "missing-call correction needs g and C");
    JM_CHECK_ARG(ld_g % 16 == 0 && ((uintptr_t)g % 16) == 0, "g rows must be 16-byte aligned");
    JM_CHECK_ARG(ld_gT % 16 == 0 && ((uintptr_t)gT % 16) == 0 && ld_gT >= ((n_snps + 15) / 16) * 16,
                 "gT rows must be 16-byte aligned and padded to a multiple of 16");
    JM_CHECK_ARG(n_missing >= 0 && n_missing < (1LL << 31),
                 "n_missing out of range (at most 2^31 - 1 missing calls per tile)");
    JM_CHECK_ARG(miss_ws && ((uintptr_t)miss_ws % 8) == 0 &&
                     miss_ws_len >= jm_kinship_miss_ws_len(n, n_snps, n_missing),
                 "missing-call correction needs an 8-byte aligned workspace of "
                 "jm_kinship_miss_ws_len(n, n_snps, n_missing) int32");
    JM_CHECK_ARG(n < (1LL << 31) && n_snps < (1LL << 31), "sizes exceed int32 indices");
    const int64_t cap = n_missing;
    int32_t* off = miss_ws;
    int32_t* list = miss_ws + n + 1;
    const int64_t ldc = codes_ld(n);
    // n + 1 + even(n_missing): the code rows start 4-byte aligned past the lists
    uint32_t* gc = (uint32_t*)(miss_ws + n + 1 + (n_missing + 1) / 2 * 2);
    const unsigned rb = (unsigned)((n + 3) / 4);
    hipLaunchKernelGGL(k_miss_count, dim3(rb), dim3(256), 0, st, gT, ld_gT, n, n_snps, off);
    JM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_miss_scan, dim3(1), dim3(1024), 0, st, off, n);
    JM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_miss_fill, dim3(rb), dim3(256), 0, st, gT, ld_gT, n, n_snps, off, list,
                       cap);
    JM_CHECK_LAUNCH();
    const int64_t blocks = ((n + C_ICHUNK - 1) / C_ICHUNK) * ((n + C_WAVES - 1) / C_WAVES);
    JM_CHECK_ARG(blocks < (1LL << 31), "n too large");
    static const bool per_snp = [] {  // JM_IMPUTE_C_PER_SNP=1: the per-SNP form (comparisons)
      const char* e = getenv("JM_IMPUTE_C_PER_SNP");
      return e && e[0] == '1';
    }();
    if (per_snp) {  // 2-bit codes, 16 samples per word
      hipLaunchKernelGGL(k_pack_codes, dim3(4096), dim3(256), 0, st, g, ld_g, n, n_snps, gc, ldc);
      JM_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_impute_C, dim3((unsigned)blocks), dim3(256), 0, st, gc, ldc, n, mu,
                         off, list, cap, C);
    } else {  // one code per 4-bit slot, 8 samples per word
      const int64_t ldc4 = codes_ld4(n);
      hipLaunchKernelGGL(k_pack_codes4, dim3(4096), dim3(256), 0, st, g, ld_g, n, n_snps, gc, ldc4);
      JM_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_impute_C2, dim3((unsigned)blocks), dim3(256), 0, st, gc, ldc4, n, mu,
                         off, list, cap, C);
    }
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
  const int64_t nt = (n + KF_T - 1) / KF_T;
  const int64_t pairs = nt * (nt + 1) / 2;
  if (pairs < (1LL << 31)) {
    hipLaunchKernelGGL(k_kin_finalize_tiled, dim3((unsigned)pairs), dim3(256), 0,
                       (hipStream_t)stream, S, u, C, s0, S_sub, u_sub, C_sub, s0_sub, n, p_snps,
                       K);
  } else {
    hipLaunchKernelGGL(k_kin_finalize, dim3(4096), dim3(256), 0, (hipStream_t)stream, S, u, C,
                       s0, S_sub, u_sub, C_sub, s0_sub, n, p_snps, K);
  }
  JM_CHECK_LAUNCH();
  return JM_OK;
}
