// Rotation of genotypes (and covariates / phenotype) into the kinship
// eigenbasis: Z = U^T X.
//
// Replaces UtG = UT @ geno_chunk (src/jamma/lmm/runner_streaming.py:392-410)
// and UtW = UT @ W, Uty = UT @ y (:349-352), which the reference runs as host
// FP64 BLAS.
//
//  * k_gemm_f64: FP64 MFMA (v_mfma_f64_16x16x4f64), 64x128 block tile,
//    4 waves of 32x64, LDS-staged A (UT rows) and B (genotype int8 or float64).
//  * k_rotate_bf16x3_packed: bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
//    accumulation over the split UT = hi + mid + lo and the genotype split
//    X = X1 + X2 (X1 = genotype or bf16(mean) at a missing call, X2 = the
//    bf16 residual of the mean at missing calls):
//        Z = lo X1 + mid X1 + hi X2 + hi X1.
//    256x256 block tile, both operands pre-blocked (k_pack_ut, k_pack_geno)
//    and DMA-copied into double-buffered LDS.
//
// Block order is XCD-aware: consecutive hardware block ids go round-robin to
// the 8 XCDs, so blocks b, b+8, b+16, ... share an XCD (and its L2). Each XCD
// takes a contiguous range of tiles; in the FP64 GEMM the range sweeps the SNP
// columns of one UT row panel, in the BF16x3 rotation it covers 4 x 8 patches
// of (row panel, SNP block), which share both operands in that XCD's L2.
//
// MFMA lane layouts (16x16 outputs): lane l supplies A[l%16][k-group l/16] and
// B[k-group l/16][l%16]. The accumulator element r (0..3) of lane l holds
//   bf16 / int8 forms:  D[4*(l/16) + r][l%16]
//   f64 16x16x4 form:   D[(l/16) + 4*r][l%16]
// (both measured on gfx950 with a single-wave layout check).

#include "jm_common.h"


namespace jm {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

// XCD-aware mapping of a linear block id onto (row block, column block).
// Hardware block bid runs on XCD bid % 8; XCD x receives the contiguous tile
// range [off_x, off_x + cnt_x) of the row-panel-major tile order (bijective).
__device__ __forceinline__ void xcd_tile(int bid, int nrb, int ncb, int& rb, int& cb) {
  const int nxcd = 8;
  const int total = nrb * ncb;
  const int x = bid % nxcd, j = bid / nxcd;
  const int base = total / nxcd, rem = total % nxcd;
  const int off = x * base + (x < rem ? x : rem);
  const int t = off + j;
  rb = t / ncb;
  cb = t % ncb;
}

// The same XCD-contiguous ranges over a 2-D order: stripes of 4 row panels,
// column-major inside a stripe, so the 32 blocks resident on one XCD cover a
// 4 x 8 patch of (row panel, column block) and share 4 row panels and 8
// column blocks in its L2 instead of 1 and 32 (bijective for any grid).
__device__ __forceinline__ void xcd_tile_2d(int bid, int nrb, int ncb, int& rb, int& cb) {
  const int total = nrb * ncb;
  const int x = bid % 8, j = bid / 8;
  const int base = total / 8, rem = total % 8;
  const int t = x * base + (x < rem ? x : rem) + j;
  const int stripe = t / (4 * ncb);
  const int h = min(4, nrb - 4 * stripe);
  const int u = t - stripe * 4 * ncb;
  cb = u / h;
  rb = 4 * stripe + u % h;
}

// ---------------------------------------------------------------------------
// FP64 GEMM: C[i][s] = sum_k A[i][k] * B(s, k)
// ---------------------------------------------------------------------------
constexpr int F_BM = 64, F_BN = 128, F_BK = 32;
constexpr int F_AS = F_BK + 1;  // padded LDS row stride (doubles)

struct BSrcGeno {  // B(s, k) = g[r(s)][k] (int8, -1 -> mean[r(s)]), r(s) = col ? col[s] : s
  const int8_t* g;
  int64_t ld;
  const double* mean;
  const int32_t* col;
};
struct BSrcF64 {  // B(s, k) = b[s][k]
  const double* b;
  int64_t ld;
};

template <typename BS>
struct BTile;

template <>
struct BTile<BSrcGeno> {
  static constexpr int kStride = F_BK + 4;  // bytes
  int8_t v[F_BN * kStride];
};
template <>
struct BTile<BSrcF64> {
  static constexpr int kStride = F_BK + 1;  // doubles
  double v[F_BN * kStride];
};

__device__ __forceinline__ void load_b_tile(BTile<BSrcGeno>& t, const BSrcGeno& src, int64_t s0,
                                            int64_t k0, int64_t n_cols, int64_t K, int tid) {
  // 128 rows x 32 bytes: 256 threads x 16 bytes
  const int r = tid >> 1;
  const int c = (tid & 1) * 16;
  const int64_t s = s0 + r;
  const int64_t sr = (s < n_cols && src.col) ? (int64_t)src.col[s] : s;
  int8_t tmp[16];
  if (s < n_cols && k0 + c + 16 <= K) {
    const int4 w = *reinterpret_cast<const int4*>(src.g + sr * src.ld + k0 + c);
    memcpy(tmp, &w, 16);
  } else {
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t k = k0 + c + e;
      tmp[e] = (s < n_cols && k < K) ? src.g[sr * src.ld + k] : (int8_t)0;
    }
  }
#pragma unroll
  for (int e = 0; e < 16; ++e) t.v[r * BTile<BSrcGeno>::kStride + c + e] = tmp[e];
}

__device__ __forceinline__ void load_b_tile(BTile<BSrcF64>& t, const BSrcF64& src, int64_t s0,
                                            int64_t k0, int64_t n_cols, int64_t K, int tid) {
  // 128 rows x 32 doubles = 4096: 16 per thread
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int idx = tid + 256 * u;
    const int r = idx / F_BK, c = idx % F_BK;
    const int64_t s = s0 + r, k = k0 + c;
    t.v[r * BTile<BSrcF64>::kStride + c] = (s < n_cols && k < K) ? src.b[s * src.ld + k] : 0.0;
  }
}

__device__ __forceinline__ double b_frag(const BTile<BSrcGeno>& t, int col, int k, double mean) {
  const int8_t g = t.v[col * BTile<BSrcGeno>::kStride + k];
  return g < 0 ? mean : (double)g;
}
__device__ __forceinline__ double b_frag(const BTile<BSrcF64>& t, int col, int k, double) {
  return t.v[col * BTile<BSrcF64>::kStride + k];
}

__device__ __forceinline__ double col_mean(const BSrcGeno& src, int64_t s, int64_t n_cols) {
  if (s >= n_cols) return 0.0;
  return src.mean[src.col ? (int64_t)src.col[s] : s];
}
__device__ __forceinline__ double col_mean(const BSrcF64&, int64_t, int64_t) { return 0.0; }

template <typename BS>
__global__ void __launch_bounds__(256) k_gemm_f64(const double* __restrict__ A, int64_t lda,
                                                  BS bsrc, int64_t M, int64_t n_cols, int64_t K,
                                                  double* __restrict__ C, int64_t ldc) {
  __shared__ double As[F_BM * F_AS];
  __shared__ BTile<BS> Bs;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nrb = (int)((M + F_BM - 1) / F_BM), ncb = (int)((n_cols + F_BN - 1) / F_BN);
  int rb, cb;
  xcd_tile(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * F_BM, s0 = (int64_t)cb * F_BN;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  double means[4];
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) means[tj] = col_mean(bsrc, s0 + wc + tj * 16 + lr, n_cols);
  d4_t acc[2][4];
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = (d4_t){0.0, 0.0, 0.0, 0.0};

  for (int64_t k0 = 0; k0 < K; k0 += F_BK) {
    // A tile: 64 x 32 doubles, 8 per thread
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int idx = tid + 256 * u;
      const int r = idx / F_BK, c = idx % F_BK;
      const int64_t i = i0 + r, k = k0 + c;
      As[r * F_AS + c] = (i < M && k < K) ? A[i * lda + k] : 0.0;
    }
    load_b_tile(Bs, bsrc, s0, k0, n_cols, K, tid);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < F_BK; kk += 4) {
      double a[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) a[ti] = As[(wr + ti * 16 + lr) * F_AS + kk + lk];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const double b = b_frag(Bs, wc + tj * 16 + lr, kk + lk, means[tj]);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ti], b, acc[ti][tj], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + lk + 4 * r;  // f64 output layout
        const int64_t s = s0 + wc + tj * 16 + lr;
        if (i < M && s < n_cols) C[i * ldc + s] = acc[ti][tj][r];
      }
}

// ---------------------------------------------------------------------------
// bf16 split of UT
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ double bf16_to_f64(uint16_t h) {
  return (double)__uint_as_float(((uint32_t)h) << 16);
}

// ---------------------------------------------------------------------------
// BF16x3 rotation
// ---------------------------------------------------------------------------
// Per 32-deep K step a 256x256 block stages the UT planes (2 or 3 x 256 x 32
// bf16 = 16 KB each) and the genotype codes (256 x 32 B = 8 KB) from global
// memory into LDS with global_load_lds_dwordx4 (no register staging). UT rows
// in LDS are 64 bytes, XOR-swizzled in 16-byte chunks so that the ds_read_b128
// fragment reads of 16 lanes hit distinct banks: chunk c of row r sits at
// c ^ ((r >> 2) & 3) (the 4-bank group of a chunk is 4 (r & 3) + c). Genotype
// rows are 32 bytes, 8-byte chunk c of SNP j at c ^ ((j >> 2) & 3), so the
// ds_read_b64 reads of 16 lanes cover all 32 bank pairs.
constexpr int V3_BM = 256, V3_BN = 256, V3_BK = 32;
constexpr int V3_APLANE = V3_BM * V3_BK * 2;      // 16 KB per plane
constexpr int V3_BBYTES = V3_BN * V3_BK;          // 8 KB: one code byte per genotype
// LDS stage of NPL planes + genotypes: 40 KB (NPL = 2) or 56 KB (NPL = 3)
template <int NPL>
constexpr int v3_stage() { return NPL * V3_APLANE + V3_BBYTES; }

// Byte-permute widening of 8 genotypes to bf16. The 2-bit genotype code c (0, 1,
// 2, missing = 3, one byte per genotype in the packed tile) indexes 4-entry byte
// tables: one v_perm_b32 looks up the low bytes of four genotypes' bf16 values,
// one the high bytes, and two more interleave them into the four bf16 (two
// dwords). Per-column tables:
//   X1: lo bytes {00, 80, 00, lo(m1)}, hi bytes {00, 3F, 40, hi(m1)}  (0, 1.0, 2.0, m1)
//   X2: lo bytes {00, 00, 00, lo(m2)}, hi bytes {00, 00, 00, hi(m2)}
// mm = m1 | m2 << 16. The 4 table words of a column are made once per launch;
// 16 permute VALU per 8 genotypes for X1 and X2.
struct widen_tab {  // byte tables of one column: X1 {lo, hi} and X2 {lo, hi}
  uint32_t t1lo, t1hi, t2lo, t2hi;
};
__device__ __forceinline__ widen_tab widen_tables(uint32_t mm) {
  widen_tab t;
  t.t1lo = __builtin_amdgcn_perm(mm, 0x00008000u, 0x04020100u);  // lo(m1) in byte 3
  t.t1hi = __builtin_amdgcn_perm(mm, 0x00403F00u, 0x05020100u);  // hi(m1) in byte 3
  t.t2lo = __builtin_amdgcn_perm(mm, 0u, 0x060C0C0Cu);           // lo(m2) in byte 3
  t.t2hi = mm & 0xFF000000u;                                     // hi(m2) in byte 3
  return t;
}
// four codes (bytes of w) -> four bf16 from the tables (lo, hi): two dwords
__device__ __forceinline__ void widen4(uint32_t w, uint32_t tlo, uint32_t thi, uint32_t& d0,
                                       uint32_t& d1) {
  const uint32_t plo = __builtin_amdgcn_perm(0u, tlo, w);  // byte i = tlo byte code_i
  const uint32_t phi = __builtin_amdgcn_perm(0u, thi, w);
  d0 = __builtin_amdgcn_perm(phi, plo, 0x05010400u);  // lo0 hi0 lo1 hi1
  d1 = __builtin_amdgcn_perm(phi, plo, 0x07030602u);  // lo2 hi2 lo3 hi3
}
__device__ __forceinline__ void widen8c(const uint32_t (&w)[2], const widen_tab& t, bf16x8_t& x1,
                                        bf16x8_t& x2) {
  uint32_t w1[4], w2[4];
  widen4(w[0], t.t1lo, t.t1hi, w1[0], w1[1]);
  widen4(w[1], t.t1lo, t.t1hi, w1[2], w1[3]);
  widen4(w[0], t.t2lo, t.t2hi, w2[0], w2[1]);
  widen4(w[1], t.t2lo, t.t2hi, w2[2], w2[3]);
  memcpy(&x1, w1, 16);
  memcpy(&x2, w2, 16);
}

// ---------------------------------------------------------------------------
// Packed operands for the BF16x3 rotation
// ---------------------------------------------------------------------------
// The rotation streams UT (three bf16 planes) and the genotype tile through
// LDS in 256 x 32 blocks. Both are stored pre-blocked so that every DMA of
// the rotation kernel copies one contiguous kilobyte:
//   UT planes: plane-major, then panel P (256 rows), then K step s (32
//     columns): a 16 KB block of 256 rows x 64 B, row r's 16-byte chunk c at
//     r * 64 + 16 * (c ^ ((r >> 2) & 3)). Rows >= n and columns >= n are 0.
//   genotypes: column block B (256 SNPs), then K step s (32 samples): an 8 KB
//     block of 256 SNPs x 32 B, SNP j's 8-byte chunk c (samples 8c .. 8c + 7,
//     one code byte each; code = decoded value & 3, missing 0xFF -> 3) at
//     j * 32 + 8 * (c ^ ((j >> 2) & 3)). SNPs >= n_snps and samples >= n hold
//     code 0.
// The swizzles are those of the LDS stage, so a block copies verbatim.
constexpr int64_t UTP_BLOCK = 16384, GP_BLOCK = 8192;

__global__ void k_pack_ut(const double* __restrict__ UT, int64_t ld_u, int64_t n, int64_t npan,
                          int64_t nks, int nplanes, uint16_t* __restrict__ planes) {
  const int64_t per_plane = npan * nks * (UTP_BLOCK / 2);  // bf16 elements per plane
  const int64_t chunks = npan * nks * 1024;                // 16-byte chunks per plane
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < chunks;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = q >> 10;
    const int pos = (int)(q & 1023);  // chunk slot within the 16 KB block
    const int r = pos >> 2, cs = pos & 3;
    const int c = cs ^ ((r >> 2) & 3);
    const int64_t P = blk / nks, sk = blk % nks;
    const int64_t i = P * 256 + r, k0 = sk * 32 + 8 * c;
    uint16_t h[3][8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      double u = (i < n && k0 + e < n) ? UT[i * ld_u + k0 + e] : 0.0;
#pragma unroll
      for (int pl = 0; pl < 3; ++pl) {  // successive residuals, each rounded to bf16
        h[pl][e] = f32_to_bf16_rne((float)u);
        u -= bf16_to_f64(h[pl][e]);
      }
    }
    const int64_t off = q * 8;
    for (int pl = 0; pl < nplanes; ++pl) memcpy(planes + pl * per_plane + off, h[pl], 16);
  }
}

__global__ void k_pack_geno(const int8_t* __restrict__ g, int64_t ld_g,
                            const int32_t* __restrict__ col, const double* __restrict__ mean,
                            int64_t n, int64_t n_snps, int64_t nks, uint8_t* __restrict__ gp,
                            uint32_t* __restrict__ mm) {
  const int64_t nblk = (n_snps + 255) / 256;
  const int64_t chunks = nblk * nks * 1024;  // 8-byte chunks (8 genotypes each)
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < chunks;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t blk = q >> 10;
    const int pos = (int)(q & 1023);  // chunk slot within the 8 KB block
    const int j = pos >> 2, cs = pos & 3;
    const int c = cs ^ ((j >> 2) & 3);
    const int64_t B = blk / nks, sk = blk % nks;
    const int64_t snp = B * 256 + j, k0 = sk * 32 + 8 * c;
    uint64_t codes = 0;  // byte e = code of sample k0 + e; padding reads as code 0
    if (snp < n_snps) {
      const int8_t* row = g + (col ? (int64_t)col[snp] : snp) * ld_g;
      if (k0 + 8 <= n) {  // rows are 8-byte aligned and k0 % 8 == 0: one 8-byte load
        codes = *reinterpret_cast<const uint64_t*>(row + k0) & 0x0303030303030303ull;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (k0 + e < n) codes |= (uint64_t)((uint8_t)row[k0 + e] & 3u) << (8 * e);
      }
      if (sk == 0 && c == 0) {
        const double mu = mean[col ? (int64_t)col[snp] : snp];
        const uint16_t h1 = f32_to_bf16_rne((float)mu);
        mm[snp] = (uint32_t)h1 | ((uint32_t)f32_to_bf16_rne((float)(mu - bf16_to_f64(h1))) << 16);
      }
    }
    *reinterpret_cast<uint64_t*>(gp + q * 8) = codes;
  }
}

// Operands of one K step as seen from a wave. ap / bp: this lane's copy sources
// (plane 0 of row panel rb, genotype block cb); a stage holds NPL UT planes then
// the genotype codes. Each wave copies NPL x 2 KB of UT and 1 KB of codes.
template <int NPL>
struct v3_ops {
  const uint8_t* ap;
  const uint8_t* bp;
  int64_t pstride;  // bytes per UT plane
  int wave, wr, wc, lr, lk;
  static constexpr int ADMA = NPL * 2;  // 1 KB UT copies per wave per stage
  static constexpr int DPS = ADMA + 1;  // copies per wave per stage
  __device__ __forceinline__ void issue(uint8_t* base, int64_t sk) const {
#pragma unroll
    for (int u = 0; u < ADMA; ++u) {
      const int q = wave * ADMA + u;  // plane q >> 4, kilobyte q & 15
      lds_copy16(ap + (q >> 4) * pstride + sk * UTP_BLOCK + (q & 15) * 1024,
                 base + (q >> 4) * V3_APLANE + (q & 15) * 1024);
    }
    lds_copy16(bp + sk * GP_BLOCK + wave * 1024, base + NPL * V3_APLANE + wave * 1024);
  }
  // UT fragment of plane pl, 16-row block ti of the wave: A[row][8 lk .. 8 lk + 7]
  __device__ __forceinline__ bf16x8_t a_frag(const uint8_t* base, int pl, int ti) const {
    const int r = wr + 16 * ti + lr;
    const int c = lk ^ ((r >> 2) & 3);
    bf16x8_t v;
    memcpy(&v, base + pl * V3_APLANE + r * 64 + 16 * c, 16);
    return v;
  }
  // codes of samples 8 lk .. 8 lk + 7 of the wave's column 16 tj + lr (one 8-byte read)
  __device__ __forceinline__ void g_codes(const uint8_t* base, int tj, uint32_t (&w)[2]) const {
    const int j = wc + 16 * tj + lr;
    const int c = lk ^ ((j >> 2) & 3);
    uint64_t v;
    memcpy(&v, base + NPL * V3_APLANE + j * 32 + 8 * c, 8);
    w[0] = (uint32_t)v;
    w[1] = (uint32_t)(v >> 32);
  }
};

// BF16x3 rotation on packed operands: 256x256 block tile, 8 waves (two per
// SIMD) in a grid of (16 WR) x (16 WC) wave tiles, WR x WC = 32 MFMA blocks of
// 16x16 (128 accumulators per lane), both operands copied block by block into
// NBUF LDS stages with global_load_lds_dwordx4 (the copies of stages kt+1 ..
// kt+NBUF-1 in flight while stage kt is multiplied), one barrier per 32-deep K
// step. Genotype fragments are widened to X1 and X2 in one pass per column
// block. Per K step and accumulator:
//   NPL = 2 (UT = hi + lo):        hi x X2, hi x X1, lo x X1   (3 products)
//   NPL = 3 (UT = hi + mid + lo):  hi x X2, hi x X1, mid x X1, lo x X1
template <int NPL, int NBUF, int WR>
__global__ void __launch_bounds__(512) k_rotate_bf16x3_packed(
    const uint16_t* __restrict__ planes, int64_t per_plane, const uint8_t* __restrict__ gp,
    const uint32_t* __restrict__ mmv, int64_t M, int64_t n_cols, int64_t nks,
    float* __restrict__ Z, int64_t ldz) {
  constexpr int WC = 32 / WR;   // 16-column blocks per wave
  constexpr int NWC = 16 / WC;  // waves across the 256 columns
  static_assert(WR * WC == 32 && (8 / NWC) * WR * 16 == V3_BM, "wave grid");
  constexpr int STAGE = v3_stage<NPL>();
  constexpr int DPS = v3_ops<NPL>::DPS;
  static_assert(NBUF >= 2 && NBUF * STAGE <= 160 * 1024, "LDS stages");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[NBUF * STAGE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nrb = (int)((M + V3_BM - 1) / V3_BM), ncb = (int)((n_cols + V3_BN - 1) / V3_BN);
  int rb, cb;
  xcd_tile_2d(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V3_BM, s0 = (int64_t)cb * V3_BN;
  v3_ops<NPL> op;
  op.ap = (const uint8_t*)planes + (int64_t)rb * nks * UTP_BLOCK + lane * 16;
  op.bp = gp + (int64_t)cb * nks * GP_BLOCK + lane * 16;
  op.pstride = per_plane * 2;
  op.wave = wave;
  op.wr = (wave / NWC) * 16 * WR;
  op.wc = (wave % NWC) * 16 * WC;
  op.lr = lane & 15;
  op.lk = lane >> 4;
  widen_tab tab[WC];
#pragma unroll
  for (int tj = 0; tj < WC; ++tj) {
    const int64_t sc = s0 + op.wc + 16 * tj + op.lr;
    tab[tj] = widen_tables(sc < n_cols ? mmv[sc] : 0u);
  }
  f4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
  auto compute = [&](const uint8_t* base) {
    // every LDS read of the step is issued up front (hi plane, the genotype
    // codes, then the other planes), so the later planes' fragments arrive
    // while the hi plane's MFMAs run
    bf16x8_t af[NPL][WR];
    uint32_t gw[WC][2];
#pragma unroll
    for (int ti = 0; ti < WR; ++ti) af[0][ti] = op.a_frag(base, 0, ti);
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) op.g_codes(base, tj, gw[tj]);
#pragma unroll
    for (int pl = 1; pl < NPL; ++pl)
#pragma unroll
      for (int ti = 0; ti < WR; ++ti) af[pl][ti] = op.a_frag(base, pl, ti);
    // hi plane first: X1 and X2 of a column block from one widening
    bf16x8_t x1[WC];
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) {
      bf16x8_t x2;
      widen8c(gw[tj], tab[tj], x1[tj], x2);
#pragma unroll
      for (int ti = 0; ti < WR; ++ti)
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0][ti], x2, acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
      for (int ti = 0; ti < WR; ++ti)
#pragma unroll
        for (int tj = 0; tj < WC; ++tj)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[pl][ti], x1[tj], acc[ti][tj], 0, 0, 0);
  };
#pragma unroll
  for (int st = 0; st < NBUF - 1; ++st)
    if (st < nks) op.issue(lds + st * STAGE, st);
  int cur = 0;  // LDS stage of K step kt
  for (int kt = 0; kt < nks; ++kt) {
    // In flight: this wave's copies of K steps kt .. min(kt + NBUF - 2, nks - 1), issued
    // in that order and completed in order. Keep the later ones, wait for step kt.
    if (kt + NBUF - 2 < nks)
      wait_vmcnt<DPS * (NBUF - 2)>();
    else
      wait_vmcnt<0>();
    __syncthreads();  // every wave's copies of step kt landed; step kt-1's stage is free
    if (kt + NBUF - 1 < nks) {
      const int nxt = cur == 0 ? NBUF - 1 : cur - 1;  // (kt + NBUF - 1) % NBUF = stage of kt-1
      op.issue(lds + nxt * STAGE, kt + NBUF - 1);
    }
    compute(lds + cur * STAGE);
    cur = cur + 1 == NBUF ? 0 : cur + 1;
  }
#pragma unroll
  for (int ti = 0; ti < WR; ++ti)
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + op.wr + ti * 16 + 4 * op.lk + r;
        const int64_t sc = s0 + op.wc + tj * 16 + op.lr;
        if (i < M && sc < n_cols) Z[i * ldz + sc] = acc[ti][tj][r];
      }
}

// The two-plane rotation with its operand copies two K steps ahead and its LDS
// reads software-pipelined behind the MFMAs. Same block tile, wave grid (2 x 4
// waves of 128 x 64), operands, products and summation order per accumulator as
// k_rotate_bf16x3_packed<2, *, 8>, so Z is the same bit for bit. What changes:
//  * four LDS stages of 40 KB (the 8 KB genotype block stores one code byte per
//    genotype): in the middle of K step kt (after row block 6) the block waits for
//    step kt + 1 and copies step kt + 3 into the stage step kt - 1 used, so a copy
//    has two K steps to land (the copies' latency under load is about one step's
//    MFMA time);
//  * a wave walks its 8 row blocks in order and holds the UT fragments of only two
//    of them: the reads of row block ti + 1 are issued before the 12 MFMAs of row
//    block ti (instruction groups fixed with sched_group_barrier), and during row
//    block 7 the next step's codes and first UT fragments are read and widened
//    into the other of two X register sets (the K loop is unrolled by two), so a
//    step starts on operands already in registers.
// Per row block the products are issued hi x X2, hi x X1, lo x X1, each over the 4
// column blocks (consecutive MFMAs write different accumulators).
__global__ void __launch_bounds__(512) k_rotate_bf16x3_pipe4(
    const uint16_t* __restrict__ planes, int64_t per_plane, const uint8_t* __restrict__ gp,
    const uint32_t* __restrict__ mmv, int64_t M, int64_t n_cols, int64_t nks,
    float* __restrict__ Z, int64_t ldz) {
  constexpr int NPL = 2, WR = 8, WC = 4, NWC = 4, NBUF = 4;
  constexpr int STAGE = v3_stage<NPL>();
  constexpr int DPS = v3_ops<NPL>::DPS;
  static_assert(NBUF * STAGE <= 160 * 1024, "LDS stages");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[NBUF * STAGE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nrb = (int)((M + V3_BM - 1) / V3_BM), ncb = (int)((n_cols + V3_BN - 1) / V3_BN);
  int rb, cb;
  xcd_tile_2d(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V3_BM, s0 = (int64_t)cb * V3_BN;
  v3_ops<NPL> op;
  op.ap = (const uint8_t*)planes + (int64_t)rb * nks * UTP_BLOCK + lane * 16;
  op.bp = gp + (int64_t)cb * nks * GP_BLOCK + lane * 16;
  op.pstride = per_plane * 2;
  op.wave = wave;
  op.wr = (wave / NWC) * 16 * WR;
  op.wc = (wave % NWC) * 16 * WC;
  op.lr = lane & 15;
  op.lk = lane >> 4;
  widen_tab tab[WC];
#pragma unroll
  for (int tj = 0; tj < WC; ++tj) {
    const int64_t sc = s0 + op.wc + 16 * tj + op.lr;
    tab[tj] = widen_tables(sc < n_cols ? mmv[sc] : 0u);
  }
  f4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};

  // prologue: steps 0, 1, 2 in flight; step 0's codes and first row block read
#pragma unroll
  for (int st = 0; st < NBUF - 1; ++st)
    if (st < nks) op.issue(lds + st * STAGE, st);
  if (nks > 2)
    wait_vmcnt<2 * DPS>();
  else if (nks > 1)
    wait_vmcnt<DPS>();
  else
    wait_vmcnt<0>();
  __syncthreads();
  bf16x8_t af[2][NPL];  // rolling pair of row blocks
  bf16x8_t xa1[WC], xa2[WC], xb1[WC], xb2[WC];  // X of even / odd K steps
  {
    uint32_t gw[WC][2];
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) op.g_codes(lds, tj, gw[tj]);
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) af[0][pl] = op.a_frag(lds, pl, 0);
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) widen8c(gw[tj], tab[tj], xa1[tj], xa2[tj]);
  }
  // one K step on X (x1, x2); widens the next step's X into (y1, y2)
  auto step = [&](int64_t kt, int cur, bf16x8_t (&x1)[WC], bf16x8_t (&x2)[WC],
                  bf16x8_t (&y1)[WC], bf16x8_t (&y2)[WC]) {
    const int nxt = cur == NBUF - 1 ? 0 : cur + 1;
    const uint8_t* bc = lds + cur * STAGE;
    const uint8_t* bn = lds + nxt * STAGE;
    auto row_block = [&](int ti) {
      const int s = ti & 1;
#pragma unroll
      for (int tj = 0; tj < WC; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s][0], x2[tj], acc[ti][tj], 0, 0, 0);
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
        for (int tj = 0; tj < WC; ++tj)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s][pl], x1[tj], acc[ti][tj], 0, 0, 0);
    };
#pragma unroll
    for (int ti = 0; ti < WR - 1; ++ti) {
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl) af[(ti & 1) ^ 1][pl] = op.a_frag(bc, pl, ti + 1);
      row_block(ti);
    }
    static_assert(WR == 8 && WC == 4 && NPL == 2, "the instruction group sizes below");
#define JM_ROW_BLOCK                                              \
  __builtin_amdgcn_sched_group_barrier(0x100, 2, 0); /* DS reads */ \
  __builtin_amdgcn_sched_group_barrier(0x008, 12, 0) /* MFMAs */
    JM_ROW_BLOCK; JM_ROW_BLOCK; JM_ROW_BLOCK; JM_ROW_BLOCK; JM_ROW_BLOCK; JM_ROW_BLOCK; JM_ROW_BLOCK;
#undef JM_ROW_BLOCK
    // step kt + 1 landed for every wave (step kt + 2 may be in flight); nobody reads
    // the stage of step kt - 1 any more: step kt + 3 goes there
    if (kt + 2 < nks)
      wait_vmcnt<DPS>();
    else
      wait_vmcnt<0>();
    __syncthreads();
    if (kt + 3 < nks) op.issue(lds + (nxt + 2) % NBUF * STAGE, kt + 3);
    // row block 7 with the next step's operands read and widened behind it (after
    // the last step these reads are harmless and unused)
    uint32_t gw[WC][2];
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) op.g_codes(bn, tj, gw[tj]);
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) af[0][pl] = op.a_frag(bn, pl, 0);
    row_block(WR - 1);
#pragma unroll
    for (int tj = 0; tj < WC; ++tj) widen8c(gw[tj], tab[tj], y1[tj], y2[tj]);
    __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);  // DS reads: codes, next row block
    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);  // MFMAs interleaved with
    __builtin_amdgcn_sched_group_barrier(0x002, 16, 0); // the widening
    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 16, 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 32, 0);
  };
  int cur = 0;  // LDS stage of K step kt
  int64_t kt = 0;
  for (; kt + 1 < nks; kt += 2) {
    step(kt, cur, xa1, xa2, xb1, xb2);
    cur = cur == NBUF - 1 ? 0 : cur + 1;
    step(kt + 1, cur, xb1, xb2, xa1, xa2);
    cur = cur == NBUF - 1 ? 0 : cur + 1;
  }
  if (kt < nks) step(kt, cur, xa1, xa2, xb1, xb2);
#pragma unroll
  for (int ti = 0; ti < WR; ++ti)
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + op.wr + ti * 16 + 4 * op.lk + r;
        const int64_t sc = s0 + op.wc + tj * 16 + op.lr;
        if (i < M && sc < n_cols) Z[i * ldz + sc] = acc[ti][tj][r];
      }
}

// The three-plane rotation with 128 x 64 wave tiles (2 x 4 waves): a wave holds the
// UT fragments of only two row blocks (the reads of row block ti + 1, 3 planes, issued
// before the 16 MFMAs of row block ti), so the 3 planes fit beside the 128
// accumulators and each column block is widened by 2 waves instead of 4. Two LDS stages
// of 56 KB (three would exceed the 160 KB); per K step the codes and the first row
// block are read after the barrier. Per accumulator and K step: hi x X2, hi x X1,
// mid x X1, lo x X1, as in k_rotate_bf16x3_packed<3, *, *>: Z is the same bit for bit.
__global__ void __launch_bounds__(512) k_rotate_bf16x3_pipe3(
    const uint16_t* __restrict__ planes, int64_t per_plane, const uint8_t* __restrict__ gp,
    const uint32_t* __restrict__ mmv, int64_t M, int64_t n_cols, int64_t nks,
    float* __restrict__ Z, int64_t ldz) {
  constexpr int NPL = 3, WR = 8, WC = 4, NWC = 4, NBUF = 2;
  constexpr int STAGE = v3_stage<NPL>();
  static_assert(NBUF * STAGE <= 160 * 1024, "LDS stages");
  __shared__ __attribute__((aligned(1024))) uint8_t lds[NBUF * STAGE];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nrb = (int)((M + V3_BM - 1) / V3_BM), ncb = (int)((n_cols + V3_BN - 1) / V3_BN);
  int rb, cb;
  xcd_tile_2d(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V3_BM, s0 = (int64_t)cb * V3_BN;
  v3_ops<NPL> op;
  op.ap = (const uint8_t*)planes + (int64_t)rb * nks * UTP_BLOCK + lane * 16;
  op.bp = gp + (int64_t)cb * nks * GP_BLOCK + lane * 16;
  op.pstride = per_plane * 2;
  op.wave = wave;
  op.wr = (wave / NWC) * 16 * WR;
  op.wc = (wave % NWC) * 16 * WC;
  op.lr = lane & 15;
  op.lk = lane >> 4;
  widen_tab tab[WC];
#pragma unroll
  for (int tj = 0; tj < WC; ++tj) {
    const int64_t sc = s0 + op.wc + 16 * tj + op.lr;
    tab[tj] = widen_tables(sc < n_cols ? mmv[sc] : 0u);
  }
  f4_t acc[WR][WC];
#pragma unroll
  for (int a = 0; a < WR; ++a)
#pragma unroll
    for (int b = 0; b < WC; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
  op.issue(lds, 0);
  int cur = 0;
  for (int64_t kt = 0; kt < nks; ++kt) {
    wait_vmcnt<0>();  // this wave's copies of step kt (the only ones in flight)
    __syncthreads();  // every wave's copies landed; nobody reads the other stage any more
    if (kt + 1 < nks) op.issue(lds + (cur ^ 1) * STAGE, kt + 1);
    const uint8_t* bc = lds + cur * STAGE;
    bf16x8_t x1[WC], x2[WC];
    bf16x8_t af[2][NPL];
    {
      uint32_t gw[WC][2];
#pragma unroll
      for (int tj = 0; tj < WC; ++tj) op.g_codes(bc, tj, gw[tj]);
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl) af[0][pl] = op.a_frag(bc, pl, 0);
#pragma unroll
      for (int tj = 0; tj < WC; ++tj) widen8c(gw[tj], tab[tj], x1[tj], x2[tj]);
    }
#pragma unroll
    for (int ti = 0; ti < WR; ++ti) {
      const int s = ti & 1;
      if (ti + 1 < WR) {
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) af[s ^ 1][pl] = op.a_frag(bc, pl, ti + 1);
      }
#pragma unroll
      for (int tj = 0; tj < WC; ++tj)
        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s][0], x2[tj], acc[ti][tj], 0, 0, 0);
#pragma unroll
      for (int pl = 0; pl < NPL; ++pl)
#pragma unroll
        for (int tj = 0; tj < WC; ++tj)
          acc[ti][tj] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[s][pl], x1[tj], acc[ti][tj], 0, 0, 0);
    }
    cur ^= 1;
  }
#pragma unroll
  for (int ti = 0; ti < WR; ++ti)
#pragma unroll
    for (int tj = 0; tj < WC; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + op.wr + ti * 16 + 4 * op.lk + r;
        const int64_t sc = s0 + op.wc + tj * 16 + op.lr;
        if (i < M && sc < n_cols) Z[i * ldz + sc] = acc[ti][tj][r];
      }
}

}  // namespace jm

using namespace jm;

namespace jm {

// ---------------------------------------------------------------------------
// FP64 rotation of a few SNP columns (n_snps <= 16): HBM-bound GEMV form
// ---------------------------------------------------------------------------
// A 64 x 128 MFMA tile spends almost all of its work on padding when only a few
// columns are rotated (the FP64 recomputation of a handful of SNPs), and then reads
// U^T at a fraction of HBM speed. Here each block owns 16 consecutive rows of U^T
// (4 per wave); the <= 16 genotype columns are staged through LDS as int8, k-major
// (16 bytes per eigen-index: one ds_read_b128 per lane gives a lane all columns of
// its k), 4096 eigen-indices per stage. Lanes stride over k, so every U^T row is
// read once, coalesced; per lane and row the partial sums of the 16 columns are
// then summed over the wavefront in a fixed order (deterministic).
constexpr int kNarrowCols = 16;
constexpr int kNarrowRowsPerWave = 4;
constexpr int kNarrowChunk = 4096;

__global__ void __launch_bounds__(256) k_rotate_f64_narrow(const double* __restrict__ UT,
                                                           int64_t ld_u, const int8_t* __restrict__ g,
                                                           int64_t ld_g,
                                                           const double* __restrict__ mean,
                                                           const int32_t* __restrict__ col,
                                                           int64_t n, int ncols,
                                                           double* __restrict__ Z, int64_t ldz) {
  __shared__ __align__(16) int8_t xs[kNarrowChunk * kNarrowCols];  // 64 KB
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * (4 * kNarrowRowsPerWave);
  double mu[kNarrowCols];
#pragma unroll
  for (int c = 0; c < kNarrowCols; ++c) {
    const int64_t r = c < ncols ? (col ? col[c] : c) : 0;
    mu[c] = c < ncols ? mean[r] : 0.0;
  }
  double acc[kNarrowRowsPerWave][kNarrowCols];
#pragma unroll
  for (int q = 0; q < kNarrowRowsPerWave; ++q)
#pragma unroll
    for (int c = 0; c < kNarrowCols; ++c) acc[q][c] = 0.0;
  for (int64_t k0 = 0; k0 < n; k0 += kNarrowChunk) {
    const int len = (int)min((int64_t)kNarrowChunk, n - k0);
    const int len4 = (len + 3) & ~3;  // rows are zero-padded to ld_g (a multiple of 16)
    __syncthreads();  // the previous stage has been read
    // stage: column c, eigen-indices k..k+3 of this chunk -> xs[(k + j) * 16 + c]
    for (int e = threadIdx.x; e < kNarrowCols * (len4 / 4); e += blockDim.x) {
      const int c = e / (len4 / 4), k = (e % (len4 / 4)) * 4;
      uint32_t v = 0;
      if (c < ncols) {
        const int64_t r = col ? col[c] : c;
        v = *(const uint32_t*)(g + r * ld_g + k0 + k);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) xs[(k + j) * kNarrowCols + c] = (int8_t)(v >> (8 * j));
    }
    __syncthreads();
    for (int k = lane; k < len; k += 64) {
      const u32x4_t raw = *(const u32x4_t*)(xs + k * kNarrowCols);
      double x[kNarrowCols];
#pragma unroll
      for (int c = 0; c < kNarrowCols; ++c) {
        const int8_t b = (int8_t)(raw[c >> 2] >> (8 * (c & 3)));
        x[c] = b < 0 ? mu[c] : (double)b;
      }
#pragma unroll
      for (int q = 0; q < kNarrowRowsPerWave; ++q) {
        const int64_t i = row0 + wave + 4 * q;
        const double u = i < n ? UT[i * ld_u + k0 + k] : 0.0;
#pragma unroll
        for (int c = 0; c < kNarrowCols; ++c) acc[q][c] = fma(u, x[c], acc[q][c]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < kNarrowRowsPerWave; ++q) {
    const int64_t i = row0 + wave + 4 * q;
#pragma unroll
    for (int c = 0; c < kNarrowCols; ++c) {
      const double v = wave_sum(acc[q][c]);
      if (lane == 0 && i < n && c < ncols) Z[i * ldz + c] = v;
    }
  }
}

}  // namespace jm

extern "C" int jm_rotate_f64(const double* UT, int64_t ld_u, const int8_t* g, int64_t ld_g,
                             const double* mean, const int32_t* col_idx, int64_t n, int64_t n_snps,
                             double* Z, int64_t ldz, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(UT && g && mean && Z, "null pointer");
  JM_CHECK_ARG(ld_u >= n && ld_g >= n && ldz >= n_snps, "bad leading dimension");
  JM_CHECK_ARG(ld_g % 16 == 0 && ((uintptr_t)g % 16) == 0, "g must be 16-byte aligned rows");
  if (n_snps <= kNarrowCols) {  // a few columns: the HBM-bound GEMV form
    const int64_t blocks = (n + 4 * kNarrowRowsPerWave - 1) / (4 * kNarrowRowsPerWave);
    hipLaunchKernelGGL(k_rotate_f64_narrow, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, UT, ld_u, g, ld_g, mean, col_idx, n, (int)n_snps, Z,
                       ldz);
    JM_CHECK_LAUNCH();
    return JM_OK;
  }
  const int nrb = ceil_div(n, F_BM), ncb = ceil_div(n_snps, F_BN);
  BSrcGeno src{g, ld_g, mean, col_idx};
  hipLaunchKernelGGL(k_gemm_f64<BSrcGeno>, dim3(nrb * ncb), dim3(256), 0, (hipStream_t)stream,
                     UT, ld_u, src, n, n_snps, n, Z, ldz);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_gemm_nt_f64(const double* A, int64_t lda, const double* B, int64_t ldb,
                              int64_t m, int64_t ncols, int64_t k, double* C, int64_t ldc,
                              void* stream) {
  JM_CHECK_ARG(m >= 1 && ncols >= 0 && k >= 1, "bad sizes");
  if (ncols == 0) return JM_OK;
  JM_CHECK_ARG(A && B && C, "null pointer");
  JM_CHECK_ARG(lda >= k && ldb >= k && ldc >= ncols, "bad leading dimension");
  const int nrb = ceil_div(m, F_BM), ncb = ceil_div(ncols, F_BN);
  BSrcF64 src{B, ldb};
  hipLaunchKernelGGL(k_gemm_f64<BSrcF64>, dim3(nrb * ncb), dim3(256), 0, (hipStream_t)stream, A,
                     lda, src, m, ncols, k, C, ldc);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int64_t jm_packed_ut_bytes(int64_t n, int nplanes) {
  if (n < 1 || nplanes < 2 || nplanes > 3) return 0;
  return nplanes * ((n + 255) / 256) * ((n + 31) / 32) * UTP_BLOCK;
}

extern "C" int64_t jm_packed_geno_bytes(int64_t n, int64_t n_snps) {
  if (n < 1 || n_snps < 1) return 0;
  return ((n_snps + 255) / 256) * ((n + 31) / 32) * GP_BLOCK;
}

extern "C" int jm_pack_ut_bf16x3(const double* UT, int64_t ld_u, int64_t n, int nplanes,
                                 uint16_t* planes, void* stream) {
  JM_CHECK_ARG(n >= 1 && ld_u >= n, "bad sizes");
  JM_CHECK_ARG(nplanes == 2 || nplanes == 3, "nplanes must be 2 or 3");
  JM_CHECK_ARG(UT && planes && ((uintptr_t)planes % 16) == 0, "null or misaligned pointer");
  const int64_t npan = (n + 255) / 256, nks = (n + 31) / 32;
  const int64_t chunks = npan * nks * 1024;
  const unsigned grid = (unsigned)(chunks < 65536 ? (chunks + 255) / 256 : 256 * 64);
  hipLaunchKernelGGL(k_pack_ut, dim3(grid), dim3(256), 0, (hipStream_t)stream, UT, ld_u, n, npan,
                     nks, nplanes, planes);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_pack_genotypes(const int8_t* g, int64_t ld_g, const int32_t* col_idx,
                                 const double* mean, int64_t n, int64_t n_snps, uint8_t* gp,
                                 uint32_t* mm, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0 && ld_g >= n, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(g && mean && gp && mm && ((uintptr_t)gp % 16) == 0, "null or misaligned pointer");
  JM_CHECK_ARG(ld_g % 8 == 0 && ((uintptr_t)g % 8) == 0, "genotype rows must be 8-byte aligned");
  const int64_t nks = (n + 31) / 32;
  const int64_t chunks = ((n_snps + 255) / 256) * nks * 1024;
  const unsigned grid = (unsigned)(chunks < 65536 ? (chunks + 255) / 256 : 256 * 64);
  hipLaunchKernelGGL(k_pack_geno, dim3(grid), dim3(256), 0, (hipStream_t)stream, g, ld_g, col_idx,
                     mean, n, n_snps, nks, gp, mm);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_rotate_bf16x3(const uint16_t* planes, int nplanes, const uint8_t* gp,
                                const uint32_t* mm, int64_t n, int64_t n_snps, float* Z,
                                int64_t ldz, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  JM_CHECK_ARG(nplanes == 2 || nplanes == 3, "nplanes must be 2 or 3");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(planes && gp && mm && Z, "null pointer");
  JM_CHECK_ARG(ldz >= n_snps, "ldz < n_snps");
  JM_CHECK_ARG(((uintptr_t)planes % 16) == 0 && ((uintptr_t)gp % 16) == 0,
               "packed operands must be 16-byte aligned");
  const int64_t nks = (n + 31) / 32;
  const int64_t per_plane = jm_packed_ut_bytes(n, nplanes) / (2 * nplanes);  // bf16 per plane
  const int nrb = ceil_div(n, V3_BM), ncb = ceil_div(n_snps, V3_BN);
  static const bool unpiped = [] {  // JM_ROTATE_UNPIPED=1: the kernel without read pipelining
    const char* e = getenv("JM_ROTATE_UNPIPED");
    return e && e[0] == '1';
  }();
  if (nplanes == 2 && !unpiped)  // 4 x 40 KB stages, copies 2 steps ahead, reads pipelined
    hipLaunchKernelGGL(k_rotate_bf16x3_pipe4, dim3(nrb * ncb), dim3(512), 0, (hipStream_t)stream,
                       planes, per_plane, gp, mm, n, n_snps, nks, Z, ldz);
  else if (nplanes == 2)  // 2 x 40 KB stages, 128 x 64 wave tiles
    hipLaunchKernelGGL((k_rotate_bf16x3_packed<2, 2, 8>), dim3(nrb * ncb), dim3(512), 0,
                       (hipStream_t)stream, planes, per_plane, gp, mm, n, n_snps, nks, Z, ldz);
  else if (!unpiped)  // 2 x 56 KB stages, 128 x 64 wave tiles, rolling UT fragments
    hipLaunchKernelGGL(k_rotate_bf16x3_pipe3, dim3(nrb * ncb), dim3(512), 0, (hipStream_t)stream,
                       planes, per_plane, gp, mm, n, n_snps, nks, Z, ldz);
  else  // 2 x 56 KB stages, 64 x 128 wave tiles (all fragments of a step in registers)
    hipLaunchKernelGGL((k_rotate_bf16x3_packed<3, 2, 4>), dim3(nrb * ncb), dim3(512), 0,
                       (hipStream_t)stream, planes, per_plane, gp, mm, n, n_snps, nks, Z, ldz);
  JM_CHECK_LAUNCH();
  return JM_OK;
}
