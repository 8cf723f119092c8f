// Rotation of genotypes (and covariates / phenotype) into the kinship
// eigenbasis: Z = U^T X.
//
// Replaces UtG = UT @ geno_chunk (src/jamma/lmm/runner_streaming.py:392-410)
// and UtW = UT @ W, Uty = UT @ y (:349-352), which the reference runs as host
// FP64 BLAS.
//
//  * k_gemm_f64: FP64 MFMA (v_mfma_f64_16x16x4f64), 64x128 block tile,
//    4 waves of 32x64, LDS-staged A (UT rows) and B (genotype int8 or float64).
//  * k_rotate_bf16x3: bf16 MFMA (v_mfma_f32_16x16x32_bf16) with FP32
//    accumulation over the split UT = hi + mid + lo and the genotype split
//    X = X1 + X2 (X1 = genotype or bf16(mean) at a missing call, X2 = the
//    bf16 residual of the mean at missing calls):
//        Z = hi X1 + mid X1 + lo X1 + hi X2.
//    128x128 block tile, 4 waves of 64x64.
//
// Block order is XCD-aware: consecutive hardware block ids go round-robin to
// the 8 XCDs, so blocks b, b+8, b+16, ... share an XCD (and its L2). Tiles are
// assigned so that the blocks of one XCD sweep the SNP columns of the same
// row panel of UT, which then stays L2/MALL-resident while it is consumed.
//
// MFMA lane layouts (16x16 outputs): lane l supplies A[l%16][k-group l/16] and
// B[k-group l/16][l%16]. The accumulator element r (0..3) of lane l holds
//   bf16 / int8 forms:  D[4*(l/16) + r][l%16]
//   f64 16x16x4 form:   D[(l/16) + 4*r][l%16]
// (both measured on gfx950 with a single-wave layout check).

#include "jm_common.h"

#include <cstdlib>

namespace jm {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

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

__global__ void k_split_bf16x3(const double* __restrict__ UT, int64_t ld_u, int64_t n,
                               uint16_t* __restrict__ hi, uint16_t* __restrict__ mid,
                               uint16_t* __restrict__ lo, int64_t ld_s) {
  const int64_t total = n * ld_s;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / ld_s, k = idx % ld_s;
    const double u = (k < n) ? UT[i * ld_u + k] : 0.0;
    const uint16_t h = f32_to_bf16_rne((float)u);
    const double r1 = u - bf16_to_f64(h);
    const uint16_t m = f32_to_bf16_rne((float)r1);
    const double r2 = r1 - bf16_to_f64(m);
    const uint16_t l = f32_to_bf16_rne((float)r2);
    hi[idx] = h;
    mid[idx] = m;
    lo[idx] = l;
  }
}

// ---------------------------------------------------------------------------
// BF16x3 rotation
// ---------------------------------------------------------------------------
constexpr int B_BM = 128, B_BN = 128, B_BK = 64;
constexpr int B_AS = B_BK + 8;   // bf16 elements per LDS row (16-byte aligned rows)
constexpr int B_GS = B_BK + 16;  // int8 bytes per genotype LDS row

__global__ void __launch_bounds__(256) k_rotate_bf16x3(
    const uint16_t* __restrict__ Ahi, const uint16_t* __restrict__ Amid,
    const uint16_t* __restrict__ Alo, int64_t lda, const int8_t* __restrict__ g, int64_t ldg,
    const double* __restrict__ mean, const int32_t* __restrict__ col, int64_t M, int64_t n_cols,
    int64_t K, float* __restrict__ Z, int64_t ldz) {
  __shared__ __attribute__((aligned(16))) uint16_t As[3][B_BM * B_AS];
  __shared__ __attribute__((aligned(16))) int8_t Gs[B_BN * B_GS];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nrb = (int)((M + B_BM - 1) / B_BM), ncb = (int)((n_cols + B_BN - 1) / B_BN);
  int rb, cb;
  xcd_tile(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * B_BM, s0 = (int64_t)cb * B_BN;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;

  // per-column bf16 split of the imputation mean: m1 = bf16(mean), m2 = bf16(mean - m1)
  uint16_t m1[4], m2[4];
#pragma unroll
  for (int tj = 0; tj < 4; ++tj) {
    const int64_t s = s0 + wc + tj * 16 + lr;
    const double mu = (s < n_cols) ? mean[col ? (int64_t)col[s] : s] : 0.0;
    m1[tj] = f32_to_bf16_rne((float)mu);
    m2[tj] = f32_to_bf16_rne((float)(mu - bf16_to_f64(m1[tj])));
  }

  f4_t acc[4][4];
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = (f4_t){0.f, 0.f, 0.f, 0.f};

  const uint16_t* Ap[3] = {Ahi, Amid, Alo};
  for (int64_t k0 = 0; k0 < K; k0 += B_BK) {
    // A planes: 128 rows x 64 bf16 = 128 B per row -> 8 x 16 B chunks; 1024 chunks / plane
#pragma unroll
    for (int pl = 0; pl < 3; ++pl) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int idx = tid + 256 * u;
        const int r = idx >> 3, c = (idx & 7) * 8;
        const int64_t i = i0 + r;
        int4 w = {0, 0, 0, 0};
        if (i < M && k0 + c + 8 <= lda) w = *reinterpret_cast<const int4*>(Ap[pl] + i * lda + k0 + c);
        *reinterpret_cast<int4*>(&As[pl][r * B_AS + c]) = w;
      }
    }
    // genotype tile: 128 SNP rows x 64 bytes -> 4 x 16 B per row; 512 chunks
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int idx = tid + 256 * u;
      const int r = idx >> 2, c = (idx & 3) * 16;
      const int64_t s = s0 + r;
      int4 w = {0, 0, 0, 0};
      if (s < n_cols) {
        const int64_t sr = col ? (int64_t)col[s] : s;
        if (k0 + c + 16 <= K) {
          w = *reinterpret_cast<const int4*>(g + sr * ldg + k0 + c);
        } else {
          int8_t tmp[16];
#pragma unroll
          for (int e = 0; e < 16; ++e) tmp[e] = (k0 + c + e < K) ? g[sr * ldg + k0 + c + e] : (int8_t)0;
          memcpy(&w, tmp, 16);
        }
      }
      *reinterpret_cast<int4*>(&Gs[r * B_GS + c]) = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < B_BK; kk += 32) {
      // B fragments: 8 consecutive k for column (wc + tj*16 + lr)
      bf16x8_t b1[4], b2[4];
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        const uint2 raw = *reinterpret_cast<const uint2*>(&Gs[(wc + tj * 16 + lr) * B_GS + kk + 8 * lk]);
        uint16_t v1[8], v2[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int8_t x = (int8_t)(((e < 4 ? raw.x : raw.y) >> (8 * (e & 3))) & 0xFF);
          const bool miss = x < 0;
          // integers 0,1,2 are exact in bf16: 0x0000, 0x3F80, 0x4000
          const uint16_t xi = (x == 0) ? (uint16_t)0 : (x == 1 ? (uint16_t)0x3F80 : (uint16_t)0x4000);
          v1[e] = miss ? m1[tj] : xi;
          v2[e] = miss ? m2[tj] : (uint16_t)0;
        }
        memcpy(&b1[tj], v1, 16);
        memcpy(&b2[tj], v2, 16);
      }
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
        const int row = wr + ti * 16 + lr;
        bf16x8_t ah, am, al;
        memcpy(&ah, &As[0][row * B_AS + kk + 8 * lk], 16);
        memcpy(&am, &As[1][row * B_AS + kk + 8 * lk], 16);
        memcpy(&al, &As[2][row * B_AS + kk + 8 * lk], 16);
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
          f4_t c = acc[ti][tj];
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b1[tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b1[tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b2[tj], c, 0, 0, 0);
          c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b1[tj], c, 0, 0, 0);
          acc[ti][tj] = c;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t s = s0 + wc + tj * 16 + lr;
        if (i < M && s < n_cols) Z[i * ldz + s] = acc[ti][tj][r];
      }
}

// ---------------------------------------------------------------------------
// BF16x3 rotation, v2: 256x128 block tile (8 waves of 64x64), BK = 32,
// genotypes converted to bf16 X1/X2 once per stage in LDS, next K-slab
// prefetched into registers while the current one is multiplied.
// ---------------------------------------------------------------------------
constexpr int V2_BM = 256, V2_BN = 128, V2_BK = 32;
constexpr int V2_AS = V2_BK + 8;  // bf16 per LDS row (80 B, 16-byte aligned)
constexpr int V2_BS = V2_BK + 8;

__device__ __forceinline__ uint16_t geno_bf16(int8_t x, uint16_t m1) {
  // 0 -> 0x0000, 1 -> 0x3F80, 2 -> 0x4000 (exact), missing -> bf16(mean)
  return x < 0 ? m1 : (x == 0 ? (uint16_t)0 : (uint16_t)(0x3F00 + 0x80 * (int)x));
}

__global__ void __launch_bounds__(512) k_rotate_bf16x3_v2(
    const uint16_t* __restrict__ Ahi, const uint16_t* __restrict__ Amid,
    const uint16_t* __restrict__ Alo, int64_t lda, const int8_t* __restrict__ g, int64_t ldg,
    const double* __restrict__ mean, const int32_t* __restrict__ col, int64_t M, int64_t n_cols,
    int64_t K, float* __restrict__ Z, int64_t ldz) {
  __shared__ __attribute__((aligned(16))) uint16_t As[3][V2_BM * V2_AS];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][V2_BN * V2_BS];
  __shared__ uint16_t Bm[2][V2_BN];  // per column bf16(mean), residual
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nrb = (int)((M + V2_BM - 1) / V2_BM), ncb = (int)((n_cols + V2_BN - 1) / V2_BN);
  int rb, cb;
  xcd_tile(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V2_BM, s0 = (int64_t)cb * V2_BN;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid < V2_BN) {
    const int64_t s = s0 + tid;
    const double mu = (s < n_cols) ? mean[col ? (int64_t)col[s] : s] : 0.0;
    const uint16_t m1 = f32_to_bf16_rne((float)mu);
    Bm[0][tid] = m1;
    Bm[1][tid] = f32_to_bf16_rne((float)(mu - bf16_to_f64(m1)));
  }
  // per-thread load slots: A: 3 planes x 256 rows x 64 B = 3072 x 16 B -> 6 per thread
  //                        B: 128 rows x 32 B -> 256 x 16 B: threads 0..255, one each
  const int b_r = tid >> 1, b_c = (tid & 1) * 16;
  const int64_t b_s = s0 + b_r;
  const int64_t b_row = (b_s < n_cols) ? (col ? (int64_t)col[b_s] : b_s) : -1;
  int4 ra[6];
  int4 rbv = {0, 0, 0, 0};
  const uint16_t* Ap[3] = {Ahi, Amid, Alo};
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int idx = tid + 512 * u;        // 0..3071
      const int pl = idx >> 10;             // plane
      const int r = (idx & 1023) >> 2;      // row 0..255
      const int c = (idx & 3) * 8;          // bf16 column 0,8,16,24
      const int64_t i = i0 + r;
      int4 w = {0, 0, 0, 0};
      if (i < M && k0 + c + 8 <= lda) w = *reinterpret_cast<const int4*>(Ap[pl] + i * lda + k0 + c);
      ra[u] = w;
    }
    if (tid < 256) {
      int4 w = {0, 0, 0, 0};
      if (b_row >= 0) {
        if (k0 + b_c + 16 <= K) {
          w = *reinterpret_cast<const int4*>(g + b_row * ldg + k0 + b_c);
        } else {
          int8_t tmp[16];
#pragma unroll
          for (int e = 0; e < 16; ++e)
            tmp[e] = (k0 + b_c + e < K) ? g[b_row * ldg + k0 + b_c + e] : (int8_t)0;
          memcpy(&w, tmp, 16);
        }
      }
      rbv = w;
    }
  };
  auto lstore = [&]() {
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int idx = tid + 512 * u;
      const int pl = idx >> 10;
      const int r = (idx & 1023) >> 2;
      const int c = (idx & 3) * 8;
      *reinterpret_cast<int4*>(&As[pl][r * V2_AS + c]) = ra[u];
    }
    if (tid < 256) {
      const uint16_t m1 = Bm[0][b_r], m2 = Bm[1][b_r];
      int8_t x[16];
      memcpy(x, &rbv, 16);
      uint16_t v1[16], v2[16];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        v1[e] = geno_bf16(x[e], m1);
        v2[e] = x[e] < 0 ? m2 : (uint16_t)0;
      }
      memcpy(&Bs[0][b_r * V2_BS + b_c], v1, 32);
      memcpy(&Bs[1][b_r * V2_BS + b_c], v2, 32);
    }
  };
  f4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};
  __syncthreads();  // Bm visible
  gload(0);
  for (int64_t k0 = 0; k0 < K; k0 += V2_BK) {
    lstore();
    __syncthreads();
    if (k0 + V2_BK < K) gload(k0 + V2_BK);  // prefetch the next slab
    bf16x8_t b1[4], b2[4];
#pragma unroll
    for (int tj = 0; tj < 4; ++tj) {
      const int cidx = (wc + tj * 16 + lr) * V2_BS + 8 * lk;
      memcpy(&b1[tj], &Bs[0][cidx], 16);
      memcpy(&b2[tj], &Bs[1][cidx], 16);
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) {
      const int aidx = (wr + ti * 16 + lr) * V2_AS + 8 * lk;
      bf16x8_t ah, am, al;
      memcpy(&ah, &As[0][aidx], 16);
      memcpy(&am, &As[1][aidx], 16);
      memcpy(&al, &As[2][aidx], 16);
#pragma unroll
      for (int tj = 0; tj < 4; ++tj) {
        f4_t c = acc[ti][tj];
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, b1[tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, b1[tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b2[tj], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, b1[tj], c, 0, 0, 0);
        acc[ti][tj] = c;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 4; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t s = s0 + wc + tj * 16 + lr;
        if (i < M && s < n_cols) Z[i * ldz + s] = acc[ti][tj][r];
      }
}


// ---------------------------------------------------------------------------
// BF16x3 rotation, 256x256 block tile (default)
// ---------------------------------------------------------------------------
// 4 waves, each owning a 128x128 output tile (8x8 MFMA tiles, accumulators in
// AGPRs). Per 32-deep K step the block stages the three UT planes (3 x 256 x
// 32 bf16 = 48 KB) and the int8 genotypes (256 x 32 B = 8 KB) straight from
// global memory into LDS with global_load_lds_dwordx4 (no register staging),
// double buffered, one barrier per step. Genotype fragments are widened to
// bf16 X1 / X2 per fragment. LDS rows are XOR-swizzled in 16-byte chunks so
// that the ds_read_b128 (A) and ds_read_b64 (B) fragment reads of 16 lanes hit
// 64 distinct banks: A chunk c of row r sits at c ^ ((r >> 2) & 3), genotype
// chunk c of column j at c ^ ((j >> 3) & 1).
constexpr int V3_BM = 256, V3_BN = 256, V3_BK = 32;
constexpr int V3_APLANE = V3_BM * V3_BK * 2;      // 16 KB per plane
constexpr int V3_BBYTES = V3_BN * V3_BK;          // 8 KB
constexpr int V3_STAGE = 3 * V3_APLANE + V3_BBYTES;  // 56 KB

typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ void lds_dma16(const void* gsrc, uint8_t* lds_dst) {
  __builtin_amdgcn_global_load_lds(gsrc, (lds_void*)lds_dst, 16, 0, 0);
}

// 8 genotype bytes -> bf16x8 X1 (0/1/2 exact, missing -> m1) and X2 (missing -> m2).
__device__ __forceinline__ void widen8(uint64_t bytes, uint32_t m1, uint32_t m2, bf16x8_t& x1,
                                       bf16x8_t& x2) {
  uint32_t w1[4], w2[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const uint32_t b0 = (uint32_t)(bytes >> (16 * e)) & 0xFFu;
    const uint32_t b1 = (uint32_t)(bytes >> (16 * e + 8)) & 0xFFu;
    const uint32_t v0 = (b0 & 0x80u) ? m1 : (b0 ? 0x3F00u + (b0 << 7) : 0u);
    const uint32_t v1 = (b1 & 0x80u) ? m1 : (b1 ? 0x3F00u + (b1 << 7) : 0u);
    const uint32_t r0 = (b0 & 0x80u) ? m2 : 0u;
    const uint32_t r1 = (b1 & 0x80u) ? m2 : 0u;
    w1[e] = v0 | (v1 << 16);
    w2[e] = r0 | (r1 << 16);
  }
  memcpy(&x1, w1, 16);
  memcpy(&x2, w2, 16);
}

__global__ void __launch_bounds__(256) k_rotate_bf16x3_v3(
    const uint16_t* __restrict__ Ahi, const uint16_t* __restrict__ Amid,
    const uint16_t* __restrict__ Alo, int64_t lda, const int8_t* __restrict__ g, int64_t ldg,
    const double* __restrict__ mean, const int32_t* __restrict__ col, int64_t M, int64_t n_cols,
    int64_t K, float* __restrict__ Z, int64_t ldz) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * V3_STAGE];
  __shared__ uint32_t Bm[2][V3_BN];  // bf16(mean), bf16 residual of the mean
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int nrb = (int)((M + V3_BM - 1) / V3_BM), ncb = (int)((n_cols + V3_BN - 1) / V3_BN);
  int rb, cb;
  xcd_tile(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V3_BM, s0 = (int64_t)cb * V3_BN;
  const int wr = (wave >> 1) * 128, wc = (wave & 1) * 128;
  const int lr = lane & 15, lk = lane >> 4;
  {
    const int64_t sc = s0 + tid;
    const double mu = (sc < n_cols) ? mean[col ? (int64_t)col[sc] : sc] : 0.0;
    const uint16_t h1 = f32_to_bf16_rne((float)mu);
    Bm[0][tid] = h1;
    Bm[1][tid] = f32_to_bf16_rne((float)(mu - bf16_to_f64(h1)));
  }
  // per-lane global sources of the DMA loads (fixed over K; the K offset is added per step)
  // A: instruction q (0..15) of plane p covers rows 16q .. 16q+15; lane -> (row, chunk)
  const uint16_t* Ap[3] = {Ahi, Amid, Alo};
  const int a_r = lane >> 2, a_cs = lane & 3;  // row within 16, stored chunk
  // B: instruction q (0..7) covers columns 32q .. 32q+31; lane -> (column, chunk)
  const int b_j = lane >> 1, b_cs = lane & 1;
  // each wave issues A instructions q = wave*4 .. wave*4+3 of every plane, B q = wave*2, +1
  int64_t a_off[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int row = 16 * (wave * 4 + u) + a_r;
    int64_t gi = i0 + row;
    if (gi >= M) gi = M - 1;  // rows past the end load a valid row; never stored
    const int c = a_cs ^ ((row >> 2) & 3);
    a_off[u] = gi * lda + 8 * c;
  }
  int64_t b_off[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int j = 32 * (wave * 2 + u) + b_j;
    int64_t sc = s0 + j;
    if (sc >= n_cols) sc = n_cols - 1;
    const int64_t grow = col ? (int64_t)col[sc] : sc;
    const int c = b_cs ^ ((j >> 3) & 1);
    b_off[u] = grow * ldg + 16 * c;
  }
  auto issue = [&](int stage, int64_t k0) {
    uint8_t* base = lds + stage * V3_STAGE;
#pragma unroll
    for (int pl = 0; pl < 3; ++pl)
#pragma unroll
      for (int u = 0; u < 4; ++u)
        lds_dma16(Ap[pl] + a_off[u] + k0, base + pl * V3_APLANE + (wave * 4 + u) * 1024);
#pragma unroll
    for (int u = 0; u < 2; ++u)
      lds_dma16(g + b_off[u] + k0, base + 3 * V3_APLANE + (wave * 2 + u) * 1024);
  };

  f4_t acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};

  const int nk = (int)((K + V3_BK - 1) / V3_BK);
  issue(0, 0);
  __syncthreads();  // Bm visible
  uint32_t m1[8], m2[8];
#pragma unroll
  for (int tj = 0; tj < 8; ++tj) {
    m1[tj] = Bm[0][wc + 16 * tj + lr];
    m2[tj] = Bm[1][wc + 16 * tj + lr];
  }
  for (int kt = 0; kt < nk; ++kt) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's DMA loads of stage kt have landed
    __syncthreads();                // ... and everyone's; stage kt-1 is free again
    if (kt + 1 < nk) issue((kt + 1) & 1, (int64_t)(kt + 1) * V3_BK);
    const uint8_t* base = lds + (kt & 1) * V3_STAGE;
    // genotype fragments: raw bytes, widened to X1 once and to X2 in the hi pass
    uint64_t gb[8];
    bf16x8_t x1[8];
#pragma unroll
    for (int tj = 0; tj < 8; ++tj) {
      const int j = wc + 16 * tj + lr;
      const int c = (lk >> 1) ^ ((j >> 3) & 1);
      memcpy(&gb[tj], base + 3 * V3_APLANE + j * 32 + 16 * c + 8 * (lk & 1), 8);
      bf16x8_t unused;
      widen8(gb[tj], m1[tj], m2[tj], x1[tj], unused);
    }
    // planes lo, mid, hi (small terms first)
#pragma unroll 1
    for (int pl = 2; pl >= 0; --pl) {
      bf16x8_t af[8];
#pragma unroll
      for (int ti = 0; ti < 8; ++ti) {
        const int r = wr + 16 * ti + lr;
        const int c = lk ^ ((r >> 2) & 3);
        memcpy(&af[ti], base + pl * V3_APLANE + r * 64 + 16 * c, 16);
      }
      if (pl == 0) {
#pragma unroll
        for (int tj = 0; tj < 8; ++tj) {
          bf16x8_t x2, unused;
          widen8(gb[tj], m1[tj], m2[tj], unused, x2);
#pragma unroll
          for (int ti = 0; ti < 8; ++ti)
            acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ti], x2, acc[ti][tj], 0, 0, 0);
        }
      }
#pragma unroll
      for (int ti = 0; ti < 8; ++ti)
#pragma unroll
        for (int tj = 0; tj < 8; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ti], x1[tj], acc[ti][tj], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ti = 0; ti < 8; ++ti)
#pragma unroll
    for (int tj = 0; tj < 8; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t sc = s0 + wc + tj * 16 + lr;
        if (i < M && sc < n_cols) Z[i * ldz + sc] = acc[ti][tj][r];
      }
}

// Same staging as k_rotate_bf16x3_v3 (256x256 block tile, DMA into double-
// buffered swizzled LDS), 8 waves in a 4 x 2 grid of 64x128 wave tiles: 128
// accumulators per lane, so two waves share each SIMD and hide each other's
// LDS latency.
__global__ void __launch_bounds__(512) k_rotate_bf16x3_v4(
    const uint16_t* __restrict__ Ahi, const uint16_t* __restrict__ Amid,
    const uint16_t* __restrict__ Alo, int64_t lda, const int8_t* __restrict__ g, int64_t ldg,
    const double* __restrict__ mean, const int32_t* __restrict__ col, int64_t M, int64_t n_cols,
    int64_t K, float* __restrict__ Z, int64_t ldz) {
  __shared__ __attribute__((aligned(1024))) uint8_t lds[2 * V3_STAGE];
  __shared__ uint32_t Bm[2][V3_BN];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;  // 8 waves
  const int nrb = (int)((M + V3_BM - 1) / V3_BM), ncb = (int)((n_cols + V3_BN - 1) / V3_BN);
  int rb, cb;
  xcd_tile(blockIdx.x, nrb, ncb, rb, cb);
  const int64_t i0 = (int64_t)rb * V3_BM, s0 = (int64_t)cb * V3_BN;
  const int wr = (wave >> 1) * 64, wc = (wave & 1) * 128;
  const int lr = lane & 15, lk = lane >> 4;
  if (tid < V3_BN) {
    const int64_t sc = s0 + tid;
    const double mu = (sc < n_cols) ? mean[col ? (int64_t)col[sc] : sc] : 0.0;
    const uint16_t h1 = f32_to_bf16_rne((float)mu);
    Bm[0][tid] = h1;
    Bm[1][tid] = f32_to_bf16_rne((float)(mu - bf16_to_f64(h1)));
  }
  const uint16_t* Ap[3] = {Ahi, Amid, Alo};
  // A: 48 DMA instructions per stage (3 planes x 16 row groups), 6 per wave;
  // source offsets are recomputed per issue to keep registers for the MFMAs
  const int a_rl = lane >> 2, a_cs = lane & 3;
  // B: 8 DMA instructions per stage (32 columns each), one per wave
  int64_t b_off;
  {
    const int j = 32 * wave + (lane >> 1);
    int64_t sc = s0 + j;
    if (sc >= n_cols) sc = n_cols - 1;
    const int64_t grow = col ? (int64_t)col[sc] : sc;
    const int c = (lane & 1) ^ ((j >> 3) & 1);
    b_off = grow * ldg + 16 * c;
  }
  auto issue = [&](int stage, int64_t k0) {
    uint8_t* base = lds + stage * V3_STAGE;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int q = wave * 6 + u;
      const int row = 16 * (q & 15) + a_rl;
      int64_t gi = i0 + row;
      if (gi >= M) gi = M - 1;  // rows past the end load a valid row; never stored
      const int c = a_cs ^ ((row >> 2) & 3);
      lds_dma16(Ap[q >> 4] + gi * lda + 8 * c + k0, base + (q >> 4) * V3_APLANE + (q & 15) * 1024);
    }
    lds_dma16(g + b_off + k0, base + 3 * V3_APLANE + wave * 1024);
  };

  f4_t acc[4][8];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = (f4_t){0.f, 0.f, 0.f, 0.f};

  const int nk = (int)((K + V3_BK - 1) / V3_BK);
  issue(0, 0);
  __syncthreads();  // Bm visible
  uint32_t mm[8];  // bf16(mean) | residual << 16
#pragma unroll
  for (int tj = 0; tj < 8; ++tj) mm[tj] = Bm[0][wc + 16 * tj + lr] | (Bm[1][wc + 16 * tj + lr] << 16);
  for (int kt = 0; kt < nk; ++kt) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (kt + 1 < nk) issue((kt + 1) & 1, (int64_t)(kt + 1) * V3_BK);
    const uint8_t* base = lds + (kt & 1) * V3_STAGE;
    auto gbytes = [&](int tj) -> uint64_t {
      const int j = wc + 16 * tj + lr;
      const int c = (lk >> 1) ^ ((j >> 3) & 1);
      uint64_t v;
      memcpy(&v, base + 3 * V3_APLANE + j * 32 + 16 * c + 8 * (lk & 1), 8);
      return v;
    };
    bf16x8_t x1[8];
#pragma unroll
    for (int tj = 0; tj < 8; ++tj) {
      bf16x8_t unused;
      widen8(gbytes(tj), mm[tj] & 0xFFFFu, mm[tj] >> 16, x1[tj], unused);
    }
    // planes lo, mid, hi x X1, and hi x X2 (small terms first)
#pragma unroll 1
    for (int pl = 2; pl >= 0; --pl) {
      bf16x8_t af[4];
#pragma unroll
      for (int ti = 0; ti < 4; ++ti) {
        const int r = wr + 16 * ti + lr;
        const int c = lk ^ ((r >> 2) & 3);
        memcpy(&af[ti], base + pl * V3_APLANE + r * 64 + 16 * c, 16);
      }
      if (pl == 0) {
#pragma unroll
        for (int tj = 0; tj < 8; ++tj) {
          bf16x8_t x2, unused;
          widen8(gbytes(tj), mm[tj] & 0xFFFFu, mm[tj] >> 16, unused, x2);
#pragma unroll
          for (int ti = 0; ti < 4; ++ti)
            acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ti], x2, acc[ti][tj], 0, 0, 0);
        }
      }
#pragma unroll
      for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 8; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[ti], x1[tj], acc[ti][tj], 0, 0, 0);
    }
  }
#pragma unroll
  for (int ti = 0; ti < 4; ++ti)
#pragma unroll
    for (int tj = 0; tj < 8; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wr + ti * 16 + 4 * lk + r;
        const int64_t sc = s0 + wc + tj * 16 + lr;
        if (i < M && sc < n_cols) Z[i * ldz + sc] = acc[ti][tj][r];
      }
}

}  // namespace jm

using namespace jm;

extern "C" int jm_rotate_f64(const double* UT, int64_t ld_u, const int8_t* g, int64_t ld_g,
                             const double* mean, const int32_t* col_idx, int64_t n, int64_t n_snps,
                             double* Z, int64_t ldz, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(UT && g && mean && Z, "null pointer");
  JM_CHECK_ARG(ld_u >= n && ld_g >= n && ldz >= n_snps, "bad leading dimension");
  JM_CHECK_ARG(ld_g % 16 == 0 && ((uintptr_t)g % 16) == 0, "g must be 16-byte aligned rows");
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

extern "C" int jm_split_bf16x3(const double* UT, int64_t ld_u, int64_t n, uint16_t* hi,
                               uint16_t* mid, uint16_t* lo, int64_t ld_s, void* stream) {
  JM_CHECK_ARG(n >= 1 && ld_u >= n && ld_s >= n && ld_s % 32 == 0, "bad sizes");
  JM_CHECK_ARG(UT && hi && mid && lo, "null pointer");
  hipLaunchKernelGGL(k_split_bf16x3, dim3(2048), dim3(256), 0, (hipStream_t)stream, UT, ld_u, n,
                     hi, mid, lo, ld_s);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_rotate_bf16x3(const uint16_t* hi, const uint16_t* mid, const uint16_t* lo,
                                int64_t ld_s, const int8_t* g, int64_t ld_g, const double* mean,
                                const int32_t* col_idx, int64_t n, int64_t n_snps, float* Z,
                                int64_t ldz, void* stream) {
  JM_CHECK_ARG(n >= 1 && n_snps >= 0, "bad sizes");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(hi && mid && lo && g && mean && Z, "null pointer");
  JM_CHECK_ARG(ld_s >= n && ld_s % 32 == 0 && ld_g >= n && ld_g % 16 == 0 && ldz >= n_snps,
               "bad leading dimension");
  JM_CHECK_ARG(((uintptr_t)g % 16) == 0 && ((uintptr_t)hi % 16) == 0 &&
                   ((uintptr_t)mid % 16) == 0 && ((uintptr_t)lo % 16) == 0,
               "operands must be 16-byte aligned");
  // tile variant (JM_ROT_BF16_V): 3 = 256x256, DMA-staged, double buffered (default);
  // 2 = 256x128 / 512 threads with register prefetch; 1 = 128x128 / 256 threads
  const char* venv = getenv("JM_ROT_BF16_V");
  const int version = venv ? atoi(venv) : 3;
  if (version == 4) {
    JM_CHECK_ARG(ld_g >= ((n + 31) / 32) * 32, "ld_g must be >= n rounded up to 32");
    const int nrb = ceil_div(n, V3_BM), ncb = ceil_div(n_snps, V3_BN);
    hipLaunchKernelGGL(k_rotate_bf16x3_v4, dim3(nrb * ncb), dim3(512), 0, (hipStream_t)stream, hi,
                       mid, lo, ld_s, g, ld_g, mean, col_idx, n, n_snps, n, Z, ldz);
  } else if (version == 3) {
    JM_CHECK_ARG(ld_g >= ((n + 31) / 32) * 32, "ld_g must be >= n rounded up to 32");
    const int nrb = ceil_div(n, V3_BM), ncb = ceil_div(n_snps, V3_BN);
    hipLaunchKernelGGL(k_rotate_bf16x3_v3, dim3(nrb * ncb), dim3(256), 0, (hipStream_t)stream, hi,
                       mid, lo, ld_s, g, ld_g, mean, col_idx, n, n_snps, n, Z, ldz);
  } else if (version == 2) {
    const int nrb = ceil_div(n, V2_BM), ncb = ceil_div(n_snps, V2_BN);
    hipLaunchKernelGGL(k_rotate_bf16x3_v2, dim3(nrb * ncb), dim3(512), 0, (hipStream_t)stream, hi,
                       mid, lo, ld_s, g, ld_g, mean, col_idx, n, n_snps, n, Z, ldz);
  } else {
    const int nrb = ceil_div(n, B_BM), ncb = ceil_div(n_snps, B_BN);
    hipLaunchKernelGGL(k_rotate_bf16x3, dim3(nrb * ncb), dim3(256), 0, (hipStream_t)stream, hi,
                       mid, lo, ld_s, g, ld_g, mean, col_idx, n, n_snps, n, Z, ldz);
  }
  JM_CHECK_LAUNCH();
  return JM_OK;
}
