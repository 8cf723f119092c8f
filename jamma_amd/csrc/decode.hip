// PLINK .bed decoding on the device, fused with the pass-1 per-SNP counts.
//
// Replaces the bed-reader windowed read used by stream_genotype_chunks
// (src/jamma/io/plink.py:239-300; bed-reader 1.0.6, count_A1 = True) and the
// pass-1 statistics (src/jamma/lmm/runner_streaming.py:248-284,
// src/jamma/kinship/compute.py:601-611): the host turns the four counts into
// the float32 nanmean, missing rate and var > 0 exactly as NumPy does.
//
// 2-bit code (low bits first, four samples per byte) -> A1 count:
//   00 -> 2, 01 -> missing (-1), 10 -> 1, 11 -> 0.
// One block per SNP row; the row's bytes are read once (through L2), the int8
// output row is written coalesced, and the counts are reduced in the block.
// The kernel is HBM-bound: ceil(n_total/4) bytes in, n_out bytes out per SNP.

#include "jm_common.h"

#include <algorithm>

namespace jm {

__global__ void __launch_bounds__(256) k_decode_bed(const uint8_t* __restrict__ bed,
                                                    int64_t bytes_per_snp,
                                                    const int32_t* __restrict__ snp_idx,
                                                    int64_t n_snps,
                                                    const int32_t* __restrict__ sample_idx,
                                                    int64_t n_out, int8_t* __restrict__ g,
                                                    int64_t ld_g, int32_t* __restrict__ counts) {
  __shared__ int red[4][4];
  const int64_t row = blockIdx.x;
  if (row >= n_snps) return;
  const int64_t src = snp_idx ? (int64_t)snp_idx[row] : row;
  const uint8_t* b = bed + src * bytes_per_snp;
  int c0 = 0, c1 = 0, c2 = 0, cm = 0;
  const int64_t lim = g ? ld_g : n_out;
  for (int64_t o = threadIdx.x; o < lim; o += blockDim.x) {
    int8_t v = 0;
    if (o < n_out) {
      const int64_t t = sample_idx ? (int64_t)sample_idx[o] : o;
      const int code = (b[t >> 2] >> (2 * (t & 3))) & 3;
      v = (code == 0) ? 2 : (code == 2 ? 1 : (code == 3 ? 0 : -1));
      c0 += (v == 0);
      c1 += (v == 1);
      c2 += (v == 2);
      cm += (v < 0);
    }
    if (g) g[row * ld_g + o] = v;
  }
  if (counts) {
    // wave reduce then block reduce (fixed order -> deterministic)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      c0 += __shfl_xor(c0, m, 64);
      c1 += __shfl_xor(c1, m, 64);
      c2 += __shfl_xor(c2, m, 64);
      cm += __shfl_xor(cm, m, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      red[w][0] = c0;
      red[w][1] = c1;
      red[w][2] = c2;
      red[w][3] = cm;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
      int acc = 0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) acc += red[k][threadIdx.x];
      counts[row * 4 + threadIdx.x] = acc;
    }
  }
}

// 64x64 LDS-tiled int8 transpose.
__global__ void __launch_bounds__(256) k_transpose_i8(const int8_t* __restrict__ in, int64_t rows,
                                                      int64_t cols, int64_t ld_in,
                                                      int8_t* __restrict__ out, int64_t ld_out,
                                                      int8_t* __restrict__ out0, int64_t ld_out0) {
  __shared__ int8_t tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) {
    const int64_t r = r0 + k, c = c0 + tx;
    tile[k][tx] = (r < rows && c < cols) ? in[r * ld_in + c] : (int8_t)0;
  }
  __syncthreads();
  for (int k = ty; k < 64; k += 4) {
    const int64_t c = c0 + k, r = r0 + tx;  // out row c, column r
    if (c < cols && r < rows) {
      const int8_t v = tile[tx][k];
      if (out) out[c * ld_out + r] = v;
      if (out0) out0[c * ld_out0 + r] = v < 0 ? (int8_t)0 : v;  // missing -> 0
    }
  }
}

}  // namespace jm

using namespace jm;

extern "C" int jm_decode_bed(const uint8_t* bed, int64_t bytes_per_snp, const int32_t* snp_idx,
                             int64_t n_snps, const int32_t* sample_idx, int64_t n_out, int8_t* g,
                             int64_t ld_g, int32_t* counts, void* stream) {
  JM_CHECK_ARG(n_snps >= 0 && n_out >= 0 && bytes_per_snp >= 0, "negative size");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(bed != nullptr, "null bed");
  JM_CHECK_ARG(g == nullptr || ld_g >= n_out, "ld_g < n_out");
  JM_CHECK_ARG(sample_idx != nullptr || (n_out + 3) / 4 <= bytes_per_snp,
               "n_out exceeds the samples of a .bed row");
  JM_CHECK_ARG(n_snps < (int64_t)1 << 31, "too many SNPs in one call");
  hipLaunchKernelGGL(k_decode_bed, dim3((unsigned)n_snps), dim3(256), 0, (hipStream_t)stream, bed,
                     bytes_per_snp, snp_idx, n_snps, sample_idx, n_out, g, ld_g, counts);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

namespace jm {

// The same transpose in 128 x 128 tiles with 16-byte global loads and stores (each
// lane moves 16 bytes per access instead of one); used when every row start is
// 16-byte aligned. Tile row stride 132 bytes in LDS (4-byte aligned rows, the
// column gathers of the store phase spread over the banks).
constexpr int TR_T = 128, TR_S = 132;
__global__ void __launch_bounds__(256) k_transpose_i8_v16(const int8_t* __restrict__ in,
                                                          int64_t rows, int64_t cols,
                                                          int64_t ld_in, int8_t* __restrict__ out,
                                                          int64_t ld_out, int8_t* __restrict__ out0,
                                                          int64_t ld_out0) {
  __shared__ __align__(16) int8_t tile[TR_T * TR_S];
  const int64_t r0 = (int64_t)blockIdx.y * TR_T, c0 = (int64_t)blockIdx.x * TR_T;
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = t + 256 * q;
    const int row = idx >> 3, ch = (idx & 7) * 16;
    const int64_t r = r0 + row, c = c0 + ch;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    if (r < rows) {
      const int8_t* src = in + r * ld_in + c;
      if (c + 16 <= cols) {
        const uint4 v = *reinterpret_cast<const uint4*>(src);
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
      } else {
        for (int e = 0; e < 16 && c + e < cols; ++e)
          w[e >> 2] |= (uint32_t)(uint8_t)src[e] << (8 * (e & 3));
      }
    }
    uint32_t* dst = reinterpret_cast<uint32_t*>(tile + row * TR_S + ch);
#pragma unroll
    for (int e = 0; e < 4; ++e) dst[e] = w[e];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = t + 256 * q;
    const int col = idx >> 3, rch = (idx & 7) * 16;
    const int64_t c = c0 + col, r = r0 + rch;
    if (c >= cols || r >= rows) continue;
    uint32_t w[4] = {0u, 0u, 0u, 0u}, w0[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const uint8_t b = (uint8_t)tile[(rch + e) * TR_S + col];
      w[e >> 2] |= (uint32_t)b << (8 * (e & 3));
      w0[e >> 2] |= (uint32_t)((b & 0x80u) ? 0u : b) << (8 * (e & 3));  // missing -> 0
    }
    if (r + 16 <= rows) {
      if (out) *reinterpret_cast<uint4*>(out + c * ld_out + r) = make_uint4(w[0], w[1], w[2], w[3]);
      if (out0)
        *reinterpret_cast<uint4*>(out0 + c * ld_out0 + r) = make_uint4(w0[0], w0[1], w0[2], w0[3]);
    } else {
      for (int e = 0; r + e < rows; ++e) {
        if (out) out[c * ld_out + r + e] = (int8_t)(w[e >> 2] >> (8 * (e & 3)));
        if (out0) out0[c * ld_out0 + r + e] = (int8_t)(w0[e >> 2] >> (8 * (e & 3)));
      }
    }
  }
}

}  // namespace jm

extern "C" int jm_transpose_i8(const int8_t* in, int64_t rows, int64_t cols, int64_t ld_in,
                               int8_t* out, int64_t ld_out, int8_t* out0, int64_t ld_out0,
                               void* stream) {
  JM_CHECK_ARG(rows >= 0 && cols >= 0 && ld_in >= cols, "bad sizes");
  JM_CHECK_ARG((out == nullptr || ld_out >= rows) && (out0 == nullptr || ld_out0 >= rows),
               "bad output leading dimension");
  if (rows == 0 || cols == 0) return JM_OK;
  JM_CHECK_ARG(in && (out || out0), "null pointer");
  auto a16 = [](const void* p, int64_t ld) { return ((uintptr_t)p % 16) == 0 && ld % 16 == 0; };
  if (a16(in, ld_in) && (!out || a16(out, ld_out)) && (!out0 || a16(out0, ld_out0))) {
    const dim3 grid((unsigned)ceil_div(cols, TR_T), (unsigned)ceil_div(rows, TR_T));
    hipLaunchKernelGGL(k_transpose_i8_v16, grid, dim3(256), 0, (hipStream_t)stream, in, rows,
                       cols, ld_in, out, ld_out, out0, ld_out0);
    JM_CHECK_LAUNCH();
    return JM_OK;
  }
  const dim3 grid((unsigned)ceil_div(cols, 64), (unsigned)ceil_div(rows, 64));
  hipLaunchKernelGGL(k_transpose_i8, grid, dim3(256), 0, (hipStream_t)stream, in, rows, cols,
                     ld_in, out, ld_out, out0, ld_out0);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

// ---------------------------------------------------------------------------
// Dosage (non-integer) genotypes in float64
// ---------------------------------------------------------------------------
namespace jm {

// impute_and_center (kinship/missing.py:22-61) of a sample-major tile X (n x cols,
// row stride ld): per column j, mu_j = mean of the non-NaN entries (0 when all are
// NaN); X[i][j] <- X[i][j] - mu_j, NaN -> 0. One thread per column, rows walked in
// order (coalesced across the threads of a wave; a fixed summation order).
__global__ void __launch_bounds__(256) k_center_columns_f64(double* __restrict__ X, int64_t n,
                                                            int64_t cols, int64_t ld,
                                                            double* __restrict__ mean_out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  double s = 0.0;
  int64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double v = X[i * ld + j];
    if (v == v) {
      s += v;
      ++cnt;
    }
  }
  const double mu = cnt > 0 ? s / (double)cnt : 0.0;
  if (mean_out) mean_out[j] = mu;
  for (int64_t i = 0; i < n; ++i) {
    const double v = X[i * ld + j];
    X[i * ld + j] = (v == v) ? v - mu : 0.0;
  }
}

// SNP-major rows X (rows x n, row stride ld): NaN -> mean[r] (mean imputation of the
// association runners, runner_jax.py / runner_streaming.py:429-434).
__global__ void __launch_bounds__(256) k_impute_rows_f64(double* __restrict__ X, int64_t rows,
                                                         int64_t n, int64_t ld,
                                                         const double* __restrict__ mean) {
  const int64_t r = blockIdx.y;
  if (r >= rows) return;
  const double m = mean[r];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = X[r * ld + i];
    if (!(v == v)) X[r * ld + i] = m;
  }
}

}  // namespace jm

extern "C" int jm_center_columns_f64(double* X, int64_t n, int64_t cols, int64_t ld,
                                     double* mean_out, void* stream) {
  JM_CHECK_ARG(n >= 1 && cols >= 0 && ld >= cols, "bad sizes");
  if (cols == 0) return JM_OK;
  JM_CHECK_ARG(X != nullptr, "null X");
  hipLaunchKernelGGL(jm::k_center_columns_f64, dim3((unsigned)((cols + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, X, n, cols, ld, mean_out);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_impute_rows_f64(double* X, int64_t rows, int64_t n, int64_t ld,
                                  const double* mean, void* stream) {
  JM_CHECK_ARG(rows >= 0 && n >= 1 && ld >= n, "bad sizes");
  JM_CHECK_ARG(rows <= 65535, "at most 65535 rows per call");
  if (rows == 0) return JM_OK;
  JM_CHECK_ARG(X && mean, "null pointer");
  const unsigned bx = (unsigned)std::min<int64_t>((n + 255) / 256, 64);
  hipLaunchKernelGGL(jm::k_impute_rows_f64, dim3(bx, (unsigned)rows), dim3(256), 0,
                     (hipStream_t)stream, X, rows, n, ld, mean);
  JM_CHECK_LAUNCH();
  return JM_OK;
}
