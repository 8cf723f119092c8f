// Device kernels of the large-cohort symmetric eigendecomposition.
//
// Replaces np.linalg.eigh (src/jamma/lmm/eigen.py:94) for kinship matrices
// whose n*n element count exceeds 2^31 - 1 (n > 46,340), where the library
// tridiagonal solvers, which take 32-bit sizes, no longer apply. Every offset
// here is 64-bit. The host side (jamma_amd/lmm/eigen.py) runs:
//
//  1. blocked Householder reduction to tridiagonal form (lower, LAPACK dsytrd /
//     dlatrd order of operations): per column the update by the panel's earlier
//     reflectors (k_panel_update), the reflector (k_larfg_unit), the symmetric
//     matrix-vector product (k_symv_lower / k_symv_reduce, the lower triangle
//     read once), the panel's corrections (k_panel_dots / k_panel_apply) and the
//     end of the step (scale by tau, alpha = -tau/2 w.v, w += alpha v:
//     k_latrd_finish); per panel the rank-2k update of the trailing matrix
//     (k_dgemm);
//  2. Cuppen divide and conquer on the tridiagonal matrix: tear it into halves
//     (recursively) until a half fits the library solver, then merge each pair
//     through the rank-one update D + rho z z^T (Bunch-Nielsen-Sorensen):
//     deflation on the host, the secular equation's roots (k_secular_roots),
//     the Gu-Eisenstat (Loewner) recomputation of z (k_secular_zhat) and the
//     normalised eigenvectors (k_secular_norms, k_secular_fill), the Givens
//     rotations of deflated pairs (k_apply_rotations), then
//     Q = blockdiag(Q1, Q2) M with k_dgemm;
//  3. back-transformation Q <- H_1 ... H_{n-1} Q in blocks of reflectors
//     (compact WY, I - V T V^T) with k_dgemm.
//
// k_dgemm is an FP64 MFMA GEMM (v_mfma_f64_16x16x4f64) in column-major BLAS
// semantics: C = alpha op(A) op(B) + beta C. 64x64 block tile (i x j), 4 waves
// of 32x32, 32-deep K steps staged through LDS in [index][k] order whatever the
// operand's storage order, so all four transpose combinations share one inner
// loop; the next step's global loads are issued into registers before the
// current step's MFMAs (one FP64 MFMA occupies a SIMD for 64 cycles, so the 32
// per wave and step cover the load latency). The MFMA's B slot takes the i side, so the 16 lanes of an output group
// store 16 consecutive rows of the column-major C.

#include "jm_common.h"

namespace jm {

typedef double d4e_t __attribute__((ext_vector_type(4)));

constexpr int DG_TI = 64, DG_TJ = 64, DG_BK = 32, DG_LS = DG_BK + 1;

// Global -> register loads of one K step: op(A) tile 64 x 32 and op(B) tile
// 32 x 64, 8 doubles each per thread, with consecutive lanes on the operand's
// contiguous dimension.
// K indices from k1 on belong to the second operand pair (A2, B2) at kk - k1
// (k1 >= k: one pair).
template <bool TA, bool TB>
__device__ __forceinline__ void dgemm_fetch(const double* __restrict__ A,
                                            const double* __restrict__ A2, int64_t lda,
                                            const double* __restrict__ B,
                                            const double* __restrict__ B2, int64_t ldb,
                                            int64_t k1, int64_t m, int64_t n, int64_t i0,
                                            int64_t j0, int64_t k0, int64_t kend, int tid,
                                            double (&ra)[8], double (&rb)[8]) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int idx = tid + 256 * u;
    const int r = TA ? (idx >> 5) : (idx & 63), c = TA ? (idx & 31) : (idx >> 6);
    const int64_t i = i0 + r, kk = k0 + c;
    const bool two = kk >= k1;
    const double* Ab = two ? A2 : A;
    const int64_t kq = two ? kk - k1 : kk;
    ra[u] = (i < m && kk < kend) ? (TA ? Ab[kq + i * lda] : Ab[i + kq * lda]) : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int idx = tid + 256 * u;
    const int r = TB ? (idx & 63) : (idx >> 5), c = TB ? (idx >> 6) : (idx & 31);
    const int64_t j = j0 + r, kk = k0 + c;
    const bool two = kk >= k1;
    const double* Bb = two ? B2 : B;
    const int64_t kq = two ? kk - k1 : kk;
    rb[u] = (j < n && kk < kend) ? (TB ? Bb[j + kq * ldb] : Bb[kq + j * ldb]) : 0.0;
  }
}

template <bool TA, bool TB>
__device__ __forceinline__ void dgemm_stash(double* As, double* Bs, int tid, const double (&ra)[8],
                                            const double (&rb)[8]) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int idx = tid + 256 * u;
    const int r = TA ? (idx >> 5) : (idx & 63), c = TA ? (idx & 31) : (idx >> 6);
    As[r * DG_LS + c] = ra[u];
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int idx = tid + 256 * u;
    const int r = TB ? (idx & 63) : (idx >> 5), c = TB ? (idx >> 6) : (idx & 31);
    Bs[r * DG_LS + c] = rb[u];
  }
}

template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_dgemm(int64_t m, int64_t n, int64_t k, double alpha,
                                               const double* __restrict__ A, int64_t lda,
                                               const double* __restrict__ B, int64_t ldb,
                                               double beta, double* __restrict__ C, int64_t ldc,
                                               int lower_only, int64_t nti, int64_t kchunk,
                                               int64_t cstride, const double* __restrict__ A2,
                                               const double* __restrict__ B2, int64_t k1) {
  __shared__ double As[DG_TI * DG_LS];  // op(A)(i, kk) at [i][kk]
  __shared__ double Bs[DG_TJ * DG_LS];  // op(B)(kk, j) at [j][kk]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t bid = blockIdx.x;
  // split K: block row blockIdx.y takes k in [kbeg, kend) and writes C + blockIdx.y * cstride
  const int64_t kbeg = (int64_t)blockIdx.y * kchunk;
  const int64_t kend = (kbeg + kchunk < k) ? kbeg + kchunk : k;
  C += (int64_t)blockIdx.y * cstride;
  const int64_t i0 = (bid % nti) * DG_TI, j0 = (bid / nti) * DG_TJ;
  if (lower_only && i0 + DG_TI <= j0) return;  // every element of the tile has i < j
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  d4e_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (d4e_t){0.0, 0.0, 0.0, 0.0};
  double ra[8], rb[8];
  if (kbeg < kend)
    dgemm_fetch<TA, TB>(A, A2, lda, B, B2, ldb, k1, m, n, i0, j0, kbeg, kend, tid, ra, rb);
  for (int64_t k0 = kbeg; k0 < kend; k0 += DG_BK) {
    dgemm_stash<TA, TB>(As, Bs, tid, ra, rb);
    __syncthreads();
    if (k0 + DG_BK < kend)  // the next step's loads fly while this step multiplies
      dgemm_fetch<TA, TB>(A, A2, lda, B, B2, ldb, k1, m, n, i0, j0, k0 + DG_BK, kend, tid, ra,
                          rb);
#pragma unroll
    for (int kk = 0; kk < DG_BK; kk += 4) {
      double bi[2], aj[2];
#pragma unroll
      for (int ti = 0; ti < 2; ++ti) bi[ti] = As[(wi + ti * 16 + lr) * DG_LS + kk + lk];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj) aj[tj] = Bs[(wj + tj * 16 + lr) * DG_LS + kk + lk];
#pragma unroll
      for (int tj = 0; tj < 2; ++tj)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
          acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(aj[tj], bi[ti], acc[ti][tj], 0, 0, 0);
    }
    __syncthreads();
  }
  // accumulator element r of lane l: D[(l / 16) + 4 r][l % 16] = C(i = l % 16, j = l / 16 + 4 r)
#pragma unroll
  for (int ti = 0; ti < 2; ++ti)
#pragma unroll
    for (int tj = 0; tj < 2; ++tj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + wi + ti * 16 + lr;
        const int64_t j = j0 + wj + tj * 16 + lk + 4 * r;
        if (i < m && j < n) {
          double v = alpha * acc[ti][tj][r];
          if (beta != 0.0) v += beta * C[i + j * ldc];
          C[i + j * ldc] = v;
        }
      }
}

// C = alpha sum_s W_s + beta C, W_s = W + s m n (column-major m x n, ld m), s in a
// fixed order.
__global__ void __launch_bounds__(256) k_splitk_sum(int64_t m, int64_t n, int splits, double alpha,
                                                    const double* __restrict__ W, double beta,
                                                    double* __restrict__ C, int64_t ldc) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m * n) return;
  const int64_t i = q % m, j = q / m;
  double s = 0.0;
  for (int t = 0; t < splits; ++t) s += W[t * m * n + q];
  double v = alpha * s;
  if (beta != 0.0) v += beta * C[i + j * ldc];
  C[i + j * ldc] = v;
}

// ---------------------------------------------------------------------------
// Tridiagonal reduction helpers
// ---------------------------------------------------------------------------
constexpr int PANEL_MAX = 128;  // reflectors per panel (panel column index i < PANEL_MAX)

// dlatrd's column update with the panel's earlier reflectors, one kernel for the
// two matrix-vector products: y[r] -= sum_{c < i} a0[r + c lda] xw[c] + w0[r + c ldw] xa[c]
// (r < len), where xw[c] = w0[c ldw] and xa[c] = a0[c lda] are the panels' first rows
// (A_loc(i, c) and W_loc(i, c)); y is column i from row i (it does not overlap them).
__global__ void __launch_bounds__(256) k_panel_update(int64_t len, int i,
                                                      const double* __restrict__ a0, int64_t lda,
                                                      const double* __restrict__ w0, int64_t ldw,
                                                      double* __restrict__ y) {
  __shared__ double xs[2][PANEL_MAX];
  if (threadIdx.x < (unsigned)i) {
    xs[0][threadIdx.x] = w0[(int64_t)threadIdx.x * ldw];
    xs[1][threadIdx.x] = a0[(int64_t)threadIdx.x * lda];
  }
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < len;
       r += (int64_t)gridDim.x * blockDim.x) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 8
    for (int c = 0; c < i; ++c) {
      s0 += a0[r + c * lda] * xs[0][c];
      s1 += w0[r + c * ldw] * xs[1][c];
    }
    y[r] -= s0 + s1;
  }
}

// Partial column dot products of the two panels with v: block (b, q) sums column q
// (q < i: Wp's column q, else Ap's column q - i) times v over rows
// [b PD_ROWS, (b + 1) PD_ROWS), 4 rows per thread, into part[b][q].
constexpr int PD_ROWS = 1024;  // rows per block
__global__ void __launch_bounds__(256) k_panel_dots(int64_t len, int i,
                                                    const double* __restrict__ ap, int64_t lda,
                                                    const double* __restrict__ wp, int64_t ldw,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ part) {
  __shared__ double red[4];
  const int q = blockIdx.y;
  const double* col = q < i ? wp + (int64_t)q * ldw : ap + (int64_t)(q - i) * lda;
  const int64_t r0 = (int64_t)blockIdx.x * PD_ROWS;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < PD_ROWS / 256; ++k) {
    const int64_t r = r0 + threadIdx.x + 256 * k;
    if (r < len) s += col[r] * v[r];
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    part[(int64_t)blockIdx.x * 2 * PANEL_MAX + q] = (red[0] + red[1]) + (red[2] + red[3]);
}

// t1 = sum of the W-panel partials, t2 = sum of the A-panel partials (block order);
// block 0 stores t2 = Ap^T v (the panel's V^T v_i, kept for the compact-WY factor);
// w[r] -= sum_c Ap(r, c) t1[c] + Wp(r, c) t2[c].
__global__ void __launch_bounds__(256) k_panel_apply(int64_t len, int i, int nblk,
                                                     const double* __restrict__ ap, int64_t lda,
                                                     const double* __restrict__ wp, int64_t ldw,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ gram_col,
                                                     double* __restrict__ w) {
  __shared__ double ts[2 * PANEL_MAX];
  if (threadIdx.x < (unsigned)(2 * i)) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[(int64_t)b * 2 * PANEL_MAX + threadIdx.x];
    ts[threadIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x >= (unsigned)i) gram_col[threadIdx.x - i] = s;
  }
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < len;
       r += (int64_t)gridDim.x * blockDim.x) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll 8
    for (int c = 0; c < i; ++c) {
      s0 += ap[r + c * lda] * ts[c];
      s1 += wp[r + c * ldw] * ts[i + c];
    }
    w[r] -= s0 + s1;
  }
}

// Householder reflector (LAPACK dlarfg) on (alpha, x[0 .. len-2]) in one block, and
// the step that follows it: e = beta, alpha's slot = 1 (the reflector's unit entry).
// ||x|| is formed as scale * sqrt(sum (x / scale)^2), scale = max |x|.
__global__ void __launch_bounds__(1024) k_larfg_unit(int64_t len, double* __restrict__ alpha_p,
                                                     double* __restrict__ x,
                                                     double* __restrict__ tau,
                                                     double* __restrict__ e) {
  __shared__ double part[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t m = len - 1;
  double mx = 0.0;
  for (int64_t q = threadIdx.x; q < m; q += blockDim.x) mx = fmax(mx, fabs(x[q]));
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) mx = fmax(mx, shfl_xor_d(mx, o));
  if (lane == 0) part[wv] = mx;
  __syncthreads();
  double scale = 0.0;
  for (int q = 0; q < nw; ++q) scale = fmax(scale, part[q]);
  __syncthreads();
  double ss = 0.0;
  if (scale > 0.0) {
    const double inv = 1.0 / scale;
    for (int64_t q = threadIdx.x; q < m; q += blockDim.x) {
      const double t = x[q] * inv;
      ss += t * t;
    }
  }
  ss = wave_sum(ss);
  if (lane == 0) part[wv] = ss;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < nw; ++q) tot += part[q];
  const double alpha = *alpha_p;
  const double xnorm = scale * sqrt(tot);
  __syncthreads();
  if (xnorm == 0.0) {  // H = I
    if (threadIdx.x == 0) {
      *tau = 0.0;
      *e = alpha;
      *alpha_p = 1.0;
    }
    return;
  }
  const double beta = -copysign(hypot(alpha, xnorm), alpha);
  const double f = 1.0 / (alpha - beta);
  for (int64_t q = threadIdx.x; q < m; q += blockDim.x) x[q] *= f;
  if (threadIdx.x == 0) {
    *tau = (beta - alpha) / beta;
    *e = beta;
    *alpha_p = 1.0;
  }
}

// One block: w *= tau; alpha = -tau/2 (w . v); w += alpha v   (len elements)
__global__ void __launch_bounds__(1024) k_latrd_finish(int64_t len, const double* __restrict__ tau,
                                                       double* __restrict__ w,
                                                       const double* __restrict__ v) {
  __shared__ double part[16];
  const double t = *tau;
  double s = 0.0;
  for (int64_t q = threadIdx.x; q < len; q += blockDim.x) {
    const double x = t * w[q];
    w[q] = x;
    s += x * v[q];
  }
  s = wave_sum(s);
  const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) part[wave] = s;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < nw; ++q) tot += part[q];
  const double alpha = -0.5 * t * tot;
  for (int64_t q = threadIdx.x; q < len; q += blockDim.x) w[q] += alpha * v[q];
}

// y = A x for a symmetric A stored in its lower triangle (column-major, lda),
// reading every stored element once. 64 x 64 tiles; block (J, c) takes column
// tile J and the row tiles I in [J + c chunk, min(J + (c + 1) chunk, nt)), so
// each wave streams its 16 columns (16 g .. 16 g + 15 of the tile) downwards,
// lane r taking row r of each tile (one load = 64 consecutive doubles):
//   y_I += A_IJ x_J: summed over the 4 waves' columns in LDS, written to
//     P[I][J] (64 values; the diagonal tile, symmetric from its lower half,
//     goes there too);
//   y_J += A_IJ^T x_I (I > J): 16 column sums per lane accumulated over the
//     block's tiles, reduced across the lanes once at the end (halving
//     exchanges) and written to Y1[J][c].
// k_symv_reduce: y_i = sum_{J <= R} P[R][J][r] + sum_c Y1[R][c][r] (i = 64 R + r)
// in a fixed order (deterministic).
constexpr int SV_T = 64;

__device__ __forceinline__ double reduce16_lanes(double (&v)[16], int lane) {
  // after the exchanges lane L holds the 64-lane sum of v[(L >> 2) & 15]
#pragma unroll
  for (int h = 8, m = 32; h >= 1; h >>= 1, m >>= 1) {
    const bool up = (lane & m) != 0;
#pragma unroll
    for (int k = 0; k < h; ++k) {
      const double keep = up ? v[k + h] : v[k];
      const double send = up ? v[k] : v[k + h];
      v[k] = keep + shfl_xor_d(send, m);
    }
  }
  double t = v[0];
  t += shfl_xor_d(t, 2);
  t += shfl_xor_d(t, 1);
  return t;
}

// The work of symv block (J, c) (see above); vx.val(i) gives x_i and vx.first()
// x_0, so that a caller can form x on the fly (the fused column step scales the
// reflector this way; its x_0 is the unit entry). Rows of off-diagonal tiles are
// never 0.
template <class XF>
__device__ __forceinline__ void symv_tile(int64_t N, const double* __restrict__ A, int64_t lda,
                                          const XF& vx, int chunk, int nchmax, int64_t J, int c,
                                          double* __restrict__ P, double* __restrict__ Y1,
                                          double (*red)[4][64]) {
  const int64_t nt = (N + SV_T - 1) / SV_T;
  const int64_t I0 = J + (int64_t)c * chunk;
  if (I0 >= nt) return;
  const int64_t I1 = (I0 + chunk < nt) ? I0 + chunk : nt;
  const int lane = threadIdx.x & 63;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int64_t col0 = J * SV_T + 16 * g;
  double xj[16], t[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    xj[k] = (col0 + k < N) ? vx.val(col0 + k) : 0.0;
    t[k] = 0.0;
  }
  if (col0 == 0) xj[0] = vx.first();
  // the diagonal tile (first tile of chunk 0): symmetric from its lower half
  if (I0 == J) {
    const int64_t row = J * SV_T + lane;
    double yr = 0.0;
#pragma unroll 4
    for (int k = 0; k < 16; ++k) {
      const int64_t col = col0 + k;
      if (row < N && col < N)
        yr += ((lane >= 16 * g + k) ? A[row + col * lda] : A[col + row * lda]) * xj[k];
    }
    red[0][g][lane] = yr;
    __syncthreads();
    if (g == 0)
      P[(J * nt + J) * SV_T + lane] = (red[0][0][lane] + red[0][1][lane]) +
                                      (red[0][2][lane] + red[0][3][lane]);
    __syncthreads();
  }
  // off-diagonal tiles, two per step: 32 loads in flight per lane, one barrier per pair
  // (the LDS sums alternate between two halves of red, so a step's writes cannot meet
  // the previous step's reads). A is read once per call: non-temporal loads keep it from
  // displacing the partials in L2 (symv alone at N = 20,000: 4.66 -> 5.03 TB/s)
  int par = 0;
  for (int64_t I = (I0 == J ? I0 + 1 : I0); I < I1; I += 2) {
    const bool two = I + 1 < I1;
    const int64_t row0 = I * SV_T + lane, row1 = row0 + SV_T;
    const bool rok0 = row0 < N, rok1 = two && row1 < N;
    double a0[16], a1[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {  // every column of these tiles is < N (I > J)
      a0[k] = rok0 ? __builtin_nontemporal_load(&A[row0 + (col0 + k) * lda]) : 0.0;
      a1[k] = rok1 ? __builtin_nontemporal_load(&A[row1 + (col0 + k) * lda]) : 0.0;
    }
    const double xi0 = rok0 ? vx.val(row0) : 0.0, xi1 = rok1 ? vx.val(row1) : 0.0;
    double yr0 = 0.0, yr1 = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      yr0 += a0[k] * xj[k];
      yr1 += a1[k] * xj[k];
      t[k] += a0[k] * xi0 + a1[k] * xi1;
    }
    double (*r0)[64] = red[2 * par], (*r1)[64] = red[2 * par + 1];
    r0[g][lane] = yr0;
    r1[g][lane] = yr1;
    __syncthreads();
    if (g == 0) {
      P[(I * nt + J) * SV_T + lane] = (r0[0][lane] + r0[1][lane]) + (r0[2][lane] + r0[3][lane]);
      if (two)
        P[((I + 1) * nt + J) * SV_T + lane] = (r1[0][lane] + r1[1][lane]) +
                                               (r1[2][lane] + r1[3][lane]);
    }
    par ^= 1;
  }
  const double tot = reduce16_lanes(t, lane);
  if ((lane & 3) == 0) Y1[(J * nchmax + c) * SV_T + 16 * g + ((lane >> 2) & 15)] = tot;
}

struct PlainVec {
  const double* x;
  __device__ double val(int64_t i) const { return x[i]; }
  __device__ double first() const { return x[0]; }
};

__global__ void __launch_bounds__(256) k_symv_lower(int64_t N, const double* __restrict__ A,
                                                    int64_t lda, const double* __restrict__ x,
                                                    int chunk, int nchmax,
                                                    double* __restrict__ P,
                                                    double* __restrict__ Y1) {
  __shared__ double red[4][4][64];  // [2 parities x 2 tiles][wave][lane]
  symv_tile(N, A, lda, PlainVec{x}, chunk, nchmax, (int64_t)blockIdx.x, (int)blockIdx.y, P, Y1,
            red);
}

// One block per row tile R (64 rows, one per lane); wave w sums P[R][J] for
// J = w, w + 16, ... <= R; the 16 wave sums and the Y1 chunks of column tile R
// are then added in a fixed order.
__global__ void __launch_bounds__(1024) k_symv_reduce(int64_t N, int64_t nt, int chunk,
                                                      int nchmax,
                                                      const double* __restrict__ P,
                                                      const double* __restrict__ Y1,
                                                      double* __restrict__ y) {
  __shared__ double part[16][64];
  const int64_t R = blockIdx.x;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const double* pr = P + R * nt * SV_T + lane;
  double s = 0.0;
  int64_t J = w;
  for (; J + 48 <= R; J += 64) {  // four independent loads in flight
    const double a0 = pr[J * SV_T], a1 = pr[(J + 16) * SV_T];
    const double a2 = pr[(J + 32) * SV_T], a3 = pr[(J + 48) * SV_T];
    s += (a0 + a1) + (a2 + a3);
  }
  for (; J <= R; J += 16) s += pr[J * SV_T];
  part[w][lane] = s;
  __syncthreads();
  const int64_t i = R * SV_T + lane;
  if (w == 0 && i < N) {
    double tt = 0.0;
    const int nch = (int)((nt - R + chunk - 1) / chunk);
    for (int c = 0; c < nch; ++c) tt += Y1[(R * nchmax + c) * SV_T + lane];
#pragma unroll
    for (int q = 0; q < 16; ++q) tt += part[q][lane];
    y[i] = tt;
  }
}

// ---------------------------------------------------------------------------
// Fused column step of the blocked reduction: three kernels per column instead of
// seven (k_panel_update, k_larfg_unit, k_symv_lower, k_symv_reduce, k_panel_dots,
// k_panel_apply, k_latrd_finish), the same LAPACK dlatrd / dlarfg arithmetic with
// the column-wide reductions split into fixed-order block partials:
//   k_col_a (column i, rows i..): dlatrd's finish of column i-1 (w <- tau w +
//     alpha v, alpha = -tau/2 (tau w).v from k_col_c's partials), then the update of
//     column i with the panel's earlier reflectors (the c = i-1 terms use the just
//     finished w), then per block max|x| and sum (x / max)^2 of the reflector's x;
//   k_col_b: every block combines those pairs into ||x|| and forms dlarfg's beta,
//     tau and f = 1/(alpha - beta) itself; the symv tiles and the blocks of the panel
//     dot products work on u = (alpha - beta, x) = v / f (the column as stored, its
//     first entry replaced: a scale on the fly would hold 34 more registers in the
//     symv), giving A22 u, W^T u, V^T u; block 0 stores tau, e = beta and f;
//   k_col_c (one 64-row tile per block): y = f A22 u from the symv partials, the dots
//     times f, w = y - V (W^T v) - W (V^T v) (stored before dlatrd's finish),
//     v = (1, f x) written into A's column, and the block's partial of (tau w).v for
//     the next k_col_a.
// Panel-relative coordinates: A_loc(r, c) = Ap[r + c lda], W_loc(r, c) = Wp[r + c ldw].
// ---------------------------------------------------------------------------
constexpr int CA_T = 64;  // rows per k_col_a block

// workspace of the fused step (doubles): symv partials, dot partials, norm pairs,
// finish partials, scalars (f, the unscaled w at the pivot row)
struct ColLayout {
  int64_t sym, dot, nrm, fin, sc, total;
};
__host__ __device__ inline ColLayout col_layout(int64_t N) {
  const int64_t nt = (N + SV_T - 1) / SV_T;
  ColLayout L;
  L.sym = 0;
  L.dot = nt * nt * SV_T + nt * (nt / 4 + 1) * SV_T;
  L.nrm = L.dot + ((N + PD_ROWS - 1) / PD_ROWS + 1) * 2 * PANEL_MAX;
  L.fin = L.nrm + 2 * ((N + CA_T - 1) / CA_T + 1);
  L.sc = L.fin + (N + SV_T - 1) / SV_T + 1;
  L.total = L.sc + 8;
  return L;
}

// block-wide sum / max of one value per thread, every thread gets the result
// (fixed order: a wave butterfly, then the waves in order)
template <int NT>
__device__ __forceinline__ double block_sum_all(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) t += red[q];
  return t;
}
template <int NT>
__device__ __forceinline__ double block_max_all(double v, double* red) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, shfl_xor_d(v, o));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < NT / 64; ++q) t = fmax(t, red[q]);
  return t;
}

// dlatrd's finish of one element: w <- tau w, then w + alpha v
__device__ __forceinline__ double finish_elem(double tau, double w, double alpha, double v) {
  const double x = tau * w;
  return x + alpha * v;
}

// 64 rows per block (lane = row), the panel columns c split over the 4 waves
// (c = wave, wave + 4, ...), their sums added in wave order.
__global__ void __launch_bounds__(256) k_col_a(int64_t len, int i, int do_norm,
                                               double* __restrict__ Acol,
                                               const double* __restrict__ a0, int64_t lda,
                                               double* __restrict__ w0, int64_t ldw,
                                               const double* __restrict__ tau_prev,
                                               const double* __restrict__ fin_part, int n_fin,
                                               const double* __restrict__ sc,
                                               double* __restrict__ nrm_part) {
  __shared__ double xs[2][PANEL_MAX];
  __shared__ double red[4];
  __shared__ double part[4][64];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t r = (int64_t)blockIdx.x * CA_T + lane;
  double tp = 0.0, alpha = 0.0;
  if (i > 0) {
    tp = *tau_prev;
    double s = 0.0;
    for (int q = threadIdx.x; q < n_fin; q += 256) s += fin_part[q];
    alpha = -0.5 * tp * block_sum_all<256>(s, red);
    if (threadIdx.x < (unsigned)i) {
      const int c = threadIdx.x;
      // W_loc(i, c), A_loc(i, c); column i-1 of W is finished here (its pivot row's
      // value before the finish comes from sc[1]: row 0 is rewritten by this kernel)
      xs[0][c] = (c == i - 1) ? finish_elem(tp, sc[1], alpha, 1.0) : w0[(int64_t)c * ldw];
      xs[1][c] = a0[(int64_t)c * lda];
    }
  }
  __syncthreads();
  double s = 0.0;
  if (i > 0 && r < len) {
    double s0 = 0.0, s1 = 0.0;
    for (int c = wv; c < i; c += 4) {
      double wc = w0[r + (int64_t)c * ldw];
      if (c == i - 1) {
        wc = finish_elem(tp, wc, alpha, a0[r + (int64_t)c * lda]);
        w0[r + (int64_t)c * ldw] = wc;
      }
      s0 += a0[r + (int64_t)c * lda] * xs[0][c];
      s1 += wc * xs[1][c];
    }
    s = s0 + s1;
  }
  part[wv][lane] = s;
  __syncthreads();
  if (wv != 0) return;
  double y = 0.0;
  if (r < len) {
    y = Acol[r];
    if (i > 0) {
      y -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
      Acol[r] = y;
    }
  }
  if (do_norm) {  // the reflector's x: rows 2 .. len-1 of the updated column
    const double ax = (r >= 2 && r < len) ? fabs(y) : 0.0;
    double mb = ax;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mb = fmax(mb, shfl_xor_d(mb, o));
    double q = 0.0;
    if (mb > 0.0) {
      const double t = ax / mb;
      q = t * t;
    }
    q = wave_sum(q);
    if (lane == 0) {
      nrm_part[2 * blockIdx.x] = mb;
      nrm_part[2 * blockIdx.x + 1] = q;
    }
  }
}

struct ReflVec {  // u = (u0, x): x = the column below the pivot, as stored
  const double* x;
  double u0;
  __device__ double val(int64_t i) const { return x[i]; }  // i >= 1
  __device__ double first() const { return u0; }
};

// dlarfg from the block partials (every block): beta, tau, f = 1/(alpha - beta) and
// u0 = alpha - beta (f = 1, u0 = 1 when x = 0: H = I)
__device__ __forceinline__ void larfg_from_parts(const double* __restrict__ nrm_part, int n_nrm,
                                                 double alpha, double* red, double& f,
                                                 double& u0, double& tau, double& e) {
  double m = 0.0;
  for (int q = threadIdx.x; q < n_nrm; q += blockDim.x) m = fmax(m, nrm_part[2 * q]);
  const double scale = block_max_all<256>(m, red);
  double ss = 0.0;
  if (scale > 0.0)
    for (int q = threadIdx.x; q < n_nrm; q += blockDim.x) {
      const double rb = nrm_part[2 * q] / scale;
      ss += nrm_part[2 * q + 1] * (rb * rb);
    }
  ss = block_sum_all<256>(ss, red);
  const double xnorm = scale * sqrt(ss);
  if (xnorm == 0.0) {  // H = I
    f = 1.0;
    u0 = 1.0;
    tau = 0.0;
    e = alpha;
    return;
  }
  const double beta = -copysign(hypot(alpha, xnorm), alpha);
  u0 = alpha - beta;
  f = 1.0 / u0;
  tau = (beta - alpha) / beta;
  e = beta;
}

// Blocks [0, nsym): symv tile (J = b % nt, chunk c = b / nt) of y = A22 v; blocks
// [nsym, nsym + nblk 2i): partial dot products of the panels' columns with v (as
// k_panel_dots).
__global__ void __launch_bounds__(256) k_col_b(
    int64_t L, int i, const double* __restrict__ A22,
                                               int64_t lda, const double* __restrict__ x,
                                               const double* __restrict__ nrm_part, int n_nrm,
                                               int64_t nt, int chunk, int nchmax, int64_t nsym,
                                               int nblk, const double* __restrict__ ap,
                                               const double* __restrict__ wp, int64_t ldw,
                                               double* __restrict__ P, double* __restrict__ Y1,
                                               double* __restrict__ dpart,
                                               double* __restrict__ tau, double* __restrict__ e,
                                               double* __restrict__ sc) {
  __shared__ double red[4][4][64];
  const int64_t b = blockIdx.x;
  if (b < nsym && b % nt + (b / nt) * chunk >= nt) return;  // a symv block past the last tile
  double f, u0, tv, ev;
  larfg_from_parts(nrm_part, n_nrm, x[0], &red[0][0][0], f, u0, tv, ev);
  __syncthreads();  // red is reused below
  if (b == 0 && threadIdx.x == 0) {
    *tau = tv;
    *e = ev;
    sc[0] = f;
  }
  const ReflVec vx{x, u0};
  if (b < nsym) {
    symv_tile(L, A22, lda, vx, chunk, nchmax, b % nt, (int)(b / nt), P, Y1, red);
    return;
  }
  const int64_t d = b - nsym;
  const int q = (int)(d / nblk);
  const int64_t bx = d % nblk;
  const double* col = q < i ? wp + (int64_t)q * ldw : ap + (int64_t)(q - i) * lda;
  const int64_t r0 = bx * PD_ROWS;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < PD_ROWS / 256; ++k) {
    const int64_t r = r0 + threadIdx.x + 256 * k;
    if (r < L) s += col[r] * (r == 0 ? vx.first() : vx.val(r));
  }
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[0][0][threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    dpart[bx * 2 * PANEL_MAX + q] = (red[0][0][0] + red[0][0][1]) + (red[0][0][2] + red[0][0][3]);
}

// One block (16 waves) per 64-row tile R of the column step (L rows).
__global__ void __launch_bounds__(1024) k_col_c(int64_t L, int i, int64_t nt, int chunk,
                                                int nchmax, int nblk,
                                                const double* __restrict__ P,
                                                const double* __restrict__ Y1,
                                                const double* __restrict__ dpart,
                                                const double* __restrict__ ap, int64_t lda,
                                                const double* __restrict__ wp, int64_t ldw,
                                                double* __restrict__ x,
                                                double* __restrict__ w,
                                                double* __restrict__ gram_col,
                                                const double* __restrict__ tau,
                                                double* __restrict__ sc,
                                                double* __restrict__ fin_part) {
  __shared__ double part[16][64];
  __shared__ double ts[2 * PANEL_MAX];
  const int64_t R = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // t1 = W^T v (c < i), t2 = V^T v (i <= c < 2i): the dot partials in block order
  if (threadIdx.x < (unsigned)(2 * i)) {
    double s = 0.0;
    for (int bb = 0; bb < nblk; ++bb) s += dpart[(int64_t)bb * 2 * PANEL_MAX + threadIdx.x];
    s *= sc[0];  // the dots were taken with u = v / f
    ts[threadIdx.x] = s;
    if (R == 0 && threadIdx.x >= (unsigned)i) gram_col[threadIdx.x - i] = s;
  }
  // y: the symv partials of row tile R (as k_symv_reduce)
  const double* pr = P + R * nt * SV_T + lane;
  double s = 0.0;
  int64_t J = wv;
  for (; J + 112 <= R; J += 128) {  // eight independent loads in flight
    double a[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = pr[(J + 16 * k) * SV_T];
    s += ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  }
  for (; J <= R; J += 16) s += pr[J * SV_T];
  __syncthreads();  // ts complete
  // the panel corrections of row 64 R + lane, columns c = wv, wv + 16, ...
  const double f = sc[0];
  const int64_t r = R * SV_T + lane;
  double cs = 0.0;
  if (r < L)
    for (int c = wv; c < i; c += 16)
      cs += ap[r + (int64_t)c * lda] * ts[c] + wp[r + (int64_t)c * ldw] * ts[i + c];
  part[wv][lane] = f * s - cs;
  __syncthreads();
  if (wv == 0) {
    double yu = 0.0;
    const int nch = (int)((nt - R + chunk - 1) / chunk);
    for (int c = 0; c < nch; ++c) yu += Y1[(R * nchmax + c) * SV_T + lane];
    double y = f * yu;
#pragma unroll
    for (int q = 0; q < 16; ++q) y += part[q][lane];
    double pf = 0.0;
    if (r < L) {
      const double v = (r == 0) ? 1.0 : x[r] * f;
      w[r] = y;
      x[r] = v;
      pf = ((*tau) * y) * v;
      if (r == 0) sc[1] = y;
    }
    pf = wave_sum(pf);
    if (lane == 0) fin_part[R] = pf;
  }
}

// ---------------------------------------------------------------------------
// Rank-one merge: eigen-pairs of D + rho w w^T (D = diag(dl) ascending, rho > 0,
// no w_i = 0 and no two equal dl after deflation).
// ---------------------------------------------------------------------------
// Root j lies in (dl[j], dl[j+1]) (j < K-1) or in (dl[K-1], dl[K-1] + rho |w|^2].
// It is stored as origin pole o_j (dorg[j] = dl[o_j]) plus offset tau_j, with
// the pole nearer to the root as origin, so that every difference
// dl[i] - lambda_j = (dl[i] - dl[o_j]) - tau_j keeps its relative accuracy.
// The secular function g(lambda) = 1/rho + sum_i w_i^2 / (dl[i] - lambda) is
// increasing between poles. With lambda = dl[o] + s t (s = +1 from the left
// pole, -1 from the right one), h(t) = s g(dl[o] + s t) increases from -inf at
// t = 0+; bisection runs over the bit patterns of t in (0, T], which orders
// positive doubles, so it finds t to the last bit in at most 64 evaluations
// whatever its magnitude.
__device__ __forceinline__ double secular_h(const double* __restrict__ dl,
                                            const double* __restrict__ w2, int64_t K, double inv_rho,
                                            double dorg, double s, double t) {
  double g = inv_rho;
  for (int64_t i = 0; i < K; ++i) g += w2[i] / ((dl[i] - dorg) - s * t);
  return s * g;
}

__global__ void __launch_bounds__(256) k_secular_roots(int64_t K, const double* __restrict__ dl,
                                                       const double* __restrict__ w2, double rho,
                                                       double w2sum, double* __restrict__ dorg,
                                                       double* __restrict__ tau,
                                                       double* __restrict__ lam) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const double inv_rho = 1.0 / rho;
  double o, s, T;
  if (j < K - 1) {
    const double gap = dl[j + 1] - dl[j];
    const double mid = 0.5 * gap;
    if (secular_h(dl, w2, K, inv_rho, dl[j], 1.0, mid) >= 0.0) {
      o = dl[j];
      s = 1.0;
      T = mid;
    } else {
      o = dl[j + 1];
      s = -1.0;
      T = gap - mid;
    }
  } else {
    o = dl[j];
    s = 1.0;
    T = rho * w2sum;
  }
  uint64_t lo = 0, hi = (uint64_t)__double_as_longlong(T);
  while (hi - lo > 1) {
    const uint64_t md = lo + (hi - lo) / 2;
    if (secular_h(dl, w2, K, inv_rho, o, s, __longlong_as_double((long long)md)) >= 0.0)
      hi = md;
    else
      lo = md;
  }
  const double t = __longlong_as_double((long long)hi);
  dorg[j] = o;
  tau[j] = s * t;
  lam[j] = o + s * t;
}

// Gu-Eisenstat: zhat_i^2 = (lambda_i - dl_i)/rho * prod_{j != i} (lambda_j - dl_i)/(dl_j - dl_i),
// the weights for which the computed roots are exact; sign from w_i. The product
// is renormalised every 32 factors (frexp) so that it cannot over- or underflow.
__global__ void __launch_bounds__(256) k_secular_zhat(int64_t K, const double* __restrict__ dl,
                                                      const double* __restrict__ w, double rho,
                                                      const double* __restrict__ dorg,
                                                      const double* __restrict__ tau,
                                                      double* __restrict__ zhat) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const double di = dl[i];
  double p = 1.0;
  int e2 = 0;
  for (int64_t j = 0; j < K; ++j) {
    const double num = (dorg[j] - di) + tau[j];  // lambda_j - dl_i
    p *= (j == i) ? num : num / (dl[j] - di);
    if ((j & 31) == 31) {
      int e;
      p = frexp(p, &e);
      e2 += e;
    }
  }
  p = fabs(p) / rho;
  int e;
  p = frexp(p, &e);
  e2 += e;
  if (e2 & 1) {
    p *= 2.0;
    e2 -= 1;
  }
  const double z = ldexp(sqrt(p), e2 / 2);
  zhat[i] = w[i] < 0.0 ? -z : z;
}

// inv_norm[j] = 1 / || (zhat_i / (dl_i - lambda_j))_i ||
__global__ void __launch_bounds__(256) k_secular_norms(int64_t K, const double* __restrict__ dl,
                                                       const double* __restrict__ zhat,
                                                       const double* __restrict__ dorg,
                                                       const double* __restrict__ tau,
                                                       double* __restrict__ inv_norm) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const double o = dorg[j], t = tau[j];
  double s = 0.0;
  for (int64_t i = 0; i < K; ++i) {
    const double v = zhat[i] / ((dl[i] - o) - t);
    s += v * v;
  }
  inv_norm[j] = 1.0 / sqrt(s);
}

// M[row[i] * ldm + col[j]] = zhat_i / (dl_i - lambda_j) * inv_norm[j]   (i, j < K)
__global__ void __launch_bounds__(256) k_secular_fill(int64_t K, const double* __restrict__ dl,
                                                      const double* __restrict__ zhat,
                                                      const double* __restrict__ dorg,
                                                      const double* __restrict__ tau,
                                                      const double* __restrict__ inv_norm,
                                                      const int64_t* __restrict__ row,
                                                      const int64_t* __restrict__ col,
                                                      double* __restrict__ M, int64_t ldm,
                                                      int64_t njb) {
  const int64_t i = (int64_t)blockIdx.x / njb;
  const int64_t j = ((int64_t)blockIdx.x % njb) * blockDim.x + threadIdx.x;
  if (i >= K || j >= K) return;
  const double v = zhat[i] / ((dl[i] - dorg[j]) - tau[j]) * inv_norm[j];
  M[row[i] * ldm + col[j]] = v;
}

// M[row[t] * ldm + col[t]] = 1   (deflated eigen-pairs keep their basis vector)
__global__ void k_scatter_ones(int64_t cnt, const int64_t* __restrict__ row,
                               const int64_t* __restrict__ col, double* __restrict__ M,
                               int64_t ldm) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < cnt) M[row[t] * ldm + col[t]] = 1.0;
}

// Left-multiply M (row-major, n_cols columns) by G_0 G_1 ... G_{r-1}: applied
// last to first, G_t mixes rows p_t and q_t:
//   M[p, :] <- c M[p, :] - s M[q, :],   M[q, :] <- s M[p, :] + c M[q, :].
// One thread per column; the rotation sequence is read in the same order by
// every lane.
__global__ void __launch_bounds__(256) k_apply_rotations(int64_t r, const int64_t* __restrict__ pq,
                                                         const double* __restrict__ cs,
                                                         int64_t n_cols, double* __restrict__ M,
                                                         int64_t ldm) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= n_cols) return;
  for (int64_t t = r - 1; t >= 0; --t) {
    const int64_t p = pq[2 * t], q = pq[2 * t + 1];
    const double c = cs[2 * t], s = cs[2 * t + 1];
    double* mp = M + p * ldm + col;
    double* mq = M + q * ldm + col;
    const double x = *mp, y = *mq;
    *mp = c * x - s * y;
    *mq = s * x + c * y;
  }
}

}  // namespace jm

using namespace jm;

static int launch_dgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc, int lo, int splits, int64_t kchunk,
                        int64_t cstride, hipStream_t st, const double* A2 = nullptr,
                        const double* B2 = nullptr, int64_t k1 = INT64_MAX) {
  const int64_t nti = (m + DG_TI - 1) / DG_TI, ntj = (n + DG_TJ - 1) / DG_TJ;
  JM_CHECK_ARG(nti * ntj < (int64_t)1 << 31, "grid too large");
  const dim3 grid((unsigned)(nti * ntj), (unsigned)splits), block(256);
  if (!ta && !tb)
    hipLaunchKernelGGL((k_dgemm<false, false>), grid, block, 0, st, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc, lo, nti, kchunk, cstride, A2, B2, k1);
  else if (!ta && tb)
    hipLaunchKernelGGL((k_dgemm<false, true>), grid, block, 0, st, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc, lo, nti, kchunk, cstride, A2, B2, k1);
  else if (ta && !tb)
    hipLaunchKernelGGL((k_dgemm<true, false>), grid, block, 0, st, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc, lo, nti, kchunk, cstride, A2, B2, k1);
  else
    hipLaunchKernelGGL((k_dgemm<true, true>), grid, block, 0, st, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc, lo, nti, kchunk, cstride, A2, B2, k1);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_dgemm(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc, int lower_only, void* stream) {
  JM_CHECK_ARG(m >= 0 && n >= 0 && k >= 0, "bad sizes");
  if (m == 0 || n == 0) return JM_OK;
  JM_CHECK_ARG(C && (k == 0 || (A && B)), "null pointer");
  JM_CHECK_ARG(lda >= (trans_a ? k : m) && lda >= 1, "lda too small");
  JM_CHECK_ARG(ldb >= (trans_b ? n : k) && ldb >= 1, "ldb too small");
  JM_CHECK_ARG(ldc >= m, "ldc < m");
  return launch_dgemm(trans_a, trans_b, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                      lower_only ? 1 : 0, 1, k, 0, (hipStream_t)stream);
}

extern "C" int jm_dgemm2(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k, double alpha,
                         const double* A1, const double* B1, const double* A2, const double* B2,
                         int64_t lda, int64_t ldb, double beta, double* C, int64_t ldc,
                         int lower_only, void* stream) {
  JM_CHECK_ARG(m >= 0 && n >= 0 && k >= 0, "bad sizes");
  if (m == 0 || n == 0) return JM_OK;
  JM_CHECK_ARG(C && (k == 0 || (A1 && B1 && A2 && B2)), "null pointer");
  JM_CHECK_ARG(lda >= (trans_a ? k : m) && lda >= 1, "lda too small");
  JM_CHECK_ARG(ldb >= (trans_b ? n : k) && ldb >= 1, "ldb too small");
  JM_CHECK_ARG(ldc >= m, "ldc < m");
  return launch_dgemm(trans_a, trans_b, m, n, 2 * k, alpha, A1, lda, B1, ldb, beta, C, ldc,
                      lower_only ? 1 : 0, 1, 2 * k, 0, (hipStream_t)stream, A2, B2, k);
}

extern "C" int jm_dgemm_splitk(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k,
                               double alpha, const double* A, int64_t lda, const double* B,
                               int64_t ldb, double beta, double* C, int64_t ldc, int splits,
                               double* work, int64_t work_doubles, void* stream) {
  JM_CHECK_ARG(m >= 0 && n >= 0 && k >= 0 && splits >= 1 && splits <= 64, "bad sizes");
  if (m == 0 || n == 0) return JM_OK;
  JM_CHECK_ARG(C && work && (k == 0 || (A && B)), "null pointer");
  JM_CHECK_ARG(lda >= (trans_a ? k : m) && lda >= 1, "lda too small");
  JM_CHECK_ARG(ldb >= (trans_b ? n : k) && ldb >= 1, "ldb too small");
  JM_CHECK_ARG(ldc >= m, "ldc < m");
  JM_CHECK_ARG(work_doubles >= (int64_t)splits * m * n, "workspace too small");
  // K chunks in whole LDS steps; splits past the end of K write zeros
  const int64_t kchunk = ((k + splits - 1) / splits + DG_BK - 1) / DG_BK * DG_BK;
  hipStream_t st = (hipStream_t)stream;
  const int rc = launch_dgemm(trans_a, trans_b, m, n, k, 1.0, A, lda, B, ldb, 0.0, work, m, 0,
                              splits, kchunk > 0 ? kchunk : DG_BK, m * n, st);
  if (rc != JM_OK) return rc;
  hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((m * n + 255) / 256)), dim3(256), 0, st, m, n,
                     splits, alpha, work, beta, C, ldc);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

static void symv_shape(int64_t N, int64_t& nt, int& chunk, int& nchmax) {
  nt = (N + SV_T - 1) / SV_T;
  const int64_t tiles = nt * (nt + 1) / 2;
  int64_t ch = tiles / 2048;  // aim at a few thousand blocks
  ch = ch < 4 ? 4 : (ch > 64 ? 64 : ch);
  chunk = (int)ch;
  nchmax = (int)((nt + chunk - 1) / chunk);
}

extern "C" int64_t jm_symv_work_doubles(int64_t N) {
  if (N < 1) return 0;
  const int64_t nt = (N + SV_T - 1) / SV_T;
  return nt * nt * SV_T + nt * (nt / 4 + 1) * SV_T;  // P + Y1 for any chunk >= 4
}

extern "C" int jm_symv_lower(int64_t N, const double* A, int64_t lda, const double* x, double* y,
                             double* work, int64_t work_doubles, void* stream) {
  JM_CHECK_ARG(N >= 0 && lda >= N, "bad sizes");
  if (N == 0) return JM_OK;
  JM_CHECK_ARG(A && x && y && work, "null pointer");
  JM_CHECK_ARG(work_doubles >= jm_symv_work_doubles(N), "workspace too small");
  int64_t nt;
  int chunk, nchmax;
  symv_shape(N, nt, chunk, nchmax);
  JM_CHECK_ARG(nt < 65536 * 32768LL, "N too large");
  double* P2 = work;
  double* Y1 = work + nt * nt * SV_T;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_symv_lower, dim3((unsigned)nt, (unsigned)nchmax), dim3(256), 0, st, N, A,
                     lda, x, chunk, nchmax, P2, Y1);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_symv_reduce, dim3((unsigned)nt), dim3(1024), 0, st, N, nt, chunk, nchmax,
                     P2, Y1, y);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_tridiag_larfg(int64_t len, double* alpha, double* x, double* tau, double* e,
                                void* stream) {
  JM_CHECK_ARG(len >= 1, "bad length");
  JM_CHECK_ARG(alpha && tau && e && (len == 1 || x), "null pointer");
  hipLaunchKernelGGL(k_larfg_unit, dim3(1), dim3(1024), 0, (hipStream_t)stream, len, alpha, x, tau,
                     e);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int64_t jm_tridiag_panel_work_doubles(int64_t N) {
  return ((N + PD_ROWS - 1) / PD_ROWS + 1) * 2 * PANEL_MAX;
}

extern "C" int jm_tridiag_panel_update(int64_t len, int i, const double* a0, int64_t lda,
                                       const double* w0, int64_t ldw, double* y, void* stream) {
  JM_CHECK_ARG(len >= 0 && i >= 0 && i < PANEL_MAX, "bad sizes");
  if (len == 0 || i == 0) return JM_OK;
  JM_CHECK_ARG(a0 && w0 && y, "null pointer");
  const unsigned grid = (unsigned)((len + 255) / 256 < 1024 ? (len + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_panel_update, dim3(grid), dim3(256), 0, (hipStream_t)stream, len, i, a0,
                     lda, w0, ldw, y);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_tridiag_panel_step(int64_t len, int i, const double* ap, int64_t lda,
                                     const double* wp, int64_t ldw, const double* v, double* w,
                                     double* gram_col, double* work, int64_t work_doubles,
                                     void* stream) {
  JM_CHECK_ARG(len >= 0 && i >= 0 && i < PANEL_MAX, "bad sizes");
  if (len == 0 || i == 0) return JM_OK;
  JM_CHECK_ARG(ap && wp && v && w && gram_col && work, "null pointer");
  const int64_t nblk = (len + PD_ROWS - 1) / PD_ROWS;
  JM_CHECK_ARG(work_doubles >= nblk * 2 * PANEL_MAX, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_panel_dots, dim3((unsigned)nblk, (unsigned)(2 * i)), dim3(256), 0, st, len,
                     i, ap, lda, wp, ldw, v, work);
  JM_CHECK_LAUNCH();
  const unsigned grid = (unsigned)((len + 255) / 256 < 1024 ? (len + 255) / 256 : 1024);
  hipLaunchKernelGGL(k_panel_apply, dim3(grid), dim3(256), 0, st, len, i, (int)nblk, ap, lda, wp,
                     ldw, work, gram_col, w);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_tridiag_finish(int64_t len, const double* tau, double* w, const double* v,
                                 void* stream) {
  JM_CHECK_ARG(len >= 0, "bad length");
  if (len == 0) return JM_OK;
  JM_CHECK_ARG(tau && w && v, "null pointer");
  hipLaunchKernelGGL(k_latrd_finish, dim3(1), dim3(1024), 0, (hipStream_t)stream, len, tau, w, v);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int64_t jm_tridiag_column_work_doubles(int64_t N) {
  return N < 1 ? 0 : col_layout(N).total;
}

extern "C" int jm_tridiag_column(int64_t Np, int i, double* Ap, int64_t lda, double* Wp,
                                 int64_t ldw, double* tau, double* e, double* work,
                                 int64_t work_doubles, void* stream) {
  JM_CHECK_ARG(Np >= 1 && i >= 0 && i < PANEL_MAX && i < Np, "bad sizes");
  JM_CHECK_ARG(Ap && Wp && tau && e && work, "null pointer");
  JM_CHECK_ARG(lda >= Np && ldw >= Np, "lda / ldw < Np");
  const ColLayout Ly = col_layout(Np);
  JM_CHECK_ARG(work_doubles >= Ly.total, "workspace too small");
  hipStream_t st = (hipStream_t)stream;
  double* sc = work + Ly.sc;
  const int64_t len = Np - i;      // column i from row i
  const bool refl = i < Np - 1;    // column i has a reflector
  const int na = (int)((len + CA_T - 1) / CA_T);
  const int n_fin = i > 0 ? (int)((len + SV_T - 1) / SV_T) : 0;  // column i-1's row tiles
  hipLaunchKernelGGL(k_col_a, dim3((unsigned)na), dim3(256), 0, st, len, i, refl ? 1 : 0,
                     Ap + i + (int64_t)i * lda, Ap + i, lda, Wp + i, ldw,
                     i > 0 ? tau + i - 1 : nullptr, work + Ly.fin, n_fin, sc, work + Ly.nrm);
  JM_CHECK_LAUNCH();
  if (!refl) return JM_OK;
  const int64_t L = Np - i - 1;
  int64_t nt;
  int chunk, nchmax;
  symv_shape(L, nt, chunk, nchmax);
  const int64_t ntN = (Np + SV_T - 1) / SV_T;
  double* P = work + Ly.sym;
  double* Y1 = P + ntN * ntN * SV_T;
  const int nblk = (int)((L + PD_ROWS - 1) / PD_ROWS);
  const int64_t nsym = nt * nchmax;
  const int64_t nb_total = nsym + (int64_t)nblk * 2 * i;
  JM_CHECK_ARG(nb_total < (1LL << 31), "grid too large");
  double* x = Ap + (i + 1) + (int64_t)i * lda;  // A_loc(i + 1, i): alpha, then x
  hipLaunchKernelGGL(k_col_b, dim3((unsigned)nb_total), dim3(256), 0, st, L, i,
                     Ap + (i + 1) + (int64_t)(i + 1) * lda, lda, x, work + Ly.nrm, na, nt, chunk,
                     nchmax, nsym, nblk, Ap + (i + 1), Wp + (i + 1), ldw, P, Y1, work + Ly.dot,
                     tau + i, e + i, sc);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_col_c, dim3((unsigned)nt), dim3(1024), 0, st, L, i, nt, chunk, nchmax,
                     nblk, P, Y1, work + Ly.dot, Ap + (i + 1), lda, Wp + (i + 1), ldw, x,
                     Wp + (i + 1) + (int64_t)i * ldw, Wp + (int64_t)i * ldw, tau + i, sc,
                     work + Ly.fin);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_secular_solve(int64_t K, const double* dl, const double* w, const double* w2,
                                double rho, double w2sum, double* dorg, double* tau, double* lam,
                                double* zhat, double* inv_norm, void* stream) {
  JM_CHECK_ARG(K >= 0 && rho > 0.0, "bad sizes or rho <= 0");
  if (K == 0) return JM_OK;
  JM_CHECK_ARG(dl && w && w2 && dorg && tau && lam && zhat && inv_norm, "null pointer");
  const dim3 grid((unsigned)((K + 255) / 256)), block(256);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_secular_roots, grid, block, 0, st, K, dl, w2, rho, w2sum, dorg, tau, lam);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_secular_zhat, grid, block, 0, st, K, dl, w, rho, dorg, tau, zhat);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_secular_norms, grid, block, 0, st, K, dl, zhat, dorg, tau, inv_norm);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_secular_vectors(int64_t K, const double* dl, const double* zhat,
                                  const double* dorg, const double* tau, const double* inv_norm,
                                  const int64_t* row, const int64_t* col, int64_t n_defl,
                                  const int64_t* row_defl, const int64_t* col_defl, double* M,
                                  int64_t ldm, void* stream) {
  JM_CHECK_ARG(K >= 0 && n_defl >= 0 && ldm >= K + n_defl, "bad sizes");
  JM_CHECK_ARG(M, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (K > 0) {
    JM_CHECK_ARG(dl && zhat && dorg && tau && inv_norm && row && col, "null pointer");
    const int64_t njb = (K + 255) / 256;
    JM_CHECK_ARG(K * njb < (int64_t)1 << 31, "K too large");
    hipLaunchKernelGGL(k_secular_fill, dim3((unsigned)(K * njb)), dim3(256), 0, st, K, dl, zhat,
                       dorg, tau, inv_norm, row, col, M, ldm, njb);
    JM_CHECK_LAUNCH();
  }
  if (n_defl > 0) {
    JM_CHECK_ARG(row_defl && col_defl, "null pointer");
    hipLaunchKernelGGL(k_scatter_ones, dim3((unsigned)((n_defl + 255) / 256)), dim3(256), 0, st,
                       n_defl, row_defl, col_defl, M, ldm);
    JM_CHECK_LAUNCH();
  }
  return JM_OK;
}

extern "C" int jm_apply_rotations(int64_t r, const int64_t* pq, const double* cs, int64_t n_cols,
                                  double* M, int64_t ldm, void* stream) {
  JM_CHECK_ARG(r >= 0 && n_cols >= 0 && ldm >= n_cols, "bad sizes");
  if (r == 0 || n_cols == 0) return JM_OK;
  JM_CHECK_ARG(pq && cs && M, "null pointer");
  hipLaunchKernelGGL(k_apply_rotations, dim3((unsigned)((n_cols + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, r, pq, cs, n_cols, M, ldm);
  JM_CHECK_LAUNCH();
  return JM_OK;
}
