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
// SNP reads coalesced rows while walking i. Clusters are cut into pieces of at
// most a fixed number of eigen-indices (jm_spectrum_plan); k_mom_seg writes
// one partial per piece and k_mom_fold adds the pieces of each cluster in
// order, so the work is balanced and the results are bitwise reproducible.

#include "jm_common.h"

#include <mutex>
#include <type_traits>

namespace jm {

constexpr int kMomMaxTerms = 16;  // MAX_TERMS_PER_CLUSTER on the host side
constexpr int kMomGroup = 3;      // partner rows accumulated per pass (48 accumulators)
constexpr int kMomTerms = kMomMaxTerms;
constexpr int kMomUnroll = 4;     // eigen-indices whose Z loads are in flight together
constexpr int kMomWaves = 4;      // row groups of one block, one wavefront each

// The LDS transpose areas are private to a wavefront: a wave-level barrier
// with an LDS fence orders its writes and reads.
__device__ __forceinline__ void wave_barrier_lds() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Wavefronts per k_mom_seg block: one per group of kMomGroup moment rows, at most
// kMomWaves (the waves then take every kMomWaves-th group). The waves of a block
// read the same Z columns at about the same time, so Z comes from HBM once.
static inline int mom_waves(int nvm) {
  const int groups = (nvm + 1 + kMomGroup - 1) / kMomGroup;
  return groups < kMomWaves ? groups : kMomWaves;
}
static inline size_t mom_lds(int nvm) {
  return (size_t)mom_waves(nvm) * kMomTerms * 65 * sizeof(double);
}

typedef const double __attribute__((address_space(4))) mom_cdbl;

// Partial moments of one cluster piece for 64 SNP columns, with the piece's
// term count M a compile-time constant so that the inner loop is straight-line
// code:
//   work[s][b][slot0 + m] = sum_{i in piece} Z[i][s] V[b][i] delta_i^m
// (row b == nv: Z[i][s]^2). The summation order is fixed (i ascending, one
// fused multiply-add per (i, b, m)), so every M gives bitwise the same sums
// as a loop over m would. Single problem: V and delta are the same for every
// lane and come through scalar loads, issued for kMomUnroll consecutive
// eigen-indices together with their Z loads.
template <typename ZT, bool MULTI, int M>
__device__ __forceinline__ void mom_piece(const ZT* __restrict__ Z, int64_t ldz, int64_t n_snps,
                                          int64_t sc, const double* V, int64_t ldv, int nv,
                                          int nvm, const double* __restrict__ delta, int i0,
                                          int len, int slot0, int n_slots, int g_first,
                                          int g_step, double* tr,
                                          double* __restrict__ work) {
  const mom_cdbl* dls = (const mom_cdbl*)delta;
  const mom_cdbl* Vs = (const mom_cdbl*)V;
  const int iend = i0 + len;
  const int lane = threadIdx.x & 63;
  for (int g0 = g_first; g0 <= nvm; g0 += g_step) {
    // rows of this group present in every column of the block (MULTI: nv
    // differs per column; missing rows are written as 0 and never read)
    int ng = nvm + 1 - g0;
    ng = ng > kMomGroup ? kMomGroup : ng;
    double acc[kMomGroup][M];
#pragma unroll
    for (int b = 0; b < kMomGroup; ++b)
#pragma unroll
      for (int m = 0; m < M; ++m) acc[b][m] = 0.0;
    // KIND (wave-uniform, fixed per group): 1 = all kMomGroup rows are V rows, 2 = V rows
    // then the Z^2 row last, 0 = anything else (short last group, MULTI's per-column nv).
    // 1 and 2 leave out the row selects (the same products as kind 0 where it has them)
    const int kind = MULTI || ng != kMomGroup ? 0
                     : g0 + kMomGroup <= nv   ? 1
                     : g0 + kMomGroup - 1 == nv ? 2
                                                : 0;
    auto run = [&](auto kind_c) {
      constexpr int KIND = decltype(kind_c)::value;
      // one eigen-index: prod_b = Z V_b (or Z^2), acc[b][m] += prod_b delta^m
      auto step = [&](double z, double dl, const double (&vb)[kMomGroup]) {
        double prod[kMomGroup];
#pragma unroll
        for (int b = 0; b < kMomGroup; ++b) {
          const int row = g0 + b;
          // the product is rounded on its own (__dmul_rn: never fused into the m = 0 sum),
          // so every kind adds the same bits
          if (KIND == 1)
            prod[b] = __dmul_rn(z, vb[b]);
          else if (KIND == 2)
            prod[b] = __dmul_rn(z, b < kMomGroup - 1 ? vb[b] : z);
          else
            prod[b] = (b < ng && row <= nv) ? __dmul_rn(z, row < nv ? vb[b] : z) : 0.0;
        }
        double pw = 1.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
#pragma unroll
          for (int b = 0; b < kMomGroup; ++b) acc[b][m] += prod[b] * pw;
          pw *= dl;
        }
      };
      auto load_v = [&](int row, int i) -> double {
        if (MULTI) return row < nv ? V[(int64_t)row * ldv + i] : 0.0;
        if (KIND == 1 || (KIND == 2 && row < g0 + kMomGroup - 1)) return Vs[(int64_t)row * ldv + i];
        if (KIND == 2) return 0.0;  // the Z^2 row reads no V
        return row < nv ? Vs[(int64_t)row * ldv + i] : 0.0;
      };
      // Z of the next kMomUnroll eigen-indices is loaded while the current ones
      // are accumulated
      int ib = i0;
      ZT zc[kMomUnroll];
      if (ib + kMomUnroll <= iend) {
#pragma unroll
        for (int u = 0; u < kMomUnroll; ++u) zc[u] = Z[(int64_t)(ib + u) * ldz + sc];
      }
      for (; ib + kMomUnroll <= iend; ib += kMomUnroll) {
        const bool more = ib + 2 * kMomUnroll <= iend;
        ZT zn[kMomUnroll];
        if (more) {
#pragma unroll
          for (int u = 0; u < kMomUnroll; ++u)
            zn[u] = Z[(int64_t)(ib + kMomUnroll + u) * ldz + sc];
        }
        double zu[kMomUnroll], dl[kMomUnroll], vv[kMomUnroll][kMomGroup];
#pragma unroll
        for (int u = 0; u < kMomUnroll; ++u) zu[u] = (double)zc[u];
#pragma unroll
        for (int u = 0; u < kMomUnroll; ++u) {
          dl[u] = MULTI ? delta[ib + u] : dls[ib + u];
#pragma unroll
          for (int b = 0; b < kMomGroup; ++b)
            vv[u][b] = (KIND != 0 || b < ng) ? load_v(g0 + b, ib + u) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kMomUnroll; ++u) step(zu[u], dl[u], vv[u]);
        if (more) {
#pragma unroll
          for (int u = 0; u < kMomUnroll; ++u) zc[u] = zn[u];
        }
      }
      for (; ib < iend; ++ib) {
        double vv[kMomGroup];
#pragma unroll
        for (int b = 0; b < kMomGroup; ++b) vv[b] = (KIND != 0 || b < ng) ? load_v(g0 + b, ib) : 0.0;
        step((double)Z[(int64_t)ib * ldz + sc], MULTI ? delta[ib] : dls[ib], vv);
      }
    };
    if (kind == 1)
      run(std::integral_constant<int, 1>{});
    else if (kind == 2)
      run(std::integral_constant<int, 2>{});
    else
      run(std::integral_constant<int, 0>{});
    // transpose through LDS, one row at a time, so that the M consecutive slots
    // of a (column, row) are written by consecutive lanes
#pragma unroll
    for (int b = 0; b < kMomGroup; ++b) {
      if (b < ng) {
#pragma unroll
        for (int m = 0; m < M; ++m) tr[m * 65 + lane] = acc[b][m];
        wave_barrier_lds();
        for (int e = lane; e < 64 * M; e += 64) {
          const int col = e / M, m = e % M;
          const int64_t sc2 = (int64_t)blockIdx.x * 64 + col;
          if (sc2 < n_snps)
            work[(sc2 * (nvm + 1) + g0 + b) * n_slots + slot0 + m] = tr[m * 65 + col];
        }
        wave_barrier_lds();
      }
    }
  }
}

// One block per (64 SNP columns, cluster piece): grid (ceil(n_snps / 64), n_seg).
// Pieces hold at most a fixed number of eigen-indices, so the blocks carry
// balanced work however the spectrum clusters; the piece's term count
// (1..kMomMaxTerms, fixed by the plan) selects the instantiation. Four waves per
// SIMD (128 VGPRs): the instantiations with many terms spill some accumulators, but
// the occupancy pays for it (n = 10k: 0.43 -> 0.41 ms per 8192-SNP tile; n = 50k:
// 2.47 -> 2.38 ms per 10,000-SNP tile; against three waves per SIMD at 167 VGPRs).
template <typename ZT, bool MULTI>
__global__ void __launch_bounds__(64 * kMomWaves) __attribute__((amdgpu_waves_per_eu(4))) k_mom_seg(const ZT* __restrict__ Z, int64_t ldz,
                                                int64_t n_snps, const double* __restrict__ V_one,
                                                int64_t ldv_one, int nv_one,
                                                const jm_mom_col* __restrict__ cols, int nv_max,
                                                jm_spectrum_plan pl, double* __restrict__ work) {
  const int64_t s = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  // wave-uniform (readfirstlane): the rows of the wave's group, and with them the V
  // addresses, stay uniform, so V comes through scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), n_waves = blockDim.x >> 6;
  const int64_t sc = s < n_snps ? s : 0;
  const int seg = blockIdx.y;
  const double* V = V_one;
  int64_t ldv = ldv_one;
  int nv = nv_one;
  if (MULTI) {
    const jm_mom_col d = cols[sc];
    V = d.V;
    ldv = d.ldv;
    nv = d.nv;
  }
  const int nvm = MULTI ? nv_max : nv_one;
  const int i0 = pl.seg_start[seg], len = pl.seg_len[seg];
  const int M = pl.cl_terms[pl.seg_cluster[seg]];
  const int slot0 = pl.seg_slot[seg];
  extern __shared__ double tr_all[];  // per wave: term x column, padded
  double* tr = tr_all + wave * kMomTerms * 65;
#define JM_MOM_CASE(k)                                                                       \
  case k:                                                                                    \
    mom_piece<ZT, MULTI, k>(Z, ldz, n_snps, sc, V, ldv, nv, nvm, pl.delta, i0, len, slot0,   \
                            pl.n_slots, wave * kMomGroup, n_waves * kMomGroup, tr, work);    \
    break;
  switch (M) {
    JM_MOM_CASE(1) JM_MOM_CASE(2) JM_MOM_CASE(3) JM_MOM_CASE(4)
    JM_MOM_CASE(5) JM_MOM_CASE(6) JM_MOM_CASE(7) JM_MOM_CASE(8)
    JM_MOM_CASE(9) JM_MOM_CASE(10) JM_MOM_CASE(11) JM_MOM_CASE(12)
    JM_MOM_CASE(13) JM_MOM_CASE(14) JM_MOM_CASE(15) JM_MOM_CASE(16)
    default: break;  // the plan caps terms per cluster at kMomMaxTerms
  }
#undef JM_MOM_CASE
}

// mom[s][b][t] = sum over the pieces p of the cluster c of term t (in order)
// of work[s][b][seg_slot[seg0_c + p] + m_t]. Thread per (column, row, term),
// term fastest: reads and writes of a (column, row) are contiguous.
template <bool MULTI>
__global__ void k_mom_fold(int64_t n_snps, int nv_one, const jm_mom_col* __restrict__ cols,
                           int nv_max, jm_spectrum_plan pl, const double* __restrict__ work,
                           double* __restrict__ mom_one, int64_t ldm_one) {
  const int nvm = MULTI ? nv_max : nv_one;
  const int T = pl.n_terms;
  const int64_t total = n_snps * (int64_t)(nvm + 1) * T;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int t = (int)(q % T);
    const int64_t sb = q / T;  // s * (nvm + 1) + b
    const int b = (int)(sb % (nvm + 1));
    const int64_t s = sb / (nvm + 1);
    double* dst;
    if (MULTI) {
      const jm_mom_col d = cols[s];
      if (b > d.nv) continue;
      dst = d.out + (int64_t)b * d.ldm;
    } else {
      dst = mom_one + sb * ldm_one;
    }
    const double* src = work + sb * pl.n_slots;
    const int c = pl.term_cluster[t], m = pl.term_m[t];
    const int sg0 = pl.cl_seg0[c], ns = pl.cl_nseg[c];
    double v = src[pl.seg_slot[sg0] + m];
    for (int p = 1; p < ns; ++p) v += src[pl.seg_slot[sg0 + p] + m];
    dst[t] = v;
  }
}

}  // namespace jm

using namespace jm;

static int check_plan(const jm_spectrum_plan* pl) {
  JM_CHECK_ARG(pl != nullptr, "null plan");
  JM_CHECK_ARG(pl->n >= 1 && pl->n_terms >= 1 && pl->n_clusters >= 1 && pl->n_seg >= 1 &&
                   pl->n_slots >= 1,
               "bad plan sizes");
  JM_CHECK_ARG(pl->cl_terms && pl->cl_toff && pl->cl_seg0 && pl->cl_nseg && pl->seg_start &&
                   pl->seg_len && pl->seg_cluster && pl->seg_slot && pl->delta &&
                   pl->term_cluster && pl->term_m,
               "missing plan array");
  return JM_OK;
}

extern "C" int64_t jm_moments_work_bytes(const jm_spectrum_plan* plan, int64_t n_snps, int nv) {
  if (!plan || n_snps < 1 || nv < 0) return 0;
  return n_snps * (int64_t)(nv + 1) * plan->n_slots * (int64_t)sizeof(double);
}

extern "C" int jm_spectrum_moments(const void* Z, int z_is_f32, int64_t ldz, int64_t n_snps,
                                   const double* V, int64_t ldv, int nv,
                                   const jm_spectrum_plan* plan, double* work, double* mom,
                                   int64_t ldm, void* stream) {
  int rc = check_plan(plan);
  if (rc) return rc;
  JM_CHECK_ARG(n_snps >= 0 && nv >= 0, "bad sizes");
  JM_CHECK_ARG(ldz >= n_snps, "ldz < n_snps");
  JM_CHECK_ARG(ldm >= plan->n_terms, "ldm < n_terms");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(Z && work && mom, "null pointer");
  JM_CHECK_ARG(nv == 0 || (V && ldv >= plan->n), "V missing or ldv < n");
  const jm_spectrum_plan pl = *plan;
  const dim3 grid((unsigned)((n_snps + 63) / 64), (unsigned)pl.n_seg);
  hipStream_t st = (hipStream_t)stream;
  if (z_is_f32)
    hipLaunchKernelGGL((k_mom_seg<float, false>), grid, dim3(64 * mom_waves(nv)), mom_lds(nv), st, (const float*)Z, ldz,
                       n_snps, V, ldv, nv, nullptr, 0, pl, work);
  else
    hipLaunchKernelGGL((k_mom_seg<double, false>), grid, dim3(64 * mom_waves(nv)), mom_lds(nv), st, (const double*)Z, ldz,
                       n_snps, V, ldv, nv, nullptr, 0, pl, work);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_mom_fold<false>), dim3(2048), dim3(256), 0, st, n_snps, nv, nullptr, 0, pl,
                     work, mom, ldm);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

extern "C" int jm_spectrum_moments_multi(const void* Z, int z_is_f32, int64_t ldz,
                                         int64_t n_snps, const jm_mom_col* cols, int nv_max,
                                         const jm_spectrum_plan* plan, double* work,
                                         void* stream) {
  int rc = check_plan(plan);
  if (rc) return rc;
  JM_CHECK_ARG(n_snps >= 0 && nv_max >= 0, "bad sizes");
  JM_CHECK_ARG(ldz >= n_snps, "ldz < n_snps");
  if (n_snps == 0) return JM_OK;
  JM_CHECK_ARG(Z && cols && work, "null pointer");
  const jm_spectrum_plan pl = *plan;
  const dim3 grid((unsigned)((n_snps + 63) / 64), (unsigned)pl.n_seg);
  hipStream_t st = (hipStream_t)stream;
  if (z_is_f32)
    hipLaunchKernelGGL((k_mom_seg<float, true>), grid, dim3(64 * mom_waves(nv_max)), mom_lds(nv_max), st, (const float*)Z, ldz,
                       n_snps, nullptr, 0, 0, cols, nv_max, pl, work);
  else
    hipLaunchKernelGGL((k_mom_seg<double, true>), grid, dim3(64 * mom_waves(nv_max)), mom_lds(nv_max), st, (const double*)Z, ldz,
                       n_snps, nullptr, 0, 0, cols, nv_max, pl, work);
  JM_CHECK_LAUNCH();
  hipLaunchKernelGGL((k_mom_fold<true>), dim3(2048), dim3(256), 0, st, n_snps, 0, cols, nv_max,
                     pl, work, nullptr, 0);
  JM_CHECK_LAUNCH();
  return JM_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the per-piece partials live in a fixed device array.
// ---------------------------------------------------------------------------
// The array is jm::scratch; every GPU the library runs on holds its own copy
// (JM_MOMENTS_SCRATCH_BYTES) until the process ends. Host code never uses the
// array's name as an address: its device address is looked up once per device
// with hipGetSymbolAddress (with that device current) and kept in g_scratch_ptr.
// k_mom_seg writes every (column, row, slot) partial that k_mom_fold then reads,
// so nothing has to start at zero; calls that overlap on one GPU share the array
// and race. The array has external linkage (inside namespace jm): with internal
// linkage the device code object keeps it as a local symbol under another name
// than the one the host registers, and the runtime's symbol lookup aborts.
namespace jm {
__device__ __align__(256) double scratch[JM_MOMENTS_SCRATCH_BYTES / sizeof(double)];
}  // namespace jm

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
  int dev = -1;
  JM_CHECK_HIP(hipGetDevice(&dev));
  JM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    JM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(jm::scratch)));
    JM_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned a null address");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return JM_OK;
}
}  // namespace

extern "C" int64_t jm_moments_scratch_bytes(void) { return JM_MOMENTS_SCRATCH_BYTES; }

extern "C" int jm_spectrum_moments_noalloc(const void* Z, int z_is_f32, int64_t ldz,
                                           int64_t n_snps, const double* V, int64_t ldv, int nv,
                                           const jm_spectrum_plan* plan, double* mom, int64_t ldm,
                                           void* stream) {
  int rc = check_plan(plan);
  if (rc) return rc;
  JM_CHECK_ARG(n_snps >= 0 && nv >= 0, "bad sizes");
  const int64_t need = jm_moments_work_bytes(plan, n_snps, nv);
  JM_CHECK_ARG(need <= (int64_t)JM_MOMENTS_SCRATCH_BYTES,
               "jm_spectrum_moments_noalloc: %lld SNP columns x %d rows x %d slots need %lld "
               "bytes of partials, more than the fixed %lld-byte scratch; pass fewer columns "
               "per call",
               (long long)n_snps, nv + 1, plan->n_slots, (long long)need,
               (long long)JM_MOMENTS_SCRATCH_BYTES);
  if (n_snps == 0) return JM_OK;
  double* work = nullptr;
  rc = scratch_for_current_device(&work);
  if (rc) return rc;
  return jm_spectrum_moments(Z, z_is_f32, ldz, n_snps, V, ldv, nv, plan, work, mom, ldm, stream);
}
