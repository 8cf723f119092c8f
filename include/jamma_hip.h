/*
 * jamma_hip.h -- C ABI of libjamma_hip.so, the MI355X (gfx950) implementation of
 * JAMMA's linear-mixed-model association hot path.
 *
 * Conventions
 *  - Every entry point returns int: JM_OK (0) or a negative JM_ERR_* code, and
 *    never throws. jm_last_error() returns the message of the calling thread's
 *    last failure.
 *  - All array arguments are DEVICE pointers to caller-owned memory unless the
 *    name says *_host. `stream` is a hipStream_t (NULL = default stream); every
 *    call is asynchronous on that stream.
 *  - Row-major storage; `ld*` arguments are leading dimensions in elements.
 *  - Genotypes are int8 A1-allele counts 0/1/2 with -1 for a missing call, in
 *    SNP-major rows ("g", one row per SNP) or sample-major rows ("gT").
 *    Padding elements must be 0.
 *
 * Reference paths below are relative to the michael-denyer/jamma repository
 * root (v2.0.1). The reference is Python/JAX; these entry points replace its
 * JAX/NumPy operators, which a Python host binds through ctypes
 * (jamma_amd/_lib.py).
 */
#ifndef JAMMA_HIP_H
#define JAMMA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JM_OK 0
#define JM_ERR_ARG (-1)
#define JM_ERR_HIP (-2)
#define JM_ERR_UNSUPPORTED (-3)

#define JM_VERSION 101

/* Library version (JM_VERSION). */
int jm_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated). */
int jm_last_error(char* buf, size_t len);

/* ------------------------------------------------------------------------ */
/* PLINK decode + pass-1 statistics                                          */
/* ------------------------------------------------------------------------ */

/*
 * Replaces bed-reader's windowed read in stream_genotype_chunks
 * (src/jamma/io/plink.py:239-300) and the pass-1 per-SNP statistics
 * (src/jamma/lmm/runner_streaming.py:248-284, kinship/compute.py:601-611).
 *
 * bed: SNP-major .bed payload (magic bytes stripped), bytes_per_snp =
 *      ceil(n_samples_total / 4). snp_idx (n_snps entries, may be NULL =
 *      0..n_snps-1) selects the SNP rows to decode; sample_idx (n_out entries,
 *      may be NULL = 0..n_out-1) selects and orders the samples.
 * g (may be NULL): n_snps x ld_g int8 output, 0/1/2 = A1 count, -1 missing,
 *      columns [n_out, ld_g) are written as 0.
 * counts (may be NULL): n_snps x 4 int32 = (#0, #1, #2, #missing) over the
 *      selected samples; mean, missing rate, var>0 and HWE follow from them.
 */
int jm_decode_bed(const uint8_t* bed, int64_t bytes_per_snp, const int32_t* snp_idx,
                  int64_t n_snps, const int32_t* sample_idx, int64_t n_out, int8_t* g,
                  int64_t ld_g, int32_t* counts, void* stream);

/* out (cols x ld_out) = transpose of in (rows x ld_in), int8, and/or out0
 * (cols x ld_out0), the same transpose with missing calls (-1) written as 0.
 * out holds the sample-major genotype tile the imputation terms read, out0 the
 * operand of the kinship GEMM (possibly a column range of a larger staging
 * buffer). Either output may be NULL; columns [rows, ld) are not written. */
int jm_transpose_i8(const int8_t* in, int64_t rows, int64_t cols, int64_t ld_in, int8_t* out,
                    int64_t ld_out, int8_t* out0, int64_t ld_out0, void* stream);

/*
 * Dosage (non-integer) genotypes, float64 (compute_centered_kinship and the
 * in-memory runner accept any float matrix with NaN for missing:
 * kinship/compute.py:93-226, runner_jax.py:52-481):
 *   jm_center_columns_f64: impute_and_center (kinship/missing.py:22-61) of a
 *     sample-major tile X (n x cols, row stride ld) in place: mu_j = nanmean of
 *     column j (0 if all NaN, written to mean_out[j] when not NULL), X <- X - mu_j,
 *     NaN -> 0. The kinship then accumulates Xc Xc^T with jm_dgemm.
 *   jm_impute_rows_f64: SNP-major rows X (rows x n, stride ld), NaN -> mean[r],
 *     before the FP64 rotation (jm_gemm_nt_f64) of the association.
 */
int jm_center_columns_f64(double* X, int64_t n, int64_t cols, int64_t ld, double* mean_out,
                          void* stream);
int jm_impute_rows_f64(double* X, int64_t rows, int64_t n, int64_t ld, const double* mean,
                       void* stream);

/*
 * Replaces _accumulate_kinship (src/jamma/kinship/compute.py:46-57) inside the
 * batch loops of compute_centered_kinship (:205-214) and
 * compute_kinship_streaming (:663-681), together with impute_and_center
 * (src/jamma/kinship/missing.py:22-61), which this library applies as exact
 * corrections (see jm_kinship_finalize).
 *
 * S (n x n int32) += X0 X0^T over n_snps SNPs. gT0 is X0^T, the sample-major
 * genotypes with missing calls as 0 (jm_transpose_i8's out0): round_up(n, 256)
 * rows of ld_gT bytes, ld_gT a multiple of 128, columns [n_snps, round_up(n_snps,
 * 128)) zero. Rows n .. round_up(n, 256) - 1 are read but never reach S (any
 * value). Exact int32 accumulation on int8 MFMA (v_mfma_i32_16x16x64_i8); one
 * call over many SNPs is faster than several over fewer. Only the lower
 * triangle (i >= k) of S is guaranteed; jm_kinship_finalize mirrors it.
 */
int jm_kinship_syrk_i8(const int8_t* gT0, int64_t ld_gT, int64_t n, int64_t n_snps, int32_t* S,
                       void* stream);

/*
 * The same S += X0 X0^T (replacing the same reference code) from a 2-bit packed
 * operand, a quarter of the int8 operand's bytes, S bit-identical to
 * jm_kinship_syrk_i8. The codes are widened in registers to e2m1 (FP4) operands of
 * v_mfma_scale_f32_16x16x128_f8f6f4 (twice the int8 MFMA rate): integer products
 * summed in FP32, exact because a launch covers at most 2^22 SNPs (sums < 2^24;
 * longer calls are split into such launches). JM_KINSHIP_SYRK_I8MFMA=1 widens to
 * int8 for v_mfma_i32_16x16x64_i8 instead (same S).
 *   jm_pack_gram_2bit: SNP-major int8 genotypes g (n_snps rows of ld_g, -1 =
 *     missing, as jm_decode_bed writes them) -> SNP columns col0 .. col0 + n_snps
 *     of gP (col0 a multiple of 16): sample-major rows of ldp bytes (ldp a multiple
 *     of 64, 4 SNPs per byte), round_up(n, 256) rows of which rows >= n must be zero
 *     (never written). The codes from col0 + n_snps up to the next multiple of 256
 *     are written as 0; require round_up(col0 + n_snps, 256) <= 4 * ldp. Within a
 *     256-SNP group the 16-SNP words are interleaved for the GEMM's fragment reads.
 *   jm_kinship_syrk_2bit: S (n x n, lower triangle) += X0 X0^T over SNP columns
 *     [0, n_snps) of gP; every 256-SNP group touched must be fully written
 *     (columns of the gaps between packed tiles read as zero genotypes).
 */
int jm_pack_gram_2bit(const int8_t* g, int64_t ld_g, int64_t n_snps, int64_t n, uint8_t* gP,
                      int64_t ldp, int64_t col0, void* stream);
int jm_kinship_syrk_2bit(const uint8_t* gP, int64_t ldp, int64_t n, int64_t n_snps, int32_t* S,
                         void* stream);

/*
 * Per-tile mean-imputation / centring terms, accumulated over SNP tiles
 * (the algebra of impute_and_center, kinship/missing.py:48-61):
 *   mu[j]    = float64 nanmean of SNP j (host computes from jm_decode_bed counts)
 *   u[i]    += sum_j (g_ij >= 0 ? g_ij mu_j : mu_j^2)
 *   C[k][i] += sum_{j: g_kj missing} mu_j * (g_ij >= 0 ? g_ij : mu_j / 2)
 * g is the SNP-major tile (n_snps x ld_g, rows 16-byte aligned), gT the
 * sample-major one (n x ld_gT, rows 16-byte aligned and zero-padded to a multiple
 * of 16). With has_missing, n_missing is the tile's number of missing calls (from
 * the decode counts) and miss_ws holds jm_kinship_miss_ws_len(n, n_snps,
 * n_missing) int32 of 8-byte aligned device scratch: the per-sample lists of
 * missing SNPs and the tile's genotypes as 2-bit codes are built there first.
 */
int64_t jm_kinship_miss_ws_len(int64_t n, int64_t n_snps, int64_t n_missing);
int jm_kinship_impute_terms(const int8_t* g, int64_t ld_g, const int8_t* gT, int64_t ld_gT,
                            int64_t n, int64_t n_snps, const double* mu, double* u, double* C,
                            int has_missing, int32_t* miss_ws, int64_t miss_ws_len,
                            int64_t n_missing, void* stream);

/*
 * K[i][k] = (S_ik - u_i - u_k + s0 + C[i][k] + C[k][i]) / p_snps
 * i.e. Xc Xc^T / p with Xc = impute_and_center(X) (kinship/compute.py:217,684).
 * S is read from the lower triangle written by jm_kinship_syrk_i8. S_sub / u_sub /
 * C_sub (may be NULL) are subtracted first: the LOCO subtraction
 * K_c = (S_full - S_c) / (p - p_c) of kinship/compute.py:898-916.
 */
int jm_kinship_finalize(const int32_t* S, const double* u, const double* C, double s0,
                        const int32_t* S_sub, const double* u_sub, const double* C_sub,
                        double s0_sub, int64_t n, double p_snps, double* K, void* stream);

/* ------------------------------------------------------------------------ */
/* Rotation into the kinship eigenbasis                                      */
/* ------------------------------------------------------------------------ */

/*
 * Replaces UtG = UT @ geno_chunk (src/jamma/lmm/runner_streaming.py:392-410;
 * also runner_jax.py:236-254 and loco.py:406-423).
 * Z (n x ldz, row i = eigen-index, column s = SNP) = U^T X, where column s of X
 * is SNP row r(s) of g with missing calls replaced by mean[r(s)] (float32-derived
 * mean, runner_streaming.py:322-329,429-434); r(s) = col_idx[s], or s when
 * col_idx is NULL (the gather lets one launch serve several requests).
 * UT = U^T, n x ld_u float64. FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 output.
 */
int jm_rotate_f64(const double* UT, int64_t ld_u, const int8_t* g, int64_t ld_g,
                  const double* mean, const int32_t* col_idx, int64_t n, int64_t n_snps,
                  double* Z, int64_t ldz, void* stream);

/* Generic C (m x ldc) = A (m x k, lda) * B^T where B is (ncols x k, ldb), all
 * float64 (UtW = UT @ W and Uty = UT @ y, runner_streaming.py:349-352). */
int jm_gemm_nt_f64(const double* A, int64_t lda, const double* B, int64_t ldb, int64_t m,
                   int64_t ncols, int64_t k, double* C, int64_t ldc, void* stream);

/*
 * Mixed-precision rotation (the same Z = U^T X as jm_rotate_f64) on bf16 MFMA
 * (v_mfma_f32_16x16x32_bf16) with FP32 accumulation. U^T is split into nplanes
 * bfloat16 planes, each the bf16 rounding of the residual left by the planes
 * before it; X1 = genotypes with missing -> bf16(mean), X2 = (mean - bf16(mean))
 * at missing calls, 0 elsewhere:
 *   nplanes = 2: UT = hi + lo (16 significant bits),
 *                Z = lo^T X1 + hi^T X2 + hi^T X1            (3 bf16 products)
 *   nplanes = 3: UT = hi + mid + lo (24 significant bits),
 *                Z = lo^T X1 + mid^T X1 + hi^T X2 + hi^T X1 (4 bf16 products)
 * Both operands are pre-blocked so that every copy into LDS is one contiguous
 * kilobyte:
 *   jm_pack_ut_bf16x3: UT (float64, n x ld_u) -> planes
 *     (jm_packed_ut_bytes(n, nplanes) bytes, 16-byte aligned), once per cohort;
 *   jm_pack_genotypes: SNP rows r(s) of g (col_idx as in jm_rotate_f64; rows
 *     8-byte aligned: g and ld_g multiples of 8) and
 *     their means -> gp (jm_packed_geno_bytes(n, n_snps) bytes) and mm
 *     (n_snps x uint32: bf16(mean) | bf16 residual << 16), per SNP tile;
 *   jm_rotate_bf16x3: Z (float32, n x ldz) from planes, gp, mm.
 */
int64_t jm_packed_ut_bytes(int64_t n, int nplanes);
int64_t jm_packed_geno_bytes(int64_t n, int64_t n_snps);
int jm_pack_ut_bf16x3(const double* UT, int64_t ld_u, int64_t n, int nplanes, uint16_t* planes,
                      void* stream);
int jm_pack_genotypes(const int8_t* g, int64_t ld_g, const int32_t* col_idx, const double* mean,
                      int64_t n, int64_t n_snps, uint8_t* gp, uint32_t* mm, void* stream);
int jm_rotate_bf16x3(const uint16_t* planes, int nplanes, const uint8_t* gp, const uint32_t* mm,
                     int64_t n, int64_t n_snps, float* Z, int64_t ldz, void* stream);

/* ------------------------------------------------------------------------ */
/* Per-SNP REML/MLE lambda search + Wald/LRT/Score                           */
/* ------------------------------------------------------------------------ */

/* Device description of a compressed spectrum (host struct of device pointers).
 * Cluster c (sorted eigen-indices, centre d_c) owns the terms toff_c ..
 * toff_c + M_c - 1 and is cut into pieces (segments) of at most a fixed number
 * of eigen-indices; piece q covers [seg_start, seg_start + seg_len) of cluster
 * seg_cluster and its partial moments occupy slots seg_slot .. seg_slot + M - 1
 * of the workspace. */
typedef struct jm_spectrum_plan {
  int32_t n;          /* eigenvalues */
  int32_t n_terms;    /* T */
  int32_t n_clusters; /* C */
  int32_t n_seg;      /* pieces */
  int32_t n_slots;    /* sum over pieces of M of their cluster */
  int32_t reserved;
  const int32_t* cl_terms;    /* C : M_c */
  const int32_t* cl_toff;     /* C : first term of c */
  const int32_t* cl_seg0;     /* C : first piece of c */
  const int32_t* cl_nseg;     /* C : number of pieces of c */
  const int32_t* seg_start;   /* n_seg */
  const int32_t* seg_len;     /* n_seg */
  const int32_t* seg_cluster; /* n_seg */
  const int32_t* seg_slot;    /* n_seg */
  const double* delta;        /* n : d_i - d_c(i) */
  const int32_t* term_cluster; /* T */
  const int32_t* term_m;       /* T */
} jm_spectrum_plan;

/* Workspace (bytes) of jm_spectrum_moments for n_snps columns and nv partner
 * rows (jm_spectrum_moments_multi: nv = nv_max). */
int64_t jm_moments_work_bytes(const jm_spectrum_plan* plan, int64_t n_snps, int nv);

/*
 * Spectrum moments. For each column s of Z (n x ldz; float64 when z_is_f32 ==
 * 0, else float32) and each term t = (cluster c, power m) of the plan:
 *   mom[s][b][t] = sum_{i in c} Z[i][s] * V[b][i] * delta_i^m   (b < nv)
 *   mom[s][nv][t] = sum_{i in c} Z[i][s]^2 * delta_i^m
 * V (nv x ldv float64) are the partner vectors (rotated covariates, then the
 * rotated phenotype); mom is (n_snps x (nv+1) x ldm); work holds
 * jm_moments_work_bytes(plan, n_snps, nv) bytes of per-piece partials. This is
 * the lambda-independent part of batch_compute_uab + calc_pab row 0
 * (likelihood_jax.py:112-186, 342-393).
 */
int jm_spectrum_moments(const void* Z, int z_is_f32, int64_t ldz, int64_t n_snps,
                        const double* V, int64_t ldv, int nv, const jm_spectrum_plan* plan,
                        double* work, double* mom, int64_t ldm, void* stream);

/*
 * Allocation-free form of jm_spectrum_moments (no counterpart in the reference:
 * for callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments and results
 * (bitwise) without `work`: the per-piece partials go to a static device array of
 * JM_MOMENTS_SCRATCH_BYTES that each GPU holds for the life of the process (its
 * address is looked up once per device at the first call). Calls needing more
 * than that (jm_moments_work_bytes > JM_MOMENTS_SCRATCH_BYTES) are rejected with
 * JM_ERR_ARG: pass fewer columns per call. The entry point allocates nothing;
 * calls that overlap on one GPU share the array and race.
 */
#define JM_MOMENTS_SCRATCH_BYTES (64ll << 20)
int64_t jm_moments_scratch_bytes(void);
int jm_spectrum_moments_noalloc(const void* Z, int z_is_f32, int64_t ldz, int64_t n_snps,
                                const double* V, int64_t ldv, int nv,
                                const jm_spectrum_plan* plan, double* mom, int64_t ldm,
                                void* stream);

/* Per-column descriptor of jm_spectrum_moments_multi: the partner rows
 * (nv x ldv float64) and the output block ((nv+1) x ldm float64) of one column. */
typedef struct jm_mom_col {
  const double* V;
  int64_t ldv;
  int32_t nv;
  int32_t reserved;
  double* out;
  int64_t ldm;
} jm_mom_col;

/* jm_spectrum_moments with a per-column descriptor (device array cols[n_snps]),
 * so columns of different requests (phenotypes / covariate sets) on the same
 * cohort share one launch. nv_max = max nv over the columns. */
int jm_spectrum_moments_multi(const void* Z, int z_is_f32, int64_t ldz, int64_t n_snps,
                              const jm_mom_col* cols, int nv_max, const jm_spectrum_plan* plan,
                              double* work, void* stream);

/* Parameters of one (cohort, phenotype) association problem. Device pointers. */
typedef struct jm_lmm_problem {
  int32_t n;           /* analysed samples */
  int32_t n_cvt;       /* covariate columns incl. intercept (c) */
  int32_t n_terms;     /* T */
  int32_t n_clusters;  /* C */
  int32_t n_grid;      /* lambda grid points (reference default 50) */
  int32_t n_iter;      /* golden-section iterations = max(n_refine, 20) */
  double l_min, l_max;
  const int32_t* term_cluster; /* T */
  const int32_t* term_m;       /* T */
  const double* cl_center;     /* C */
  const double* logdet_mom;    /* T : sum_{i in c} delta_i^m */
  const double* shared_mom;    /* n_index x T : moments of the lambda-independent
                                  (covariate/phenotype) pairs, packed GEMMA order;
                                  rows of genotype pairs are ignored */
  const int32_t* pair_row;     /* n_index : >= 0 -> row of the per-SNP moments
                                  (pair involves the genotype), -1 -> shared row */
  const double* grid_loglam;   /* n_grid : numpy.linspace(log l_min, log l_max) */
  /* tables filled by jm_lmm_prepare */
  double* grid_kappa;   /* T x n_grid */
  double* grid_shared;  /* n_grid x n_index */
  double* grid_logdet;  /* n_grid */
  double* null_kappa;   /* T */
  double* null_shared;  /* n_index */
  double lambda_null;   /* null-model MLE lambda (Score test) */
  double logl_null;     /* null-model MLE log-likelihood (LRT) */
  const double* cl_weight;     /* C : cluster sizes (weight of log|1+lambda d_c| in
                                  logdet_h; equals logdet_mom of the m = 0 term) */
  double* iab_shared;   /* n_index : shared pair sums at lambda = 0 (Iab), filled by
                           jm_lmm_prepare */
} jm_lmm_problem;

/* Output columns of jm_lmm_assoc (n_snps x JM_NOUT float64). */
#define JM_OUT_BETA 0
#define JM_OUT_SE 1
#define JM_OUT_P_WALD 2
#define JM_OUT_L_REMLE 3
#define JM_OUT_LOGL_H1 4
#define JM_OUT_L_MLE 5
#define JM_OUT_LOGL_MLE 6
#define JM_OUT_P_LRT 7
#define JM_OUT_P_SCORE 8
#define JM_OUT_BETA_SCORE 9
#define JM_OUT_SE_SCORE 10
#define JM_OUT_EVALS 11
#define JM_NOUT 12

/*
 * Null model MLE (compute_null_model_mle, src/jamma/lmm/likelihood.py:571-607,
 * scalar golden section :395-474 over mle_log_likelihood_null :349-392) on one
 * wavefront. Writes out_host-free device doubles out[0] = lambda0,
 * out[1] = logl_H0, out[2] = null REML lambda, out[3] = null REML logl.
 */
int jm_lmm_null(const jm_lmm_problem* prob, double* out, void* stream);

/* Fills the grid / null tables of prob (needs prob->lambda_null). */
int jm_lmm_prepare(const jm_lmm_problem* prob, void* stream);

/*
 * Per-SNP association for n_snps SNPs from their spectrum moments
 * (xmom: n_snps x (n_cvt+2) x ldm, jm_spectrum_moments with V = [UtW | Uty]).
 * Replaces, per chunk, batch_compute_iab, golden_section_optimize_lambda
 * (+_mle), batch_calc_wald_stats, batch_calc_score_stats and
 * calc_lrt_pvalue_jax (likelihood_jax.py:231-883; call sites
 * runner_streaming.py:471-554). mode: 1 Wald, 2 LRT, 3 Score, 4 all.
 * 16 lanes per SNP when the problem has at most 256 spectrum terms and its
 * tables fit 64 KB of LDS, else one wavefront per SNP (the environment
 * variable JM_LMM_WAVE_PER_SNP=1 forces the latter); the reference's 50-point
 * log grid + golden-section schedule is kept exactly.
 */
int jm_lmm_assoc(const jm_lmm_problem* prob, int mode, const double* xmom, int64_t ldm,
                 int64_t n_snps, double* out, void* stream);

/*
 * Batched form of jm_lmm_assoc for the serving loop: SNP s is tested under
 * problem probs[snp_prob[s]] (device array of n_probs problems, all on one
 * cohort so they share n_terms) with mode modes[snp_prob[s]], from the moments
 * block starting at snp_xmom[s] (row stride ldm); results go to row s of out.
 * max_cvt = the widest problem's n_cvt (sizes the LDS layouts); n_clusters = the
 * cohort plan's cluster count. width_cvt (0: max_cvt; else in [max_cvt, 20]) picks
 * the kernel variant, and with it the summation order: launches with one
 * width_cvt give a SNP the same result bit for bit whatever else is in the batch
 * (the serving runtime passes its configured max_cvt). With n_terms <= 256 the
 * 16-lanes-per-SNP kernel runs when its LDS fits at width_cvt (the shared pair
 * rows from global memory when they do not fit in LDS: same arithmetic), each
 * block taking 16 consecutive SNPs and the rows of their problems (fastest when a
 * problem's SNPs are contiguous).
 */
int jm_lmm_assoc_multi(const jm_lmm_problem* probs, const int32_t* modes, int n_probs, int max_cvt,
                       int width_cvt, int n_terms, int n_clusters, const int32_t* snp_prob,
                       const double* const* snp_xmom, int64_t ldm, int64_t n_snps, double* out,
                       void* stream);

/* ------------------------------------------------------------------------ */
/* Native serving runtime. No reference counterpart: the reference runs one    */
/* association per call (run_lmm_association_streaming,                        */
/* lmm/runner_streaming.py:94-628); this is the continuous-batching loop that  */
/* keeps cohorts resident and serves requests as they arrive.                  */
/* ------------------------------------------------------------------------ */
#define JM_SERVE_F64 0   /* cohort precision: FP64 rotation (UT)            */
#define JM_SERVE_MIXED 1 /* BF16x3 rotation (packed planes)                 */
#define JM_SERVE_STUB 1  /* jm_serve_config.flags: no HIP call at all       */
#define JM_SERVE_MANUAL 2 /* with STUB: no thread, jm_serve_run_once drives it */

typedef struct jm_serve_config {
  int64_t slice_snps;   /* SNPs a request contributes to one step */
  int64_t max_batch;    /* SNPs per step (>= slice_snps) */
  int32_t max_reqs;     /* requests per step */
  int32_t max_inflight; /* steps queued on a cohort's stream ahead of the device */
  int32_t max_skips;    /* aging: overtakes a waiting request tolerates */
  int32_t max_cvt;      /* widest problem (covariate columns incl. intercept) accepted */
  int32_t flags;        /* JM_SERVE_STUB */
  int32_t reserved;
} jm_serve_config;

typedef struct jm_serve_cohort_desc {
  int32_t n;            /* samples */
  int32_t precision;    /* JM_SERVE_F64 or JM_SERVE_MIXED */
  int64_t n_snps;       /* genotype rows */
  const int8_t* g;      /* device, n_snps x ld_g int8 SNP-major, -1 missing, 8-byte rows */
  int64_t ld_g;
  const double* means;  /* device, n_snps imputation means */
  const double* UT;     /* device U^T (n x ld_u), FP64 cohorts */
  int64_t ld_u;
  const uint16_t* planes; /* device jm_pack_ut_bf16x3 output, mixed cohorts */
  int32_t nplanes;
  int32_t n_terms;      /* of the cohort's spectrum plan */
  int32_t n_clusters;
  int32_t max_batch;    /* SNPs per step of this cohort, slice_snps .. the server's
                           max_batch (0: the server's max_batch); a smaller cap keeps a
                           large cohort's steps short, so other work queued behind them
                           on the device waits less */
} jm_serve_cohort_desc;

typedef struct jm_serve_done {
  int64_t id;
  int32_t status;       /* JM_OK, or the error code of the step that failed it */
  int32_t reserved;
  char message[240];    /* the error message, empty on success */
} jm_serve_done;

typedef struct jm_serve_stats_t {
  int64_t queued, active, steps, snps, arena_used, arena_bytes;
  int64_t max_active;   /* most requests admitted at once so far */
  int64_t reserved;     /* arena bytes of the cohorts' step workspaces */
} jm_serve_stats_t;

typedef struct jm_serve jm_serve;

/* Arena bytes of a cohort's step workspace (carved from the arena by
 * jm_serve_add_cohort) and of one admitted request (its moments slice). */
int64_t jm_serve_cohort_ws_bytes(const jm_serve_config* cfg, const jm_serve_cohort_desc* c,
                                 const jm_spectrum_plan* plan);
int64_t jm_serve_request_ws_bytes(const jm_serve_config* cfg, int n_terms, int nv);

/* Starts the scheduler thread over a caller-owned device arena (256-byte aligned;
 * every device buffer the runtime uses comes from it, so serving never calls
 * hipMalloc). With JM_SERVE_STUB the arena may be NULL: no HIP call is made and
 * steps complete as they are formed (admission and batching logic without a GPU). */
int jm_serve_create(const jm_serve_config* cfg, void* arena, int64_t arena_bytes, jm_serve** out);
/* Registers a resident cohort (its arrays stay caller-owned and must outlive the
 * runtime): carves its step workspace, creates its stream and pinned slots. */
int jm_serve_add_cohort(jm_serve* s, const jm_serve_cohort_desc* c, const jm_spectrum_plan* plan,
                        int* cohort_id);
/* Queues a request: a prepared problem (jm_lmm_prepare and jm_lmm_null done; its
 * arrays must live until the request completes) with its partner rows V
 * ((n_cvt + 1) x ldv), a test mode 1..4 and SNP rows [snp_start, snp_end) of the
 * cohort; results go to out ((snp_end - snp_start) x JM_NOUT float64, device memory
 * or pinned host memory from hipHostMalloc: each step copies its rows there on the
 * cohort's stream), complete once the request's completion record has been polled. */
int jm_serve_submit(jm_serve* s, int cohort, const jm_lmm_problem* prob, const double* V,
                    int64_t ldv, int mode, int64_t snp_start, int64_t snp_end, double* out,
                    int64_t* request_id);
/* Completion records (non-blocking): up to max_recs into recs, count in *n_recs. */
int jm_serve_poll(jm_serve* s, jm_serve_done* recs, int max_recs, int* n_recs);
/* Blocks up to timeout_ms for a completion record; returns how many are waiting. */
int jm_serve_wait(jm_serve* s, int timeout_ms);
int jm_serve_stats(jm_serve* s, jm_serve_stats_t* st);
/* paused != 0: no admission and no new step until resumed (steps already launched
 * complete and are reported); requests submitted meanwhile wait in the queue, so a
 * caller can hand over a set of requests that then start together. Destroy resumes. */
int jm_serve_pause(jm_serve* s, int paused);
/* The cohort's stream (NULL with JM_SERVE_STUB): a caller that prepares requests on
 * it adds no stream of its own, so its work never waits behind another cohort's
 * step in a shared hardware queue; it runs in order with the cohort's steps. */
int jm_serve_cohort_stream(jm_serve* s, int cohort, void** stream);
/* JM_SERVE_MANUAL only: one pass of the loop (admission, a step per cohort,
 * completions); returns launched steps * 65536 + completed requests. */
int64_t jm_serve_run_once(jm_serve* s);
/* Failure-handling hook: the cohort's next step fails on the host before any
 * launch (its active requests fail with JM_ERR_UNSUPPORTED, the loop goes on). */
int jm_serve_inject_step_error(jm_serve* s, int cohort);
/* Drains every queued and admitted request, then stops the thread and frees the
 * streams, events and pinned slots (the arena stays the caller's). */
int jm_serve_destroy(jm_serve* s);

/* ------------------------------------------------------------------------ */
/* Large-cohort eigendecomposition (n*n > 2^31 - 1). Together with rocBLAS'   */
/* 64-bit level-2 routines and rocSOLVER dstedc on the halves, these replace  */
/* np.linalg.eigh (lmm/eigen.py:94) where the library's 32-bit tridiagonal    */
/* solvers stop. All matrices column-major, all offsets 64-bit.              */
/* ------------------------------------------------------------------------ */

/* C (m x n, ldc) = alpha op(A) op(B) + beta C, op(X) = X or X^T (trans_* != 0);
 * FP64 MFMA. lower_only != 0 skips the tiles with every element above the
 * diagonal (i < j); elements above the diagonal in the tiles it does compute are
 * written too. */
int jm_dgemm(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc, int lower_only, void* stream);

/* y = A x for a symmetric A (N x N, lda) stored in its lower triangle, column-
 * major: every stored element is read once (64 x 64 tiles; each off-diagonal
 * tile serves A_IJ x_J and A_IJ^T x_I). work: jm_symv_work_doubles(N) doubles of
 * device scratch (partial sums, summed in a fixed order: deterministic). The
 * reduction's symv (dlatrd, dsymv), with 64-bit offsets. */
int64_t jm_symv_work_doubles(int64_t N);
int jm_symv_lower(int64_t N, const double* A, int64_t lda, const double* x, double* y,
                  double* work, int64_t work_doubles, void* stream);

/* Two products in one pass: C = alpha (op(A1) op(B1) + op(A2) op(B2)) + beta C, both
 * pairs of the same shapes and leading dimensions (the rank-2k trailing update
 * A22 -= V W^T + W V^T reads and writes C once). */
int jm_dgemm2(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k, double alpha,
              const double* A1, const double* B1, const double* A2, const double* B2, int64_t lda,
              int64_t ldb, double beta, double* C, int64_t ldc, int lower_only, void* stream);

/* jm_dgemm with K split over `splits` block rows (1..64) whose partial products
 * go to work (splits * m * n doubles, column-major m x n each) and are then summed
 * in a fixed order: C = alpha sum + beta C. For outputs too small to fill the
 * device with one block per tile (V^T Q in the back-transformation). */
int jm_dgemm_splitk(int trans_a, int trans_b, int64_t m, int64_t n, int64_t k, double alpha,
                    const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                    double* C, int64_t ldc, int splits, double* work, int64_t work_doubles,
                    void* stream);

/* Householder reflector of a dlatrd column step (LAPACK dlarfg on alpha and the
 * len - 1 elements of x, tau and x written as dlarfg does), then *e = beta and
 * *alpha = 1 (the reflector's unit entry). One block; deterministic. */
int jm_tridiag_larfg(int64_t len, double* alpha, double* x, double* tau, double* e, void* stream);

/* dlatrd's update of column i by the panel's earlier i reflectors:
 * y[r] -= sum_c a0[r + c lda] w0[c ldw] + w0[r + c ldw] a0[c lda] for r < len
 * (the vectors are the panels' first rows; y does not overlap them). */
int jm_tridiag_panel_update(int64_t len, int i, const double* a0, int64_t lda, const double* w0,
                            int64_t ldw, double* y, void* stream);

/* dlatrd's corrections after the symmetric product w = A22 v (len rows, i < 128
 * panel columns ap (lda) and wp (ldw)): t1 = wp^T v, t2 = ap^T v,
 * w -= ap t1 + wp t2, t2 stored to gram_col (i values: the panel's V^T v). work:
 * jm_tridiag_panel_work_doubles(len) doubles; block partial sums added in a
 * fixed order. */
int64_t jm_tridiag_panel_work_doubles(int64_t N);
int jm_tridiag_panel_step(int64_t len, int i, const double* ap, int64_t lda, const double* wp,
                          int64_t ldw, const double* v, double* w, double* gram_col, double* work,
                          int64_t work_doubles, void* stream);

/* End of a dlatrd column step on w, v (len elements): w *= tau,
 * alpha = -tau/2 (w . v), w += alpha v (tau read from device memory). */
int jm_tridiag_finish(int64_t len, const double* tau, double* w, const double* v, void* stream);

/* One fused column step i of the blocked reduction of a panel (LAPACK dlatrd,
 * lower, with dlarfg), replacing the seven launches of jm_tridiag_panel_update,
 * jm_tridiag_larfg, jm_symv_lower, jm_tridiag_panel_step and jm_tridiag_finish by
 * three: the finish of column i-1 and the update of column i, then the symv with
 * the panel dot products (the reflector formed on the fly), then w and the stored
 * reflector. Ap = A_loc(0, 0) and Wp = W_loc(0, 0) are the panel's top-left
 * corners (column-major, lda / ldw >= Np; W's rows are panel-relative), tau and e
 * the panel's entries. Columns 0 .. nbp-1 are stepped in order; after the last
 * one, if it has a reflector (nbp - 1 < Np - 1), jm_tridiag_finish completes its
 * w. work: jm_tridiag_column_work_doubles(Np) doubles, kept between the steps of
 * a panel (block partials in a fixed order: deterministic). Replaces the host
 * LAPACK dsyevd's reduction under the reference's eigh (lmm/eigen.py:28-128). */
int64_t jm_tridiag_column_work_doubles(int64_t N);
int jm_tridiag_column(int64_t Np, int i, double* Ap, int64_t lda, double* Wp, int64_t ldw,
                      double* tau, double* e, double* work, int64_t work_doubles, void* stream);

/* Rank-one update D + rho w w^T, D = diag(dl) strictly ascending, rho > 0, no w_i
 * = 0 (the deflated problem of a divide-and-conquer merge; K eigen-pairs). Root j
 * = dorg[j] + tau[j] (lam[j] holds the sum) with dorg[j] the pole dl[j] or
 * dl[j+1] nearer to it; zhat = Gu-Eisenstat weights (exact for the computed
 * roots); inv_norm[j] = 1 / || zhat / (dl - lambda_j) ||. w2 = w^2, w2sum = sum
 * w2. Device arrays of K doubles. */
int jm_secular_solve(int64_t K, const double* dl, const double* w, const double* w2, double rho,
                     double w2sum, double* dorg, double* tau, double* lam, double* zhat,
                     double* inv_norm, void* stream);

/* Writes the rank-one update's eigenvectors into the merge matrix M (row-major,
 * leading dimension ldm, zero elsewhere): M[row[i]][col[j]] = zhat_i /
 * (dl_i - lambda_j) * inv_norm[j] for i, j < K, and M[row_defl[t]][col_defl[t]]
 * = 1 for the n_defl deflated pairs. */
int jm_secular_vectors(int64_t K, const double* dl, const double* zhat, const double* dorg,
                       const double* tau, const double* inv_norm, const int64_t* row,
                       const int64_t* col, int64_t n_defl, const int64_t* row_defl,
                       const int64_t* col_defl, double* M, int64_t ldm, void* stream);

/* M (row-major, n_cols columns, ldm) <- G_0 G_1 ... G_{r-1} M, where G_t rotates
 * rows p = pq[2t], q = pq[2t+1] by (c, s) = (cs[2t], cs[2t+1]):
 * M[p] <- c M[p] - s M[q], M[q] <- s M[p] + c M[q]. */
int jm_apply_rotations(int64_t r, const int64_t* pq, const double* cs, int64_t n_cols, double* M,
                       int64_t ldm, void* stream);

/* ------------------------------------------------------------------------ */
/* Results files (host code, several threads into one caller buffer)         */
/* ------------------------------------------------------------------------ */

/*
 * GEMMA .assoc.txt rows, byte-identical to the reference's writers
 * (src/jamma/lmm/io.py:12-147 via IncrementalAssocWriter :172-262): row r is
 *   pre[pre_off[r] .. pre_off[r+1])  (the row's "chr\trs\tps\t")
 *   n_miss[r] "\t"
 *   mid[mid_off[r] .. mid_off[r+1])  ("allele1\tallele0\t")
 *   af[r] as %.3f, then n_stats values of stats row r (row stride ld_stats) as
 *   "\t%.6e" each, then "\n".
 * NaN prints "nan" (Python's spelling). Returns the bytes written, or minus the
 * bytes needed when out_cap is too small (nothing written), INT64_MIN on bad
 * arguments. n_threads host threads (<= 64) format row ranges in parallel.
 */
int64_t jm_format_assoc_rows(int64_t n_rows, const char* pre, const int64_t* pre_off,
                             const int64_t* n_miss, const char* mid, const int64_t* mid_off,
                             const double* af, const double* stats, int n_stats,
                             int64_t ld_stats, char* out, int64_t out_cap, int n_threads);

/* Rows of a matrix as tab-separated %.10g values, one line each: the GEMMA
 * .cXX.txt / .eigenU.txt (one row per line) and .eigenD.txt (n_cols = 1)
 * formats (kinship/io.py, lmm/eigen_io.py:98-137). Same return convention. */
int64_t jm_format_g10_rows(int64_t n_rows, int64_t n_cols, const double* a, int64_t lda, char* out,
                           int64_t out_cap, int n_threads);

/* Parses whitespace-separated numbers, one matrix row per non-blank line (the
 * GEMMA text files above), into out (row-major, *n_rows x *n_cols; the column
 * count is the first line's). Returns the number of values, or -2 when they exceed
 * out_cap, -3 for a ragged row, -4 for a token that is not a number, -1 on bad
 * arguments. Conversion by strtod (correctly rounded), rows split over threads. */
int64_t jm_parse_text_rows(const char* buf, int64_t len, double* out, int64_t out_cap,
                           int64_t* n_rows, int64_t* n_cols, int n_threads);

/* ------------------------------------------------------------------------ */
/* Host-callable special functions (the device code compiled for the CPU),   */
/* exported so the CPU test-suite can check them against SciPy.              */
/* ------------------------------------------------------------------------ */

/* p = I_x(a, b), regularised incomplete beta (jax.scipy.special.betainc). */
double jm_host_betainc(double a, double b, double x);
/* chi-square(1) survival function (jax.scipy.stats.chi2.sf(x, 1)). */
double jm_host_chi2_sf1(double x);
/* Wald / Score p-value from an F(1, df) statistic with the reference's guards
 * (likelihood_jax.py:668-673): 1 for f_stat <= 0, NaN for a NaN statistic. */
double jm_host_f_pvalue(double df, double f_stat);
/* LRT p-value, calc_lrt_pvalue_jax (likelihood_jax.py:231-251). */
double jm_host_lrt_pvalue(double logl_h1, double logl_h0);

#ifdef __cplusplus
}
#endif

#endif /* JAMMA_HIP_H */
