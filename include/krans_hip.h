/*
 * krans_hip.h - C ABI of the KraNS-MI null-space engine (libkrans_hip.so).
 *
 * The reference (Deep-SANE/krans) exposes no C ABI: its boundary is the C++
 * plugin API of include/krans/{MatrixInterface,VectorInterface,
 * VectorFactory}.hpp and the two solver entry points. This header is the
 * flat layer under the C++ adapters of include/krans/adapters/hip/, and what
 * a ctypes / FFI binding calls. Every function returns a status code
 * (KH_OK = 0); kh_last_error() gives the thread's last message. The codes map
 * to the reference's exception types: KH_EINVAL <-> std::invalid_argument,
 * KH_ERUNTIME <-> std::runtime_error, KH_ENOMEM <-> std::bad_alloc.
 *
 * Threading: one context per device; handles are not thread-safe, calls on
 * one context are serialised by the caller (the reference is single-threaded:
 * its solvers take const refs and never synchronise).
 */
#ifndef KRANS_HIP_H
#define KRANS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KH_OK 0
#define KH_EINVAL 1
#define KH_ERUNTIME 2
#define KH_ENOMEM 3
#define KH_EDEVICE 4

/* precision of a solve: FP64 everywhere, or FP32 Krylov/LOBPCG blocks with
 * FP64 inner products, projected problems and polish residuals */
#define KH_PREC_FP64 0
#define KH_PREC_MIXED 1

typedef struct kh_context_s kh_context;
typedef struct kh_matrix_s kh_matrix;
typedef struct kh_rng_s kh_rng;

/* Field for field krans::solvers::CyclicLanczosOptions
 * (include/krans/Solvers/CyclicLanczosSolver.hpp:84-146); the two non-POD
 * members are replaced by `y0` (initial_right_space[0]) in
 * kh_cyclic_lanczos and by `precond_jacobi` (M^{-1} = diag(A A^T)^{-1}).
 * `graphs` is an engine choice with no reference counterpart: 1 replays each
 * Golub-Kahan cycle from a HIP graph captured once per basis size when the
 * operator only enqueues kernels (CSR, shifted difference; not callbacks),
 * -1 launches every kernel directly, 0 (default) decides by problem size.
 * Results are bitwise the same either way. `screen` (mixed precision with
 * prescreen_factor < 0.1, also an engine choice) picks how candidates are
 * screened before their FP64 test: 0 (default) from the singular values of
 * the projected problem when use_harmonic is set and target_tau is zero,
 * else as 1; 1 from FP32 candidate vectors and residuals; -1 not at all.
 * `second_pass` (reorth = 2, engine choice) sets the second pass of each
 * Golub-Kahan step's Gram-Schmidt pair: 1 always (the reference's double
 * reorthogonalisation); 2 only where the first pass cancelled more than
 * half of ||w||^2 (the Daniel-Gragg-Kaufman-Stewart test with
 * eta = 1/sqrt(2), as in ARPACK, decided on the device); 0 (default) 2 for
 * mixed precision and 1 for FP64. Selective steps run one pass each; a
 * cycle in which a step's test asks for the second pass is queued again from
 * that step with both passes there, so each step ends with the passes its
 * own test asks for. */
typedef struct kh_cyclic_options {
  int64_t k, ncv, max_restarts;
  double tol_abs, tol_rel;
  int32_t reorth;       /* 0 none, 1 single, 2 double (ReorthMode) */
  int32_t use_harmonic; /* refined extraction when nonzero */
  double target_tau;
  int32_t nullity_known, rank_known;
  int64_t cg_maxit;
  double cg_rel_tol, cg_breakdown_tol;
  int32_t projector_passes, grad_max_iters, frob_samples;
  double prescreen_factor;
  int64_t stagnation_restarts;
  int32_t return_left_space, verbose;
  int32_t precond_jacobi, graphs;
  int32_t screen, second_pass;
  double second_pass_eta; /* eta of the test, 0 < eta <= 1 (0: 1/sqrt(2)) */
} kh_cyclic_options;

/* krans::solvers::CyclicLanczosResult (CyclicLanczosSolver.hpp:149-167) plus
 * operator-application counts and phase times. */
typedef struct kh_cyclic_result {
  int64_t n_right, n_left, restarts, krylov_steps_total;
  double frob_norm_est;
  int64_t converged_to_target, cg_solves, cg_maxit_hits, spmv, spmtv, cg_iters;
  double t_gk, t_tiny, t_cand, t_polish, t_fallback, t_rng;
  int64_t screened; /* mixed precision: candidates rejected by the FP32 screen */
  int64_t graph_captures, graph_replays; /* Golub-Kahan cycles as HIP graphs */
  /* mixed precision: locked vectors replaced by the final batched FP64
   * projector pass (those above max(1e-14, min(1e-10, 0.01 thr))) */
  int64_t refined;
  /* serving only: the number of requests in the pack whose shared launches
   * ran this request's Golub-Kahan cycles (0: solved on its own) */
  int64_t packed;
  /* mixed precision: projector CGs whose FP32 recurrence did not converge
   * (stalled, cg_maxit or breakdown) and were solved again in FP64; after
   * 4 batches in a row of which every column failed, the solve's later
   * polishes run in FP64 directly and count here too */
  int64_t escalated;
  /* Gram-Schmidt pairs of the Golub-Kahan steps and how many of them ran
   * their second pass (all of them unless the second pass is selective);
   * selective: the steps are queued with one pass, and a cycle is queued
   * again from a step whose test asks for the second pass (gk_requeued) */
  int64_t cgs_pairs, second_passes, gk_requeued;
} kh_cyclic_result;

/* krans::solvers::LOBPCGOptions (LOBPCGSolver.hpp:71-115); M_inv and
 * initial_block are replaced by `initial_block` in kh_lobpcg. */
typedef struct kh_lobpcg_options {
  int64_t nev, block_size, max_iters;
  double tol_abs, tol_rel;
  int32_t soft_locking, gs_reorth;
  double restart, gs_tol;
  int32_t frob_samples, nullity_known, rank_known, verbose;
  int64_t max_tiny_dim, progress_every;
} kh_lobpcg_options;

/* krans::solvers::LOBPCGResult (LOBPCGSolver.hpp:117-139). */
typedef struct kh_lobpcg_result {
  int64_t n_right, iters, locked;
  double frob_norm_est, accept_threshold, best_unaccepted_residual;
  int64_t block_size, converged_to_target, spmv, spmtv;
  double t_gram, t_sparse, t_host;
  /* FP32 blocks orthonormalised by Cholesky QR, and those sent on to MGS */
  int64_t cholqr_blocks, cholqr_fallbacks;
} kh_lobpcg_result;

const char *kh_last_error(void);
int kh_version(void);
int kh_device_count(int *n);

/* --- context ------------------------------------------------------------ */
int kh_context_create(int device, kh_context **out);
int kh_context_destroy(kh_context *ctx);
int kh_context_sync(kh_context *ctx);

/* --- matrices ------------------------------------------------------------ */
/* Canonical CSR on the host (rowptr[0]=0, columns strictly ascending per
 * row, int32 indices); the device copy gets a CSC twin built from it.
 * Replaces EigenSparseMatrixAdapter(const SparseMat&)
 * (include/krans/adapters/eigen/EigenSparseMatrixAdapter.hpp:20-24). */
int kh_matrix_create_csr(kh_context *ctx, int64_t m, int64_t n,
                         const int32_t *rowptr, const int32_t *col,
                         const double *val, kh_matrix **out);
/* Matrix-free operators (device analogues of the reference's
 * StdMatrixAdapter with y = A x / y = A^T x callbacks,
 * include/krans/adapters/universal/StdMatrixAdapter.hpp:24-87):
 *   kh_matrix_create_shifted_difference: (A x)_i = x_{i+N} - c x_i, A is
 *     (n-N) x n (examples/quickstart.cpp:25-37), as HIP kernels with no
 *     matrix in memory; products are bit-identical to the materialised CSR.
 *   kh_matrix_create_operator: `apply(user, stream, precision, transpose, s,
 *     x, ldx, y, ldy)` must compute y(:, j) = op(A) x(:, j) for j < s (op = A,
 *     or A^T when transpose != 0) on device columns of type double
 *     (precision KH_PREC_FP64) or float (KH_PREC_MIXED), enqueued on `stream`
 *     (a hipStream_t) or finished before it returns, and return 0 on success
 *     (anything else surfaces as KH_ERUNTIME). It is called with all earlier
 *     work on the stream complete; `user` must outlive the matrix.
 * kh_matrix_download_* and kh_server_submit_resident need a CSR matrix
 * (KH_EINVAL for these operators); kh_matrix_info reports nnz = 2(n-N) for
 * the stencil and -1 for a callback operator, whose ||A||_F reads 0. */
typedef int (*kh_apply_fn)(void *user, void *stream, int precision, int transpose, int s,
                           const void *x, int64_t ldx, void *y, int64_t ldy);
int kh_matrix_create_shifted_difference(kh_context *ctx, int64_t n, int64_t N, double c,
                                        kh_matrix **out);
int kh_matrix_create_operator(kh_context *ctx, int64_t m, int64_t n, kh_apply_fn apply,
                              void *user, kh_matrix **out);
int kh_matrix_destroy(kh_matrix *A);
/* --- matrix ingestion (host only; no device or context needed) ---------- *
 * Canonical CSR from an IJA triplet file, a Matrix Market file or raw 0-based
 * triplets: rows in order, columns strictly ascending per row, repeated
 * (i, j) summed in input order, explicit zeros kept — what Eigen's
 * setFromTriplets produces (src/IO/MatrixLoader.cpp:53-54), laid out as CSR.
 *   kh_load_ija:            load_ija_triplet_matrix(path, rows, cols)
 *                           (include/krans/IO/MatrixLoader.hpp:12-13,
 *                           src/IO/MatrixLoader.cpp:21-56); rows/cols < 0
 *                           take the header's.
 *   kh_load_matrix_market:  load_matrix_market(path) (MatrixLoader.hpp:9,
 *                           MatrixLoader.cpp:12-18, Eigen 3.4 loadMarket).
 *   kh_csr_from_triplets:   setFromTriplets on caller arrays; indices outside
 *                           m x n are KH_EINVAL.
 * A file that cannot be opened, or has no valid header, is KH_ERUNTIME (the
 * reference's std::runtime_error). declared_nnz is the header's entry count
 * (-1 for raw triplets), entries_read the triplets taken. */
typedef struct kh_host_csr_s kh_host_csr;
int kh_load_ija(const char *path, int64_t rows, int64_t cols, kh_host_csr **out);
int kh_load_matrix_market(const char *path, kh_host_csr **out);
int kh_csr_from_triplets(int64_t m, int64_t n, int64_t count, const int32_t *I,
                         const int32_t *J, const double *V, kh_host_csr **out);
int kh_host_csr_info(const kh_host_csr *h, int64_t *m, int64_t *n, int64_t *nnz,
                     int64_t *declared_nnz, int64_t *entries_read);
/* Pointers stay valid until kh_host_csr_free; rowptr has m+1 entries. */
int kh_host_csr_data(const kh_host_csr *h, const int32_t **rowptr, const int32_t **col,
                     const double **val);
int kh_host_csr_free(kh_host_csr *h);
/* rows(), cols() (MatrixInterface.hpp:20-23), nnz and exact ||A||_F */
int kh_matrix_info(const kh_matrix *A, int64_t *m, int64_t *n, int64_t *nnz,
                   double *fro);
/* Download the device structures (bit-exact parity checks). */
int kh_matrix_download_csr(const kh_matrix *A, int32_t *rowptr, int32_t *col,
                           double *val);
int kh_matrix_download_csc(const kh_matrix *A, int32_t *colptr,
                           int32_t *rowidx, double *val, int32_t *perm);
/* y = A x or y = A^T x on host vectors (MatrixInterface::multiply /
 * multiply_transpose, MatrixInterface.hpp:14-19); precision 0/1 = FP64/FP32. */
int kh_matrix_multiply(kh_matrix *A, int transpose, int precision,
                       const double *x, double *y);

/* --- reference-compatible random stream (StdVectorFactory.hpp:16-58) ---- */
int kh_rng_create(uint32_t seed, kh_rng **out);
int kh_rng_destroy(kh_rng *r);
int kh_rng_normal(kh_rng *r, int64_t n, double *out);

/* --- tiny projected problem (host FP64, no device needed) -----------------
 * The right singular system of the refined Golub-Kahan matrix that
 * solve_tiny_refined hands to JacobiSVD (src/Solvers/CyclicLanczosSolver.cpp:
 * 246-262), as the solver computes it: one-sided Jacobi on the column-major
 * rows x cols matrix `a` (rows >= cols), singular values descending in sv
 * (cols), right singular vectors in the columns of v (cols x cols, column-
 * major). mode 0 sweeps the whole matrix; mode 1 sweeps the two parity
 * blocks of a matrix whose entries with i + j even are zero (bitwise the
 * same result); mode 2 as 1 but forms only the even block's vectors (the
 * others are unit vectors on odd rows). Modes 1 and 2 fall back to 0 for
 * other matrices. */
int kh_tiny_svd_right(const double *a, int32_t rows, int32_t cols, int32_t mode, double *sv,
                      double *v);

/* --- solvers --------------------------------------------------------------
 * find_nullspace_cyclic_lanczos_ex (CyclicLanczosSolver.hpp:170-178).
 * Bases are written column-major: right (n x right_cap), left (m x
 * left_cap); counts in the result. `y0` may be NULL. */
int kh_cyclic_lanczos(kh_matrix *A, kh_rng *rng, const kh_cyclic_options *opt,
                      int precision, const double *y0, double *right,
                      int64_t right_cap, double *left, int64_t left_cap,
                      kh_cyclic_result *res);
/* kh_cyclic_lanczos with the bases written to device memory of the
 * matrix's device (right_dev n x right_cap, left_dev m x left_cap, column-
 * major; either may be NULL): the reference's result vectors of a device
 * VectorFactory (CyclicLanczosSolver.hpp:170-174, HipVectorFactory here),
 * with no host copy. `y0` stays a host vector. */
int kh_cyclic_lanczos_dev(kh_matrix *A, kh_rng *rng, const kh_cyclic_options *opt,
                          int precision, const double *y0, double *right_dev,
                          int64_t right_cap, double *left_dev, int64_t left_cap,
                          kh_cyclic_result *res);
/* find_nullspace_lobpcg_ex (LOBPCGSolver.hpp:142-149); theta and
 * residuals_Av per accepted vector; initial_block n x n_initial or NULL. */
int kh_lobpcg(kh_matrix *A, kh_rng *rng, const kh_lobpcg_options *opt,
              int precision, const double *initial_block, int64_t n_initial,
              double *right, int64_t right_cap, double *theta,
              double *residuals, kh_lobpcg_result *res);

/* --- serving: continuous batching of null-space requests -----------------
 * No counterpart in the reference (which solves one matrix per call). A
 * server owns `slots` solver slots (one device stream each) on `device`;
 * requests are admitted in arrival order when a slot is free and their
 * device footprint (matrix + Krylov state) fits in `hbm_budget_bytes`
 * (<= 0: 85% of the free HBM) beside the running ones. Each request is
 * find_nullspace_cyclic_lanczos_ex with its own seeded stream; its result
 * does not depend on the other requests. */
typedef struct kh_server_s kh_server;
int kh_server_create(int device, int slots, double hbm_budget_bytes, kh_server **out);
int kh_server_destroy(kh_server *s);
/* Host CSR, copied at submit. */
int kh_server_submit_cyclic(kh_server *s, int64_t m, int64_t n, const int32_t *rowptr,
                            const int32_t *col, const double *val,
                            const kh_cyclic_options *opt, int precision, uint32_t seed,
                            int64_t *id);
/* A resident device matrix (must outlive the request). */
int kh_server_submit_resident(kh_server *s, kh_matrix *A, const kh_cyclic_options *opt,
                              int precision, uint32_t seed, int64_t *id);
/* Block until request `id` finishes; basis n x cap column-major. */
int kh_server_wait(kh_server *s, int64_t id, double *right, int64_t cap, kh_cyclic_result *res,
                   double *latency_s, double *queued_s);
/* Completed requests, returned vectors, latency percentiles (s), wall time
 * from creation (or the last reset) to the last completion. */
int kh_server_stats(kh_server *s, int64_t *done, int64_t *vectors, double *p50, double *p99,
                    double *wall_s);
int kh_server_reset_clock(kh_server *s);
/* Batched launches for small requests: a request with m + n <= pack_len
 * (host CSR, reorth 2) is solved in a pack with up to pack_max - 1
 * compatible queued requests (same precision and ncv), whose Golub-Kahan
 * cycles run in shared segmented launches; alone it runs in a pack of one,
 * so its result is bitwise the same alone and in any mix. pack_max = 1 (the
 * default, with pack_len 2^19) solves every request on its own. No
 * counterpart in the reference, which has no server. */
int kh_server_set_packing(kh_server *s, int pack_max, int64_t pack_len);

/* --- device buffers and kernel entry points (tests, benchmarks) ---------- */
int kh_dev_alloc(kh_context *ctx, size_t bytes, void **out);
int kh_dev_free(kh_context *ctx, void *p);
int kh_h2d(kh_context *ctx, void *dst, const void *src, size_t bytes);
int kh_d2h(kh_context *ctx, void *dst, const void *src, size_t bytes);
/* y = alpha x + beta y on FP64 device vectors; x may be NULL (y = beta y,
 * beta = 0 writes zeros) -- the level-1 work of VectorInterface
 * (VectorInterface.hpp:11-27) for HipVector. */
int kh_vec_axpby(kh_context *ctx, int64_t len, double alpha, const double *x,
                 double beta, double *y);
/* y = r ./ diag(A A^T) (rows with a zero diagonal copy r): the Jacobi
 * preconditioner for CyclicLanczosOptions::precond_AAt. Device FP64. */
int kh_matrix_jacobi_AAt(kh_matrix *A, const double *r, double *y);
/* y = op(A) x on device vectors, s columns (ld = vector length) */
int kh_spmm_dev(kh_matrix *A, int transpose, int precision, int s,
                const void *x, void *y);
/* H = V^T W (s x q, column-major, FP64 on the host) for device blocks with
 * leading dimension len: the tall-skinny product of the block Gram-Schmidt
 * steps (FP64 MFMA for s, q <= 128 in FP64; FP64-accumulated tiles
 * otherwise). */
int kh_gemm_tn_dev(kh_context *ctx, int precision, const void *V, int s, const void *W,
                   int q, int64_t len, double *H);
/* W = V H (subtract == 0) or W -= V H (subtract != 0) for device blocks
 * with leading dimension len and H (s x q, column-major) on the host: the
 * candidate formation, combine and block update products (FP64 MFMA for
 * 2 <= q <= 128 when H^T fits 64 KB of LDS). */
int kh_gemm_nn_dev(kh_context *ctx, int precision, const void *V, int s, const double *H,
                   int q, void *W, int64_t len, int subtract);
/* The same product on row-major blocks (element (r, c) at r * s + c), the
 * layout the LOBPCG normal products use for blocks of 4 or more columns. */
int kh_spmm_rows_dev(kh_matrix *A, int transpose, int precision, int s,
                     const void *x, void *y);
/* out = a.b (FP64 accumulate); precision of a, b */
int kh_dot_dev(kh_context *ctx, int precision, const void *a, const void *b,
               int64_t len, double *out);
/* Two classical Gram-Schmidt passes of w against the s columns of V
 * (len x s, column-major): the reorthogonalisation of
 * CyclicLanczosSolver.cpp:399-407 in CGS2 form. h_out gets the summed
 * coefficients of both passes, norm_out ||w|| afterwards. */
int kh_cgs2_dev(kh_context *ctx, int precision, const void *V, int s,
                void *w, int64_t len, double *h_out, double *norm_out);
/* The same with the pass-1 update and pass-2 projection fused into one read
 * of V (fused != 0, s <= 1024; what kh_cgs2_dev and the Golub-Kahan steps
 * use) or as four separate sweeps (fused == 0); the results are identical. */
int kh_cgs2_dev_ex(kh_context *ctx, int precision, const void *V, int s, void *w,
                   int64_t len, int fused, double *h_out, double *norm_out);
/* The selective pair of the mixed-precision Golub-Kahan steps (1 <= s < 512):
 * the first pass always, the second only where the first cancelled more than
 * half of ||w||^2 (Daniel-Gragg-Kaufman-Stewart, eta = 1/sqrt(2)), decided on
 * the device. h_out gets the coefficients removed (h1, plus h2 when the
 * second pass ran), norm_out ||w|| afterwards, second_out 1 when the second
 * pass ran and 0 otherwise. With the second pass the results are bitwise
 * those of kh_cgs2_dev. */
int kh_cgs2_dev_select(kh_context *ctx, int precision, const void *V, int s, void *w,
                       int64_t len, double *h_out, double *norm_out, int32_t *second_out);
/* kh_cgs2_dev without allocating. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same kernels as
 * kh_cgs2_dev (bitwise the same w, h and norm), with their reduction partials
 * and second-pass coefficients in a static 32 MiB device array per GPU
 * instead of allocated buffers; h_out (s doubles) and norm_out (1 double)
 * are DEVICE pointers, and the call only enqueues work on the context's
 * stream (wait with kh_context_sync). Limits: 0 <= s <= 1024 and the call's
 * partials within the array (len x s about 4 M doubles; larger calls fail
 * with KH_EINVAL before any work); kernel tracing must be off. All calls on
 * one GPU share its array: calls that overlap on one GPU (other contexts,
 * other threads) race and must be serialised by the caller. */
int kh_cgs2_dev_noalloc(kh_context *ctx, int precision, const void *V, int s, void *w,
                        int64_t len, double *h_out, double *norm_out);
/* G = Z^T [Z AZ]: the S and G Gram blocks of LOBPCGSolver.cpp:67-80 on
 * FP32 MFMA (precision 1) or FP64 (precision 0); Z, AZ len x k column-major,
 * S and G k x k FP64 out (host). */
int kh_gram_dev(kh_context *ctx, int precision, const void *Z, const void *AZ,
                int k, int64_t len, double *S, double *G);
/* Kernel timing with HIP events on the context's stream: between begin and
 * end every launch of the class (1 = reorthogonalisation V^T w products,
 * 2 = w -= V h updates, 3 = sparse products, 4 = FP32-MFMA Gram tiles) is
 * bracketed by an event pair (class 1, the FP32 V^T w products: the start
 * and end its dispatch stamps, i.e. the kernel's own duration; the other
 * classes: events recorded on the stream around the launch, i.e. including
 * its dispatch); end returns the launch count, their summed
 * algorithmic bytes (each array touched once; for class 4 the summed flops,
 * 2 len s q per product) and summed device time in ms. */
int kh_trace_begin(kh_context *ctx, int kernel_class);
int kh_trace_end(kh_context *ctx, int64_t *launches, double *bytes, double *ms);
/* Several classes at once: bit c of `mask` traces class c (1..4);
 * kh_trace_class reads one class's totals before kh_trace_end clears all. */
int kh_trace_begin_mask(kh_context *ctx, unsigned mask);
int kh_trace_class(kh_context *ctx, int kernel_class, int64_t *launches, double *bytes,
                   double *ms);
/* Average device time (ms) of `reps` launches of y = op(A) x, measured with
 * HIP events on the library stream. */
int kh_time_spmv(kh_matrix *A, int transpose, int precision, int reps,
                 double *avg_ms);
/* The gather kernel of y = op(A) x for a CSR matrix in `precision`: *path
 * receives 0 (not chosen yet: the first product times the variants), 1
 * (gathers through L2), 2 or 3 (x staged in LDS per block of rows, first or
 * second block size). With set != 0 the kernel is pinned to `*path` first
 * (0 times again at the next product; a window the operator has no table
 * for, or one larger than a workgroup's LDS in that precision, is refused).
 * Every kernel gives bitwise the same y. Tuning control of this build; the
 * reference has no counterpart. */
int kh_spmv_kernel(kh_matrix *A, int transpose, int precision, int set, int *path);
/* Launch shapes of one traced class (call before kh_trace_end): the distinct
 * (len, s) pairs its launches recorded (rows and columns of the V^T w
 * products of class 1; classes that record no shape report none), with the
 * number of launches of each, at most `max` pairs in order of first use.
 * Observability of this build; the reference has no counterpart. */
int kh_trace_shapes(kh_context *ctx, int kernel_class, int max, int64_t *len, int32_t *s,
                    int64_t *count, int32_t *n_shapes);
/* Average device time (ms) of one V^T w partial-product launch (the kernel of
 * class 1 without its finishing sum) on len x s zero blocks of precision
 * 0 (FP64) or 1 (FP32), over `reps` back-to-back launches between two HIP
 * events on the context's stream, rotating over copies of the blocks that
 * total at least 1 GiB (none is served from the Infinity Cache); 0 < s <= 512. Observability of this build;
 * the reference has no counterpart. */
int kh_time_vtw(kh_context *ctx, int precision, int64_t len, int s, int reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif /* KRANS_HIP_H */
