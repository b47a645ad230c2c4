// Internal C++ API of the KraNS-MI engine: device context, device CSR with its
// CSC twin, device buffers and the launchers of the HIP kernels in
// kernels.hip. Host code above this layer (solvers, adapters, C ABI) never
// touches hip* calls except through these helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <atomic>
#include <stdexcept>
#include <string>
#include <vector>

namespace kh {

// ---------------------------------------------------------------------------
// Errors: every HIP failure becomes a DeviceError; the C ABI maps it to a
// status code, the C++ adapters re-throw the reference's exception types.
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceOOM : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);
#define KH_CHECK(expr) ::kh::hip_check((expr), #expr)

enum class Prec : int { F64 = 0, F32 = 1 };

// ---------------------------------------------------------------------------
// Context: one device, one stream, a slab of device scalars and a pinned
// host mirror used to read scalars back with one small copy.
// Optional per-launch timing of kernel classes with HIP events on the
// context's stream (for roofline reporting): algorithmic bytes of each traced
// launch, its class and its event pair. trace_mask has bit c set for every
// traced class c.
enum TraceClass : int {
  kTraceNone = 0,
  kTraceVtW = 1,
  kTraceVhSub = 2,
  kTraceSpmv = 3,
  kTraceGram = 4,  // FP32-MFMA Gram tiles; the recorded amount is flops, not bytes
  kTraceOther = 5,  // other launches with exact algorithmic bytes (FP64 V^T w,
                    // normalised copies, window products)
};
struct TraceRec {
  double bytes;
  hipEvent_t a, b;
  int cls;
  int64_t len = 0;  // launch shape, where the class records one (rows, columns)
  int s = 0;
};

struct Context {
  unsigned trace_mask = 0;
  std::vector<TraceRec> trace;
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  void trace_reset();
  int device = 0;
  hipStream_t stream = nullptr;
  double *dscal = nullptr;     // device scalar slab (kScalars doubles)
  double *hscal = nullptr;     // pinned host mirror
  // Pinned staging for host-to-device uploads: each upload takes the next
  // slice and is queued without waiting; slices are reused only after a
  // wait on the stream (sync()), which rewinds the area.
  char *hstage = nullptr;
  std::size_t stage_off = 0;
  static constexpr std::size_t kStage = std::size_t(4) << 20;
  double *dpart = nullptr;     // reduction partials
  std::size_t part_cap = 0;    // doubles
  // A caller-supplied partial buffer that replaces dpart for the duration of
  // one call (set and cleared by that call; nullptr otherwise).
  double *part_ext = nullptr;
  std::size_t part_ext_cap = 0;  // doubles
  std::size_t part_capacity() const { return part_ext ? part_ext_cap : part_cap; }
  int cu_count = 0;
  int lds_max = 0;  // bytes of LDS one workgroup may allocate
  static constexpr int kScalars = 1 << 16;
  // part_doubles: capacity of the reduction partials (default 16 Mi doubles)
  explicit Context(int dev, std::size_t part_doubles = std::size_t(1) << 24);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  void sync();  // waits for the stream and rewinds the upload staging
  // queues a copy of `bytes` host bytes to device `dst`; the host source may
  // change as soon as this returns
  void upload(void *dst, const void *src, std::size_t bytes);
  // copy `n` device scalars starting at slab index `i0` into hscal[i0..]
  void fetch_scalars(int i0, int n);
  double *partials(std::size_t need);
};

// One instantiated HIP graph of work queued on a context's stream, kept
// under an integer key and captured again when the key changes. Capture
// records launches without running them (thread-local mode, so other
// threads' allocations stay legal); the queued code must not synchronise,
// allocate, or clear memory with hipMemsetAsync.
struct CapturedWork {
  hipGraphExec_t exec = nullptr;
  int64_t key = -1;
  int64_t captures = 0, replays = 0;
  CapturedWork() = default;
  ~CapturedWork() { reset(); }
  CapturedWork(const CapturedWork &) = delete;
  CapturedWork &operator=(const CapturedWork &) = delete;
  void reset();
  bool holds(int64_t k) const { return exec != nullptr && key == k; }
  void begin(Context &ctx);
  // ends the capture begun by begin(); ok = false discards it
  void end(Context &ctx, int64_t k, bool ok);
  void launch(Context &ctx);
};

// Device allocation with an owner; moves only.
struct DBuf {
  void *p = nullptr;
  std::size_t bytes = 0;
  DBuf() = default;
  explicit DBuf(std::size_t b);
  ~DBuf();
  DBuf(DBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// ---------------------------------------------------------------------------
// Device sparse matrix: canonical CSR (int32 pointers/indices, columns
// ascending within each row) and its CSC twin (rows ascending within each
// column) with the permutation perm[q] = CSR position of CSC entry q. Values
// are held in FP64 and, for the mixed-precision path, rounded to FP32.
struct DeviceCsr {
  int64_t m = 0, n = 0, nnz = 0;
  DBuf rp, ci, va, va32;  // CSR
  DBuf cp, ri, vt, vt32;  // CSC twin
  DBuf perm;
  int lpr_rows = 1, lpr_cols = 1;  // lanes per row / per column (autotuned)
  int max_row = 0, max_col = 0;
  // entries in every row / every column when all have the same count, else 0
  int uniform_rows = 0, uniform_cols = 0;
  double fro = 0.0;  // exact ||A||_F from the CSR values
  std::vector<double> diag_AAt_host;
  DBuf diag_AAt;  // row sums of squares, for the Jacobi preconditioner
  // Block products of two or more columns: row-major (one gather of a
  // contiguous row per nonzero, plus two transposes) or column by column,
  // whichever was faster when first timed, per [FP64|FP32][A|A^T][q <= 8|q > 8]
  // (0 = not timed yet, 1 = column by column, 2 = row-major). Both give
  // bitwise the same products.
  mutable std::atomic<signed char> block_path[2][2][2]{};
  // Row-major block products whose gathered block is larger than the MALL
  // run in column chunks (the gathered rows of one chunk then stay in the
  // MALL): the chunk width, per [FP64|FP32][A|A^T], picked by timing on
  // first use (0 = not timed yet; widths >= the block's mean no chunking).
  // Every width gives bitwise the same products.
  mutable std::atomic<int> rows_chunk[2][2]{};
  // LDS-staged gathers (operators whose rows all hold the same 1-4 or 8
  // entries), per [A|A^T] and two block sizes k: rows in blocks of
  // span_rb[d][k] rows, block b reading x only within [span[d][k][b],
  // span[d][k][b] + span_len[d][k]) (starts multiples of 4). span_rb 0: no
  // table (irregular rows, fewer than 65,536 rows, or windows above 160 KB of
  // FP32).
  int span_rb[2][2] = {};
  int span_len[2][2] = {};
  DBuf span[2][2];
  // y = op(A) x per [FP64|FP32][A|A^T]: 0 not timed yet, 1 gathers from L2,
  // 2 + k the LDS window of block size k; the fastest when first timed (all
  // give bitwise the same y).
  mutable std::atomic<signed char> spmv_path[2][2]{};
};

// Checks a host canonical CSR (rowptr[0] = 0, monotone within [0, nnz],
// columns in range and strictly ascending per row) without reading outside
// rowptr[0 .. m] and col/val[0 .. nnz); returns the longest row. Throws
// std::invalid_argument on malformed input.
int validate_csr(int64_t m, int64_t n, const int32_t *rowptr, const int32_t *col,
                 const double *val);
// Build from host canonical CSR (validated). Throws std::invalid_argument
// on malformed input.
DeviceCsr *csr_create(Context &ctx, int64_t m, int64_t n,
                      const int32_t *rowptr, const int32_t *col,
                      const double *val);

// ---------------------------------------------------------------------------
// Kernel launchers (kernels.hip). All run on ctx.stream, asynchronously.
// Vectors are contiguous; multi-vector blocks are column-major with leading
// dimension `ld` (elements).

// y = A x (trans=false, CSR) or y = A^T x (trans=true, CSC twin).
// With z != nullptr: y = A x - (*beta) z (beta on device).
void spmv(Context &ctx, const DeviceCsr &A, bool trans, Prec p, const void *x,
          void *y, const void *z = nullptr, const double *beta = nullptr);
// The kernel spmv picked for op(A) in precision p: 0 not timed yet, 1 gathers
// from L2, 2 / 3 the LDS window of the first / second block size.
int spmv_path(const DeviceCsr &A, bool trans, Prec p);
void set_spmv_path(const DeviceCsr &A, bool trans, Prec p, int path);
// Same operator on `s` columns at once (x: len_in x s, y: len_out x s).
// Columns whose active flag is 0 are skipped (active may be nullptr).
void spmm(Context &ctx, const DeviceCsr &A, bool trans, Prec p, int s,
          const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *active);

// Row-major block product: Y(r, c) = sum_p op(A)(r, p) X(p, c) with X and Y
// stored row-major (stride s); same per-entry rounding as spmm.
void spmm_rows(Context &ctx, const DeviceCsr &A, bool trans, Prec p, int s, const void *X,
               void *Y);
// Shifted-difference stencil (A x)_i = x_{i+N} - c x_i, A (n-N) x n, applied
// without a matrix: y = op(A) x (+ optional y -= beta z), and on blocks
// (column-major with active mask, or row-major with stride s).
void sd_apply(Context &ctx, int64_t n, int64_t N, double c, bool trans, Prec p, const void *x,
              void *y, const void *z, const double *beta);
void sd_block(Context &ctx, int64_t n, int64_t N, double c, bool trans, Prec p, int s,
              const void *x, int64_t ldx, void *y, int64_t ldy, const int *act, bool row_major);
// y = y + (-beta) z, beta on the device.
void sub_beta(Context &ctx, Prec p, int64_t len, const double *beta, const void *z, void *y);
// Y(:, c) = X(:, c) where act[c] (device mask).
void copy_active(Context &ctx, Prec p, int64_t len, int s, const void *X, int64_t ldx, void *Y,
                 int64_t ldy, const int *act);
// Column-major (leading dimension ld) <-> row-major (stride s) block copies.
void cols_to_rows(Context &ctx, Prec p, int64_t len, int s, const void *X, int64_t ldx, void *R);
void rows_to_cols(Context &ctx, Prec p, int64_t len, int s, const void *R, void *X, int64_t ldx);
// out[0] = sum_i a_i b_i (FP64 accumulate, fixed reduction tree); with
// do_sqrt: out[0] = sqrt(max(0,sum)).
void dot(Context &ctx, Prec pa, const void *a, Prec pb, const void *b,
         int64_t len, double *out, bool do_sqrt = false);
// Average duration (ms) of one V^T w partial launch (k_vtw_partial, no
// finishing sum) at len x s in precision p, over reps back-to-back launches
// on the context's stream (V, w zero-filled, allocated here).
double time_vtw(Context &ctx, Prec p, int64_t len, int s, int reps);
// out[j] = V(:,j) . w for j < s (FP64 accumulate). With wsq (s < 512) also
// out[s] = ||w||^2 from the same read of w: the input of the selective
// second pass below.
void gemv_t(Context &ctx, Prec pv, const void *V, int64_t ld, int s, Prec pw,
            const void *w, int64_t len, double *out, bool wsq = false);
// Reduction partials (doubles) that a CGS2 of w (length len) against s
// columns can take in the launches of gemv_t, gemv_n_sub_t and gemv_n_sub:
// an upper bound over every kernel those launchers may choose.
std::size_t cgs2_partials_bound(int64_t len, int s);
// R(:, j) = AX(:, j) - theta_j X(:, j), j < q (neg_theta: q device doubles,
// -theta_j), columns with leading dimension ld: bitwise a copy of AX followed
// by axpby(-theta_j, X, 1) per column, in one launch.
void resid_block(Context &ctx, Prec p, int64_t len, int q, const void *AX, const void *X,
                 int64_t ld, const double *neg_theta, void *R);
// Selective CGS pair (g non-null): g holds the first pass's s coefficients
// and ||w||^2 (gemv_t with wsq). When the first pass cancelled less than half
// of ||w||^2 (the DGKS test, eta = 1/sqrt(2)), gemv_n_sub_t does nothing,
// its out is zero, and the following gemv_n_sub / gemv_n_sub_normalise apply
// the first pass's update w -= V g instead of w -= V h: one CGS pass, decided
// on the device. *ran (optional) gets 1.0 when the second pass ran, else 0.
// `cancel` (gemv_n_sub / gemv_n_sub_normalise) is 1 - eta^2 of that test:
// the second pass is due when ||h1||^2 >= cancel ||w||^2. With h = g the
// launch is a speculative single pass whose *ran reports the test.
//
// w -= V h, then ||w|| into *nrm and dst = w / ||w|| (bitwise as gemv_n_sub
// with the norm followed by scale_copy_dev), the norm and copy in one launch.
void gemv_n_sub_normalise(Context &ctx, Prec pv, const void *V, int64_t ld, int s,
                          const double *h, Prec pw, void *w, int64_t len, double *nrm,
                          void *dst, const double *g = nullptr, double *ran = nullptr,
                          double cancel = 0.5);
// w -= V h  (h: s FP64 device scalars); with nrm2_out: also ||w_new||.
void gemv_n_sub(Context &ctx, Prec pv, const void *V, int64_t ld, int s,
                const double *h, Prec pw, void *w, int64_t len,
                double *nrm2_out = nullptr, const double *g = nullptr, double *ran = nullptr,
                double cancel = 0.5);
// w -= V h, then out = V^T w for the updated w (FP64): the middle of a CGS
// pair with one read of V; the update as gemv_n_sub, out as gemv_t up to the
// grouping of its partial sums. With nrm1, also ||w|| after the update.
void gemv_n_sub_t(Context &ctx, Prec pv, const void *V, int64_t ld, int s, const double *h,
                  Prec pw, void *w, int64_t len, double *out, double *nrm1 = nullptr,
                  const double *g = nullptr);
// out = V h (overwrite)
void gemv_n(Context &ctx, Prec pv, const void *V, int64_t ld, int s,
            const double *h, Prec po, void *out, int64_t len);
// Multi-column: OUT(:,c) = V * H(:,c) for c < q, H is s x q FP64 on device.
void gemm_nn(Context &ctx, Prec pv, const void *V, int64_t ldv, int s,
             const double *H, int q, Prec po, void *OUT, int64_t ldo,
             int64_t len);

// H(:,c) = V^T W(:,c) for c < q; H is s x q (column-major) FP64 on device.
void gemm_tn(Context &ctx, Prec pv, const void *V, int64_t ldv, int s, Prec pw,
             const void *W, int64_t ldw, int q, int64_t len, double *H);
// H = V^T W for FP32 V, W on the matrix cores (FP32 MFMA accumulation within
// a row chunk, FP64 sums across chunks): the LOBPCG Gram blocks.
// S = Z^T Z and G = Z^T AZ (q <= 128) in one fused pass: SG[0:q*q] = S,
// SG[q*q:2*q*q] = G, column-major, FP64 (FP32 MFMA, FP64 across row chunks).
void gram2_mfma(Context &ctx, const float *Z, int64_t ldz, const float *AZ, int64_t ldaz, int q,
                int64_t len, double *SG);
// FP64 blocks: the same pass on FP64 MFMA (v_mfma_f64_16x16x4f64).
void gram2_mfma(Context &ctx, const double *Z, int64_t ldz, const double *AZ, int64_t ldaz,
                int q, int64_t len, double *SG);
void gram_mfma(Context &ctx, const float *V, int64_t ldv, int s, const float *W, int64_t ldw,
               int q, int64_t len, double *H);
// W(:,c) -= V H(:,c) for c < q; optional norms[c] = ||W(:,c)|| afterwards.
void gemm_nn_sub(Context &ctx, Prec pv, const void *V, int64_t ldv, int s,
                 const double *H, int q, Prec pw, void *W, int64_t ldw,
                 int64_t len, double *norms = nullptr);
// out[c] = ||W(:,c)||_2 (FP64 accumulate, fixed tree).
void col_norms(Context &ctx, Prec p, const void *W, int64_t ldw, int q,
               int64_t len, double *out);
// W(:,c) *= scale[c] (device scalars).
void scale_cols(Context &ctx, Prec p, int64_t len, int q, void *W, int64_t ldw,
                const double *scale, bool skip_zero);
// OUT(:,c) = X(:,c) - Y(:,c) where mask[c] (mask may be null).
void sub_cols(Context &ctx, Prec p, int64_t len, int q, const void *X,
              int64_t ldx, const void *Y, int64_t ldy, void *OUT, int64_t ldo,
              const int *mask);

// y = alpha*x + beta*y (host scalars); x may be null (y = beta*y).
void axpby(Context &ctx, Prec p, int64_t len, double alpha, const void *x,
           double beta, void *y);
void scale_dev(Context &ctx, Prec p, int64_t len, const double *num,
               bool reciprocal, void *y);  // y *= (*num) or y *= 1/(*num)
void copy(Context &ctx, Prec p, int64_t len, const void *x, void *y);
// y = x * (1 / *den) with den on the device
void scale_copy_dev(Context &ctx, Prec p, int64_t len, const double *den, const void *x,
                    void *y);
void convert(Context &ctx, Prec from, const void *x, Prec to, void *y,
             int64_t len);
void fill_zero(Context &ctx, Prec p, int64_t len, void *y);

// Jacobi preconditioner y = r ./ d (d>0) else r.
void diag_solve(Context &ctx, Prec p, int64_t len, const double *d,
                const void *r, void *y);

// ---------------------------------------------------------------------------
// CG on (A A^T) w = rhs with s independent right-hand sides (column-major,
// length m each), FP64 or FP32 arithmetic. Every column follows exactly the
// scalar recurrence of cg_normal_eq_prec, with its own alpha/beta and stop
// flags kept on the device; the host launches iterations in chunks and
// polls the per-column state once per chunk.
struct CgState {  // per column, device resident
  double rz, rnorm, tol, pAp, rz_new, rhsn;
  int32_t k, done, reason, pad;
  // stall test: the smallest residual norm so far that halved the one
  // before it, and the iteration that reached it
  double best;
  int32_t kbest, pad2;
};
struct CgBatch {
  int s = 0;
  int64_t m = 0, n = 0;
  Prec p = Prec::F64;
  DBuf r, z, pv, Ap, tmp_n, st, act;
  DBuf rows_m;  // row-major m x s block of the row-major product path
  bool have_prec = false;
  void ensure(int s_, int64_t m_, int64_t n_, Prec p_);
};
// Initialise: w = 0, r = rhs, z = M^{-1} r, p = z, per-column rho/tol.
void cg_init(Context &ctx, CgBatch &B, const void *rhs, void *w,
             const double *diag, double rel_tol, int64_t maxit);
// The scalar part of one CG iteration, after the caller has formed
// B.Ap = A (A^T B.pv) for the active columns (B.act): pAp, alpha, the w/r/z
// updates, rho, beta and the new direction. No-ops for stopped columns.
// stall > 0: a column whose residual norm has not halved within `stall`
// iterations stops with reason 5 (w and r updated, as at maxit).
void cg_after_products(Context &ctx, CgBatch &B, void *w, const double *diag,
                       int64_t maxit, double breakdown_tol, int64_t stall = 0);

// The device scratch of kh_cgs2_dev_noalloc (noalloc.hip): noalloc_prepare
// looks up its address on `device` (once per device, that device current);
// noalloc_scratch returns it and its capacity in doubles.
void noalloc_prepare(int device);
double *noalloc_scratch(int device, std::size_t *cap_doubles);

// ---------------------------------------------------------------------------
// Packed Golub-Kahan steps of several cyclic solves (the serving loop's
// batched launches). The requests' vectors are stacked in one array, each
// request's rows padded to a multiple of kPackRows (the row block of the
// V^T w and update kernels), so every such block belongs to one request, its
// segment. A launch serves every active segment; a segment's arithmetic
// depends on its own data only, never on the other segments of the pack.
constexpr int64_t kPackRows = 2048;
struct PackSeg {
  double *bnorm = nullptr;    // the request's beta_j slots (device)
  int32_t nb = 0;             // its locked-basis size (y-side columns beyond are zero)
  int32_t active = 0;         // 0: every kernel skips the segment
  int64_t blk0[2] = {0, 0};   // first block on the m side [0] and the n side [1]
  int64_t nblk[2] = {0, 0};   // block counts
};
struct PackLayout {           // device tables of one pack
  const PackSeg *segs = nullptr;
  int nseg = 0;
  const int32_t *seg_of_blk[2] = {nullptr, nullptr};  // block -> segment, per side
  int64_t nblk[2] = {0, 0};   // total blocks per side (lengths nblk * kPackRows)
};
// y = op(A) x - beta_{seg, jb} z (z may be null; one lane per row, the
// rounding of the solo lane / ELL kernels) on the packed block-diagonal CSR;
// side of the output = 0 for A x, 1 for A^T x.
void pack_spmv(Context &ctx, const DeviceCsr &A, bool trans, Prec p, const PackLayout &L,
               const void *x, void *y, const void *z, int jb);
// V^T w block partials over the whole packed length of `side` (part[c * nblk + b]).
void pack_vtw(Context &ctx, Prec p, const PackLayout &L, int side, const void *V, int64_t ld,
              int s, const void *w, double *part);
// per segment and column: H[seg * hs + c] = sum of the segment's partials
void pack_finish(Context &ctx, const PackLayout &L, int side, const double *part, int s,
                 double *H, int64_t hs);
// w -= V h_seg in every active segment (h_seg = H + seg * hs); with norm_part,
// the per-block partials of ||w_new||^2
void pack_vh_sub(Context &ctx, Prec p, const PackLayout &L, int side, const void *V, int64_t ld,
                 int s, const double *H, int64_t hs, void *w, double *norm_part);
// beta_{seg, j} = ||w_seg|| from the norm partials, into the segment's bnorm[j]
void pack_norm(Context &ctx, const PackLayout &L, int side, const double *norm_part, int j);
// dst = w * (1 / beta_{seg, j}) in every active segment
void pack_scale_copy(Context &ctx, Prec p, const PackLayout &L, int side, int j, const void *w,
                     void *dst);

// Partial-buffer size needed for the reductions of `len` elements.
int reduce_blocks(int64_t len);

}  // namespace kh
