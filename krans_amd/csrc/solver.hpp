// Device-resident null-space solvers of the KraNS-MI engine: the operator
// abstraction they run on, the reference-compatible random stream, options,
// results and the two drivers (cyclic Lanczos / Golub-Kahan and LOBPCG).
#pragma once
#include <stdexcept>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "refstream.hpp"

namespace kh {

// ---------------------------------------------------------------------------
// A linear operator on the device: A (rows x cols) with A x, A^T x and their
// multi-column forms. `mv` optionally fuses y = A x - (*beta) z.
struct DeviceOp {
  virtual ~DeviceOp() = default;
  virtual int64_t rows() const = 0;
  virtual int64_t cols() const = 0;
  virtual void mv(Context &c, Prec p, const void *x, void *y, const void *z = nullptr,
                  const double *beta = nullptr) const = 0;
  virtual void mtv(Context &c, Prec p, const void *x, void *y, const void *z = nullptr,
                   const double *beta = nullptr) const = 0;
  virtual void mm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y,
                  int64_t ldy, const int *act) const = 0;
  virtual void mtm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y,
                   int64_t ldy, const int *act) const = 0;
  // Block products on row-major blocks (stride s): Y = op(A) X, op = A or
  // A^T. Operators without them report false and callers use mm/mtm.
  virtual bool has_rows() const { return false; }
  // True when mv / mtv only enqueue kernels on the context's stream (no
  // synchronisation, no host callback), so they can be captured in a graph.
  virtual bool capturable() const { return false; }
  virtual void mm_rows(Context &c, Prec p, int s, const void *X, void *Y, bool trans) const {
    (void)c, (void)p, (void)s, (void)X, (void)Y, (void)trans;
    throw std::logic_error("row-major block products not provided by this operator");
  }
  // The cached choice between row-major and column-by-column block products
  // for op(A) in precision p on q columns, or nullptr (no cache: row-major).
  virtual std::atomic<signed char> *block_path(Prec p, bool trans, int q) const {
    (void)p, (void)trans, (void)q;
    return nullptr;
  }
  // diag(A A^T) on the device (Jacobi preconditioner), or nullptr.
  virtual const double *diag_AAt() const { return nullptr; }
  virtual double fro() const { return 0.0; }
};

// CSR matrix (or its transpose view, which swaps the CSR and CSC roles).
struct CsrOp final : DeviceOp {
  const DeviceCsr &A;
  bool t;
  CsrOp(const DeviceCsr &a, bool transposed) : A(a), t(transposed) {}
  int64_t rows() const override { return t ? A.n : A.m; }
  int64_t cols() const override { return t ? A.m : A.n; }
  void mv(Context &c, Prec p, const void *x, void *y, const void *z,
          const double *beta) const override {
    spmv(c, A, t, p, x, y, z, beta);
  }
  void mtv(Context &c, Prec p, const void *x, void *y, const void *z,
           const double *beta) const override {
    spmv(c, A, !t, p, x, y, z, beta);
  }
  void mm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *act) const override {
    spmm(c, A, t, p, s, x, ldx, y, ldy, act);
  }
  void mtm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
           const int *act) const override {
    spmm(c, A, !t, p, s, x, ldx, y, ldy, act);
  }
  bool has_rows() const override { return true; }
  bool capturable() const override { return true; }
  void mm_rows(Context &c, Prec p, int s, const void *X, void *Y, bool trans) const override {
    spmm_rows(c, A, t != trans, p, s, X, Y);
  }
  std::atomic<signed char> *block_path(Prec p, bool trans, int q) const override {
    return &A.block_path[p == Prec::F64 ? 0 : 1][t != trans ? 1 : 0][q > 8 ? 1 : 0];
  }
  const double *diag_AAt() const override { return t ? nullptr : A.diag_AAt.as<double>(); }
  double fro() const override { return A.fro; }
};

// A caller's preconditioner on FP64 device vectors of length len: z = M^{-1} r
// (the reference's AAtPreconditioner / NormalPreconditioner std::functions,
// CyclicLanczosSolver.hpp:77-78, LOBPCGSolver.hpp:68-69). Called with all
// earlier work on the context's stream complete; its work is complete, or
// enqueued on that stream, when it returns; it must not write r.
struct DevicePreconditioner {
  virtual ~DevicePreconditioner() = default;
  virtual void apply(Context &c, const double *r, double *z, int64_t len) const = 0;
};

// ---------------------------------------------------------------------------
// Options: field for field CyclicLanczosOptions (CyclicLanczosSolver.hpp:84-146)
// minus the two non-POD members, which are passed separately.
struct CyclicParams {
  int64_t k = 1, ncv = 48, max_restarts = 80;
  double tol_abs = 1e-12, tol_rel = 1e-10;
  int reorth = 2;
  bool use_harmonic = true;
  double target_tau = 0.0;
  int nullity_known = -1, rank_known = -1;
  int64_t cg_maxit = 80;
  double cg_rel_tol = 1e-10, cg_breakdown_tol = 1e-30;
  int projector_passes = 1, grad_max_iters = 3000, frob_samples = 4;
  double prescreen_factor = 0.8;
  int64_t stagnation_restarts = 30;
  bool return_left_space = false, verbose = false;
  // engine choices (not in the reference struct)
  bool jacobi_AAt = false;  // M^{-1} = diag(A A^T)^{-1} in the polish CG
  // precond_AAt as a callback (CyclicLanczosSolver.hpp:115): the polish CGs
  // then run in FP64, one column at a time, calling it once per iteration
  const DevicePreconditioner *precond = nullptr;
  Prec krylov = Prec::F64;  // Golub-Kahan vectors and products
  Prec polish = Prec::F64;  // CG of the projector polish (F32 = refinement)
  const double *y0_host = nullptr;  // initial_right_space[0], host, length n
  // Golub-Kahan cycles replayed from a captured HIP graph, for operators
  // that only enqueue kernels and with tracing off: 1 always, 0 automatic
  // (vectors short enough for the cycle to be launch-bound), -1 never
  int graphs = 0;
  // mixed precision, prescreen_factor < 0.1: how candidates are screened
  // before their FP64 evaluation. 0: from the projected problem when
  // refined extraction has a zero target (no vector work), else 1; 1: FP32
  // candidate vectors and their residuals; -1: no screen
  int screen = 0;
  // second pass of each Golub-Kahan step's classical Gram-Schmidt pair
  // (reorth = 2): 1 always (the reference's double reorthogonalisation);
  // 2 only where the first pass cancelled half of ||w||^2 or more (the DGKS
  // test with eta = 1/sqrt(2); steps are queued with one pass and a cycle is
  // queued again from a step whose test asks for the second); 0 automatic:
  // 2 for FP32 Krylov vectors, 1 for FP64
  int second_pass = 0;
  // eta of that test (0 < eta <= 1; 1 asks for every second pass)
  double second_pass_eta = 0.70710678118654752;
};

struct PhaseTimes {
  double gk = 0, tiny = 0, cand = 0, polish = 0, fallback = 0, rng = 0;
};

struct CyclicStats {
  int64_t restarts = 0, krylov_steps_total = 0;
  double frob_norm_est = 0.0;
  bool converged_to_target = false;
  int64_t cg_solves = 0, cg_maxit_hits = 0;
  int64_t spmv = 0, spmtv = 0;    // operator applications (column count)
  int64_t cg_iters = 0;           // CG iterations over all columns
  int64_t screened = 0;           // candidates rejected by the FP32 prescreen (mixed)
  int64_t refined = 0;            // locked vectors replaced by the final FP64 pass (mixed)
  int64_t escalated = 0;          // FP32 projector CGs solved again in FP64 (mixed)
  int64_t graph_captures = 0, graph_replays = 0;  // Golub-Kahan cycles as HIP graphs
  int64_t cgs_pairs = 0, second_passes = 0;  // Golub-Kahan CGS pairs, second passes run
  int64_t gk_requeued = 0;  // cycles queued again from a step that needed its second pass
  PhaseTimes t;
};

// (Preconditioned) CG on (A A^T) W = RHS for q columns (column-major, length
// rows), each column the recurrence of cg_normal_eq_prec
// (CyclicLanczosSolver.cpp:48-113); the host polls the device-resident
// states once per chunk of iterations. Returns the per-column states.
// Y = op(A) X on q columns (column-major, leading dimensions ldx / ldy; op =
// A or A^T). With two or more columns and row-major products available the
// block goes either through row-major copies (rin / rout grown as needed;
// every column is computed, so columns that `mask` marks inactive may be
// overwritten) or column by column through op's mm / mtm with the mask,
// whichever the operator's cache says was faster; the first product of a
// kind runs and times both (bitwise the same results). Otherwise mm / mtm.
void block_product(Context &c, const DeviceOp &A, Prec p, int q, const void *X, int64_t ldx,
                   void *Y, int64_t ldy, bool trans, const int *mask, DBuf &rin, DBuf &rout);
// Batched projector CG A A^T w = rhs on q columns; stall > 0 stops a column
// whose residual norm has not halved within that many iterations (reason 5).
void cg_solve_ops(Context &c, const DeviceOp &A, CgBatch &B, int q, const void *rhs, void *w,
                  Prec p, const double *diag, int64_t maxit, double rel_tol,
                  double breakdown_tol, std::vector<CgState> &st, int64_t stall = 0);

// A device block of `count` FP64 vectors of length `len` (column-major).
struct DeviceBasis {
  int64_t len = 0, count = 0, cap = 0;
  DBuf buf;
  double *col(int64_t j) const { return buf.as<double>() + j * len; }
};

// Right null space of op; left space computed by the caller with the
// transposed operator.
DeviceBasis cyclic_right(Context &ctx, const DeviceOp &op, RefStream &rng,
                         const CyclicParams &P, CyclicStats &S);

// Device bytes the cyclic driver allocates for an m x n operator (the
// admission footprint of the serving loop, excluding the matrix).
double cyclic_footprint(int64_t m, int64_t n, const CyclicParams &P);

// LOBPCGOptions (LOBPCGSolver.hpp:71-115).
struct LobpcgParams {
  int64_t nev = 1, block_size = 0, max_iters = 1000;
  double tol_abs = 1e-12, tol_rel = 1e-10;
  bool soft_locking = true;
  double restart = 0.9, gs_tol = 1e-14;
  bool gs_reorth = true;
  int frob_samples = 4;
  int64_t max_tiny_dim = 0;
  int nullity_known = -1, rank_known = -1;
  bool verbose = false;
  int64_t progress_every = 100;
  Prec block = Prec::F64;  // X/W/P blocks and their products
  const DevicePreconditioner *M_inv = nullptr;  // W = M^{-1} R (LOBPCGSolver.cpp:654-659)
  const double *initial_block_host = nullptr;  // n x n_initial, column-major
  int64_t n_initial = 0;
};

struct LobpcgStats {
  std::vector<double> theta, residuals_Av;
  int64_t iters = 0, locked = 0, block_size = 0;
  double frob_norm_est = 0, accept_threshold = 0, best_unaccepted_residual = 0;
  bool converged_to_target = false;
  int64_t spmv = 0, spmtv = 0;
  int64_t p_drops = 0;  // FP32 steps restarted without P (dependent [X W P])
  int64_t deferred = 0;  // FP32 iterations whose candidates waited for a batch
  int64_t random_cols = 0;  // columns drawn from the random stream
  int64_t cholqr_blocks = 0, cholqr_fallbacks = 0;  // FP32 blocks: Cholesky QR / MGS after it
  double t_rng = 0, t_gep = 0;  // host time: random columns, tiny GEPs
  double t_gram = 0, t_sparse = 0, t_host = 0;
};

DeviceBasis lobpcg(Context &ctx, const DeviceOp &op, RefStream &rng,
                   const LobpcgParams &P, LobpcgStats &S);

}  // namespace kh
