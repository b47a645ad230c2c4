// The device-resident cyclic Lanczos (Golub-Kahan) driver of one solve
// (implemented in cyclic.cpp): find_right_nullspace
// (src/Solvers/CyclicLanczosSolver.cpp:291-616) with its restart loop split
// into phases, so that the packed serving loop can build the Golub-Kahan
// cycles of several solves in shared launches while each solve keeps its own
// replay, candidates, polish and restart logic.
#pragma once
#include <chrono>
#include <vector>

#include "solver.hpp"

namespace kh {

struct Group {  // a block of vectors to orthogonalise against
  const void *V;
  int64_t ld;
  int count;
  Prec p;
};

class CyclicDriver {
public:
  CyclicDriver(Context &ctx, const DeviceOp &op, RefStream &rng, const CyclicParams &P,
               CyclicStats &S)
      : c_(ctx), A_(op), rng_(rng), P_(P), S_(S), m_(op.rows()), n_(op.cols()) {}

  DeviceBasis run();

  // The restart loop in phases, for a caller that runs the Golub-Kahan build
  // itself (the packed server): prepare(); while (cycle_begin()) { build the
  // cycle; st = cycle_end(..); if (st == kDone) return finish(); if (st ==
  // kStop) break; } return after_loop(). run() is exactly this with
  // queue_cycle() as the build.
  enum { kContinue = 0, kDone = 1, kStop = 2 };
  struct CycleQueued {
    int64_t queued, nmv, nmtv;  // steps queued, A x and A^T x products
  };
  int prepare();
  bool cycle_begin();
  // The cycle from step j0 (0: the whole cycle); step force2 runs both
  // Gram-Schmidt passes whatever the second-pass setting.
  CycleQueued queue_cycle(int64_t j0 = 0, int64_t force2 = -1);
  int cycle_end(int64_t queued, int64_t nmv, int64_t nmtv);
  DeviceBasis after_loop();
  DeviceBasis finish();
  // The Golub-Kahan vectors in storage the caller owns (set before
  // prepare): Y and the Krylov-precision locked basis in `vb` (ycap + k
  // columns, leading dimension ldy), X in `x` (xcap columns, ldx), the work
  // vectors wx (m) and wy (n).
  void use_storage(void *vb, int64_t ldy, void *x, int64_t ldx, void *wx, void *wy) {
    ext_ = true;
    vbp_ = vb, ldy_ = ldy, xp_ = x, ldx_ = ldx, wxp_ = wx, wyp_ = wy;
  }
  static int64_t ycap_for(int64_t ncv) { return ncv / 2 + 2; }
  int64_t basis_count() const { return basis_.count; }
  // the (empty) basis of a solve for which prepare() found nothing to compute
  DeviceBasis take_basis() { return std::move(basis_); }
  double *beta_slots() const { return bnorm_; }
  int64_t ncv_eff() const { return ncv_; }

private:
  Context &c_;
  const DeviceOp &A_;
  RefStream &rng_;
  const CyclicParams &P_;
  CyclicStats &S_;
  int64_t m_, n_;
  Prec kp_ = Prec::F64;  // Krylov precision

  // device state
  DeviceBasis basis_;   // FP64 locked basis (n x k_target)
  // [Y reversed | basis] in Krylov precision: Golub-Kahan vector y_j sits in
  // column ycap-1-j and locked vector b_i in column ycap+i, so the y-side
  // reorthogonalisation set [basis, Y] is one contiguous block.
  DBuf VB_;
  DBuf X_, wx_, wy_;
  DBuf y0_, best_v_, tmp_m_, tmp_n_;
  DBuf small_;          // device scalars of this solve
  double *bnorm_ = nullptr, *h_ = nullptr, *wsc_ = nullptr, *C_ = nullptr, *H_ = nullptr;
  int64_t ycap_ = 0, xcap_ = 0;
  int64_t ldy_ = 0, ldx_ = 0;  // padded leading dimensions of VB_ and X_
  int64_t k_target_ = 0;
  double accept_thr_ = 0, prescreen_thr_ = 0;
  std::vector<double> hbuf_;
  // window blocks
  DBuf Wv_, AV_, Vc_, AVc_, Vs_, Rs_, Vp_, AVp_, Wm_, Tn_, Wm64_;
  DBuf Wv32_, AV32_;      // FP32 screen of a window (mixed precision)
  bool screen_ = false;   // screen candidate windows in FP32 first
  int screen_miss_ = 0;   // consecutive windows the screen did not thin out
  // mixed polish: consecutive FP32 CG batches none of whose columns
  // converged; from kPolish32Misses on the solve polishes in FP64 directly
  int polish32_miss_ = 0;
  bool polish32_off_ = false;
  int wcap_ = 0;
  CgBatch cg_;
  CapturedWork gk_graph_;  // the queued Golub-Kahan cycle, keyed by the basis size
  bool graphs_ = false;
  bool select2_ = false;  // selective second CGS pass in the Golub-Kahan steps
  double dgks_cancel_ = 0.5;  // 1 - eta^2 of its test
  // per step: 1 when the DGKS test of its speculative single pass asks for
  // the second pass
  double *pass2_ = nullptr;
  // per queued step of the current cycle: 0 no CGS pair, 1 pair with both
  // passes, 3 speculative single pass (its test in pass2_)
  std::vector<char> stepkind_;
  std::vector<char> gk_kinds_;  // stepkind_ of the captured cycle
  int64_t gk_nyc_ = 0, gk_nxc_ = 0, gk_nmv_ = 0, gk_nmtv_ = 0;  // counts of the captured cycle

  void alloc();
  void ensure_window(int w);
  void random_vec(double *d);
  double norm_dev(Prec p, const void *x, int64_t len, double *slot);
  void gs_basis_vec(double *v);           // 2 CGS passes vs basis (FP64)
  void gs_basis_block(double *W, int q);  // same on q columns
  void cgs_pass(void *w, Prec pw, int64_t len, const std::vector<Group> &groups,
                double *nrm_out);
  double estimate_fro();
  void push_basis(const double *v);
  // candidate window processing
  struct ColOut {
    bool valid = false, hit = false;  // hit: some pass stopped at cg_maxit
    double nv = 0, res_raw = 0, res = 0;
    int passes = 0, hits = 0;         // CG solves, and those that hit cg_maxit
  };
  void polish_batch(const std::vector<int> &cols, std::vector<ColOut> &out);
  bool fallback_one();
  void refine_basis();
  void random_right();
  void *Ycol(int64_t j) const;
  void *Xcol(int64_t j) const;
  // restart-loop state
  int64_t ncv_ = 0, restart_ = 0, restarts_since_accept_ = 0;
  std::vector<double> alpha_, beta_, bh_;
  std::chrono::steady_clock::time_point tcyc_;
  // Golub-Kahan storage: owned (VB_, X_, wx_, wy_) or the caller's
  bool ext_ = false;
  void *vbp_ = nullptr, *xp_ = nullptr, *wxp_ = nullptr, *wyp_ = nullptr;
  void cg_solve(int q, const void *rhs, void *w, Prec pc, const double *diag,
                int64_t maxit, double rel_tol, std::vector<CgState> &st, int64_t stall = 0);
  CgState cg_callback(const double *rhs, double *w, int64_t maxit, double rel_tol);
  // the projector CGs of the polish passes on q columns (rhs, w: m x q)
  void polish_cg(int q, const double *rhs, double *w, Prec pc, const double *diag,
                 std::vector<CgState> &st);
  void upload_mask(const std::vector<int> &mask);
  DBuf mask_;
  DBuf rin_, rout_;  // row-major copies of block_product
  std::vector<int> slot_of_;
};

}  // namespace kh
