// C++ entry points of the KraNS plugin API on the HIP backend:
// krans::hip::find_nullspace_cyclic_lanczos[_ex] and find_nullspace_lobpcg[_ex]
// (the reference's krans::solvers entry points, CyclicLanczosSolver.hpp:170-178
// and LOBPCGSolver.hpp:142-149, under names of their own so that this library
// links beside the reference's).
//
// The factory must be a HipVectorFactory: its context (device stream), seeded
// stream and precision drive the solve. The matrix may be any
// MatrixInterface: a HipMatrix on the factory's context runs directly on the
// device (CSR with its CSC twin, stencil, device callback); any other operand
// (a counting decorator, a user operator) is applied column by column
// through HipVector views of the solver's device columns (InterfaceOp).
// Preconditioner callbacks receive HipVector views as well. Results are
// HipVectors owning HBM. Errors keep the reference's exception types;
// device OOM is std::bad_alloc.
#include <cstring>

#include "../../include/krans/adapters/hip/HipSolvers.hpp"
#include "capi_internal.hpp"

namespace krans::hip {
namespace {

// A MatrixInterface applied to device columns: FP64 columns are passed as
// HipVector views; FP32 columns (mixed precision) through FP64 copies. The
// operator's multiply is synchronous from its point of view (the
// reference's single-threaded model): the solver's stream is drained before
// each call, and the operator's own work is either finished on return or
// enqueued on the factory context's stream, which the solver keeps using.
class InterfaceOp final : public kh::DeviceOp {
public:
  InterfaceOp(const MatrixInterface &A, kh_context *ctx) : A_(A), ctx_(ctx) {}
  int64_t rows() const override { return (int64_t)A_.rows(); }
  int64_t cols() const override { return (int64_t)A_.cols(); }
  void mv(kh::Context &c, kh::Prec p, const void *x, void *y, const void *z,
          const double *beta) const override {
    apply(c, p, false, x, y);
    if (z)
      kh::sub_beta(c, p, rows(), beta, z, y);
  }
  void mtv(kh::Context &c, kh::Prec p, const void *x, void *y, const void *z,
           const double *beta) const override {
    apply(c, p, true, x, y);
    if (z)
      kh::sub_beta(c, p, cols(), beta, z, y);
  }
  void mm(kh::Context &c, kh::Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *act) const override {
    block(c, p, false, s, x, ldx, y, ldy, act);
  }
  void mtm(kh::Context &c, kh::Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
           const int *act) const override {
    block(c, p, true, s, x, ldx, y, ldy, act);
  }

private:
  void apply(kh::Context &c, kh::Prec p, bool t, const void *x, void *y) const {
    const int64_t nin = t ? rows() : cols(), nout = t ? cols() : rows();
    const double *xd = static_cast<const double *>(x);
    double *yd = static_cast<double *>(y);
    if (p == kh::Prec::F32) {
      if (tin_.bytes < (std::size_t)nin * 8)
        tin_ = kh::DBuf((std::size_t)nin * 8);
      if (tout_.bytes < (std::size_t)nout * 8)
        tout_ = kh::DBuf((std::size_t)nout * 8);
      kh::convert(c, kh::Prec::F32, x, kh::Prec::F64, tin_.p, nin);
      xd = tin_.as<double>();
      yd = tout_.as<double>();
    }
    c.sync();
    const HipVector vx(ctx_, (std::size_t)nin, const_cast<double *>(xd), HipVector::View{});
    HipVector vy(ctx_, (std::size_t)nout, yd, HipVector::View{});
    if (t)
      A_.multiply_transpose(vx, vy);
    else
      A_.multiply(vx, vy);
    if (p == kh::Prec::F32)
      kh::convert(c, kh::Prec::F64, yd, kh::Prec::F32, y, nout);
  }
  void block(kh::Context &c, kh::Prec p, bool t, int s, const void *x, int64_t ldx, void *y,
             int64_t ldy, const int *act) const {
    if (s <= 0)
      return;
    std::vector<int> on((std::size_t)s, 1);
    if (act) {
      KH_CHECK(hipMemcpyAsync(on.data(), act, sizeof(int) * (std::size_t)s,
                              hipMemcpyDeviceToHost, c.stream));
      KH_CHECK(hipStreamSynchronize(c.stream));
    }
    const std::size_t e = p == kh::Prec::F64 ? 8 : 4;
    for (int j = 0; j < s; ++j)
      if (on[(std::size_t)j])
        apply(c, p, t, static_cast<const char *>(x) + (std::size_t)j * ldx * e,
              static_cast<char *>(y) + (std::size_t)j * ldy * e);
  }
  const MatrixInterface &A_;
  kh_context *ctx_;
  mutable kh::DBuf tin_, tout_;
};

// precond_AAt / M_inv called on HipVector views of the CG's device vectors.
template <class Fn> class CallbackPrecond final : public kh::DevicePreconditioner {
public:
  CallbackPrecond(const Fn &f, kh_context *ctx) : f_(f), ctx_(ctx) {}
  void apply(kh::Context &, const double *r, double *z, int64_t len) const override {
    const HipVector vr(ctx_, (std::size_t)len, const_cast<double *>(r), HipVector::View{});
    HipVector vz(ctx_, (std::size_t)len, z, HipVector::View{});
    f_(vr, vz);
  }

private:
  const Fn &f_;
  kh_context *ctx_;
};

// The solver operand of (A, factory): the device operator and, for HIP
// matrices, the handle.
struct Operand {
  const HipVectorFactory *F = nullptr;
  kh_context *ctx = nullptr;
  const HipMatrix *H = nullptr;  // A itself when it is a HipMatrix on the factory's context
  std::unique_ptr<InterfaceOp> iface;
  kh_matrix_s::View view;
  std::unique_ptr<kh::TransposedOpView> tview;
  const kh::DeviceOp &op() const { return *view; }
  const kh::DeviceOp &transposed() {
    if (H)
      return *(view = H->handle()->view(true));
    tview = std::make_unique<kh::TransposedOpView>(*iface);
    return *tview;
  }
};

Operand operand(const MatrixInterface &A, const VectorFactory &vf, const char *who) {
  Operand o;
  o.F = dynamic_cast<const HipVectorFactory *>(&vf);
  if (!o.F)
    throw std::invalid_argument(std::string(who) +
                                ": backend mismatch (a HipVectorFactory required)");
  o.ctx = o.F->context();
  const auto *H = dynamic_cast<const HipMatrix *>(&A);
  if (H && H->context() == o.ctx) {
    o.H = H;
    o.view = H->handle()->view(false);
  } else {
    o.iface = std::make_unique<InterfaceOp>(A, o.ctx);
    o.view.p = o.iface.get();
  }
  return o;
}

std::vector<std::unique_ptr<VectorInterface>> wrap(kh_context *ctx, const kh::DeviceBasis &B) {
  std::vector<std::unique_ptr<VectorInterface>> out;
  auto &c = *ctx->ctx;
  for (int64_t j = 0; j < B.count; ++j) {
    void *d = nullptr;
    check(kh_dev_alloc(ctx, (std::size_t)B.len * 8, &d));
    KH_CHECK(hipMemcpyAsync(d, B.col(j), (std::size_t)B.len * 8, hipMemcpyDeviceToDevice,
                            c.stream));
    out.push_back(std::make_unique<HipVector>(ctx, (std::size_t)B.len,
                                              static_cast<double *>(d), HipVector::Adopt{}));
  }
  c.sync();
  return out;
}

kh::Prec prec(const HipVectorFactory *F) {
  return F->precision() == Precision::Mixed ? kh::Prec::F32 : kh::Prec::F64;
}

std::vector<double> host_copy(const VectorInterface &v, std::size_t n, const char *what) {
  const auto *hv = dynamic_cast<const HipVector *>(&v);
  if (!hv)
    throw std::invalid_argument(std::string(what) + ": backend mismatch");
  if (hv->size() != n)
    throw std::invalid_argument(std::string(what) + ": size mismatch");
  return hv->to_host();
}

// Device OOM is std::bad_alloc at this boundary; other device failures stay
// std::runtime_error (kh::DeviceError derives from it).
template <class F> auto translate(F &&fn) -> decltype(fn()) {
  try {
    return fn();
  } catch (const kh::DeviceOOM &) {
    throw std::bad_alloc();
  }
}

}  // namespace

solvers::CyclicLanczosResult find_nullspace_cyclic_lanczos_ex(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::CyclicLanczosOptions &opt) {
  return translate([&] {
    Operand o = operand(A, vf, "find_nullspace_cyclic_lanczos");
    kh::CyclicParams P;
    P.k = (int64_t)opt.k;
    P.ncv = (int64_t)opt.ncv;
    P.max_restarts = (int64_t)opt.max_restarts;
    P.tol_abs = opt.tol_abs;
    P.tol_rel = opt.tol_rel;
    P.reorth = (int)opt.reorth;
    P.use_harmonic = opt.use_harmonic;
    P.target_tau = opt.target_tau;
    P.nullity_known = opt.nullity_known;
    P.rank_known = opt.rank_known;
    P.cg_maxit = (int64_t)opt.cg_maxit;
    P.cg_rel_tol = opt.cg_rel_tol;
    P.cg_breakdown_tol = opt.cg_breakdown_tol;
    P.projector_passes = opt.projector_passes;
    P.grad_max_iters = opt.grad_max_iters;
    P.frob_samples = opt.frob_samples;
    P.prescreen_factor = opt.prescreen_factor;
    P.stagnation_restarts = (int64_t)opt.stagnation_restarts;
    P.return_left_space = opt.return_left_space;
    P.verbose = opt.verbose;
    P.krylov = P.polish = prec(o.F);
    // precond_AAt: the Jacobi preconditioner of the solve's own CSR matrix
    // runs inside the CG kernels; every other callback is called on views
    using Fn = solvers::AAtPreconditioner;
    std::unique_ptr<CallbackPrecond<Fn>> cb;
    if (opt.precond_AAt) {
      const auto *J = opt.precond_AAt.target<JacobiAAtPreconditioner>();
      if (J && o.H && J->A && J->A->handle() == o.H->handle() && o.op().diag_AAt())
        P.jacobi_AAt = true;
      else
        P.precond = (cb = std::make_unique<CallbackPrecond<Fn>>(opt.precond_AAt, o.ctx)).get();
    }
    std::vector<double> y0;
    if (!opt.initial_right_space.empty() && opt.initial_right_space.front()) {
      y0 = host_copy(*opt.initial_right_space.front(), A.cols(), "initial_right_space");
      P.y0_host = y0.data();
    }
    auto &c = *o.ctx->ctx;
    kh::CyclicStats S;
    kh::DeviceBasis R = kh::cyclic_right(c, o.op(), o.F->rng()->s, P, S);
    solvers::CyclicLanczosResult out;
    out.right_null = wrap(o.ctx, R);
    out.restarts = (std::size_t)S.restarts;
    out.krylov_steps_total = (std::size_t)S.krylov_steps_total;
    out.frob_norm_est = S.frob_norm_est;
    out.converged_to_target = S.converged_to_target;
    out.cg_solves = (std::size_t)S.cg_solves;
    out.cg_maxit_hits = (std::size_t)S.cg_maxit_hits;
    if (opt.return_left_space && A.rows() > 0 && A.cols() > 0) {
      // CyclicLanczosSolver.cpp:629-659: no preconditioner for A^T A
      kh::CyclicParams L = P;
      L.y0_host = nullptr;
      L.jacobi_AAt = false;
      L.precond = nullptr;
      L.nullity_known = -1;
      if (P.rank_known >= 0) {
        const int left_nullity = (int)A.rows() - P.rank_known;
        L.nullity_known = std::max(0, std::min(left_nullity, (int)P.k));
      }
      const kh::DeviceOp &opt_t = o.transposed();
      kh::CyclicStats SL;
      kh::DeviceBasis Lb = kh::cyclic_right(c, opt_t, o.F->rng()->s, L, SL);
      out.left_null = wrap(o.ctx, Lb);
      out.cg_solves += (std::size_t)SL.cg_solves;
      out.cg_maxit_hits += (std::size_t)SL.cg_maxit_hits;
    }
    return out;
  });
}

std::vector<std::unique_ptr<VectorInterface>> find_nullspace_cyclic_lanczos(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::CyclicLanczosOptions &opt) {
  auto R = krans::hip::find_nullspace_cyclic_lanczos_ex(A, vf, opt);
  return std::move(R.right_null);
}

solvers::LOBPCGResult find_nullspace_lobpcg_ex(const MatrixInterface &A, const VectorFactory &vf,
                                               const solvers::LOBPCGOptions &opt) {
  return translate([&] {
    Operand o = operand(A, vf, "find_nullspace_lobpcg");
    kh::LobpcgParams P;
    P.nev = (int64_t)opt.nev;
    P.block_size = (int64_t)opt.block_size;
    P.max_iters = (int64_t)opt.max_iters;
    P.tol_abs = opt.tol_abs;
    P.tol_rel = opt.tol_rel;
    P.soft_locking = opt.soft_locking;
    P.restart = opt.restart;
    P.gs_tol = opt.gs_tol;
    P.gs_reorth = opt.gs_reorth;
    P.frob_samples = opt.frob_samples;
    P.max_tiny_dim = (int64_t)opt.max_tiny_dim;
    P.nullity_known = opt.nullity_known;
    P.rank_known = opt.rank_known;
    P.verbose = opt.verbose;
    P.progress_every = (int64_t)opt.progress_every;
    P.block = prec(o.F);
    using Fn = solvers::NormalPreconditioner;
    std::unique_ptr<CallbackPrecond<Fn>> cb;
    if (opt.M_inv)
      P.M_inv = (cb = std::make_unique<CallbackPrecond<Fn>>(opt.M_inv, o.ctx)).get();
    std::vector<double> ib;
    for (const auto &v : opt.initial_block) {
      if (!v)
        throw std::invalid_argument("initial_block: null vector");
      auto h = host_copy(*v, A.cols(), "initial_block");
      ib.insert(ib.end(), h.begin(), h.end());
    }
    P.initial_block_host = ib.empty() ? nullptr : ib.data();
    P.n_initial = (int64_t)opt.initial_block.size();
    auto &c = *o.ctx->ctx;
    kh::LobpcgStats S;
    kh::DeviceBasis R = kh::lobpcg(c, o.op(), o.F->rng()->s, P, S);
    solvers::LOBPCGResult out;
    out.right_null = wrap(o.ctx, R);
    out.theta = S.theta;
    out.residuals_Av = S.residuals_Av;
    out.iters = (std::size_t)S.iters;
    out.locked = (std::size_t)S.locked;
    out.frob_norm_est = S.frob_norm_est;
    out.accept_threshold = S.accept_threshold;
    out.best_unaccepted_residual = S.best_unaccepted_residual;
    out.block_size = (std::size_t)S.block_size;
    out.converged_to_target = S.converged_to_target;
    return out;
  });
}

std::vector<std::unique_ptr<VectorInterface>> find_nullspace_lobpcg(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::LOBPCGOptions &opt) {
  auto R = krans::hip::find_nullspace_lobpcg_ex(A, vf, opt);
  return std::move(R.right_null);
}

}  // namespace krans::hip
