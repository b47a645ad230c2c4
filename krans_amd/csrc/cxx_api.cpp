// C++ entry points of the KraNS plugin API on the HIP backend:
// krans::solvers::find_nullspace_cyclic_lanczos[_ex] and
// find_nullspace_lobpcg[_ex] (reference: CyclicLanczosSolver.hpp:170-178,
// LOBPCGSolver.hpp:142-149). They accept HipMatrix (CSR or matrix-free) +
// HipVectorFactory operands, run the device-resident drivers and return
// HipVector results owning HBM; any other backend is a backend mismatch
// (std::invalid_argument), as in the reference's adapters.
#include <cstring>

#include "../../include/krans/Solvers/CyclicLanczosSolver.hpp"
#include "../../include/krans/Solvers/LOBPCGSolver.hpp"
#include "../../include/krans/adapters/hip/HipAdapters.hpp"
#include "capi_internal.hpp"

namespace krans::solvers {
namespace {

struct HipOperands {
  const hip::HipMatrix *A;
  const hip::HipVectorFactory *F;
};

HipOperands require_hip(const MatrixInterface &A, const VectorFactory &vf, const char *who) {
  const auto *H = dynamic_cast<const hip::HipMatrix *>(&A);
  const auto *F = dynamic_cast<const hip::HipVectorFactory *>(&vf);
  if (!H || !F)
    throw std::invalid_argument(std::string(who) +
                                ": backend mismatch (a HIP matrix and a "
                                "HipVectorFactory required)");
  if (H->context() != F->context())
    throw std::invalid_argument(std::string(who) + ": matrix and factory on different contexts");
  return {H, F};
}

std::vector<std::unique_ptr<VectorInterface>> wrap(kh_context *ctx, const kh::DeviceBasis &B) {
  std::vector<std::unique_ptr<VectorInterface>> out;
  auto &c = *ctx->ctx;
  for (int64_t j = 0; j < B.count; ++j) {
    void *d = nullptr;
    hip::check(kh_dev_alloc(ctx, (std::size_t)B.len * 8, &d));
    KH_CHECK(hipMemcpyAsync(d, B.col(j), (std::size_t)B.len * 8, hipMemcpyDeviceToDevice,
                            c.stream));
    out.push_back(std::make_unique<hip::HipVector>(ctx, (std::size_t)B.len,
                                                   static_cast<double *>(d),
                                                   hip::HipVector::Adopt{}));
  }
  c.sync();
  return out;
}

kh::Prec prec(const hip::HipVectorFactory *F) {
  return F->precision() == hip::Precision::Mixed ? kh::Prec::F32 : kh::Prec::F64;
}

}  // namespace

CyclicLanczosResult find_nullspace_cyclic_lanczos_ex(const MatrixInterface &A,
                                                     const VectorFactory &vf,
                                                     const CyclicLanczosOptions &opt) {
  auto ops = require_hip(A, vf, "find_nullspace_cyclic_lanczos");
  kh_context *ctx = ops.A->context();
  kh_matrix *M = ops.A->handle();
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
  P.krylov = P.polish = prec(ops.F);
  if (opt.precond_AAt) {
    if (!opt.precond_AAt.target<hip::JacobiAAtPreconditioner>())
      throw std::invalid_argument(
          "find_nullspace_cyclic_lanczos: on the HIP backend precond_AAt must be nullptr or "
          "krans::hip::JacobiAAtPreconditioner");
    P.jacobi_AAt = true;
  }
  std::vector<double> y0;
  if (!opt.initial_right_space.empty() && opt.initial_right_space.front()) {
    const auto *hv = dynamic_cast<const hip::HipVector *>(opt.initial_right_space.front().get());
    if (!hv)
      throw std::invalid_argument("initial_right_space: backend mismatch");
    if (hv->size() != ops.A->cols())
      throw std::invalid_argument("initial_right_space: size mismatch");
    y0 = hv->to_host();
    P.y0_host = y0.data();
  }
  auto &c = *ctx->ctx;
  const auto op = M->view(false);
  kh::CyclicStats S;
  kh::DeviceBasis R = kh::cyclic_right(c, *op, ops.F->rng()->s, P, S);
  CyclicLanczosResult out;
  out.right_null = wrap(ctx, R);
  out.restarts = (std::size_t)S.restarts;
  out.krylov_steps_total = (std::size_t)S.krylov_steps_total;
  out.frob_norm_est = S.frob_norm_est;
  out.converged_to_target = S.converged_to_target;
  out.cg_solves = (std::size_t)S.cg_solves;
  out.cg_maxit_hits = (std::size_t)S.cg_maxit_hits;
  if (opt.return_left_space && M->rows() > 0 && M->cols() > 0) {
    kh::CyclicParams L = P;
    L.y0_host = nullptr;
    L.jacobi_AAt = false;
    L.nullity_known = -1;
    if (P.rank_known >= 0) {
      const int left_nullity = (int)M->rows() - P.rank_known;
      L.nullity_known = std::max(0, std::min(left_nullity, (int)P.k));
    }
    const auto opt_t = M->view(true);
    kh::CyclicStats SL;
    kh::DeviceBasis Lb = kh::cyclic_right(c, *opt_t, ops.F->rng()->s, L, SL);
    out.left_null = wrap(ctx, Lb);
    out.cg_solves += (std::size_t)SL.cg_solves;
    out.cg_maxit_hits += (std::size_t)SL.cg_maxit_hits;
  }
  return out;
}

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_cyclic_lanczos(const MatrixInterface &A, const VectorFactory &vf,
                              const CyclicLanczosOptions &opt) {
  auto R = find_nullspace_cyclic_lanczos_ex(A, vf, opt);
  return std::move(R.right_null);
}

LOBPCGResult find_nullspace_lobpcg_ex(const MatrixInterface &A, const VectorFactory &vf,
                                      const LOBPCGOptions &opt) {
  auto ops = require_hip(A, vf, "find_nullspace_lobpcg");
  if (opt.M_inv)
    throw std::invalid_argument("find_nullspace_lobpcg: M_inv must be nullptr on the HIP backend");
  kh_context *ctx = ops.A->context();
  kh_matrix *M = ops.A->handle();
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
  P.block = prec(ops.F);
  std::vector<double> ib;
  for (const auto &v : opt.initial_block) {
    const auto *hv = dynamic_cast<const hip::HipVector *>(v.get());
    if (!hv)
      throw std::invalid_argument("initial_block: backend mismatch");
    auto h = hv->to_host();
    ib.insert(ib.end(), h.begin(), h.end());
  }
  P.initial_block_host = ib.empty() ? nullptr : ib.data();
  P.n_initial = (int64_t)opt.initial_block.size();
  auto &c = *ctx->ctx;
  const auto op = M->view(false);
  kh::LobpcgStats S;
  kh::DeviceBasis R = kh::lobpcg(c, *op, ops.F->rng()->s, P, S);
  LOBPCGResult out;
  out.right_null = wrap(ctx, R);
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
}

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_lobpcg(const MatrixInterface &A, const VectorFactory &vf, const LOBPCGOptions &opt) {
  auto R = find_nullspace_lobpcg_ex(A, vf, opt);
  return std::move(R.right_null);
}

}  // namespace krans::solvers
