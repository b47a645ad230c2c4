// The reference's solver test scenarios on the HIP backend, through the
// unchanged KraNS C++ API: tests/test_cyclic_lanczos.cpp (tests 1-6) and
// tests/test_lobpcg.cpp (tests 1-5) with their options, operands and CHECK
// thresholds. What changes is the backend: HipCsrMatrixAdapter in place of
// EigenSparseMatrixAdapter, HipVectorFactory in place of EigenVectorFactory.
// As in the reference, every operand is wrapped in a counting decorator (a
// plain MatrixInterface, so the HIP solvers apply it through HipVector
// views), and the preconditioner callbacks (identity and steepest-descent
// precond_AAt, steepest-descent M_inv) run on HipVectors.
//
// usage: reference_scenarios_hip <data dir with mat1.ija, mat2.ija> [fp64|mixed]
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "krans/Solvers/CyclicLanczosSolver.hpp"
#include "krans/Solvers/LOBPCGSolver.hpp"
#include "krans/adapters/hip/HipAdapters.hpp"

using namespace krans;
using krans::hip::HipContext;
using krans::hip::HipVectorFactory;
using krans::hip::Precision;
using Clock = std::chrono::steady_clock;

#define CHECK(cond)                                                                   \
  do {                                                                                \
    if (!(cond)) {                                                                    \
      std::fprintf(stderr, "CHECK failed: %s (%s:%d)\n", #cond, __FILE__, __LINE__); \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

namespace {

// Delegates to another operand and counts its products.
class Counting final : public MatrixInterface {
public:
  explicit Counting(const MatrixInterface &inner) : in_(inner) {}
  std::size_t rows() const override { return in_.rows(); }
  std::size_t cols() const override { return in_.cols(); }
  void multiply(const VectorInterface &x, VectorInterface &y) const override {
    ++n_;
    in_.multiply(x, y);
  }
  void multiply_transpose(const VectorInterface &x, VectorInterface &y) const override {
    ++nt_;
    in_.multiply_transpose(x, y);
  }
  std::size_t products() const { return n_ + nt_; }

private:
  const MatrixInterface &in_;
  mutable std::size_t n_ = 0, nt_ = 0;
};

using Basis = std::vector<std::unique_ptr<VectorInterface>>;

struct Small {  // a CSR held on the host, uploaded per test
  std::size_t m, n;
  std::vector<int32_t> rp, ci;
  std::vector<double> va;
};

Small two_by_three(double s) {  // [[1,0,1],[0,1,1]] * s (test_cyclic_lanczos.cpp:146-150)
  return {2, 3, {0, 2, 4}, {0, 2, 1, 2}, {s, s, s, s}};
}
Small ones_2x2() { return {2, 2, {0, 2, 4}, {0, 1, 0, 1}, {1, 1, 1, 1}}; }
Small identity3() { return {3, 3, {0, 1, 2, 3}, {0, 1, 2}, {1, 1, 1}}; }

std::unique_ptr<hip::HipCsrMatrixAdapter> upload(kh_context *ctx, const Small &s) {
  return std::make_unique<hip::HipCsrMatrixAdapter>(ctx, s.m, s.n, s.rp.data(), s.ci.data(),
                                                    s.va.data());
}

void check_unit_norms(const Basis &B, double tol) {
  for (const auto &q : B)
    CHECK(std::fabs(q->norm() - 1.0) <= tol);
}

double max_pairwise_abs_dot(const Basis &B) {  // GramSchmidt.cpp:84-91
  double m = 0.0;
  for (std::size_t i = 0; i < B.size(); ++i)
    for (std::size_t j = i + 1; j < B.size(); ++j)
      m = std::max(m, std::fabs(B[i]->dot(*B[j])));
  return m;
}

double worst_residual(const MatrixInterface &A, const Basis &B, const VectorFactory &F) {
  auto t = F.create_zero_of_size(A.rows());
  double w = 0.0;
  for (const auto &q : B) {
    A.multiply(*q, *t);
    w = std::max(w, t->norm());
  }
  return w;
}

// ||A Proj_N v|| for a random unit v
double projected_residual(const MatrixInterface &A, const Basis &B, const VectorFactory &F,
                          bool normalise) {
  auto v = F.create_random_of_size(A.cols());
  if (normalise)
    v->scale(1.0 / v->norm());
  auto p = F.create_zero_of_size(A.cols());
  for (const auto &q : B)
    p->add_scaled(*q, v->dot(*q));
  auto t = F.create_zero_of_size(A.rows());
  A.multiply(*p, *t);
  return t->norm();
}

double seconds(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

// ---- tests/test_cyclic_lanczos.cpp ---------------------------------------

void cyclic_small(kh_context *ctx, Precision prec, double scale) {
  auto M = upload(ctx, two_by_three(scale));
  Counting C(*M);
  HipVectorFactory F(ctx, 3, 1, prec);
  solvers::CyclicLanczosOptions opt;
  opt.k = 1;
  opt.ncv = 20;
  opt.max_restarts = 50;
  opt.tol_abs = 1e-12;
  opt.tol_rel = 1e-10;
  opt.nullity_known = 1;
  opt.rank_known = 2;
  opt.cg_maxit = 40;
  opt.cg_rel_tol = 1e-10;
  auto R = solvers::find_nullspace_cyclic_lanczos_ex(C, F, opt);
  CHECK(R.right_null.size() == 1);
  check_unit_norms(R.right_null, 1e-12);
  const double res = worst_residual(C, R.right_null, F);
  if (scale == 1.0) {
    CHECK(res <= 1e-10);  // test 1
  } else {
    CHECK(res / std::max(1.0, R.frob_norm_est) <= 1e-8);  // test 6
  }
  std::printf("cyclic 2x3 (x%g): ||A v|| = %.3e, %zu products\n", scale, res, C.products());
}

void cyclic_mat1(kh_context *ctx, Precision prec, const std::string &dir) {
  auto M = hip::load_matrix(ctx, dir + "/mat1.ija");
  Counting C(*M);
  HipVectorFactory F(ctx, M->cols(), 1, prec);
  solvers::CyclicLanczosOptions opt;
  opt.k = 5;
  opt.ncv = std::max<std::size_t>(opt.k + 40, 60);
  opt.max_restarts = 200;
  opt.tol_abs = 1e-12;
  opt.tol_rel = 1e-10;
  opt.nullity_known = 5;
  opt.rank_known = 100;
  opt.cg_maxit = 60;
  opt.cg_rel_tol = 1e-8;
  opt.initial_right_space.push_back(F.create_random_of_size(M->cols()));
  opt.initial_right_space.push_back(F.create_random_of_size(M->cols()));
  opt.precond_AAt = [&](const VectorInterface &r, VectorInterface &y) { y.copy_from(r); };
  const auto t0 = Clock::now();
  auto B = solvers::find_nullspace_cyclic_lanczos(C, F, opt);
  const double sec = seconds(t0);
  CHECK(B.size() == 5);
  check_unit_norms(B, 1e-10);
  const double res = worst_residual(C, B, F);
  CHECK(res <= 1e-8);
  CHECK(max_pairwise_abs_dot(B) <= 1e-8);
  CHECK(projected_residual(C, B, F, true) <= 1e-8);
  std::printf("cyclic mat1: 5 vectors, worst ||A v|| = %.3e, %.3f s, %zu products\n", res, sec,
              C.products());
}

void cyclic_identity(kh_context *ctx, Precision prec) {
  auto M = upload(ctx, identity3());
  Counting C(*M);
  HipVectorFactory F(ctx, 3, 1, prec);
  solvers::CyclicLanczosOptions opt;
  opt.k = 1;
  opt.ncv = 20;
  opt.max_restarts = 30;
  opt.nullity_known = 0;
  opt.rank_known = 3;
  CHECK(solvers::find_nullspace_cyclic_lanczos(C, F, opt).empty());
  std::printf("cyclic identity: 0 vectors\n");
}

void cyclic_mat2(kh_context *ctx, Precision prec, const std::string &dir) {
  auto M = hip::load_matrix(ctx, dir + "/mat2.ija");
  Counting C2(*M);
  HipVectorFactory F2(ctx, M->cols(), 1, prec);
  const int rank_known = 5425, nullity_known = 15;
  solvers::CyclicLanczosOptions opt;
  opt.k = nullity_known;
  opt.ncv = std::max<std::size_t>(nullity_known + 80, 128);
  opt.max_restarts = 400;
  opt.projector_passes = 2;
  opt.cg_maxit = 200;
  opt.cg_rel_tol = 1e-10;
  opt.prescreen_factor = 4.0;
  opt.stagnation_restarts = 0;
  opt.tol_abs = 1e-12;
  opt.tol_rel = 1e-10;
  opt.nullity_known = nullity_known;
  opt.rank_known = rank_known;
  opt.initial_right_space.push_back(F2.create_random_of_size(M->cols()));
  opt.initial_right_space.push_back(F2.create_random_of_size(M->cols()));
  // a few steepest-descent steps on (A A^T) y = r, matrix-free through C2
  const int k_sd = 8;
  opt.precond_AAt = [&](const VectorInterface &r, VectorInterface &y) {
    auto p = F2.create_zero_of_size(M->rows());
    auto Ap = F2.create_zero_of_size(M->rows());
    auto tn = F2.create_zero_of_size(M->cols());
    auto res = F2.create_zero_of_size(M->rows());
    y.set_zero();
    res->copy_from(r);
    p->copy_from(*res);
    for (int it = 0; it < k_sd; ++it) {
      C2.multiply_transpose(*p, *tn);
      C2.multiply(*tn, *Ap);
      const double pAp = p->dot(*Ap);
      if (std::fabs(pAp) < 1e-30)
        break;
      const double alpha = p->dot(*p) / pAp;
      y.add_scaled(*p, alpha);
      res->add_scaled(*Ap, -alpha);
      if (res->dot(*res) <= 1e-30)
        break;
      p->copy_from(*res);
    }
  };
  const auto t0 = Clock::now();
  auto R = solvers::find_nullspace_cyclic_lanczos_ex(C2, F2, opt);
  const double sec = seconds(t0);
  CHECK(R.right_null.size() == (std::size_t)nullity_known);
  check_unit_norms(R.right_null, 1e-12);
  const double thr = std::max(opt.tol_abs, opt.tol_rel * std::max(1.0, R.frob_norm_est));
  const double res = worst_residual(C2, R.right_null, F2);
  CHECK(res <= thr);
  CHECK(max_pairwise_abs_dot(R.right_null) <= 1e-8);
  CHECK(projected_residual(C2, R.right_null, F2, true) <=
        std::sqrt((double)R.right_null.size()) * thr);
  std::printf("cyclic mat2: 15 vectors, worst ||A v|| = %.3e (thr %.3e), restarts %zu, "
              "cg_solves %zu, %.2f s, %zu products\n",
              res, thr, R.restarts, R.cg_solves, sec, C2.products());
}

void cyclic_rank1_both(kh_context *ctx, Precision prec) {
  auto M = upload(ctx, ones_2x2());
  Counting C(*M);
  HipVectorFactory F(ctx, 2, 1, prec);
  solvers::CyclicLanczosOptions opt;
  opt.k = 1;
  opt.ncv = 20;
  opt.max_restarts = 50;
  opt.nullity_known = 1;
  opt.rank_known = 1;
  opt.return_left_space = true;
  auto R = solvers::find_nullspace_cyclic_lanczos_ex(C, F, opt);
  CHECK(R.right_null.size() == 1);
  check_unit_norms(R.right_null, 1e-12);
  CHECK(worst_residual(C, R.right_null, F) <= 1e-10);
  CHECK(R.left_null.size() == 1);
  check_unit_norms(R.left_null, 1e-12);
  auto t = F.create_zero_of_size(2);
  C.multiply_transpose(*R.left_null[0], *t);
  CHECK(t->norm() <= 1e-10);
  std::printf("cyclic 2x2 rank 1: right and left vectors\n");
}

// ---- tests/test_lobpcg.cpp -----------------------------------------------

solvers::LOBPCGOptions lobpcg_small_opts(std::size_t bs, std::size_t iters) {
  solvers::LOBPCGOptions opt;
  opt.nev = 1;
  opt.block_size = bs;
  opt.max_iters = iters;
  opt.tol_abs = 1e-12;
  opt.tol_rel = 1e-10;
  opt.soft_locking = true;
  opt.restart = 0.9;
  opt.gs_reorth = true;
  opt.nullity_known = 1;
  opt.rank_known = 2;
  return opt;
}

void lobpcg_small(kh_context *ctx, Precision prec, double scale) {
  auto M = upload(ctx, two_by_three(scale));
  Counting C(*M);
  HipVectorFactory F(ctx, 3, 1, prec);
  auto opt = scale == 1.0 ? lobpcg_small_opts(4, 200) : lobpcg_small_opts(6, 300);
  auto R = solvers::find_nullspace_lobpcg_ex(C, F, opt);
  CHECK(R.right_null.size() == 1);
  check_unit_norms(R.right_null, 1e-12);
  const double res = worst_residual(C, R.right_null, F);
  if (scale == 1.0)
    CHECK(res <= 1e-10);  // test 1
  else
    CHECK(res / std::max(1.0, R.frob_norm_est) <= 1e-8);  // test 5
  std::printf("lobpcg 2x3 (x%g): ||A v|| = %.3e\n", scale, res);
}

// k-step steepest descent on N = A^T A through the counting operand
template <class Op>
solvers::NormalPreconditioner sd_normal(const Op &C, const HipVectorFactory &F, int k_sd) {
  return [&C, &F, k_sd](const VectorInterface &r, VectorInterface &y) {
    auto p = F.create_zero_of_size(C.cols());
    auto Np = F.create_zero_of_size(C.cols());
    auto tm = F.create_zero_of_size(C.rows());
    y.set_zero();
    p->copy_from(r);
    for (int it = 0; it < k_sd; ++it) {
      C.multiply(*p, *tm);
      C.multiply_transpose(*tm, *Np);
      const double pNp = p->dot(*Np);
      if (std::fabs(pNp) < 1e-30)
        break;
      const double alpha = p->dot(*p) / pNp;
      y.add_scaled(*p, alpha);
      p->add_scaled(*Np, -alpha);  // r - alpha N p, the next direction
      if (p->dot(*p) <= 1e-30)
        break;
    }
  };
}

void lobpcg_mat(kh_context *ctx, Precision prec, const std::string &dir, int which) {
  auto M = hip::load_matrix(ctx, dir + (which == 1 ? "/mat1.ija" : "/mat2.ija"));
  Counting C(*M);
  HipVectorFactory F(ctx, M->cols(), 1, prec);
  solvers::LOBPCGOptions opt;
  const std::size_t nev = which == 1 ? 5 : 15;
  opt.nev = nev;
  opt.block_size = which == 1 ? 12 : std::max<std::size_t>(nev, 24);
  opt.max_iters = which == 1 ? 500 : 1200;
  opt.tol_abs = 1e-12;
  opt.tol_rel = 1e-10;
  opt.soft_locking = true;
  opt.restart = which == 1 ? 0.9 : 0.85;
  opt.gs_reorth = true;
  opt.nullity_known = (int)nev;
  opt.rank_known = which == 1 ? 100 : 5425;
  opt.initial_block.push_back(F.create_random_of_size(M->cols()));
  opt.initial_block.push_back(F.create_random_of_size(M->cols()));
  opt.M_inv = sd_normal(C, F, which == 1 ? 6 : 10);
  const auto t0 = Clock::now();
  auto R = solvers::find_nullspace_lobpcg_ex(C, F, opt);
  const double sec = seconds(t0);
  std::printf("lobpcg mat%d: %zu of %zu vectors, iters %zu, thr %.3e, best unaccepted %.3e, "
              "%.2f s\n", which, R.right_null.size(), nev, R.iters, R.accept_threshold,
              R.best_unaccepted_residual, sec);
  CHECK(R.right_null.size() == nev);
  check_unit_norms(R.right_null, which == 1 ? 1e-10 : 1e-12);
  const double res = worst_residual(C, R.right_null, F);
  std::printf("lobpcg mat%d: worst ||A v|| = %.3e\n", which, res);
  std::fflush(stdout);
  CHECK(res <= (which == 1 ? 1e-8 : 1e-9));
  CHECK(max_pairwise_abs_dot(R.right_null) <= 1e-8);
  CHECK(projected_residual(C, R.right_null, F, false) <= 1e-8);
  std::printf("lobpcg mat%d: %zu vectors, worst ||A v|| = %.3e, iters %zu, %.2f s, "
              "%zu products\n", which, nev, res, R.iters, sec, C.products());
}

void lobpcg_identity(kh_context *ctx, Precision prec) {
  auto M = upload(ctx, identity3());
  Counting C(*M);
  HipVectorFactory F(ctx, 3, 1, prec);
  auto opt = lobpcg_small_opts(4, 100);
  opt.nullity_known = 0;
  opt.rank_known = 3;
  CHECK(solvers::find_nullspace_lobpcg(C, F, opt).empty());
  std::printf("lobpcg identity: 0 vectors\n");
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <data dir> [fp64|mixed] [test ...]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  const Precision prec =
      (argc > 2 && std::string(argv[2]) == "mixed") ? Precision::Mixed : Precision::FP64;
  std::vector<std::string> only(argv + std::min(argc, 3), argv + argc);
  auto want = [&](const char *name) {
    return only.empty() || std::find(only.begin(), only.end(), name) != only.end();
  };
  HipContext ctx(0);
  kh_context *c = ctx.get();
  if (want("cyclic_small")) cyclic_small(c, prec, 1.0);
  if (want("cyclic_mat1")) cyclic_mat1(c, prec, dir);
  if (want("cyclic_identity")) cyclic_identity(c, prec);
  if (want("cyclic_mat2")) cyclic_mat2(c, prec, dir);
  if (want("cyclic_rank1")) cyclic_rank1_both(c, prec);
  if (want("cyclic_scale")) cyclic_small(c, prec, 1e6);
  if (want("lobpcg_small")) lobpcg_small(c, prec, 1.0);
  if (want("lobpcg_mat1")) lobpcg_mat(c, prec, dir, 1);
  if (want("lobpcg_identity")) lobpcg_identity(c, prec);
  if (want("lobpcg_mat2")) lobpcg_mat(c, prec, dir, 2);
  if (want("lobpcg_scale")) lobpcg_small(c, prec, 1e6);
  std::printf("ALL OK\n");
  return 0;
}
