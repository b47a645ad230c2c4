// Oracle restatement of find_nullspace_cyclic_lanczos_ex
// (src/Solvers/CyclicLanczosSolver.cpp:17-668) and of the Gram-Schmidt and
// projection utilities it calls (src/Utils/GramSchmidt.cpp:25-91,
// src/Utils/Projection.cpp:8-17, include/krans/Utils/Projection.hpp:13-21).
// Control flow, thresholds, counters and the order of random draws follow the
// reference statement by statement; Eigen's JacobiSVD / SelfAdjointEigenSolver
// are replaced by dense.cpp.
#include "oracle.hpp"

#include <algorithm>
#include <iostream>
#include <limits>
#include <numeric>

namespace orc {

// ---------------- GramSchmidt.cpp ------------------------------------------
static double mgs_pass(Vec &v, const std::vector<Vec> &basis) { // :25-33
  for (const auto &q : basis) {
    const double c = dot(v, q);
    axpy(v, q, -c);
  }
  return nrm2(v);
}

double orthogonalize_against_basis(Vec &v, const std::vector<Vec> &basis,
                                   double tol, bool reorth) { // :35-53
  double n1 = mgs_pass(v, basis);
  if (n1 < tol)
    return 0.0;
  if (reorth) {
    double n2 = mgs_pass(v, basis);
    if (n2 < tol)
      return 0.0;
    n1 = n2;
  }
  scal(v, 1.0 / n1);
  return n1;
}

bool modified_gram_schmidt(std::vector<Vec> &vecs, double tol,
                           bool reorth) { // :55-82
  if (!vecs.empty()) {
    const std::size_t n = vecs.front().size();
    for (const auto &p : vecs)
      if (p.size() != n)
        throw std::invalid_argument("GramSchmidt: vector size mismatch");
  }
  for (std::size_t i = 0; i < vecs.size(); ++i) {
    for (std::size_t j = 0; j < i; ++j) {
      const double c = dot(vecs[i], vecs[j]);
      axpy(vecs[i], vecs[j], -c);
    }
    double n1 = nrm2(vecs[i]);
    if (n1 < tol)
      return false;
    if (reorth) {
      for (std::size_t j = 0; j < i; ++j) {
        const double c = dot(vecs[i], vecs[j]);
        axpy(vecs[i], vecs[j], -c);
      }
      double n2 = nrm2(vecs[i]);
      if (n2 < tol)
        return false;
      n1 = n2;
    }
    scal(vecs[i], 1.0 / n1);
  }
  return true;
}

double max_pairwise_abs_dot(const std::vector<Vec> &basis) { // :84-91
  double m = 0.0;
  for (std::size_t i = 0; i < basis.size(); ++i)
    for (std::size_t j = i + 1; j < basis.size(); ++j)
      m = std::max(m, std::fabs(dot(basis[i], basis[j])));
  return m;
}

namespace {

// Counting decorator, the CountingMatrixAdapter pattern of
// tests/test_cyclic_lanczos.cpp:44-71.
struct Counted {
  const Op &A;
  std::size_t spmv = 0, spmtv = 0;
  explicit Counted(const Op &a) : A(a) {}
  std::size_t rows() const { return A.rows(); }
  std::size_t cols() const { return A.cols(); }
  void multiply(const Vec &x, Vec &y) {
    ++spmv;
    A.mv(x, y);
  }
  void multiply_transpose(const Vec &x, Vec &y) {
    ++spmtv;
    A.mtv(x, y);
  }
};

// :26-33
void orth_pass(Vec &w, const std::vector<Vec> &S) {
  for (const auto &s : S) {
    const double c = dot(s, w);
    axpy(w, s, -c);
  }
}

// Projection.cpp:8-17
double residual_Av(Counted &A, const Vec &v, Vec &tmp_m) {
  if (A.rows() != tmp_m.size())
    throw std::invalid_argument("residual_norm_Av: tmp_m size mismatch");
  if (A.cols() != v.size())
    throw std::invalid_argument("residual_norm_Av: v size mismatch");
  A.multiply(v, tmp_m);
  return nrm2(tmp_m);
}

// :41-46
void apply_AAt(Counted &A, const Vec &x_m, Vec &y_m, Vec &tmp_n) {
  A.multiply_transpose(x_m, tmp_n);
  A.multiply(tmp_n, y_m);
}

// :48-113
std::pair<std::size_t, double>
cg_normal_eq_prec(Counted &A, Factory &vf, const Vec &rhs_m, Vec &w_m,
                  const Precond &M_inv, std::size_t maxit, double rel_tol,
                  double breakdown_tol) {
  const std::size_t m = A.rows(), n = A.cols();
  Vec r = vf.zeros(m), z = vf.zeros(m), p = vf.zeros(m), Ap = vf.zeros(m);
  Vec tmp_n = vf.zeros(n);

  std::fill(w_m.begin(), w_m.end(), 0.0);
  r = rhs_m;

  const double rhsn = std::sqrt(std::max(0.0, dot(r, r)));
  if (!(rhsn > 0.0) || !std::isfinite(rhsn))
    return {0, rhsn};

  if (M_inv)
    M_inv(r, z);
  else
    z = r;
  p = z;

  double rz = dot(r, z);
  if (!(rz > 0.0) || !std::isfinite(rz))
    return {0, rhsn};
  double rnorm = std::sqrt(rz);
  const double tol = std::max(rel_tol * rnorm, 1e-30);
  if (rnorm <= tol)
    return {0, rnorm};

  std::size_t k = 0;
  for (; k < maxit; ++k) {
    apply_AAt(A, p, Ap, tmp_n);
    const double pAp = dot(p, Ap);
    if (!(std::isfinite(pAp)) || std::fabs(pAp) <= breakdown_tol)
      break;
    const double alpha = rz / pAp;
    axpy(w_m, p, alpha);
    axpy(r, Ap, -alpha);
    if (M_inv)
      M_inv(r, z);
    else
      z = r;
    const double rz_new = dot(r, z);
    if (!std::isfinite(rz_new))
      break;
    rnorm = std::sqrt(std::max(0.0, rz_new));
    if (rnorm <= tol) {
      ++k;
      break;
    }
    const double beta = rz_new / std::max(1e-300, rz);
    rz = rz_new;
    scal(p, beta);
    axpy(p, z, 1.0);
  }
  return {k, rnorm};
}

// :115-127
double estimate_fro_norm(Counted &A, Factory &vf, std::size_t samples) {
  samples = std::max<std::size_t>(1, samples);
  double sum2 = 0.0;
  for (std::size_t s = 0; s < samples; ++s) {
    Vec r = vf.randn(A.cols());
    Vec Ar = vf.zeros(A.rows());
    A.multiply(r, Ar);
    const double nn = nrm2(Ar);
    sum2 += nn * nn;
  }
  return std::sqrt(sum2 / static_cast<double>(samples));
}

// Projection.hpp:13-21
void gradient_nullspace_step(Counted &A, Vec &v, Vec &tmp_m, Vec &tmp_n,
                             double tau) {
  A.multiply(v, tmp_m);
  A.multiply_transpose(tmp_m, tmp_n);
  axpy(v, tmp_n, -tau);
}

// :140-214
bool try_project_random_to_nullspace(Counted &A, Factory &vf,
                                     double tol_accept,
                                     std::vector<Vec> &basis,
                                     int grad_max_iters, int frob_samples) {
  const std::size_t m = A.rows(), n = A.cols();
  Vec v = vf.randn(n);
  for (auto &q : basis) {
    const double c = dot(v, q);
    axpy(v, q, -c);
  }
  double nv = nrm2(v);
  if (nv < 1e-14)
    return false;
  scal(v, 1.0 / nv);

  Vec tmp_m = vf.zeros(m), w_m = vf.zeros(m), tmp_n = vf.zeros(n);
  A.multiply(v, tmp_m);
  auto [its, rnorm] =
      cg_normal_eq_prec(A, vf, tmp_m, w_m, nullptr, 500, 1e-12, 1e-30);
  if (its > 0 || rnorm < nrm2(tmp_m)) {
    A.multiply_transpose(w_m, tmp_n);
    axpy(v, tmp_n, -1.0);
  }
  for (auto &q : basis) {
    const double c = dot(v, q);
    axpy(v, q, -c);
  }
  nv = nrm2(v);
  if (nv >= 1e-14)
    scal(v, 1.0 / nv);

  double res = residual_Av(A, v, tmp_m);
  if (res <= tol_accept) {
    basis.push_back(std::move(v));
    return true;
  }

  const double fn = estimate_fro_norm(A, vf, std::max(1, frob_samples));
  const double L2 = std::max(1e-30, fn * fn);
  const double tau = 0.9 / L2;
  for (int it = 0; it < grad_max_iters && res > tol_accept; ++it) {
    gradient_nullspace_step(A, v, tmp_m, tmp_n, tau);
    for (auto &q : basis) {
      const double c = dot(v, q);
      axpy(v, q, -c);
    }
    nv = nrm2(v);
    if (nv < 1e-14)
      return false;
    scal(v, 1.0 / nv);
    res = residual_Av(A, v, tmp_m);
  }
  if (res <= tol_accept) {
    basis.push_back(std::move(v));
    return true;
  }
  return false;
}

struct TinyEV {
  std::vector<double> evals;
  Mat evecs;
};

// :223-237
TinyEV solve_tiny_tridiag(const std::vector<double> &alpha,
                          const std::vector<double> &beta) {
  const int l = (int)alpha.size();
  Mat T(l, l);
  for (int i = 0; i < l; ++i) {
    T(i, i) = alpha[(std::size_t)i];
    if (i + 1 < l) {
      const double b = beta[(std::size_t)i];
      T(i, i + 1) = b;
      T(i + 1, i) = b;
    }
  }
  TinyEV r;
  sym_eig(T, r.evals, r.evecs);
  return r;
}

// :246-262
TinyEV solve_tiny_refined(const std::vector<double> &alpha,
                          const std::vector<double> &beta, double beta_last,
                          double tau) {
  const int l = (int)alpha.size();
  Mat Tbar(l + 1, l);
  for (int i = 0; i < l; ++i) {
    Tbar(i, i) = alpha[(std::size_t)i] - tau;
    if (i + 1 < l) {
      const double b = beta[(std::size_t)i];
      Tbar(i, i + 1) = b;
      Tbar(i + 1, i) = b;
    }
  }
  Tbar(l, l - 1) = beta_last;
  TinyEV r;
  svd_jacobi(Tbar, r.evals, r.evecs);
  return r;
}

bool all_finite(const std::vector<double> &v) {
  for (double x : v)
    if (!std::isfinite(x))
      return false;
  return true;
}

// :291-616
CyclicResult find_right_nullspace(Counted &A, Factory &vf,
                                  const CyclicOpts &opt) {
  const std::size_t m = A.rows(), n = A.cols();
  CyclicResult R;
  if (n == 0 || m == 0)
    return R;

  const std::size_t k_target = static_cast<std::size_t>(
      (opt.nullity_known >= 0) ? opt.nullity_known : static_cast<int>(opt.k));
  if (k_target == 0) {
    R.converged_to_target = true;
    return R;
  }
  if (opt.rank_known >= 0 && static_cast<int>(n) <= opt.rank_known)
    return R;
  const std::size_t ncv =
      std::max<std::size_t>(std::max<std::size_t>(2, opt.ncv), k_target + 2);

  R.frob_norm_est = estimate_fro_norm(A, vf, std::max(1, opt.frob_samples));
  const double accept_thresh =
      std::max(opt.tol_abs, opt.tol_rel * std::max(1.0, R.frob_norm_est));
  const double prescreen_thresh =
      std::max(0.0, opt.prescreen_factor) * std::max(1.0, R.frob_norm_est);

  std::vector<Vec> basis;
  basis.reserve(k_target);
  auto gs_against_basis = [&](Vec &v) {
    for (int pass = 0; pass < 2; ++pass)
      for (auto &q : basis) {
        const double c = dot(v, q);
        axpy(v, q, -c);
      }
  };
  auto random_right = [&]() {
    Vec y = vf.randn(n);
    gs_against_basis(y);
    return y;
  };
  Vec y0;
  bool have_y0 = true;
  if (!opt.initial_right_space.empty() && !opt.initial_right_space.front().empty())
    y0 = opt.initial_right_space.front();
  else
    y0 = vf.randn(n);

  std::vector<Vec> Y, X;
  std::vector<double> alpha, beta;
  Vec tmp_m = vf.zeros(m), w_m = vf.zeros(m), tmp_n = vf.zeros(n);
  std::size_t restarts_since_accept = 0;

  for (std::size_t restart = 0; restart < opt.max_restarts; ++restart) {
    R.restarts = restart;
    alpha.clear();
    beta.clear();

    gs_against_basis(y0);
    double ny = nrm2(y0);
    if (!std::isfinite(ny) || ny < 1e-14) {
      y0 = random_right();
      ny = nrm2(y0);
      if (!(ny >= 1e-14))
        break;
    }
    scal(y0, 1.0 / ny);
    Y.clear();
    X.clear();
    Y.push_back(std::move(y0));
    have_y0 = false;

    Vec w_x = vf.zeros(m), w_y = vf.zeros(n);
    std::size_t local_steps = 0;
    double beta_last = 0.0;
    for (std::size_t j = 0; j < ncv; ++j) {
      const bool from_y = (Y.size() > X.size());
      Vec &w = from_y ? w_x : w_y;
      if (from_y) {
        A.multiply(Y.back(), w);
        if (!beta.empty())
          axpy(w, X.back(), -beta.back());
      } else {
        A.multiply_transpose(X.back(), w);
        axpy(w, Y.back(), -beta.back());
      }
      const int reorth_passes = static_cast<int>(opt.reorth);
      for (int pass = 0; pass < reorth_passes; ++pass) {
        if (from_y) {
          orth_pass(w, X);
        } else {
          orth_pass(w, basis);
          orth_pass(w, Y);
        }
      }
      const double bj = std::sqrt(dot(w, w));
      alpha.push_back(0.0);
      beta_last = bj;
      if (!std::isfinite(bj) || bj < 1e-14 || Y.size() + X.size() >= ncv)
        break;
      beta.push_back(bj);
      Vec vnew = w;
      scal(vnew, 1.0 / bj);
      (from_y ? X : Y).push_back(std::move(vnew));
      ++local_steps;
    }
    R.krylov_steps_total += local_steps;

    const std::size_t l = alpha.size();
    if (l == 0 || !std::isfinite(beta_last))
      break;

    TinyEV tev;
    std::vector<int> ord;
    if (opt.use_harmonic) {
      tev = solve_tiny_refined(alpha, beta, beta_last, opt.target_tau);
      ord.resize(tev.evals.size());
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return tev.evals[a] < tev.evals[b];
      });
    } else {
      tev = solve_tiny_tridiag(alpha, beta);
      ord.resize(tev.evals.size());
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        return std::abs(tev.evals[a] - opt.target_tau) <
               std::abs(tev.evals[b] - opt.target_tau);
      });
    }
    if (!all_finite(tev.evals) || !all_finite(tev.evecs.a)) {
      y0 = random_right();
      continue;
    }

    const std::size_t test_cap =
        std::min<std::size_t>(l, std::max<std::size_t>(k_target, 8));
    double best_res = std::numeric_limits<double>::infinity();
    Vec best_v;
    bool have_best = false;
    bool accepted_this_cycle = false;
    std::size_t tested = 0;
    bool polish_stalled = false;
    for (std::size_t t = 0; t < ord.size() && tested < test_cap; ++t) {
      const int jbest = ord[t];
      double c_even2 = 0.0;
      for (std::size_t i = 0; i < l; i += 2)
        c_even2 += tev.evecs((int)i, jbest) * tev.evecs((int)i, jbest);
      if (!(c_even2 >= 1e-20))
        continue;
      Vec v = vf.zeros(n);
      for (std::size_t i = 0; i < l; i += 2)
        axpy(v, Y[i / 2], tev.evecs((int)i, jbest));

      gs_against_basis(v);
      double nv = nrm2(v);
      if (!std::isfinite(nv) || nv < 1e-8)
        continue;
      scal(v, 1.0 / nv);
      ++tested;

      A.multiply(v, tmp_m);
      double res = nrm2(tmp_m);
      if (!std::isfinite(res))
        continue;
      if (res > prescreen_thresh || polish_stalled) {
        if (res < best_res) {
          best_res = res;
          best_v = v;
          have_best = true;
        }
        continue;
      }

      const double raw_res = res;
      std::size_t cg_its = 0;
      bool hit_maxit = false;
      for (int pass = 0;
           pass < std::max(1, opt.projector_passes) && res > accept_thresh;
           ++pass) {
        cg_its = cg_normal_eq_prec(A, vf, tmp_m, w_m, opt.precond_AAt,
                                   opt.cg_maxit, opt.cg_rel_tol,
                                   opt.cg_breakdown_tol)
                     .first;
        ++R.cg_solves;
        if (cg_its >= opt.cg_maxit) {
          ++R.cg_maxit_hits;
          hit_maxit = true;
        }
        A.multiply_transpose(w_m, tmp_n);
        Vec vp = v;
        axpy(vp, tmp_n, -1.0);
        gs_against_basis(vp);
        const double np = nrm2(vp);
        if (!std::isfinite(np) || np < 1e-14)
          break;
        scal(vp, 1.0 / np);
        A.multiply(vp, w_m);
        const double res_p = nrm2(w_m);
        if (!(res_p < res))
          break;
        v = std::move(vp);
        res = res_p;
        tmp_m = w_m;
      }
      if (opt.verbose)
        std::cout << "[oracle cyclic] restart=" << restart << " raw=" << raw_res
                  << " polished=" << res << " thr=" << accept_thresh
                  << " cg_its=" << cg_its << "\n";

      if (res <= accept_thresh) {
        basis.push_back(std::move(v));
        accepted_this_cycle = true;
        if (basis.size() >= k_target) {
          R.right_null = std::move(basis);
          R.converged_to_target = true;
          R.spmv = A.spmv;
          R.spmtv = A.spmtv;
          return R;
        }
      } else {
        polish_stalled = polish_stalled || hit_maxit;
        if (res < best_res) {
          best_res = res;
          best_v = std::move(v);
          have_best = true;
        }
      }
    }

    if (!accepted_this_cycle) {
      ++restarts_since_accept;
      if (opt.stagnation_restarts > 0 &&
          restarts_since_accept >= opt.stagnation_restarts)
        break;
    } else {
      restarts_since_accept = 0;
    }
    y0 = have_best ? std::move(best_v) : random_right();
    have_y0 = true;
  }
  (void)have_y0;

  if (basis.size() < k_target) {
    const std::size_t max_consecutive_failures = 8;
    std::size_t failures = 0;
    while (basis.size() < k_target && failures < max_consecutive_failures) {
      if (try_project_random_to_nullspace(A, vf, accept_thresh, basis,
                                          opt.grad_max_iters, opt.frob_samples))
        failures = 0;
      else
        ++failures;
    }
    R.converged_to_target = basis.size() >= k_target;
  }
  R.right_null = std::move(basis);
  R.spmv = A.spmv;
  R.spmtv = A.spmtv;
  return R;
}

} // namespace

// :618-660
CyclicResult cyclic_lanczos(const Op &A, Factory &vf, const CyclicOpts &opt) {
  Counted CA(A);
  CyclicResult R = find_right_nullspace(CA, vf, opt);
  if (!opt.return_left_space || A.rows() == 0 || A.cols() == 0)
    return R;
  CyclicOpts lopt;
  lopt.k = opt.k;
  lopt.ncv = opt.ncv;
  lopt.max_restarts = opt.max_restarts;
  lopt.tol_abs = opt.tol_abs;
  lopt.tol_rel = opt.tol_rel;
  lopt.reorth = opt.reorth;
  lopt.use_harmonic = opt.use_harmonic;
  lopt.target_tau = opt.target_tau;
  lopt.cg_maxit = opt.cg_maxit;
  lopt.cg_rel_tol = opt.cg_rel_tol;
  lopt.cg_breakdown_tol = opt.cg_breakdown_tol;
  lopt.projector_passes = opt.projector_passes;
  lopt.grad_max_iters = opt.grad_max_iters;
  lopt.frob_samples = opt.frob_samples;
  lopt.prescreen_factor = opt.prescreen_factor;
  lopt.stagnation_restarts = opt.stagnation_restarts;
  lopt.verbose = opt.verbose;
  lopt.rank_known = opt.rank_known;
  if (opt.rank_known >= 0) {
    const int left_nullity = static_cast<int>(A.rows()) - opt.rank_known;
    lopt.nullity_known =
        std::max(0, std::min(left_nullity, static_cast<int>(opt.k)));
  }
  TransposedOp At(A);
  Counted CAt(At);
  CyclicResult L = find_right_nullspace(CAt, vf, lopt);
  R.left_null = std::move(L.right_null);
  R.cg_solves += L.cg_solves;
  R.cg_maxit_hits += L.cg_maxit_hits;
  R.spmv += L.spmtv;
  R.spmtv += L.spmv;
  return R;
}

} // namespace orc
