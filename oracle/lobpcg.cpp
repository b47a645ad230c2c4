// Oracle restatement of find_nullspace_lobpcg_ex
// (src/Solvers/LOBPCGSolver.cpp:18-842). Block algebra on plain vectors;
// Eigen's LLT / inverse / SelfAdjointEigenSolver replaced by dense.cpp.
#include "oracle.hpp"

#include <algorithm>
#include <iostream>
#include <limits>

namespace orc {
namespace {

using Block = std::vector<Vec>;

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

Block make_block(std::size_t n, std::size_t cols) { // :26-33
  return Block(cols, Vec(n, 0.0));
}

Block hstack(const Block &X, const Block &Y) { // :38-46
  Block Z;
  Z.reserve(X.size() + Y.size());
  for (const auto &p : X)
    Z.push_back(p);
  for (const auto &p : Y)
    Z.push_back(p);
  return Z;
}

Block apply_normal(Counted &A, const Block &X) { // :54-64
  const std::size_t m = A.rows(), n = A.cols();
  Block AX = make_block(n, X.size());
  Vec tmp_m(m, 0.0);
  for (std::size_t j = 0; j < X.size(); ++j) {
    A.multiply(X[j], tmp_m);
    A.multiply_transpose(tmp_m, AX[j]);
  }
  return AX;
}

void form_gram(const Block &Z, const Block &AZ, Mat &S, Mat &G) { // :67-80
  const int k = (int)Z.size();
  S = Mat(k, k);
  G = Mat(k, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= j; ++i) {
      const double sij = dot(Z[(std::size_t)i], Z[(std::size_t)j]);
      const double gij = dot(Z[(std::size_t)i], AZ[(std::size_t)j]);
      S(i, j) = S(j, i) = sij;
      G(i, j) = G(j, i) = gij;
    }
}

struct TinyGEP {
  std::vector<double> theta;
  Mat Y;
};

TinyGEP solve_tiny_gep(Mat &G, Mat &S) { // :88-119
  const int k = G.r;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < k; ++i) {
      // symmetrise (upper and lower computed from the originals)
      if (i < j) {
        const double g = 0.5 * (G(i, j) + G(j, i));
        G(i, j) = G(j, i) = g;
        const double s = 0.5 * (S(i, j) + S(j, i));
        S(i, j) = S(j, i) = s;
      }
    }
  double fro = 0.0;
  for (double v : S.a)
    fro += v * v;
  const double s_norm = std::max(1.0, std::sqrt(fro));
  const double jitters[] = {0.0, 1e-14, 1e-12, 1e-10};
  for (double jv : jitters) {
    Mat Sj = S;
    if (jv > 0.0)
      for (int i = 0; i < k; ++i)
        Sj(i, i) += jv * s_norm;
    Mat L;
    if (!cholesky(Sj, L))
      continue;
    Mat Li = lower_inverse(L);
    // C = Li G Li^T
    Mat T(k, k), C(k, k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = 0; t < k; ++t)
          s += Li(i, t) * G(t, j);
        T(i, j) = s;
      }
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = 0; t < k; ++t)
          s += T(i, t) * Li(j, t);
        C(i, j) = s;
      }
    TinyGEP out;
    Mat Q;
    sym_eig(C, out.theta, Q);
    out.Y = Mat(k, k); // Y = Li^T Q
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = 0; t < k; ++t)
          s += Li(t, i) * Q(t, j);
        out.Y(i, j) = s;
      }
    return out;
  }
  throw std::runtime_error("LOBPCG tiny GEP: S not SPD (after jitter)");
}

// :122-134, with Y restricted to its first p columns and rows [r0, r0+|Z|)
Block combine_columns(const Block &Z, const Mat &Y, int row0, int rows,
                      std::size_t p) {
  const std::size_t n = Z.front().size();
  Block X = make_block(n, p);
  for (std::size_t i = 0; i < p; ++i)
    for (int j = 0; j < rows; ++j) {
      const double c = Y(row0 + j, (int)i);
      if (c != 0.0)
        axpy(X[i], Z[(std::size_t)j], c);
    }
  return X;
}

bool orthonormalize_block(Block &X, const std::vector<Vec> &vs_basis,
                          double tol, bool reorth) { // :137-160
  for (auto &x : X) {
    for (const auto &q : vs_basis) {
      const double c = dot(x, q);
      axpy(x, q, -c);
    }
    const double nrm = nrm2(x);
    if (nrm < tol)
      return false;
    scal(x, 1.0 / nrm);
  }
  Block tmp = X;
  if (!modified_gram_schmidt(tmp, tol, reorth))
    return false;
  X = tmp;
  return true;
}

double estimate_fro_norm(Counted &A, Factory &vf, std::size_t samples) {
  samples = std::max<std::size_t>(1, samples); // :163-178
  const std::size_t m = A.rows(), n = A.cols();
  Vec r(n, 0.0), Ar(m, 0.0);
  double sum2 = 0.0;
  for (std::size_t s = 0; s < samples; ++s) {
    r = vf.randn(n);
    A.multiply(r, Ar);
    const double nrm = nrm2(Ar);
    sum2 += nrm * nrm;
  }
  return std::sqrt(sum2 / (double)samples);
}

double residual_Av(Counted &A, const Vec &v, Vec &tmp_m) {
  if (A.rows() != tmp_m.size())
    throw std::invalid_argument("residual_norm_Av: tmp_m size mismatch");
  if (A.cols() != v.size())
    throw std::invalid_argument("residual_norm_Av: v size mismatch");
  A.multiply(v, tmp_m);
  return nrm2(tmp_m);
}

void project_out(const std::vector<Vec> &basis, Vec &x) { // :188-195
  for (const auto &q : basis) {
    const double c = dot(x, q);
    if (c != 0.0)
      axpy(x, q, -c);
  }
}

// :199-278
bool two_vec_rr_sharpen(Counted &A, std::vector<Vec> &V, std::size_t idx,
                        double target_residual, const Precond *M_inv_opt,
                        int max_iters) {
  if (idx >= V.size())
    return true;
  const std::size_t m = A.rows();
  Vec Av(m, 0.0);
  Vec t = V[idx];
  Vec d = V[idx];
  Vec Ad(m, 0.0);
  std::vector<Vec> basis;
  for (std::size_t j = 0; j < V.size(); ++j)
    if (j != idx)
      basis.push_back(V[j]);
  for (int it = 0; it < max_iters; ++it) {
    A.multiply(V[idx], Av);
    const double res = nrm2(Av);
    if (res <= target_residual)
      return true;
    A.multiply_transpose(Av, t);
    if (M_inv_opt && *M_inv_opt)
      (*M_inv_opt)(t, d);
    else
      d = t;
    project_out(basis, d);
    double nd = nrm2(d);
    if (nd <= 1e-14)
      return false;
    scal(d, 1.0 / std::max(nd, 1e-300));
    A.multiply(d, Ad);
    Vec Nd = t;
    A.multiply_transpose(Ad, Nd);
    const double vNv = dot(V[idx], t);
    const double vNd = dot(V[idx], Nd);
    const double dNv = dot(d, t);
    const double dNd = dot(d, Nd);
    Mat G(2, 2);
    G(0, 0) = vNv;
    G(0, 1) = vNd;
    G(1, 0) = dNv;
    G(1, 1) = dNd;
    const double off = 0.5 * (G(0, 1) + G(1, 0));
    G(0, 1) = G(1, 0) = off;
    std::vector<double> w;
    Mat Q;
    sym_eig(G, w, Q);
    Vec v_new = V[idx];
    scal(v_new, Q(0, 0));
    axpy(v_new, d, Q(1, 0));
    project_out(basis, v_new);
    const double nv = nrm2(v_new);
    if (nv <= 1e-14)
      return false;
    scal(v_new, 1.0 / nv);
    V[idx] = v_new;
  }
  return false;
}

// :286-440
void block_rr_polish(Counted &A, Factory &vf, std::vector<Vec> &V,
                     double target_residual, int max_sweeps,
                     const Precond *M_inv_opt) {
  if (V.empty())
    return;
  const std::size_t m = A.rows();
  const std::size_t k = V.size();
  Vec r(m, 0.0);
  Vec g = V.front();
  Vec tmp(m, 0.0);
  (void)modified_gram_schmidt(V, 1e-14, true);

  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double max_res = 0.0;
    for (std::size_t i = 0; i < k; ++i) {
      A.multiply(V[i], r);
      max_res = std::max(max_res, nrm2(r));
    }
    if (max_res <= target_residual)
      break;

    Block D;
    std::vector<Vec> V_basis = V;
    for (std::size_t j = 0; j < k; ++j) {
      A.multiply(V[j], r);
      const double res = nrm2(r);
      if (res <= target_residual)
        continue;
      A.multiply_transpose(r, g);
      if (M_inv_opt && *M_inv_opt) {
        Vec g_prec = g;
        (*M_inv_opt)(g, g_prec);
        g.swap(g_prec);
      }
      project_out(V_basis, g);
      const double gn = nrm2(g);
      if (gn <= 1e-14)
        continue;
      scal(g, 1.0 / gn);
      bool dup = false;
      for (const auto &dcol : D)
        if (std::fabs(dot(g, dcol)) >= 1.0 - 1e-8) {
          dup = true;
          break;
        }
      if (!dup)
        D.push_back(g);
    }

    if (D.empty()) {
      for (std::size_t i = 0; i < k; ++i) {
        for (int it = 0; it < 50; ++it) {
          A.multiply(V[i], r);
          const double res = nrm2(r);
          if (res <= target_residual)
            break;
          A.multiply_transpose(r, g);
          for (std::size_t t = 0; t < k; ++t) {
            if (t == i)
              continue;
            const double c = dot(g, V[t]);
            if (c != 0.0)
              axpy(g, V[t], -c);
          }
          const double gg = dot(g, g);
          if (gg <= 0.0)
            break;
          A.multiply(g, tmp);
          const double Ag2 = dot(tmp, tmp);
          if (Ag2 <= 0.0)
            break;
          const double alpha = gg / Ag2;
          axpy(V[i], g, -alpha);
          for (std::size_t t = 0; t < k; ++t)
            if (t != i) {
              const double c = dot(V[i], V[t]);
              if (c != 0.0)
                axpy(V[i], V[t], -c);
            }
          const double nv = nrm2(V[i]);
          if (nv > 0.0)
            scal(V[i], 1.0 / nv);
        }
      }
      (void)modified_gram_schmidt(V, 1e-14, true);
      const int enrich = std::min<int>(2, (int)k);
      for (int e = 0; e < enrich; ++e) {
        Vec q = vf.randn(V.front().size());
        project_out(V_basis, q);
        const double nq = nrm2(q);
        if (nq > 1e-14) {
          scal(q, 1.0 / nq);
          D.push_back(std::move(q));
        }
      }
      if (!D.empty())
        (void)orthonormalize_block(D, V_basis, 1e-14, true);
      else
        continue;
    }

    if (!orthonormalize_block(D, V_basis, 1e-14, true))
      continue;

    Block Z = hstack(V, D);
    Block AZ = apply_normal(A, Z);
    Mat S, G;
    form_gram(Z, AZ, S, G);
    TinyGEP t = solve_tiny_gep(G, S);
    Block Vnew = combine_columns(Z, t.Y, 0, (int)Z.size(), k);
    std::vector<Vec> Vtmp = Vnew;
    (void)modified_gram_schmidt(Vtmp, 1e-14, true);
    for (std::size_t i = 0; i < k; ++i)
      V[i] = Vtmp[i];

    max_res = 0.0;
    for (std::size_t i = 0; i < k; ++i) {
      A.multiply(V[i], r);
      max_res = std::max(max_res, nrm2(r));
    }
    if (max_res <= target_residual)
      break;
  }
  (void)modified_gram_schmidt(V, 1e-14, true);
}

} // namespace

// :443-834
LobpcgResult lobpcg(const Op &Aop, Factory &vf, const LobpcgOpts &opt) {
  Counted A(Aop);
  LobpcgResult R;
  const std::size_t m = A.rows(), n = A.cols();
  if (n == 0 || m == 0)
    return R;

  std::size_t k_target = std::max<std::size_t>(1, opt.nev);
  if (opt.nullity_known >= 0) {
    const std::size_t nk = (std::size_t)opt.nullity_known;
    if (nk > n)
      throw std::invalid_argument("LOBPCG: nullity_known > n");
    k_target = nk;
  }
  if (opt.rank_known >= 0) {
    const std::size_t rk = (std::size_t)opt.rank_known;
    if (rk > n)
      throw std::invalid_argument("LOBPCG: rank_known > n");
    const std::size_t implied_null = n - rk;
    if (opt.nullity_known >= 0 && implied_null != k_target)
      throw std::invalid_argument(
          "LOBPCG: inconsistent rank_known vs nullity_known");
    if (opt.nullity_known < 0)
      k_target = implied_null;
    if (n <= rk) {
      R.frob_norm_est = 0.0;
      R.converged_to_target = (k_target == 0);
      return R;
    }
  }

  std::size_t bs = opt.block_size > 0
                       ? opt.block_size
                       : std::min<std::size_t>(16, std::max<std::size_t>(1, k_target));
  bs = std::min<std::size_t>(bs, n);
  R.block_size = bs;
  const std::size_t guards =
      (bs == 1) ? 0
                : std::min<std::size_t>(
                      (std::size_t)((1.0 - opt.restart) * bs + 0.45), bs - 1);

  R.frob_norm_est = estimate_fro_norm(A, vf, std::max(1, opt.frob_samples));
  const double accept_thr =
      std::max(opt.tol_abs, opt.tol_rel * std::max(1.0, R.frob_norm_est));
  R.accept_threshold = accept_thr;

  std::vector<Vec> Y_locked;
  Block X = make_block(n, bs);
  std::size_t filled = 0;
  const std::size_t seeds = std::min<std::size_t>(bs, opt.initial_block.size());
  for (; filled < seeds; ++filled)
    X[filled] = opt.initial_block[filled];
  for (; filled < bs; ++filled)
    X[filled] = vf.randn(n);
  if (!orthonormalize_block(X, Y_locked, opt.gs_tol, opt.gs_reorth))
    throw std::runtime_error("LOBPCG: initial block is dependent");

  Block AX = apply_normal(A, X);
  {
    Mat S0, G0;
    form_gram(X, AX, S0, G0);
    TinyGEP init = solve_tiny_gep(G0, S0);
    Block Xrr = combine_columns(X, init.Y, 0, (int)X.size(), bs);
    X.swap(Xrr);
    AX = apply_normal(A, X);
    std::vector<double> th(bs, 0.0);
    R.theta.clear();
    const int lim = std::min<int>((int)init.theta.size(), (int)bs);
    for (int j = 0; j < lim; ++j)
      th[(std::size_t)j] = std::max(0.0, init.theta[(std::size_t)j]);
    R.theta.swap(th); // temporarily holds the running theta
  }
  std::vector<double> theta = R.theta;
  R.theta.clear();

  Block Rb = make_block(n, bs), W = make_block(n, bs), P = make_block(n, bs);
  Vec tmp_m_v(m, 0.0);
  std::size_t its = 0, locked_total = 0;

  while (locked_total < k_target && its < opt.max_iters) {
    ++its;
    std::size_t nconv = 0;
    double best_unaccepted = std::numeric_limits<double>::infinity();
    for (std::size_t j = 0; j < bs; ++j) {
      Rb[j] = AX[j];
      axpy(Rb[j], X[j], -theta[j]);
      const double Av_orig = residual_Av(A, X[j], tmp_m_v);
      if (Av_orig <= accept_thr) {
        Vec v_acc = X[j];
        for (const auto &q : Y_locked) {
          const double c = dot(v_acc, q);
          axpy(v_acc, q, -c);
        }
        double nv = nrm2(v_acc);
        bool reject = false;
        if (nv <= std::max(1e-14, 10.0 * opt.gs_tol)) {
          reject = true;
        } else {
          scal(v_acc, 1.0 / nv);
          for (const auto &q : Y_locked)
            if (std::fabs(dot(v_acc, q)) >= 1.0 - 1e-8) {
              reject = true;
              break;
            }
        }
        double Av_acc = std::numeric_limits<double>::infinity();
        if (!reject) {
          Av_acc = residual_Av(A, v_acc, tmp_m_v);
          if (Av_acc > accept_thr)
            reject = true;
        }
        if (!reject) {
          Y_locked.push_back(v_acc);
          R.right_null.push_back(std::move(v_acc));
          R.theta.push_back(theta[j]);
          R.residuals_Av.push_back(Av_acc);
          ++locked_total;
          ++nconv;
          if (locked_total >= k_target)
            break;
        } else {
          best_unaccepted = std::min(best_unaccepted, Av_orig);
        }
      } else {
        best_unaccepted = std::min(best_unaccepted, Av_orig);
      }
    }
    R.locked = locked_total;
    R.iters = its;
    R.best_unaccepted_residual =
        std::isfinite(best_unaccepted) ? best_unaccepted : 0.0;
    if (opt.verbose && opt.progress_every > 0 &&
        (its == 1 || its % opt.progress_every == 0))
      std::cout << "[oracle LOBPCG] it=" << its << " locked=" << locked_total
                << " best=" << R.best_unaccepted_residual << "\n";
    if (locked_total >= k_target)
      break;

    const std::size_t ini = opt.soft_locking ? nconv : 0;

    if (nconv >= bs - guards) { // :624-651
      const std::size_t remain = bs - nconv;
      for (std::size_t j = 0; j < remain; ++j) {
        X[j] = X[nconv + j];
        AX[j] = AX[nconv + j];
        theta[j] = theta[nconv + j];
        P[j] = P[nconv + j];
      }
      for (std::size_t j = remain; j < bs; ++j)
        X[j] = vf.randn(n);
      if (!orthonormalize_block(X, Y_locked, opt.gs_tol, opt.gs_reorth))
        throw std::runtime_error("LOBPCG: restart orthogonalization failed");
      AX = apply_normal(A, X);
      Mat S1, G1;
      form_gram(X, AX, S1, G1);
      TinyGEP t1 = solve_tiny_gep(G1, S1);
      std::fill(theta.begin(), theta.end(), 0.0);
      const int lim = std::min<int>((int)t1.theta.size(), (int)bs);
      for (int j = 0; j < lim; ++j)
        theta[(std::size_t)j] = std::max(0.0, t1.theta[(std::size_t)j]);
      continue;
    }

    for (std::size_t j = ini; j < bs; ++j) { // :653-659
      if (opt.M_inv)
        opt.M_inv(Rb[j], W[j]);
      else
        W[j] = Rb[j];
    }
    {
      Block Wact(W.begin() + (long)ini, W.end());
      if (!orthonormalize_block(Wact, Y_locked, opt.gs_tol, opt.gs_reorth)) {
        for (std::size_t j = ini; j < bs; ++j)
          W[j] = vf.randn(n);
        for (std::size_t j = ini; j < bs; ++j)
          for (const auto &q : Y_locked) {
            const double c = dot(W[j], q);
            axpy(W[j], q, -c);
          }
      } else {
        for (std::size_t t = 0; t < Wact.size(); ++t)
          W[ini + t] = Wact[t];
      }
    }
    if (its > 1) {
      Block Pact(P.begin() + (long)ini, P.end());
      (void)orthonormalize_block(Pact, Y_locked, opt.gs_tol, opt.gs_reorth);
      for (std::size_t t = 0; t < Pact.size(); ++t)
        P[ini + t] = Pact[t];
    }

    Block Xact(X.begin() + (long)ini, X.end());
    Block Wact(W.begin() + (long)ini, W.end());
    Block Pact;
    if (its > 1)
      Pact.assign(P.begin() + (long)ini, P.end());
    Block Z = (its > 1) ? hstack(hstack(Xact, Wact), Pact) : hstack(Xact, Wact);
    const std::size_t nv = Z.size();
    if (nv == 0)
      continue;
    if (opt.max_tiny_dim > 0 && nv > opt.max_tiny_dim)
      throw std::runtime_error("LOBPCG: tiny problem too large (nv cap)");

    Block AZ = apply_normal(A, Z);
    Mat S, G;
    form_gram(Z, AZ, S, G);
    TinyGEP sol = solve_tiny_gep(G, S);

    const std::size_t take = std::min<std::size_t>(bs - ini, sol.theta.size());
    const int nx = (int)(bs - ini);
    const int nw = (int)(bs - ini);
    const int np = (its > 1) ? (int)(bs - ini) : 0;

    // P_new = [W P] * Y[nx : nx+nw+np, :take]
    Block Zwp = (its > 1) ? hstack(Wact, Pact) : Wact;
    Block Pact_new = combine_columns(Zwp, sol.Y, nx, nw + np, take);
    for (std::size_t j = 0; j < take; ++j)
      P[ini + j] = Pact_new[j];
    Block Xxy = combine_columns(Xact, sol.Y, 0, nx, take);
    for (std::size_t j = 0; j < take; ++j) {
      X[ini + j] = Xxy[j];
      axpy(X[ini + j], P[ini + j], 1.0);
    }
    {
      Block Xslice(X.begin() + (long)ini, X.begin() + (long)(ini + take));
      if (!orthonormalize_block(Xslice, Y_locked, opt.gs_tol, opt.gs_reorth))
        throw std::runtime_error("LOBPCG: X slice orthonormalization failed");
      for (std::size_t t = 0; t < take; ++t)
        X[ini + t] = Xslice[t];
    }
    AX = apply_normal(A, X); // :773-787
    {
      Mat S2, G2;
      form_gram(X, AX, S2, G2);
      TinyGEP t2 = solve_tiny_gep(G2, S2);
      Block Xrr2 = combine_columns(X, t2.Y, 0, (int)X.size(), bs);
      X.swap(Xrr2);
      AX = apply_normal(A, X);
      std::fill(theta.begin(), theta.end(), 0.0);
      const int lim = std::min<int>((int)t2.theta.size(), (int)bs);
      for (int j = 0; j < lim; ++j)
        theta[(std::size_t)j] = std::max(0.0, t2.theta[(std::size_t)j]);
    }
  }

  R.converged_to_target = (R.locked >= k_target);
  if (!R.right_null.empty()) { // :806-831
    const double desired =
        std::max(opt.tol_abs, opt.tol_rel * std::max(1.0, R.frob_norm_est));
    const double target_residual =
        std::max(1e-14, std::min(1e-10, 0.01 * desired));
    const Precond *Mptr = opt.M_inv ? &opt.M_inv : nullptr;
    block_rr_polish(A, vf, R.right_null, target_residual, 80, Mptr);
    Vec tmp(m, 0.0);
    for (std::size_t i = 0; i < R.right_null.size(); ++i) {
      A.multiply(R.right_null[i], tmp);
      if (nrm2(tmp) <= target_residual)
        continue;
      (void)two_vec_rr_sharpen(A, R.right_null, i, target_residual, Mptr, 120);
    }
    (void)modified_gram_schmidt(R.right_null, 1e-14, true);
  }
  R.spmv = A.spmv;
  R.spmtv = A.spmtv;
  return R;
}

} // namespace orc
