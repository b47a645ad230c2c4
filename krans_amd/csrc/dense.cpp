#include "dense.hpp"

#include <algorithm>
#include <cmath>
#include <exception>
#include <limits>
#include <numeric>
#include <thread>

namespace kh::dense {

static constexpr double kEps = std::numeric_limits<double>::epsilon();

// One-sided Jacobi (Hestenes) sweeps over the n columns (length m) of U,
// column-major; V (n x n, on entry the identity; may be null) accumulates the
// rotations and nrm receives the squared column norms. The norms and U do not
// depend on V.
static void jacobi_sweeps(std::vector<double> &U, int m, int n, Mat *V,
                          std::vector<double> &nrm) {
  auto col = [&](int j) { return U.data() + (std::size_t)j * m; };
  auto sq = [&](int j) {
    const double *u = col(j);
    double s = 0.0;
    for (int i = 0; i < m; ++i)
      s += u[i] * u[i];
    return s;
  };
  for (int j = 0; j < n; ++j)
    nrm[j] = sq(j);
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotations = 0;
    for (int p = 0; p + 1 < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double *up = col(p), *uq = col(q);
        double g = 0.0;
        for (int i = 0; i < m; ++i)
          g += up[i] * uq[i];
        const double a = nrm[p], b = nrm[q];
        if (g == 0.0 || std::fabs(g) <= kEps * std::sqrt(a) * std::sqrt(b))
          continue;
        ++rotations;
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::hypot(1.0, zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        double *wp = col(p), *wq = col(q);
        for (int i = 0; i < m; ++i) {
          const double x = wp[i], y = wq[i];
          wp[i] = c * x - s * y;
          wq[i] = s * x + c * y;
        }
        if (V)
          for (int i = 0; i < n; ++i) {
            const double x = (*V)(i, p), y = (*V)(i, q);
            (*V)(i, p) = c * x - s * y;
            (*V)(i, q) = s * x + c * y;
          }
        // recompute the two norms exactly (keeps tiny singular values accurate)
        nrm[p] = sq(p);
        nrm[q] = sq(q);
      }
    }
    if (rotations == 0)
      break;
  }
}

// A matrix whose entries with i + j even are all zero (the refined
// Golub-Kahan matrix with a zero diagonal) maps even columns to odd rows and
// odd columns to even rows: two independent blocks. Their columns have
// disjoint supports, so the sweeps over the whole matrix never rotate a pair
// across blocks (g == 0 exactly) and never change an entry outside a
// column's own rows; sweeping each block on its own rows only drops
// additions and rotations of exact zeros, with the same rotation order
// inside each block, and gives bitwise the same norms and V. With
// even_only, the rotations of the odd-column block are not accumulated: its
// columns of V stay unit vectors on odd rows.
static bool jacobi_parity_blocks(const Mat &A, Mat &V, std::vector<double> &nrm,
                                 bool even_only) {
  const int m = A.r, n = A.c;
  if (n < 2)
    return false;
  for (int j = 0; j < n; ++j)
    for (int i = j % 2; i < m; i += 2)
      if (A(i, j) != 0.0)
        return false;
  // the two blocks write disjoint columns of V and entries of nrm; from 16
  // columns each, the odd block runs on a second thread
  auto sweep_block = [&](int b) {
    std::vector<int> cols, rows;
    for (int j = b; j < n; j += 2)
      cols.push_back(j);
    for (int i = 1 - b; i < m; i += 2)
      rows.push_back(i);
    const int nb = (int)cols.size(), mb = (int)rows.size();
    std::vector<double> U((std::size_t)mb * nb);
    for (int t = 0; t < nb; ++t)
      for (int r = 0; r < mb; ++r)
        U[(std::size_t)t * mb + r] = A(rows[r], cols[t]);
    Mat Vb(nb, nb);
    for (int t = 0; t < nb; ++t)
      Vb(t, t) = 1.0;
    std::vector<double> nb2(nb);
    const bool vec = b == 0 || !even_only;
    jacobi_sweeps(U, mb, nb, vec ? &Vb : nullptr, nb2);
    for (int t = 0; t < nb; ++t) {
      nrm[cols[t]] = nb2[t];
      for (int r = 0; r < nb; ++r)
        V(cols[r], cols[t]) = Vb(r, t);
    }
  };
  if (n >= 32) {
    std::exception_ptr err;
    std::thread odd([&] {
      try {
        sweep_block(1);
      } catch (...) {
        err = std::current_exception();
      }
    });
    try {
      sweep_block(0);
    } catch (...) {
      odd.join();
      throw;
    }
    odd.join();
    if (err)
      std::rethrow_exception(err);
  } else {
    sweep_block(0);
    sweep_block(1);
  }
  return true;
}

void svd_right(const Mat &A, std::vector<double> &sv, Mat &V, bool split, bool even_only) {
  const int m = A.r, n = A.c;
  V = Mat(n, n);
  std::vector<double> nrm(n);
  if (!(split && jacobi_parity_blocks(A, V, nrm, even_only))) {
    std::vector<double> U(A.a);  // columns of A, orthogonalised in place
    for (int i = 0; i < n; ++i)
      V(i, i) = 1.0;
    jacobi_sweeps(U, m, n, &V, nrm);
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return nrm[x] > nrm[y]; });
  sv.assign(n, 0.0);
  Mat Vs(n, n);
  for (int j = 0; j < n; ++j) {
    sv[j] = std::sqrt(nrm[ord[j]]);
    for (int i = 0; i < n; ++i)
      Vs(i, j) = V(i, ord[j]);
  }
  V = std::move(Vs);
}

bool sym_eig(const Mat &Ain, std::vector<double> &w, Mat &Q) {
  const int n = Ain.r;
  Mat T = Ain;
  Q = Mat(n, n);
  for (int i = 0; i < n; ++i)
    Q(i, i) = 1.0;
  if (n == 0) {
    w.clear();
    return true;
  }
  // Householder tridiagonalisation: T <- H T H, Q <- Q H
  std::vector<double> v(n), tmp(n);
  for (int k = 0; k + 2 < n; ++k) {
    const int len = n - k - 1;
    double sigma = 0.0;
    for (int i = 0; i < len; ++i)
      sigma += T(k + 1 + i, k) * T(k + 1 + i, k);
    sigma = std::sqrt(sigma);
    double tail = 0.0;
    for (int i = 1; i < len; ++i)
      tail = std::max(tail, std::fabs(T(k + 1 + i, k)));
    if (sigma == 0.0 || tail == 0.0)
      continue;
    const double x0 = T(k + 1, k);
    const double alpha = x0 >= 0 ? -sigma : sigma;
    for (int i = 0; i < len; ++i)
      v[i] = T(k + 1 + i, k);
    v[0] -= alpha;
    double vn = 0.0;
    for (int i = 0; i < len; ++i)
      vn += v[i] * v[i];
    vn = std::sqrt(vn);
    for (int i = 0; i < len; ++i)
      v[i] /= vn;
    // rows: T[k+1:, :] -= 2 v (v^T T[k+1:, :])
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int i = 0; i < len; ++i)
        s += v[i] * T(k + 1 + i, j);
      s *= 2.0;
      for (int i = 0; i < len; ++i)
        T(k + 1 + i, j) -= s * v[i];
    }
    // cols: T[:, k+1:] -= 2 (T[:, k+1:] v) v^T
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < len; ++j)
        s += T(i, k + 1 + j) * v[j];
      s *= 2.0;
      for (int j = 0; j < len; ++j)
        T(i, k + 1 + j) -= s * v[j];
    }
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < len; ++j)
        s += Q(i, k + 1 + j) * v[j];
      s *= 2.0;
      for (int j = 0; j < len; ++j)
        Q(i, k + 1 + j) -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i)
      T(i, k) = T(k, i) = 0.0;
  }
  // keep the tridiagonal exactly symmetric
  for (int i = 0; i + 1 < n; ++i) {
    const double e = 0.5 * (T(i + 1, i) + T(i, i + 1));
    T(i + 1, i) = T(i, i + 1) = e;
  }
  auto rot = [&](int k, double c, double s) {
    // T <- R T R^T with R = [c s; -s c] on (k, k+1); touches columns k-1..k+2
    const int j0 = std::max(0, k - 1), j1 = std::min(n - 1, k + 2);
    for (int j = j0; j <= j1; ++j) {
      const double x = T(k, j), y = T(k + 1, j);
      T(k, j) = c * x + s * y;
      T(k + 1, j) = -s * x + c * y;
    }
    for (int i = j0; i <= j1; ++i) {
      const double x = T(i, k), y = T(i, k + 1);
      T(i, k) = c * x + s * y;
      T(i, k + 1) = -s * x + c * y;
    }
    for (int i = 0; i < n; ++i) {
      const double x = Q(i, k), y = Q(i, k + 1);
      Q(i, k) = c * x + s * y;
      Q(i, k + 1) = -s * x + c * y;
    }
  };
  int hi = n - 1, guard = 0;
  while (hi > 0) {
    if (++guard > 60 * n + 100)
      return false;
    for (int i = 0; i < hi; ++i)
      if (std::fabs(T(i + 1, i)) <= kEps * (std::fabs(T(i, i)) + std::fabs(T(i + 1, i + 1))))
        T(i + 1, i) = T(i, i + 1) = 0.0;
    if (T(hi, hi - 1) == 0.0) {
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && T(lo, lo - 1) != 0.0)
      --lo;
    const double a = T(hi - 1, hi - 1), b = T(hi, hi - 1), c = T(hi, hi);
    const double d = 0.5 * (a - c);
    const double mu = c - b * b / (d + (d >= 0 ? 1.0 : -1.0) * std::hypot(d, b));
    double x = T(lo, lo) - mu, z = T(lo + 1, lo);
    for (int k = lo; k < hi; ++k) {
      const double r = std::hypot(x, z);
      const double cs = r == 0.0 ? 1.0 : x / r, sn = r == 0.0 ? 0.0 : z / r;
      rot(k, cs, sn);
      if (k + 1 < hi) {
        x = T(k + 1, k);
        z = T(k + 2, k);
      }
    }
    // bulge chase leaves exact zeros outside the band only up to rounding
    for (int i = lo; i + 2 <= hi; ++i)
      T(i + 2, i) = T(i, i + 2) = 0.0;
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return T(x, x) < T(y, y); });
  w.assign(n, 0.0);
  Mat Qs(n, n);
  for (int j = 0; j < n; ++j) {
    w[j] = T(ord[j], ord[j]);
    for (int i = 0; i < n; ++i)
      Qs(i, j) = Q(i, ord[j]);
  }
  Q = std::move(Qs);
  return true;
}

bool cholesky(const Mat &A, Mat &L) {
  const int n = A.r;
  L = Mat(n, n);
  for (int j = 0; j < n; ++j) {
    double d = A(j, j);
    for (int k = 0; k < j; ++k)
      d -= L(j, k) * L(j, k);
    if (!(d > 0.0))
      return false;
    const double l = std::sqrt(d);
    L(j, j) = l;
    for (int i = j + 1; i < n; ++i) {
      double s = A(i, j);
      for (int k = 0; k < j; ++k)
        s -= L(i, k) * L(j, k);
      L(i, j) = s / l;
    }
  }
  return true;
}

Mat lower_inverse(const Mat &L) {
  const int n = L.r;
  Mat X(n, n);
  for (int j = 0; j < n; ++j) {
    X(j, j) = 1.0 / L(j, j);
    for (int i = j + 1; i < n; ++i) {
      double s = 0.0;
      for (int k = j; k < i; ++k)
        s += L(i, k) * X(k, j);
      X(i, j) = -s / L(i, i);
    }
  }
  return X;
}

}  // namespace kh::dense
