// Tiny dense FP64 kernels for the oracle. They replace the Eigen calls of the
// reference (JacobiSVD at CyclicLanczosSolver.cpp:246-262, LLT + inverse +
// SelfAdjointEigenSolver at LOBPCGSolver.cpp:88-119 and :255-263) with
// textbook algorithms (Golub & Van Loan, 4th ed., 8.5 and 8.6.3): one-sided
// Jacobi SVD, cyclic Jacobi symmetric eigensolver, unblocked Cholesky.
// Pinned against LAPACK (numpy) on committed fixtures in tests/golden/.
#include "oracle.hpp"

#include <algorithm>
#include <limits>
#include <numeric>

namespace orc {

static const double kEps = std::numeric_limits<double>::epsilon();

void svd_jacobi(const Mat &A, std::vector<double> &sv, Mat &V) {
  const int m = A.r, n = A.c;
  Mat U = A;
  V = Mat(n, n);
  for (int i = 0; i < n; ++i)
    V(i, i) = 1.0;
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int i = 0; i < m; ++i) {
          const double up = U(i, p), uq = U(i, q);
          alpha += up * up;
          beta += uq * uq;
          gamma += up * uq;
        }
        if (gamma == 0.0 || std::fabs(gamma) <= kEps * std::sqrt(alpha * beta))
          continue;
        rotated = true;
        const double zeta = (beta - alpha) / (2.0 * gamma);
        const double t =
            std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t);
        const double s = c * t;
        for (int i = 0; i < m; ++i) {
          const double up = U(i, p), uq = U(i, q);
          U(i, p) = c * up - s * uq;
          U(i, q) = s * up + c * uq;
        }
        for (int i = 0; i < n; ++i) {
          const double vp = V(i, p), vq = V(i, q);
          V(i, p) = c * vp - s * vq;
          V(i, q) = s * vp + c * vq;
        }
      }
    }
    if (!rotated)
      break;
  }
  std::vector<double> s(n);
  for (int j = 0; j < n; ++j) {
    double t = 0.0;
    for (int i = 0; i < m; ++i)
      t += U(i, j) * U(i, j);
    s[j] = std::sqrt(t);
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
  sv.resize(n);
  Mat Vs(n, n);
  for (int j = 0; j < n; ++j) {
    sv[j] = s[ord[j]];
    for (int i = 0; i < n; ++i)
      Vs(i, j) = V(i, ord[j]);
  }
  V = Vs;
}

void sym_eig(const Mat &Ain, std::vector<double> &w, Mat &Q) {
  const int n = Ain.r;
  Mat A = Ain;
  Q = Mat(n, n);
  for (int i = 0; i < n; ++i)
    Q(i, i) = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A(p, q);
        if (apq == 0.0)
          continue;
        const double app = A(p, p), aqq = A(q, q);
        if (std::fabs(apq) <= 0.5 * kEps * std::sqrt(std::fabs(app) * std::fabs(aqq)) ||
            std::fabs(apq) < 1e-300)
          continue;
        rotated = true;
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = std::copysign(1.0, theta) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0);
        const double s = t * c;
        // A <- J^T A J with J the (p,q) rotation [c s; -s c]
        for (int k = 0; k < n; ++k) {
          const double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        A(p, q) = 0.0;
        A(q, p) = 0.0;
        for (int k = 0; k < n; ++k) {
          const double qkp = Q(k, p), qkq = Q(k, q);
          Q(k, p) = c * qkp - s * qkq;
          Q(k, q) = s * qkp + c * qkq;
        }
      }
    }
    if (!rotated)
      break;
  }
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return A(a, a) < A(b, b); });
  w.resize(n);
  Mat Qs(n, n);
  for (int j = 0; j < n; ++j) {
    w[j] = A(ord[j], ord[j]);
    for (int i = 0; i < n; ++i)
      Qs(i, j) = Q(i, ord[j]);
  }
  Q = Qs;
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
    const double ljj = std::sqrt(d);
    L(j, j) = ljj;
    for (int i = j + 1; i < n; ++i) {
      double s = A(i, j);
      for (int k = 0; k < j; ++k)
        s -= L(i, k) * L(j, k);
      L(i, j) = s / ljj;
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

} // namespace orc
