// Host FP64 kernels for the tiny projected problems: the (l+1) x l refined
// Golub-Kahan matrix (CyclicLanczosSolver.cpp:246-262) and the LOBPCG
// Rayleigh-Ritz generalized eigenproblem (LOBPCGSolver.cpp:88-119). They run
// on the host in FP64: at most a few hundred rows, O(l^3).
#pragma once
#include <vector>

namespace kh::dense {

// Column-major dense matrix.
struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() = default;
  Mat(int rr, int cc) : r(rr), c(cc), a((std::size_t)rr * cc, 0.0) {}
  double &operator()(int i, int j) { return a[(std::size_t)j * r + i]; }
  double operator()(int i, int j) const { return a[(std::size_t)j * r + i]; }
};

// Right singular system of A (rows >= cols): singular values descending and
// V (cols x cols). One-sided Jacobi (Hestenes) with column-norm updates; with
// `split`, a matrix whose entries with i + j even are zero is swept as its
// two independent blocks (bitwise the same result, half the work). With
// even_only (and such a matrix) only singular vectors supported on even rows
// are formed; the other columns of V are unit vectors on odd rows, and all
// singular values are unchanged.
void svd_right(const Mat &A, std::vector<double> &sv, Mat &V, bool split = true,
               bool even_only = false);

// Symmetric eigen-decomposition, eigenvalues ascending, eigenvectors in the
// columns of Q: Householder tridiagonalisation + implicit QR with Wilkinson
// shifts (Golub & Van Loan, 4th ed., 8.3). Returns false if it fails to
// converge.
bool sym_eig(const Mat &A, std::vector<double> &w, Mat &Q);

// Lower Cholesky factor; false on a non-positive (or NaN) pivot.
bool cholesky(const Mat &A, Mat &L);
Mat lower_inverse(const Mat &L);

}  // namespace kh::dense
