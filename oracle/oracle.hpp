// CPU oracle: an FP64, single-threaded restatement of KraNS's cyclic-Lanczos
// and LOBPCG null-space paths, used ONLY as a checker by tests/, the smoke
// check and the CPU-baseline leg of bench.py. It is test infrastructure, not
// part of the product: the HIP library never links or calls it.
//
// Every routine cites the reference file:line it restates. Vectors are plain
// std::vector<double> with the arithmetic of StdVectorAdapter
// (include/krans/adapters/universal/StdVectorAdapter.hpp:23-98): sequential
// dot/norm in ascending index order, no FMA contraction (built with
// -ffp-contract=off).
//
// Parity status of each part (how it is pinned) is listed in the header of
// oracle/capi.cpp.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using Vec = std::vector<double>;

// ---- level-1 arithmetic, StdVectorAdapter.hpp:30-98 ------------------------
inline double dot(const Vec &a, const Vec &b) {
  if (a.size() != b.size())
    throw std::invalid_argument("oracle dot: size mismatch");
  double r = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i)
    r += a[i] * b[i];
  return r;
}
inline void axpy(Vec &y, const Vec &x, double alpha) { // y += alpha x
  if (y.size() != x.size())
    throw std::invalid_argument("oracle add_scaled: size mismatch");
  for (std::size_t i = 0; i < y.size(); ++i)
    y[i] += alpha * x[i];
}
inline void scal(Vec &y, double alpha) {
  for (double &v : y)
    v *= alpha;
}
inline double nrm2(const Vec &a) {
  double s = 0.0;
  for (double v : a)
    s += v * v;
  return std::sqrt(s);
}

// ---- operators (MatrixInterface.hpp:10-24) ----------------------------------
struct Op {
  virtual ~Op() = default;
  virtual std::size_t rows() const = 0;
  virtual std::size_t cols() const = 0;
  virtual void mv(const Vec &x, Vec &y) const = 0;  // y = A x
  virtual void mtv(const Vec &x, Vec &y) const = 0; // y = A^T x
};

// CSR matrix with its CSC twin. Products accumulate every output entry
// sequentially in ascending inner index from 0.0, which is the order of
// Eigen's column-major sparse x dense product used by
// EigenSparseMatrixAdapter::multiply/_transpose
// (EigenSparseMatrixAdapter.hpp:34-57; A_ is CSC, AT_ = A_^T).
struct CsrOp final : Op {
  std::size_t m = 0, n = 0;
  std::vector<int32_t> rp, ci;  // CSR
  Vec va;
  std::vector<int32_t> cp, ri;  // CSC twin
  Vec vt;
  std::vector<int32_t> perm;    // CSC position -> CSR position
  CsrOp(std::size_t m_, std::size_t n_, const int32_t *rowptr,
        const int32_t *col, const double *val);
  std::size_t rows() const override { return m; }
  std::size_t cols() const override { return n; }
  void mv(const Vec &x, Vec &y) const override;
  void mtv(const Vec &x, Vec &y) const override;
};

// Matrix-free shifted difference (A x)_i = x_{i+N} - c x_i, (n-N) x n,
// exactly the lambdas of examples/quickstart.cpp:33-40.
struct ShiftDiffOp final : Op {
  std::size_t n, N;
  double c;
  ShiftDiffOp(std::size_t n_, std::size_t N_, double c_) : n(n_), N(N_), c(c_) {}
  std::size_t rows() const override { return n - N; }
  std::size_t cols() const override { return n; }
  void mv(const Vec &x, Vec &y) const override {
    for (std::size_t i = 0; i + N < n; ++i)
      y[i] = x[i + N] - c * x[i];
  }
  void mtv(const Vec &x, Vec &y) const override {
    for (std::size_t j = 0; j < n; ++j)
      y[j] = (j >= N ? x[j - N] : 0.0) - (j + N < n ? c * x[j] : 0.0);
  }
};

// CyclicLanczosSolver.cpp:266-281
struct TransposedOp final : Op {
  const Op &A;
  explicit TransposedOp(const Op &a) : A(a) {}
  std::size_t rows() const override { return A.cols(); }
  std::size_t cols() const override { return A.rows(); }
  void mv(const Vec &x, Vec &y) const override { A.mtv(x, y); }
  void mtv(const Vec &x, Vec &y) const override { A.mv(x, y); }
};

// ---- vector factory: StdVectorFactory.hpp:16-58 ------------------------------
// std::mt19937 seeded with `seed` and one std::normal_distribution<double>
// whose cached second value persists across calls (libstdc++ polar method).
struct Factory {
  std::mt19937 gen;
  std::normal_distribution<double> dist;
  explicit Factory(unsigned seed) : gen(seed), dist(0.0, 1.0) {}
  Vec randn(std::size_t n) {
    Vec v(n);
    for (std::size_t i = 0; i < n; ++i)
      v[i] = dist(gen);
    return v;
  }
  Vec zeros(std::size_t n) { return Vec(n, 0.0); }
};

// ---- tiny dense FP64 kernels (replace Eigen's JacobiSVD/LLT/SAES) -----------
// Column-major dense matrix.
struct Mat {
  int r = 0, c = 0;
  std::vector<double> a;
  Mat() = default;
  Mat(int r_, int c_) : r(r_), c(c_), a((std::size_t)r_ * c_, 0.0) {}
  double &operator()(int i, int j) { return a[(std::size_t)j * r + i]; }
  double operator()(int i, int j) const { return a[(std::size_t)j * r + i]; }
};
// One-sided (Hestenes) Jacobi SVD: singular values descending (Eigen's
// JacobiSVD order) and the full right factor V (c x c).
void svd_jacobi(const Mat &A, std::vector<double> &sv, Mat &V);
// Cyclic Jacobi symmetric eigensolver, eigenvalues ascending (Eigen's
// SelfAdjointEigenSolver order), eigenvectors as columns.
void sym_eig(const Mat &A, std::vector<double> &w, Mat &Q);
// Cholesky A = L L^T; false if a pivot is not positive (Eigen LLT::info()).
bool cholesky(const Mat &A, Mat &L);
// Inverse of a lower-triangular matrix.
Mat lower_inverse(const Mat &L);

// ---- options / results (CyclicLanczosSolver.hpp:84-167) ----------------------
enum class Reorth { None = 0, Single = 1, Double = 2 };
using Precond = std::function<void(const Vec &r, Vec &y)>;

struct CyclicOpts {
  std::size_t k = 1, ncv = 48, max_restarts = 80;
  double tol_abs = 1e-12, tol_rel = 1e-10;
  Reorth reorth = Reorth::Double;
  bool use_harmonic = true;
  double target_tau = 0.0;
  int nullity_known = -1, rank_known = -1;
  std::vector<Vec> initial_right_space;
  Precond precond_AAt = nullptr;
  std::size_t cg_maxit = 80;
  double cg_rel_tol = 1e-10, cg_breakdown_tol = 1e-30;
  int projector_passes = 1;
  int grad_max_iters = 3000;
  int frob_samples = 4;
  double prescreen_factor = 0.8;
  std::size_t stagnation_restarts = 30;
  bool return_left_space = false;
  bool verbose = false;
};
struct CyclicResult {
  std::vector<Vec> right_null, left_null;
  std::size_t restarts = 0, krylov_steps_total = 0;
  double frob_norm_est = 0.0;
  bool converged_to_target = false;
  std::size_t cg_solves = 0, cg_maxit_hits = 0;
  // oracle-only counters (operator applications), like CountingMatrixAdapter
  // (tests/test_cyclic_lanczos.cpp:44-71)
  std::size_t spmv = 0, spmtv = 0;
};
CyclicResult cyclic_lanczos(const Op &A, Factory &vf, const CyclicOpts &opt);

// LOBPCGSolver.hpp:71-139
struct LobpcgOpts {
  std::size_t nev = 1, block_size = 0, max_iters = 1000;
  double tol_abs = 1e-12, tol_rel = 1e-10;
  bool soft_locking = true;
  double restart = 0.9, gs_tol = 1e-14;
  bool gs_reorth = true;
  Precond M_inv = nullptr;
  std::vector<Vec> initial_block;
  int frob_samples = 4;
  std::size_t max_tiny_dim = 0;
  int nullity_known = -1, rank_known = -1;
  bool verbose = false;
  std::size_t progress_every = 100;
};
struct LobpcgResult {
  std::vector<Vec> right_null;
  std::vector<double> theta, residuals_Av;
  std::size_t iters = 0, locked = 0;
  double frob_norm_est = 0.0, accept_threshold = 0.0,
         best_unaccepted_residual = 0.0;
  std::size_t block_size = 0;
  bool converged_to_target = false;
  std::size_t spmv = 0, spmtv = 0;
};
LobpcgResult lobpcg(const Op &A, Factory &vf, const LobpcgOpts &opt);

// ---- Gram-Schmidt, GramSchmidt.cpp:25-91 -------------------------------------
bool modified_gram_schmidt(std::vector<Vec> &vecs, double tol, bool reorth);
double orthogonalize_against_basis(Vec &v, const std::vector<Vec> &basis,
                                   double tol, bool reorth);
double max_pairwise_abs_dot(const std::vector<Vec> &basis);

// ---- basis comparison metrics ------------------------------------------------
// SubspaceCompare.hpp:17-110 (rows = basis vectors).
void compute_rref(std::vector<Vec> &rows, double tol);
double subspace_diff(const std::vector<Vec> &b1, const std::vector<Vec> &b2,
                     double tol);

// Matrix loaders (src/IO/MatrixLoader.cpp) producing Eigen's compressed CSC.
struct Csc {
  int64_t rows = 0, cols = 0, entries_read = 0;
  std::vector<int64_t> colptr;
  std::vector<int> rowidx;
  std::vector<double> val;
};
Csc set_from_triplets(int64_t rows, int64_t cols, const std::vector<int> &I,
                      const std::vector<int> &J, const std::vector<double> &V);
Csc load_ija(const std::string &path, int rows, int cols);
Csc load_mtx(const std::string &path);

} // namespace orc
