// Matrix-free device operators: the device analogue of the reference's
// StdMatrixAdapter (include/krans/adapters/universal/StdMatrixAdapter.hpp:
// 24-87), whose y = A x and y = A^T x are host std::functions.
//  - StencilOp: the shifted difference (A x)_i = x_{i+N} - c x_i of
//    examples/quickstart.cpp as HIP kernels, with no matrix in HBM;
//  - FunctorOp: a caller-supplied C callback that enqueues its own device
//    work on the solver's stream;
//  - TransposedOpView: op(A) = A^T of any operator (the solvers' left space).
#pragma once
#include <cstdint>

#include "solver.hpp"

namespace kh {

struct StencilOp final : DeviceOp {
  int64_t n, N;
  double c;
  DBuf diag;  // diag(A A^T), m doubles
  double fro_ = 0.0;
  StencilOp(Context &ctx, int64_t n, int64_t N, double c);
  int64_t rows() const override { return n - N; }
  int64_t cols() const override { return n; }
  void mv(Context &ctx, Prec p, const void *x, void *y, const void *z,
          const double *beta) const override;
  void mtv(Context &ctx, Prec p, const void *x, void *y, const void *z,
           const double *beta) const override;
  void mm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *act) const override;
  void mtm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
           const int *act) const override;
  bool has_rows() const override { return true; }
  bool capturable() const override { return true; }
  void mm_rows(Context &ctx, Prec p, int s, const void *X, void *Y, bool trans) const override;
  const double *diag_AAt() const override { return diag.as<double>(); }
  double fro() const override { return fro_; }
};

// apply(user, stream, precision, transpose, s, x, ldx, y, ldy): y(:, j) =
// op(A) x(:, j) for j < s, op = A (transpose = 0) or A^T, FP64 (precision 0)
// or FP32 (1) device columns with leading dimensions ldx / ldy; the work is
// enqueued on `stream` (a hipStream_t) or finished before returning; 0 means
// success. It is called with all earlier work on the stream complete.
typedef int (*ApplyFn)(void *user, void *stream, int precision, int transpose, int s,
                       const void *x, int64_t ldx, void *y, int64_t ldy);

struct FunctorOp final : DeviceOp {
  int64_t m, n;
  ApplyFn fn;
  void *user;
  mutable DBuf scratch;
  FunctorOp(int64_t m_, int64_t n_, ApplyFn f, void *u) : m(m_), n(n_), fn(f), user(u) {}
  int64_t rows() const override { return m; }
  int64_t cols() const override { return n; }
  void mv(Context &ctx, Prec p, const void *x, void *y, const void *z,
          const double *beta) const override;
  void mtv(Context &ctx, Prec p, const void *x, void *y, const void *z,
           const double *beta) const override;
  void mm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *act) const override;
  void mtm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
           const int *act) const override;

private:
  void call(Context &ctx, Prec p, bool trans, int s, const void *x, int64_t ldx, void *y,
            int64_t ldy) const;
  void block(Context &ctx, Prec p, bool trans, int s, const void *x, int64_t ldx, void *y,
             int64_t ldy, const int *act) const;
};

struct TransposedOpView final : DeviceOp {
  const DeviceOp &A;
  explicit TransposedOpView(const DeviceOp &a) : A(a) {}
  int64_t rows() const override { return A.cols(); }
  int64_t cols() const override { return A.rows(); }
  void mv(Context &c, Prec p, const void *x, void *y, const void *z,
          const double *beta) const override {
    A.mtv(c, p, x, y, z, beta);
  }
  void mtv(Context &c, Prec p, const void *x, void *y, const void *z,
           const double *beta) const override {
    A.mv(c, p, x, y, z, beta);
  }
  void mm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
          const int *act) const override {
    A.mtm(c, p, s, x, ldx, y, ldy, act);
  }
  void mtm(Context &c, Prec p, int s, const void *x, int64_t ldx, void *y, int64_t ldy,
           const int *act) const override {
    A.mm(c, p, s, x, ldx, y, ldy, act);
  }
  bool has_rows() const override { return A.has_rows(); }
  bool capturable() const override { return A.capturable(); }
  void mm_rows(Context &c, Prec p, int s, const void *X, void *Y, bool trans) const override {
    A.mm_rows(c, p, s, X, Y, !trans);
  }
  double fro() const override { return A.fro(); }
};

}  // namespace kh
