// HIP backend adapters for the KraNS plugin API: HipVector (VectorInterface on
// device memory), HipVectorFactory (VectorFactory with the reference's seeded
// stream, StdVectorFactory.hpp:16-58) and HipCsrMatrixAdapter
// (MatrixInterface over a device CSR with its CSC twin; it replaces
// EigenSparseMatrixAdapter, include/krans/adapters/eigen/
// EigenSparseMatrixAdapter.hpp:12-68). Header-only over the C ABI of
// include/krans_hip.h; link libkrans_hip.so.
//
// data() returns nullptr (allowed by VectorInterface.hpp:30-33): vectors live
// in HBM; use to_host() to download.
#pragma once
#include <cmath>
#include <cstring>
#include <random>
#include <functional>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "krans/MatrixInterface.hpp"
#include "krans/VectorFactory.hpp"
#include "krans/VectorInterface.hpp"
#include "krans_hip.h"

namespace krans::hip {

inline void check(int rc) {
  if (rc == KH_OK)
    return;
  const std::string msg = kh_last_error();
  if (rc == KH_EINVAL)
    throw std::invalid_argument(msg);
  if (rc == KH_ENOMEM)
    throw std::bad_alloc();
  throw std::runtime_error(msg);
}

enum class Precision { FP64 = KH_PREC_FP64, Mixed = KH_PREC_MIXED };

/** Owns one kh_context (device + stream). */
class HipContext {
public:
  explicit HipContext(int device = 0) { check(kh_context_create(device, &c_)); }
  ~HipContext() { kh_context_destroy(c_); }
  HipContext(const HipContext &) = delete;
  HipContext &operator=(const HipContext &) = delete;
  kh_context *get() const { return c_; }

private:
  kh_context *c_ = nullptr;
};

class HipVector final : public VectorInterface {
public:
  HipVector(kh_context *ctx, std::size_t n) : ctx_(ctx), n_(n) {
    check(kh_dev_alloc(ctx_, n_ * sizeof(double), reinterpret_cast<void **>(&d_)));
    set_zero();
  }
  HipVector(kh_context *ctx, std::size_t n, const double *host) : ctx_(ctx), n_(n) {
    check(kh_dev_alloc(ctx_, n_ * sizeof(double), reinterpret_cast<void **>(&d_)));
    if (n_)
      check(kh_h2d(ctx_, d_, host, n_ * sizeof(double)));
  }
  /** Adopt a device allocation made with kh_dev_alloc. */
  struct Adopt {};
  HipVector(kh_context *ctx, std::size_t n, double *dev, Adopt) : ctx_(ctx), n_(n), d_(dev) {}
  /** A non-owning view of n device doubles (valid while the memory is):
   * how the solvers hand their device columns to user operators and
   * preconditioners. */
  struct View {};
  HipVector(kh_context *ctx, std::size_t n, double *dev, View)
      : ctx_(ctx), n_(n), d_(dev), own_(false) {}
  ~HipVector() override {
    if (d_ && own_)
      kh_dev_free(ctx_, d_);
  }
  HipVector(const HipVector &) = delete;
  HipVector &operator=(const HipVector &) = delete;

  double dot(const VectorInterface &other) const override {
    const HipVector &o = same(other, "dot");
    double r = 0.0;
    check(kh_dot_dev(ctx_, KH_PREC_FP64, d_, o.d_, (int64_t)n_, &r));
    return r;
  }
  void add_scaled(const VectorInterface &other, double alpha) override {
    const HipVector &o = same(other, "add_scaled");
    check(kh_vec_axpby(ctx_, (int64_t)n_, alpha, o.d_, 1.0, d_));
  }
  void scale(double alpha) override {
    check(kh_vec_axpby(ctx_, (int64_t)n_, 0.0, nullptr, alpha, d_));
  }
  void normalize() override {
    const double nn = norm();
    if (nn > 0.0)
      scale(1.0 / nn);
  }
  double norm() const override {
    double r = 0.0;
    check(kh_dot_dev(ctx_, KH_PREC_FP64, d_, d_, (int64_t)n_, &r));
    return r > 0.0 ? std::sqrt(r) : 0.0;
  }
  std::size_t size() const override { return n_; }
  void set_zero() override { check(kh_vec_axpby(ctx_, (int64_t)n_, 0.0, nullptr, 0.0, d_)); }
  void copy_from(const VectorInterface &other) override {
    const HipVector &o = same(other, "copy_from");
    check(kh_vec_axpby(ctx_, (int64_t)n_, 1.0, o.d_, 0.0, d_));
  }
  std::unique_ptr<VectorInterface> clone() const override {
    auto v = std::make_unique<HipVector>(ctx_, n_);
    check(kh_vec_axpby(ctx_, (int64_t)n_, 1.0, d_, 0.0, v->d_));
    return v;
  }
  double *data() override { return nullptr; }
  const double *data() const override { return nullptr; }

  double *device_ptr() { return d_; }
  const double *device_ptr() const { return d_; }
  kh_context *context() const { return ctx_; }
  std::vector<double> to_host() const {
    std::vector<double> h(n_);
    if (n_)
      check(kh_d2h(ctx_, h.data(), d_, n_ * sizeof(double)));
    return h;
  }

private:
  const HipVector &same(const VectorInterface &other, const char *what) const {
    const auto *o = dynamic_cast<const HipVector *>(&other);
    if (!o)
      throw std::invalid_argument(std::string("HipVector::") + what + " backend mismatch");
    if (o->n_ != n_)
      throw std::invalid_argument(std::string("HipVector::") + what + " size mismatch");
    return *o;
  }
  kh_context *ctx_;
  std::size_t n_;
  double *d_ = nullptr;
  bool own_ = true;
};

/** VectorFactory with the reference's seeded N(0,1) stream; `precision`
 * selects the solvers' arithmetic for solves that use this factory. */
class HipVectorFactory final : public VectorFactory {
public:
  HipVectorFactory(kh_context *ctx, std::size_t default_size, unsigned seed = 0,
                   Precision precision = Precision::FP64)
      : ctx_(ctx), n_(default_size), prec_(precision) {
    if (seed == 0)
      seed = std::random_device{}();
    seed_ = seed;
    check(kh_rng_create(seed, &rng_));
  }
  ~HipVectorFactory() override { kh_rng_destroy(rng_); }
  HipVectorFactory(const HipVectorFactory &) = delete;
  HipVectorFactory &operator=(const HipVectorFactory &) = delete;

  std::unique_ptr<VectorInterface> create_zero() const override { return create_zero_of_size(n_); }
  std::unique_ptr<VectorInterface> create_random() const override {
    return create_random_of_size(n_);
  }
  std::unique_ptr<VectorInterface> create_zero_of_size(std::size_t n) const override {
    return std::make_unique<HipVector>(ctx_, n);
  }
  std::unique_ptr<VectorInterface> create_random_of_size(std::size_t n) const override {
    std::vector<double> h(n);
    check(kh_rng_normal(rng_, (int64_t)n, h.data()));
    return std::make_unique<HipVector>(ctx_, n, h.data());
  }
  unsigned seed() const { return seed_; }
  kh_rng *rng() const { return rng_; }
  kh_context *context() const { return ctx_; }
  Precision precision() const { return prec_; }
  void set_precision(Precision p) { prec_ = p; }

private:
  kh_context *ctx_;
  std::size_t n_;
  unsigned seed_ = 0;
  Precision prec_;
  kh_rng *rng_ = nullptr;
};

/** Common base of the HIP matrix operands: a kh_matrix handle (CSR or
 * matrix-free) on one context. The solvers accept any HipMatrix. */
class HipMatrix : public MatrixInterface {
public:
  ~HipMatrix() override { kh_matrix_destroy(A_); }
  HipMatrix(const HipMatrix &) = delete;
  HipMatrix &operator=(const HipMatrix &) = delete;

  void multiply(const VectorInterface &x, VectorInterface &y) const override {
    apply(x, y, 0, "multiply");
  }
  void multiply_transpose(const VectorInterface &x, VectorInterface &y) const override {
    apply(x, y, 1, "multiply_transpose");
  }
  std::size_t rows() const override { return m_; }
  std::size_t cols() const override { return n_; }
  /** Stored entries (CSR), the entries a CSR would hold (stencil), or
   * SIZE_MAX when unknown (callback operator). */
  std::size_t nnz() const { return nnz_; }
  double frobenius_norm() const { return fro_; }
  kh_matrix *handle() const { return A_; }
  kh_context *context() const { return ctx_; }

protected:
  explicit HipMatrix(kh_context *ctx) : ctx_(ctx) {}
  void adopt(kh_matrix *A) {
    A_ = A;
    int64_t mm = 0, nn = 0, nz = 0;
    check(kh_matrix_info(A_, &mm, &nn, &nz, &fro_));
    m_ = (std::size_t)mm;
    n_ = (std::size_t)nn;
    nnz_ = nz < 0 ? (std::size_t)-1 : (std::size_t)nz;
  }

private:
  void apply(const VectorInterface &x, VectorInterface &y, int t, const char *what) const {
    const auto *hx = dynamic_cast<const HipVector *>(&x);
    auto *hy = dynamic_cast<HipVector *>(&y);
    if (!hx || !hy)
      throw std::invalid_argument(std::string("HipMatrix::") + what + " backend mismatch");
    const std::size_t nin = t ? m_ : n_, nout = t ? n_ : m_;
    if (hx->size() != nin || hy->size() != nout)
      throw std::invalid_argument(std::string("HipMatrix::") + what + " size mismatch");
    // vectors of another context live on another stream: order the product
    // after their pending work and finish it before returning
    const bool other = hx->context() != ctx_ || hy->context() != ctx_;
    if (other) {
      check(kh_context_sync(hx->context()));
      if (hy->context() != hx->context())
        check(kh_context_sync(hy->context()));
    }
    check(kh_spmm_dev(A_, t, KH_PREC_FP64, 1, hx->device_ptr(), hy->device_ptr()));
    if (other)
      check(kh_context_sync(ctx_));
  }
  kh_context *ctx_;
  kh_matrix *A_ = nullptr;
  std::size_t m_ = 0, n_ = 0, nnz_ = 0;
  double fro_ = 0.0;
};

/** Replaces EigenSparseMatrixAdapter: canonical CSR on the host (see
 * kh_matrix_create_csr), held in HBM with its CSC twin. */
class HipCsrMatrixAdapter final : public HipMatrix {
public:
  HipCsrMatrixAdapter(kh_context *ctx, std::size_t m, std::size_t n, const int32_t *rowptr,
                      const int32_t *col, const double *val)
      : HipMatrix(ctx) {
    kh_matrix *A = nullptr;
    check(kh_matrix_create_csr(ctx, (int64_t)m, (int64_t)n, rowptr, col, val, &A));
    adopt(A);
  }
};

/** A matrix file loaded into HBM: the IJA triplet format of the reference's
 * load_ija_triplet_matrix (include/krans/IO/MatrixLoader.hpp; rows/cols < 0
 * take the sizes from the file) or Matrix Market (load_matrix_market), with
 * repeated entries summed in file order. Throws std::runtime_error when the
 * file cannot be read, as the reference's loaders do. */
inline std::unique_ptr<HipCsrMatrixAdapter> load_matrix(kh_context *ctx, const std::string &path,
                                                        long rows = -1, long cols = -1) {
  kh_host_csr *h = nullptr;
  const bool mtx = path.size() > 4 && path.compare(path.size() - 4, 4, ".mtx") == 0;
  const int rc = mtx ? kh_load_matrix_market(path.c_str(), &h)
                     : kh_load_ija(path.c_str(), rows, cols, &h);
  if (rc != KH_OK)
    throw std::runtime_error(kh_last_error());
  int64_t m = 0, n = 0, nnz = 0;
  const int32_t *rp = nullptr, *ci = nullptr;
  const double *va = nullptr;
  kh_host_csr_info(h, &m, &n, &nnz, nullptr, nullptr);
  kh_host_csr_data(h, &rp, &ci, &va);
  try {
    auto A = std::make_unique<HipCsrMatrixAdapter>(ctx, (std::size_t)m, (std::size_t)n, rp, ci, va);
    kh_host_csr_free(h);
    return A;
  } catch (...) {
    kh_host_csr_free(h);
    throw;
  }
}

/** The quick start's matrix-free operator (examples/quickstart.cpp:25-37,
 * a StdMatrixAdapter of two lambdas): (A x)_i = x_{i+N} - c x_i, A is
 * (n-N) x n, applied by HIP kernels with no matrix in memory. */
class HipShiftedDifferenceOperator final : public HipMatrix {
public:
  HipShiftedDifferenceOperator(kh_context *ctx, std::size_t n, std::size_t N, double c)
      : HipMatrix(ctx) {
    kh_matrix *A = nullptr;
    check(kh_matrix_create_shifted_difference(ctx, (int64_t)n, (int64_t)N, c, &A));
    adopt(A);
  }
};

/** The device analogue of StdMatrixAdapter (include/krans/adapters/
 * universal/StdMatrixAdapter.hpp:24-87): a user operator that launches its
 * own device work. apply(stream, precision, transpose, s, x, ldx, y, ldy)
 * computes y(:, j) = op(A) x(:, j), j < s, on device columns of double
 * (Precision::FP64) or float (Precision::Mixed), enqueued on `stream` (a
 * hipStream_t) or finished before returning; throwing reports failure. It
 * is called with all earlier work on the stream complete. */
class HipOperatorAdapter final : public HipMatrix {
public:
  using Apply = std::function<void(void *stream, Precision p, bool transpose, int s,
                                   const void *x, int64_t ldx, void *y, int64_t ldy)>;
  HipOperatorAdapter(kh_context *ctx, std::size_t m, std::size_t n, Apply apply)
      : HipMatrix(ctx), apply_(std::move(apply)) {
    kh_matrix *A = nullptr;
    check(kh_matrix_create_operator(ctx, (int64_t)m, (int64_t)n, &HipOperatorAdapter::call,
                                    this, &A));
    adopt(A);
  }

private:
  static int call(void *user, void *stream, int precision, int transpose, int s, const void *x,
                  int64_t ldx, void *y, int64_t ldy) {
    try {
      static_cast<HipOperatorAdapter *>(user)->apply_(
          stream, precision == KH_PREC_FP64 ? Precision::FP64 : Precision::Mixed,
          transpose != 0, s, x, ldx, y, ldy);
      return 0;
    } catch (...) {
      return 1;
    }
  }
  Apply apply_;
};

/** M^{-1} = diag(A A^T)^{-1} for the polish CG (CyclicLanczosOptions::
 * precond_AAt). The HIP solver recognises it and applies it inside its CG
 * kernels; called directly it runs the same device kernel. */
struct JacobiAAtPreconditioner {
  const HipMatrix *A;
  void operator()(const VectorInterface &r, VectorInterface &y) const {
    const auto *hr = dynamic_cast<const HipVector *>(&r);
    auto *hy = dynamic_cast<HipVector *>(&y);
    if (!hr || !hy)
      throw std::invalid_argument("JacobiAAtPreconditioner: backend mismatch");
    check(kh_matrix_jacobi_AAt(A->handle(), hr->device_ptr(), hy->device_ptr()));
  }
};

}  // namespace krans::hip
