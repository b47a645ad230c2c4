// Device-resident LOBPCG null-space driver.
//
// Follows find_nullspace_lobpcg_ex (src/Solvers/LOBPCGSolver.cpp:443-834)
// step for step: block size and guard rules, Hutchinson threshold, soft and
// hard locking, the [X W P] Rayleigh-Ritz with the jittered Cholesky GEP, the
// full RR for ordering, and the final block RR polish and 2x2 sharpening
// (:199-440), with the same order of random draws. Blocks live in HBM as
// column-major n x q arrays; the Gram blocks S = Z^T Z and G = Z^T (A^T A Z)
// are one batched V^T W product each; the tiny GEP runs on the host in FP64
// (dense.cpp).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>

#include "dense.hpp"
#include "solver.hpp"

namespace kh {
namespace {

using Clock = std::chrono::steady_clock;
inline double since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}
inline std::size_t esz(Prec p) { return p == Prec::F64 ? 8 : 4; }

void rb(Context &ctx, const double *d, int n, double *out) {
  if (n <= 0)
    return;
  if (n > Context::kScalars)
    throw std::length_error("readback too large");
  KH_CHECK(hipMemcpyAsync(ctx.hscal, d, sizeof(double) * (std::size_t)n,
                          hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  std::copy(ctx.hscal, ctx.hscal + n, out);
}

void up(Context &ctx, const double *h, int n, double *d) {
  if (n <= 0)
    return;
  if (n > Context::kScalars)
    throw std::length_error("upload too large");
  ctx.upload(d, h, sizeof(double) * (std::size_t)n);  // queued, no wait
}

// A block of `cap` columns of length n in precision p.
struct Blk {
  DBuf buf;
  int64_t n = 0;
  int cap = 0;
  Prec p = Prec::F64;
  void init(int64_t n_, int cap_, Prec p_) {
    n = n_;
    cap = cap_;
    p = p_;
    buf = DBuf((std::size_t)std::max(1, cap_) * (std::size_t)n_ * esz(p_));
  }
  void *col(int j) const { return static_cast<char *>(buf.p) + (std::size_t)j * n * esz(p); }
};

struct TinyGEP {
  std::vector<double> theta;
  dense::Mat Y;
};

// :88-119. FP64 blocks use the reference's jitter ladder. FP32 blocks carry
// Gram errors near the FP32 unit roundoff (6e-8), so their ladder starts
// there: a smaller ridge leaves a nearly dependent [X W P] with a negative
// pivot, or with a Cholesky factor whose inverse amplifies the FP32 noise in
// G. The ridge moves each Ritz value theta by about theta * ridge, which is
// negligible for the near-zero values this solver targets.
bool try_tiny_gep(dense::Mat &G, dense::Mat &S, bool f32, TinyGEP &out);

TinyGEP solve_tiny_gep(dense::Mat &G, dense::Mat &S, bool f32 = false) {
  TinyGEP out;
  if (!try_tiny_gep(G, S, f32, out))
    throw std::runtime_error("LOBPCG tiny GEP: S not SPD (after jitter)");
  return out;
}

bool try_tiny_gep(dense::Mat &G, dense::Mat &S, bool f32, TinyGEP &out) {
  const int k = G.r;
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < j; ++i) {
      const double g = 0.5 * (G(i, j) + G(j, i));
      G(i, j) = G(j, i) = g;
      const double s = 0.5 * (S(i, j) + S(j, i));
      S(i, j) = S(j, i) = s;
    }
  double fro = 0.0;
  for (double v : S.a)
    fro += v * v;
  const double s_norm = std::max(1.0, std::sqrt(fro));
  static const double j64[] = {0.0, 1e-14, 1e-12, 1e-10};
  static const double j32[] = {2e-7, 1e-6, 1e-5, 1e-4};
  for (double jv : f32 ? j32 : j64) {
    dense::Mat Sj = S;
    if (jv > 0.0)
      for (int i = 0; i < k; ++i)
        Sj(i, i) += jv * s_norm;
    dense::Mat L;
    if (!dense::cholesky(Sj, L))
      continue;
    dense::Mat Li = dense::lower_inverse(L);
    dense::Mat T(k, k), C(k, k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = 0; t <= i; ++t)
          s += Li(i, t) * G(t, j);
        T(i, j) = s;
      }
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = 0; t <= j; ++t)
          s += T(i, t) * Li(j, t);
        C(i, j) = s;
      }
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < j; ++i) {
        const double c = 0.5 * (C(i, j) + C(j, i));
        C(i, j) = C(j, i) = c;
      }
    dense::Mat Q;
    if (!dense::sym_eig(C, out.theta, Q))
      continue;
    out.Y = dense::Mat(k, k);
    for (int j = 0; j < k; ++j)
      for (int i = 0; i < k; ++i) {
        double s = 0.0;
        for (int t = i; t < k; ++t)
          s += Li(t, i) * Q(t, j);
        out.Y(i, j) = s;
      }
    return true;
  }
  return false;
}

class LobpcgDriver {
public:
  LobpcgDriver(Context &c, const DeviceOp &A, RefStream &rng, const LobpcgParams &P,
               LobpcgStats &S)
      : c_(c), A_(A), rng_(rng), P_(P), S_(S), m_(A.rows()), n_(A.cols()), bp_(P.block) {}
  DeviceBasis run();

private:
  Context &c_;
  const DeviceOp &A_;
  RefStream &rng_;
  const LobpcgParams &P_;
  LobpcgStats &S_;
  int64_t m_, n_;
  Prec bp_;
  DeviceBasis locked_;  // FP64 accepted vectors (Y_locked and R.right_null)
  DBuf small_, tmpm_, hbuf_dev_, v64_, v64b_, tm64_, tn64_, vc_, w64_;
  CgBatch cg_;
  double *sc_ = nullptr;
  std::vector<double> hbuf_;
  Blk tmpq_;  // m x q scratch for apply_normal
  int defer_ = 0;  // consecutive iterations whose few candidates wait to be refined together
  DBuf zr_, tr_;  // row-major n x q and m x q blocks of the row-major path
  DBuf vblk_, rhsblk_, wblk_, tnblk_;  // FP64 candidate refinement blocks
  DBuf hproj_;  // locked^T X coefficients of a block projection

  void random_col(void *dst, Prec p);
  double nrm(Prec p, const void *x, int64_t len);
  void project_out(void *x, Prec p);  // one pass vs locked (LOBPCGSolver.cpp:188-195)
  bool mgs(Blk &X, int i0, int cnt, double tol, bool reorth);
  bool orthonormalize(Blk &X, int i0, int cnt, double tol, bool reorth, Blk &scratch);
  void mgs_dev(Blk &X, int cnt, bool reorth, double *dn);
  bool cholqr2(Blk &X, int i0, int cnt, Blk &scratch);
  void apply_normal(const Blk &Z, int q, Blk &AZ);
  // Y = op(A) X for q FP64 columns (ld = input / output length), through
  // row-major copies for two or more columns
  void block_apply(int q, const void *X, void *Y, bool trans);
  void gram(const Blk &Z, const Blk &AZ, int q, dense::Mat &S, dense::Mat &G);
  void combine(const Blk &Z, int z0, int q, const dense::Mat &Y, int row0, int p,
               Blk &out, int o0);
  void to64(const void *x, Prec p, double *d) { convert(c_, p, x, Prec::F64, d, n_); }
  // x <- M^{-1} x for an FP64 device vector of length n (P_.M_inv)
  void precondition64(double *x) {
    if (minv_.bytes < (std::size_t)n_ * 8)
      minv_ = DBuf((std::size_t)n_ * 8);
    c_.sync();
    P_.M_inv->apply(c_, x, minv_.as<double>(), n_);
    copy(c_, Prec::F64, n_, minv_.p, x);
  }
  DBuf minv_;
  double estimate_fro();
  bool orthonormalize_vs(Blk &B, int k, int nd, Blk &scratch);
  void block_rr_polish(double target, int max_sweeps);
  void projector_polish(double target);
  bool two_vec_sharpen(int idx, double target, int max_iters);
  void mgs_locked(double tol);
};

void LobpcgDriver::random_col(void *dst, Prec p) {
  auto t0 = Clock::now();
  rng_.fill(hbuf_.data(), (std::size_t)n_);
  KH_CHECK(hipMemcpyAsync(v64_.p, hbuf_.data(), (std::size_t)n_ * 8, hipMemcpyHostToDevice,
                          c_.stream));
  convert(c_, Prec::F64, v64_.p, p, dst, n_);
  c_.sync();
  S_.t_host += since(t0);
  S_.t_rng += since(t0);
  ++S_.random_cols;
}

double LobpcgDriver::nrm(Prec p, const void *x, int64_t len) {
  dot(c_, p, x, p, x, len, sc_, true);
  double v;
  rb(c_, sc_, 1, &v);
  return v;
}

void LobpcgDriver::project_out(void *x, Prec p) {
  const int nb = (int)locked_.count;
  if (nb == 0)
    return;
  gemv_t(c_, Prec::F64, locked_.col(0), n_, nb, p, x, n_, sc_ + 16);
  gemv_n_sub(c_, Prec::F64, locked_.col(0), n_, nb, sc_ + 16, p, x, n_);
}

// modified_gram_schmidt on columns [i0, i0+cnt) (GramSchmidt.cpp:55-82), in
// place; each column against the already processed ones of the range.
bool LobpcgDriver::mgs(Blk &X, int i0, int cnt, double tol, bool reorth) {
  for (int i = 0; i < cnt; ++i) {
    void *xi = X.col(i0 + i);
    double n1;
    if (i > 0) {
      gemv_t(c_, X.p, X.col(i0), n_, i, X.p, xi, n_, sc_ + 16);
      gemv_n_sub(c_, X.p, X.col(i0), n_, i, sc_ + 16, X.p, xi, n_, sc_);
      rb(c_, sc_, 1, &n1);
    } else {
      n1 = nrm(X.p, xi, n_);
    }
    if (n1 < tol)
      return false;
    if (reorth) {
      double n2;
      if (i > 0) {
        gemv_t(c_, X.p, X.col(i0), n_, i, X.p, xi, n_, sc_ + 16);
        gemv_n_sub(c_, X.p, X.col(i0), n_, i, sc_ + 16, X.p, xi, n_, sc_);
        rb(c_, sc_, 1, &n2);
      } else {
        n2 = nrm(X.p, xi, n_);
      }
      if (n2 < tol)
        return false;
      n1 = n2;
    }
    axpby(c_, X.p, n_, 0.0, nullptr, 1.0 / n1, xi);
  }
  return true;
}

// modified_gram_schmidt (GramSchmidt.cpp:55-82) on columns [0, cnt) of X
// with every norm left on the device: dn[2i] = ||x_i|| after the first pass,
// dn[2i+1] after the re-orthogonalisation pass. Columns are scaled by the
// device norm without a host round trip; the caller replays the reference's
// `norm < tol` tests on the downloaded norms afterwards (a failed column only
// leaves non-finite values in X, which the caller then discards).
void LobpcgDriver::mgs_dev(Blk &X, int cnt, bool reorth, double *dn) {
  for (int i = 0; i < cnt; ++i) {
    void *xi = X.col(i);
    const int passes = reorth ? 2 : 1;
    if (i > 0 && i <= 496 && reorth) {
      // both passes with three reads of the previous columns: h1 = X^T x;
      // [x -= X h1, ||x||, h2 = X^T x] in one sweep; x -= X h2 with ||x||
      gemv_t(c_, X.p, X.col(0), n_, i, X.p, xi, n_, sc_ + 16);
      // h1 in sc_[16, 528), h2 in sc_[528, 1040): i <= 512 columns
      gemv_n_sub_t(c_, X.p, X.col(0), n_, i, sc_ + 16, X.p, xi, n_, sc_ + 528, dn + 2 * i);
      gemv_n_sub(c_, X.p, X.col(0), n_, i, sc_ + 528, X.p, xi, n_, dn + 2 * i + 1);
    } else {
      for (int ps = 0; ps < passes; ++ps) {
        if (i > 0) {
          gemv_t(c_, X.p, X.col(0), n_, i, X.p, xi, n_, sc_ + 16);
          gemv_n_sub(c_, X.p, X.col(0), n_, i, sc_ + 16, X.p, xi, n_, dn + 2 * i + ps);
        } else {
          dot(c_, X.p, xi, X.p, xi, n_, dn + 2 * i + ps, true);
        }
      }
    }
    scale_dev(c_, X.p, n_, dn + 2 * i + passes - 1, true, xi);
  }
}

// Phase 2 of orthonormalize for FP32 blocks of 4..128 columns as Cholesky QR
// applied twice instead of column-by-column MGS with re-orthogonalisation:
// G = Z^T Z accumulated in FP64 (products of FP32 values are exact in FP64),
// G = L L^T on the host, Z <- Z L^-T on the matrix cores, twice. Two passes
// over the block and two block products replace about 3 cnt^2 / 2 column
// sweeps. The first factor's diagonal is, in exact arithmetic, the norms the
// MGS pass would test against gs_tol (:150-157). A pivot below
// kCholQrMinPivot marks nearly dependent columns, for which a Gram-based
// factor loses the accuracy MGS keeps: then nothing has been written and the
// caller runs MGS. On success X(:, i0:i0+cnt) holds the orthonormal block.
constexpr double kCholQrMinPivot = 0.05;

bool LobpcgDriver::cholqr2(Blk &X, int i0, int cnt, Blk &scratch) {
  ++S_.cholqr_blocks;
  double *G = sc_ + 20000, *Ri = sc_ + 40000;  // cnt^2 <= 16384 each
  std::vector<double> g((std::size_t)cnt * cnt);
  for (int pass = 0; pass < 2; ++pass) {
    // pass 0: scratch -> X(:, i0..); pass 1: X(:, i0..) -> scratch
    const void *in = pass == 0 ? scratch.col(0) : X.col(i0);
    void *out = pass == 0 ? X.col(i0) : scratch.col(0);
    gemm_tn(c_, Prec::F32, in, n_, cnt, Prec::F32, in, n_, cnt, n_, G);
    rb(c_, G, cnt * cnt, g.data());
    dense::Mat S(cnt, cnt), L;
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < cnt; ++i)  // symmetrise (both triangles are summed alike)
        S(i, j) = 0.5 * (g[(std::size_t)j * cnt + i] + g[(std::size_t)i * cnt + j]);
    bool ok = dense::cholesky(S, L);
    for (int i = 0; ok && i < cnt; ++i)
      ok = std::isfinite(L(i, i)) && L(i, i) >= (pass == 0 ? kCholQrMinPivot : 0.5);
    if (!ok) {
      ++S_.cholqr_fallbacks;
      if (pass == 1)  // MGS continues from the first pass's block
        KH_CHECK(hipMemcpyAsync(scratch.col(0), X.col(i0), (std::size_t)cnt * n_ * 4,
                                hipMemcpyDeviceToDevice, c_.stream));
      return false;
    }
    const dense::Mat Li = dense::lower_inverse(L);
    std::vector<double> r((std::size_t)cnt * cnt);  // L^-T, column-major
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < cnt; ++i)
        r[(std::size_t)j * cnt + i] = Li(j, i);
    up(c_, r.data(), cnt * cnt, Ri);
    gemm_nn(c_, Prec::F32, in, n_, cnt, Ri, cnt, Prec::F32, out, n_, n_);
  }
  KH_CHECK(hipMemcpyAsync(X.col(i0), scratch.col(0), (std::size_t)cnt * n_ * 4,
                          hipMemcpyDeviceToDevice, c_.stream));
  return true;
}

// :137-160. Phase 1 (each column against the locked basis, then normalised)
// runs on all columns at once in the scratch block; phase 2 (MGS with
// re-orthogonalisation on the normalised copy, as the reference's clones)
// keeps its norms on the device. The host reads the norms twice per call and
// replays the reference's sequential tests, so X ends exactly as the
// reference leaves it: on a phase-1 failure at column i, columns < i are
// projected and normalised, column i projected, later columns untouched; on
// a phase-2 failure X holds the phase-1 result.
bool LobpcgDriver::orthonormalize(Blk &X, int i0, int cnt, double tol, bool reorth,
                                  Blk &scratch) {
  if (cnt == 0)
    return true;
  const std::size_t cbytes = (std::size_t)n_ * esz(X.p);
  if (scratch.p != X.p || scratch.cap < cnt)
    throw std::logic_error("orthonormalize: scratch block precision or width mismatch");
  KH_CHECK(hipMemcpyAsync(scratch.col(0), X.col(i0), (std::size_t)cnt * cbytes,
                          hipMemcpyDeviceToDevice, c_.stream));
  double *dn = sc_ + 60000;
  const int nb = (int)locked_.count;
  if (nb > 0) {
    const std::size_t need = (std::size_t)nb * cnt * 8;
    if (hproj_.bytes < need)
      hproj_ = DBuf(need);
    gemm_tn(c_, Prec::F64, locked_.col(0), n_, nb, X.p, scratch.col(0), n_, cnt, n_,
            hproj_.as<double>());
    gemm_nn_sub(c_, Prec::F64, locked_.col(0), n_, nb, hproj_.as<double>(), cnt, X.p,
                scratch.col(0), n_, n_, dn);
  } else {
    col_norms(c_, X.p, scratch.col(0), n_, cnt, n_, dn);
  }
  std::vector<double> nn((std::size_t)cnt), inv((std::size_t)cnt);
  rb(c_, dn, cnt, nn.data());
  int fail = -1;
  for (int i = 0; i < cnt && fail < 0; ++i)
    if (nn[(std::size_t)i] < tol)
      fail = i;
  const int good = fail < 0 ? cnt : fail;
  for (int i = 0; i < cnt; ++i)
    inv[(std::size_t)i] = i < good ? 1.0 / nn[(std::size_t)i] : 1.0;
  up(c_, inv.data(), cnt, dn + 512);
  scale_cols(c_, X.p, n_, good, scratch.col(0), n_, dn + 512, false);
  const int back = fail < 0 ? cnt : fail + 1;
  KH_CHECK(hipMemcpyAsync(X.col(i0), scratch.col(0), (std::size_t)back * cbytes,
                          hipMemcpyDeviceToDevice, c_.stream));
  if (fail >= 0)
    return false;
  if (X.p == Prec::F32 && reorth && cnt >= 4 && cnt <= 128 && cholqr2(X, i0, cnt, scratch))
    return true;
  mgs_dev(scratch, cnt, reorth, dn);
  std::vector<double> n12((std::size_t)2 * cnt);
  rb(c_, dn, 2 * cnt, n12.data());
  for (int i = 0; i < cnt; ++i) {
    if (n12[(std::size_t)2 * i] < tol)
      return false;
    if (reorth && n12[(std::size_t)2 * i + 1] < tol)
      return false;
  }
  KH_CHECK(hipMemcpyAsync(X.col(i0), scratch.col(0), (std::size_t)cnt * cbytes,
                          hipMemcpyDeviceToDevice, c_.stream));
  return true;
}

// :54-64  AZ = A^T (A Z), q columns. Blocks of 4 or more columns go through
// row-major copies, where one gather reads a whole q-wide row of the block
// (the column-major gathers would fetch a cache line per column and nonzero
// once the block outgrows the caches); per-entry arithmetic is unchanged.
void LobpcgDriver::apply_normal(const Blk &Z, int q, Blk &AZ) {
  auto t0 = Clock::now();
  if (AZ.p != Z.p || AZ.cap < q || Z.cap < q)
    throw std::logic_error("apply_normal: block precision or width mismatch");
  if (q >= 4 && A_.has_rows()) {
    const std::size_t e = esz(Z.p);
    if (zr_.bytes < (std::size_t)n_ * q * e)
      zr_ = DBuf((std::size_t)n_ * q * e);
    if (tr_.bytes < (std::size_t)m_ * q * e)
      tr_ = DBuf((std::size_t)m_ * q * e);
    cols_to_rows(c_, Z.p, n_, q, Z.col(0), n_, zr_.p);
    A_.mm_rows(c_, Z.p, q, zr_.p, tr_.p, false);
    A_.mm_rows(c_, Z.p, q, tr_.p, zr_.p, true);
    rows_to_cols(c_, Z.p, n_, q, zr_.p, AZ.col(0), n_);
    S_.spmv += q;
    S_.spmtv += q;
    S_.t_sparse += since(t0);
    return;
  }
  if (tmpq_.cap < q || tmpq_.p != Z.p)
    tmpq_.init(m_, std::max(q, 8), Z.p);
  A_.mm(c_, Z.p, q, Z.col(0), n_, tmpq_.col(0), m_, nullptr);
  A_.mtm(c_, Z.p, q, tmpq_.col(0), m_, AZ.col(0), n_, nullptr);
  S_.spmv += q;
  S_.spmtv += q;
  S_.t_sparse += since(t0);
}

void LobpcgDriver::block_apply(int q, const void *X, void *Y, bool trans) {
  const int64_t nin = trans ? m_ : n_, nout = trans ? n_ : m_;
  block_product(c_, A_, Prec::F64, q, X, nin, Y, nout, trans, nullptr, zr_, tr_);
}

// :67-80  S = Z^T Z, G = Z^T AZ (full q x q, symmetric to rounding)
void LobpcgDriver::gram(const Blk &Z, const Blk &AZ, int q, dense::Mat &S, dense::Mat &G) {
  auto t0 = Clock::now();
  const int qq = q * q;
  if (Z.p == Prec::F32 && AZ.p == Prec::F32 && q <= 128) {  // fused FP32 MFMA pass
    gram2_mfma(c_, (const float *)Z.col(0), n_, (const float *)AZ.col(0), n_, q, n_, sc_ + 1024);
  } else if (Z.p == Prec::F64 && AZ.p == Prec::F64 && q <= 128) {  // fused FP64 MFMA pass
    gram2_mfma(c_, (const double *)Z.col(0), n_, (const double *)AZ.col(0), n_, q, n_,
               sc_ + 1024);
  } else if (Z.p == Prec::F32 && AZ.p == Prec::F32) {  // FP32 MFMA tiles
    gram_mfma(c_, (const float *)Z.col(0), n_, q, (const float *)Z.col(0), n_, q, n_,
              sc_ + 1024);
    gram_mfma(c_, (const float *)Z.col(0), n_, q, (const float *)AZ.col(0), n_, q, n_,
              sc_ + 1024 + qq);
  } else {
    gemm_tn(c_, Z.p, Z.col(0), n_, q, Z.p, Z.col(0), n_, q, n_, sc_ + 1024);
    gemm_tn(c_, Z.p, Z.col(0), n_, q, AZ.p, AZ.col(0), n_, q, n_, sc_ + 1024 + qq);
  }
  std::vector<double> h((std::size_t)2 * qq);
  rb(c_, sc_ + 1024, 2 * qq, h.data());
  S = dense::Mat(q, q);
  G = dense::Mat(q, q);
  // the reference computes the upper triangle and mirrors it
  for (int j = 0; j < q; ++j)
    for (int i = 0; i <= j; ++i) {
      S(i, j) = S(j, i) = h[(std::size_t)j * q + i];
      G(i, j) = G(j, i) = h[(std::size_t)qq + (std::size_t)j * q + i];
    }
  S_.t_gram += since(t0);
}

// out(:, o0+i) = sum_j Z(:, z0+j) Y(row0+j, i), i < p, j < q  (:122-134)
void LobpcgDriver::combine(const Blk &Z, int z0, int q, const dense::Mat &Y, int row0, int p,
                           Blk &out, int o0) {
  std::vector<double> H((std::size_t)q * p);
  for (int i = 0; i < p; ++i)
    for (int j = 0; j < q; ++j)
      H[(std::size_t)i * q + j] = Y(row0 + j, i);
  up(c_, H.data(), q * p, sc_ + 40000);
  gemm_nn(c_, Z.p, Z.col(z0), n_, q, sc_ + 40000, p, out.p, out.col(o0), n_, n_);
}

double LobpcgDriver::estimate_fro() {  // :163-178
  const std::size_t samples = (std::size_t)std::max(1, P_.frob_samples);
  double sum2 = 0.0;
  for (std::size_t s = 0; s < samples; ++s) {
    random_col(v64b_.p, Prec::F64);
    A_.mv(c_, Prec::F64, v64b_.p, tm64_.p);
    ++S_.spmv;
    const double nn = nrm(Prec::F64, tm64_.p, m_);
    sum2 += nn * nn;
  }
  return std::sqrt(sum2 / (double)samples);
}

void LobpcgDriver::mgs_locked(double tol) {
  Blk view;  // non-owning view of the locked basis
  // run MGS directly on the locked buffer
  const int k = (int)locked_.count;
  for (int i = 0; i < k; ++i) {
    double *xi = locked_.col(i);
    double n1;
    for (int pass = 0; pass < 2; ++pass) {
      if (i > 0) {
        gemv_t(c_, Prec::F64, locked_.col(0), n_, i, Prec::F64, xi, n_, sc_ + 16);
        gemv_n_sub(c_, Prec::F64, locked_.col(0), n_, i, sc_ + 16, Prec::F64, xi, n_, sc_);
        rb(c_, sc_, 1, &n1);
      } else {
        n1 = nrm(Prec::F64, xi, n_);
      }
      if (n1 < tol)
        return;  // dependence: the reference ignores the return value
    }
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / n1, xi);
  }
  (void)view;
}

// :199-278 on the locked FP64 vectors
bool LobpcgDriver::two_vec_sharpen(int idx, double target, int max_iters) {
  const int k = (int)locked_.count;
  if (idx >= k)
    return true;
  double *v = locked_.col(idx);
  double *t = tn64_.as<double>(), *d = v64_.as<double>(), *Nd = v64b_.as<double>();
  double *Av = tm64_.as<double>();
  // basis = all other locked vectors, contiguous copy
  Blk B;
  B.init(n_, std::max(1, k - 1), Prec::F64);
  for (int j = 0, o = 0; j < k; ++j)
    if (j != idx) {
      KH_CHECK(hipMemcpyAsync(B.col(o++), locked_.col(j), (std::size_t)n_ * 8,
                              hipMemcpyDeviceToDevice, c_.stream));
    }
  auto proj = [&](double *x) {
    if (k - 1 <= 0)
      return;
    gemv_t(c_, Prec::F64, B.col(0), n_, k - 1, Prec::F64, x, n_, sc_ + 16);
    gemv_n_sub(c_, Prec::F64, B.col(0), n_, k - 1, sc_ + 16, Prec::F64, x, n_);
  };
  for (int it = 0; it < max_iters; ++it) {
    A_.mv(c_, Prec::F64, v, Av);
    ++S_.spmv;
    const double res = nrm(Prec::F64, Av, m_);
    if (res <= target)
      return true;
    A_.mtv(c_, Prec::F64, Av, t);
    ++S_.spmtv;
    if (P_.M_inv) {  // d = M^{-1} t (:232-237)
      c_.sync();
      P_.M_inv->apply(c_, t, d, n_);
    } else {
      copy(c_, Prec::F64, n_, t, d);
    }
    proj(d);
    const double nd = nrm(Prec::F64, d, n_);
    if (nd <= 1e-14)
      return false;
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / std::max(nd, 1e-300), d);
    A_.mv(c_, Prec::F64, d, Av);
    A_.mtv(c_, Prec::F64, Av, Nd);
    ++S_.spmv;
    ++S_.spmtv;
    dot(c_, Prec::F64, v, Prec::F64, t, n_, sc_ + 2, false);
    dot(c_, Prec::F64, v, Prec::F64, Nd, n_, sc_ + 3, false);
    dot(c_, Prec::F64, d, Prec::F64, t, n_, sc_ + 4, false);
    dot(c_, Prec::F64, d, Prec::F64, Nd, n_, sc_ + 5, false);
    double g[4];
    rb(c_, sc_ + 2, 4, g);
    dense::Mat G(2, 2);
    G(0, 0) = g[0];
    G(1, 1) = g[3];
    G(0, 1) = G(1, 0) = 0.5 * (g[1] + g[2]);
    std::vector<double> w;
    dense::Mat Q;
    if (!dense::sym_eig(G, w, Q))
      return false;
    // v_new = y0 v + y1 d, projected, normalised
    axpby(c_, Prec::F64, n_, Q(1, 0), d, Q(0, 0), v);
    proj(v);
    const double nv = nrm(Prec::F64, v, n_);
    if (nv <= 1e-14)
      return false;
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, v);
  }
  return false;
}

// orthonormalize_block(D, V_basis) (:137-160) with V_basis = B(:, 0:k) and
// D = B(:, k:k+nd): one projection pass, normalise, then MGS on a copy.
bool LobpcgDriver::orthonormalize_vs(Blk &B, int k, int nd, Blk &scratch) {
  for (int i = 0; i < nd; ++i) {
    void *x = B.col(k + i);
    gemv_t(c_, Prec::F64, B.col(0), n_, k, Prec::F64, x, n_, sc_ + 16);
    gemv_n_sub(c_, Prec::F64, B.col(0), n_, k, sc_ + 16, Prec::F64, x, n_, sc_);
    double nn;
    rb(c_, sc_, 1, &nn);
    if (nn < 1e-14)
      return false;
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nn, x);
  }
  KH_CHECK(hipMemcpyAsync(scratch.col(0), B.col(k), (std::size_t)nd * n_ * 8,
                          hipMemcpyDeviceToDevice, c_.stream));
  if (!mgs(scratch, 0, nd, 1e-14, true))
    return false;
  KH_CHECK(hipMemcpyAsync(B.col(k), scratch.col(0), (std::size_t)nd * n_ * 8,
                          hipMemcpyDeviceToDevice, c_.stream));
  return true;
}

// Mixed precision only, ahead of the reference's polish: one FP64 projector
// pass v <- v - A^T (A A^T)^{-1} A v on all locked vectors at once (one
// multi-column CG, row-major products), aimed at a tenth of the polish target.
// A vector is replaced only if its projection keeps most of its norm (an
// exact null vector moves by about its residual; a vector of a numerical null
// space, whose singular values sit above the target, would be projected
// away) and its residual drops. The reference's block and 2x2 polish then
// find nothing to do for the vectors that reached the target.
void LobpcgDriver::projector_polish(double target) {
  const int k = (int)locked_.count;
  if (k == 0 || k > 1000)  // norms live in the device scalar slab
    return;
  double *nrm_d = sc_ + 60000, *inv_d = sc_ + 62000;
  const std::size_t nm = (std::size_t)m_ * k * 8, nn = (std::size_t)n_ * k * 8;
  DBuf AV(nm), Wk(nm), T(nn), Vn(nn);
  block_apply(k, locked_.col(0), AV.p, false);
  S_.spmv += k;
  col_norms(c_, Prec::F64, AV.p, m_, k, m_, nrm_d);
  std::vector<double> r0((std::size_t)k);
  rb(c_, nrm_d, k, r0.data());
  double rel = 1e-2;
  bool any = false;
  for (double a : r0)
    if (a > target) {
      any = true;
      rel = std::min(rel, std::max(1e-12, 0.1 * target / a));
    }
  if (!any)
    return;
  std::vector<CgState> st;
  cg_solve_ops(c_, A_, cg_, k, AV.p, Wk.p, Prec::F64, A_.diag_AAt(), 1000, rel, 1e-30, st);
  for (const auto &x : st) {
    S_.spmv += x.k;
    S_.spmtv += x.k;
  }
  block_apply(k, Wk.p, T.p, true);
  S_.spmtv += k;
  sub_cols(c_, Prec::F64, n_, k, locked_.col(0), n_, T.p, n_, Vn.p, n_, nullptr);
  col_norms(c_, Prec::F64, Vn.p, n_, k, n_, nrm_d);
  std::vector<double> nv((std::size_t)k), inv((std::size_t)k);
  rb(c_, nrm_d, k, nv.data());
  for (int j = 0; j < k; ++j)
    inv[(std::size_t)j] = nv[(std::size_t)j] > 0.5 ? 1.0 / nv[(std::size_t)j] : 0.0;
  up(c_, inv.data(), k, inv_d);
  scale_cols(c_, Prec::F64, n_, k, Vn.p, n_, inv_d, false);
  block_apply(k, Vn.p, AV.p, false);
  S_.spmv += k;
  col_norms(c_, Prec::F64, AV.p, m_, k, m_, nrm_d);
  std::vector<double> r1((std::size_t)k);
  rb(c_, nrm_d, k, r1.data());
  for (int j = 0; j < k; ++j)
    if (nv[(std::size_t)j] > 0.5 && r1[(std::size_t)j] < r0[(std::size_t)j])
      KH_CHECK(hipMemcpyAsync(locked_.col(j), static_cast<double *>(Vn.p) + (std::size_t)j * n_,
                              (std::size_t)n_ * 8, hipMemcpyDeviceToDevice, c_.stream));
}

// :286-440 on the locked FP64 vectors
void LobpcgDriver::block_rr_polish(double target, int max_sweeps) {
  const int k = (int)locked_.count;
  if (k == 0)
    return;
  mgs_locked(1e-14);
  Blk Vb;  // [V | D], FP64
  Vb.init(n_, 2 * k + 2, Prec::F64);
  Blk AZ;
  AZ.init(n_, 2 * k + 2, Prec::F64);
  Blk scratch;
  scratch.init(n_, 2 * k + 2, Prec::F64);
  double *r = tm64_.as<double>(), *g = v64_.as<double>();
  auto max_res = [&]() {
    KH_CHECK(hipMemcpyAsync(Vb.col(0), locked_.col(0), (std::size_t)k * n_ * 8,
                            hipMemcpyDeviceToDevice, c_.stream));
    if (tmpq_.cap < k || tmpq_.p != Prec::F64) {
      tmpq_.init(m_, std::max(k, 8), Prec::F64);
    }
    A_.mm(c_, Prec::F64, k, Vb.col(0), n_, tmpq_.col(0), m_, nullptr);
    S_.spmv += k;
    col_norms(c_, Prec::F64, tmpq_.col(0), m_, k, m_, sc_ + 100);
    std::vector<double> rr((std::size_t)k);
    rb(c_, sc_ + 100, k, rr.data());
    double mx = 0.0;
    for (double x : rr)
      mx = std::max(mx, x);
    return std::make_pair(mx, rr);
  };
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    auto [mr, rr] = max_res();
    if (mr <= target)
      break;
    // D columns go to Vb(:, k + nd)
    int nd = 0;
    for (int j = 0; j < k; ++j) {
      if (rr[(std::size_t)j] <= target)
        continue;
      A_.mv(c_, Prec::F64, locked_.col(j), r);
      A_.mtv(c_, Prec::F64, r, g);
      ++S_.spmv;
      ++S_.spmtv;
      if (P_.M_inv)  // g = M^{-1} N v_j (:327-331)
        precondition64(g);
      // project vs V (one pass, :334)
      gemv_t(c_, Prec::F64, Vb.col(0), n_, k, Prec::F64, g, n_, sc_ + 16);
      gemv_n_sub(c_, Prec::F64, Vb.col(0), n_, k, sc_ + 16, Prec::F64, g, n_, sc_);
      double gn;
      rb(c_, sc_, 1, &gn);
      if (gn <= 1e-14)
        continue;
      axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / gn, g);
      bool dup = false;
      for (int t = 0; t < nd && !dup; ++t) {
        dot(c_, Prec::F64, g, Prec::F64, Vb.col(k + t), n_, sc_ + 1, false);
        double dd;
        rb(c_, sc_ + 1, 1, &dd);
        dup = std::fabs(dd) >= 1.0 - 1e-8;
      }
      if (!dup) {
        copy(c_, Prec::F64, n_, g, Vb.col(k + nd));
        ++nd;
      }
    }
    if (nd == 0) {
      // projected steepest descent per vector (:351-389)
      for (int i = 0; i < k; ++i) {
        double *vi = locked_.col(i);
        for (int it = 0; it < 50; ++it) {
          A_.mv(c_, Prec::F64, vi, r);
          ++S_.spmv;
          const double res = nrm(Prec::F64, r, m_);
          if (res <= target)
            break;
          A_.mtv(c_, Prec::F64, r, g);
          ++S_.spmtv;
          for (int t = 0; t < k; ++t) {
            if (t == i)
              continue;
            dot(c_, Prec::F64, g, Prec::F64, locked_.col(t), n_, sc_ + 1, false);
            double cc;
            rb(c_, sc_ + 1, 1, &cc);
            if (cc != 0.0)
              axpby(c_, Prec::F64, n_, -cc, locked_.col(t), 1.0, g);
          }
          dot(c_, Prec::F64, g, Prec::F64, g, n_, sc_ + 1, false);
          double gg;
          rb(c_, sc_ + 1, 1, &gg);
          if (gg <= 0.0)
            break;
          A_.mv(c_, Prec::F64, g, r);
          ++S_.spmv;
          dot(c_, Prec::F64, r, Prec::F64, r, m_, sc_ + 1, false);
          double ag2;
          rb(c_, sc_ + 1, 1, &ag2);
          if (ag2 <= 0.0)
            break;
          axpby(c_, Prec::F64, n_, -(gg / ag2), g, 1.0, vi);
          for (int t = 0; t < k; ++t)
            if (t != i) {
              dot(c_, Prec::F64, vi, Prec::F64, locked_.col(t), n_, sc_ + 1, false);
              double cc;
              rb(c_, sc_ + 1, 1, &cc);
              if (cc != 0.0)
                axpby(c_, Prec::F64, n_, -cc, locked_.col(t), 1.0, vi);
            }
          const double nv = nrm(Prec::F64, vi, n_);
          if (nv > 0.0)
            axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, vi);
        }
      }
      mgs_locked(1e-14);
      const int enrich = std::min(2, k);
      for (int e = 0; e < enrich; ++e) {
        double *q = static_cast<double *>(Vb.col(k + nd));
        random_col(q, Prec::F64);
        // project vs V_basis (the vectors before this sweep: still in Vb)
        gemv_t(c_, Prec::F64, Vb.col(0), n_, k, Prec::F64, q, n_, sc_ + 16);
        gemv_n_sub(c_, Prec::F64, Vb.col(0), n_, k, sc_ + 16, Prec::F64, q, n_, sc_);
        double nq;
        rb(c_, sc_, 1, &nq);
        if (nq > 1e-14) {
          axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nq, q);
          ++nd;
        }
      }
      if (nd == 0)
        continue;
      // orthonormalize_block(D, V_basis) result is ignored here (:404)
      (void)orthonormalize_vs(Vb, k, nd, scratch);
    }
    if (!orthonormalize_vs(Vb, k, nd, scratch))
      continue;
    // Z = [V | D] with the current V
    KH_CHECK(hipMemcpyAsync(Vb.col(0), locked_.col(0), (std::size_t)k * n_ * 8,
                            hipMemcpyDeviceToDevice, c_.stream));
    const int q = k + nd;
    apply_normal(Vb, q, AZ);
    dense::Mat S, G;
    gram(Vb, AZ, q, S, G);
    TinyGEP t = solve_tiny_gep(G, S);
    Blk Vn;
    Vn.init(n_, k, Prec::F64);
    combine(Vb, 0, q, t.Y, 0, k, Vn, 0);
    if (!mgs(Vn, 0, k, 1e-14, true)) {
      // the reference ignores the result and copies the partially processed
      // vectors back
    }
    KH_CHECK(hipMemcpyAsync(locked_.col(0), Vn.col(0), (std::size_t)k * n_ * 8,
                            hipMemcpyDeviceToDevice, c_.stream));
    auto [mr2, rr2] = max_res();
    (void)rr2;
    if (mr2 <= target)
      break;
  }
  mgs_locked(1e-14);
}

DeviceBasis LobpcgDriver::run() {
  locked_.len = n_;
  if (n_ == 0 || m_ == 0)
    return std::move(locked_);
  int64_t k_target = std::max<int64_t>(1, P_.nev);
  if (P_.nullity_known >= 0) {
    if (P_.nullity_known > n_)
      throw std::invalid_argument("LOBPCG: nullity_known > n");
    k_target = P_.nullity_known;
  }
  if (P_.rank_known >= 0) {
    if (P_.rank_known > n_)
      throw std::invalid_argument("LOBPCG: rank_known > n");
    const int64_t implied = n_ - P_.rank_known;
    if (P_.nullity_known >= 0 && implied != k_target)
      throw std::invalid_argument("LOBPCG: inconsistent rank_known vs nullity_known");
    if (P_.nullity_known < 0)
      k_target = implied;
    if (n_ <= P_.rank_known) {
      S_.frob_norm_est = 0.0;
      S_.converged_to_target = (k_target == 0);
      return std::move(locked_);
    }
  }
  int64_t bs = P_.block_size > 0 ? P_.block_size
                                 : std::min<int64_t>(16, std::max<int64_t>(1, k_target));
  bs = std::min<int64_t>(bs, n_);
  S_.block_size = bs;
  const int64_t guards =
      (bs == 1) ? 0 : std::min<int64_t>((int64_t)((1.0 - P_.restart) * bs + 0.45), bs - 1);
  const int B = (int)bs;

  // buffers
  hbuf_.resize((std::size_t)n_);
  small_ = DBuf((std::size_t)Context::kScalars * 8);
  sc_ = small_.as<double>();
  v64_ = DBuf((std::size_t)n_ * 8);
  v64b_ = DBuf((std::size_t)n_ * 8);
  tn64_ = DBuf((std::size_t)n_ * 8);
  tm64_ = DBuf((std::size_t)m_ * 8);
  vc_ = DBuf((std::size_t)n_ * 8);
  locked_.cap = std::max<int64_t>(1, k_target);
  locked_.count = 0;
  locked_.buf = DBuf((std::size_t)locked_.cap * n_ * 8);
  if (3 * B * 3 * B > 40000 || (int64_t)3 * B * B > 20000)
    throw std::invalid_argument("LOBPCG: block too large for the device scalar slab");

  S_.frob_norm_est = estimate_fro();
  const double accept_thr = std::max(P_.tol_abs, P_.tol_rel * std::max(1.0, S_.frob_norm_est));
  S_.accept_threshold = accept_thr;
  // FP32 blocks cannot resolve ||A x|| below ~1e-6 ||A||: such iterates become
  // candidates and are refined in FP64 before the acceptance test
  const double cand_thr =
      bp_ == Prec::F32 ? std::max(accept_thr, 1e-5 * std::max(1.0, S_.frob_norm_est)) : accept_thr;
  w64_ = DBuf((std::size_t)m_ * 8);
  if (bp_ == Prec::F32) {
    vblk_ = DBuf((std::size_t)B * n_ * 8);
    tnblk_ = DBuf((std::size_t)B * n_ * 8);
    rhsblk_ = DBuf((std::size_t)B * m_ * 8);
    wblk_ = DBuf((std::size_t)B * m_ * 8);
  }

  Blk X, AX, W, Pb, Z, AZ, scratch, Xn;
  X.init(n_, B, bp_);
  AX.init(n_, B, bp_);
  W.init(n_, B, bp_);
  Pb.init(n_, B, bp_);
  Z.init(n_, 3 * B, bp_);
  AZ.init(n_, 3 * B, bp_);
  scratch.init(n_, 3 * B, bp_);
  Xn.init(n_, B, bp_);
  fill_zero(c_, bp_, (int64_t)B * n_, Pb.col(0));
  fill_zero(c_, bp_, (int64_t)B * n_, W.col(0));
  const std::size_t cb = (std::size_t)n_ * esz(bp_);
  auto dcopy = [&](void *dst, const void *src, std::size_t bytes) {
    KH_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c_.stream));
  };

  // initial block (:504-513)
  int filled = 0;
  const int seeds = (int)std::min<int64_t>(bs, P_.n_initial);
  for (; filled < seeds; ++filled) {
    KH_CHECK(hipMemcpyAsync(v64_.p, P_.initial_block_host + (std::size_t)filled * n_,
                            (std::size_t)n_ * 8, hipMemcpyHostToDevice, c_.stream));
    convert(c_, Prec::F64, v64_.p, bp_, X.col(filled), n_);
    c_.sync();
  }
  for (; filled < B; ++filled)
    random_col(X.col(filled), bp_);
  if (!orthonormalize(X, 0, B, P_.gs_tol, P_.gs_reorth, scratch))
    throw std::runtime_error("LOBPCG: initial block is dependent");

  std::vector<double> theta((std::size_t)B, 0.0);
  auto full_rr = [&]() {  // X <- X Y, AX = N X, theta (:516-528 and :773-787)
    apply_normal(X, B, AX);
    dense::Mat S0, G0;
    gram(X, AX, B, S0, G0);
    auto th = Clock::now();
    TinyGEP t = solve_tiny_gep(G0, S0, bp_ == Prec::F32);
    S_.t_host += since(th);
    combine(X, 0, B, t.Y, 0, B, Xn, 0);
    dcopy(X.col(0), Xn.col(0), (std::size_t)B * cb);
    apply_normal(X, B, AX);
    std::fill(theta.begin(), theta.end(), 0.0);
    const int lim = std::min<int>((int)t.theta.size(), B);
    for (int j = 0; j < lim; ++j)
      theta[(std::size_t)j] = std::max(0.0, t.theta[(std::size_t)j]);
  };
  full_rr();

  Blk AXres;  // A X (m x B) for the residual norms
  AXres.init(m_, B, bp_);
  int64_t its = 0, locked_total = 0;
  while (locked_total < k_target && its < P_.max_iters) {
    ++its;
    // residuals R = AX - theta X (:545-547), straight into W (W = R below,
    // no preconditioner; W is not read before that)
    {
      std::vector<double> nt((std::size_t)B);
      for (int j = 0; j < B; ++j)
        nt[(std::size_t)j] = -theta[(std::size_t)j];
      up(c_, nt.data(), B, sc_ + 50000);
      resid_block(c_, bp_, n_, B, AX.col(0), X.col(0), n_, sc_ + 50000, W.col(0));
    }
    // ||A x_j|| for all j at once (X does not change within the loop)
    if (B >= 4 && A_.has_rows()) {
      const std::size_t e = esz(bp_);
      if (zr_.bytes < (std::size_t)n_ * B * e)
        zr_ = DBuf((std::size_t)n_ * B * e);
      if (tr_.bytes < (std::size_t)m_ * B * e)
        tr_ = DBuf((std::size_t)m_ * B * e);
      cols_to_rows(c_, bp_, n_, B, X.col(0), n_, zr_.p);
      A_.mm_rows(c_, bp_, B, zr_.p, tr_.p, false);
      rows_to_cols(c_, bp_, m_, B, tr_.p, AXres.col(0), m_);
    } else {
      A_.mm(c_, bp_, B, X.col(0), n_, AXres.col(0), m_, nullptr);
    }
    S_.spmv += B;
    col_norms(c_, bp_, AXres.col(0), m_, B, m_, sc_ + 200);
    std::vector<double> avo((std::size_t)B);
    rb(c_, sc_ + 200, B, avo.data());
    int64_t nconv = 0;
    double best_unacc = std::numeric_limits<double>::infinity();
    // the reference's acceptance of one FP64 vector v (:560-605): residual
    // against the threshold, then lock
    auto try_accept = [&](double *v, int j, double Av_orig) {
      A_.mv(c_, Prec::F64, v, tm64_.p);
      ++S_.spmv;
      const double Av_acc = nrm(Prec::F64, tm64_.p, m_);
      if (Av_acc > accept_thr) {
        best_unacc = std::min(best_unacc, Av_orig);
        return;
      }
      dcopy(locked_.col(locked_.count), v, (std::size_t)n_ * 8);
      locked_.count += 1;
      S_.theta.push_back(theta[(std::size_t)j]);
      S_.residuals_Av.push_back(Av_acc);
      ++locked_total;
      ++nconv;
      if (P_.verbose)
        std::fprintf(stderr, "[LOBPCG-HIP] accept: ||A v||=%.3e locked=%lld/%lld\n", Av_acc,
                     (long long)locked_total, (long long)k_target);
    };
    // v = x_j in FP64, projected against the locked vectors and normalised;
    // false if it is (numerically) in their span
    auto prepare = [&](double *v, const void *xj) {
      if (xj)
        to64(xj, bp_, v);
      project_out(v, Prec::F64);
      const double nv = nrm(Prec::F64, v, n_);
      if (!(nv > std::max(1e-14, 10.0 * P_.gs_tol)))
        return false;
      axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, v);
      if (locked_.count > 0) {
        gemv_t(c_, Prec::F64, locked_.col(0), n_, (int)locked_.count, Prec::F64, v, n_,
               sc_ + 300);
        std::vector<double> dd((std::size_t)locked_.count);
        rb(c_, sc_ + 300, (int)locked_.count, dd.data());
        for (double x : dd)
          if (std::fabs(x) >= 1.0 - 1e-8)
            return false;
      }
      return true;
    };
    if (bp_ == Prec::F64) {
      for (int j = 0; j < B; ++j) {
        const double Av_orig = avo[(std::size_t)j];
        double *v = vc_.as<double>();
        if (Av_orig <= cand_thr && prepare(v, X.col(j)))
          try_accept(v, j, Av_orig);
        else
          best_unacc = std::min(best_unacc, Av_orig);
        if (locked_total >= k_target)
          break;
      }
    } else {
      // FP32 iterates become candidates at cand_thr and are refined in FP64
      // by one projector pass v <- v - A^T (A A^T)^{-1} A v (iterative
      // refinement) before the reference's acceptance test. The candidates of
      // an iteration are refined together: one multi-column CG (row-major
      // products, one gather per nonzero for all columns), diag(A A^T)
      // preconditioned when the operator provides it. A v' is the CG
      // residual, so CG only has to bring ||A v|| to a tenth of the threshold.
      //
      // An iteration with fewer than kBatch candidates leaves them in the
      // block (they keep converging in FP32) for up to kMaxDefer iterations,
      // so that later candidates join the same refinement: on matrices whose
      // gathers are random (the [R | RG] family) a product costs about the
      // same for one column as for several. Candidates that would complete
      // the target, and those of the last allowed iteration, are refined at
      // once.
      constexpr int kBatch = 4, kMaxDefer = 3;
      int ncand = 0;
      for (int j = 0; j < B; ++j)
        ncand += avo[(std::size_t)j] <= cand_thr;
      const bool defer = ncand > 0 && ncand < kBatch && defer_ < kMaxDefer &&
                         locked_total + ncand < k_target && its < P_.max_iters;
      defer_ = defer ? defer_ + 1 : 0;
      std::vector<int> cand;
      for (int j = 0; j < B; ++j) {
        const double Av_orig = avo[(std::size_t)j];
        if (!defer && Av_orig <= cand_thr &&
            prepare(vblk_.as<double>() + (std::size_t)cand.size() * n_, X.col(j)))
          cand.push_back(j);
        else
          best_unacc = std::min(best_unacc, Av_orig);
      }
      S_.deferred += defer ? 1 : 0;
      const int nc = (int)cand.size();
      if (nc > 0) {
        double *Vb64 = vblk_.as<double>();
        block_apply(nc, Vb64, rhsblk_.p, false);
        S_.spmv += nc;
        col_norms(c_, Prec::F64, rhsblk_.p, m_, nc, m_, sc_ + 200);
        std::vector<double> av((std::size_t)nc);
        rb(c_, sc_ + 200, nc, av.data());
        // CG target: a tenth of the acceptance threshold. A tighter target
        // (say the final polish's 0.01 thr) would also resolve singular
        // directions just above the threshold when the null space is only
        // numerical (mat3 of the reference's data: sigma ~ 6.5e-10), and the
        // projection would then remove the candidate altogether.
        double rel = 1e-2;
        for (double a : av)
          rel = std::min(rel, a > 0.0 ? std::max(1e-12, 0.1 * accept_thr / a) : 1e-12);
        std::vector<CgState> st;
        cg_solve_ops(c_, A_, cg_, nc, rhsblk_.p, wblk_.p, Prec::F64, A_.diag_AAt(), 1000, rel,
                     1e-30, st);
        for (const auto &x : st) {
          S_.spmv += x.k;
          S_.spmtv += x.k;
        }
        block_apply(nc, wblk_.p, tnblk_.p, true);
        S_.spmtv += nc;
        sub_cols(c_, Prec::F64, n_, nc, Vb64, n_, tnblk_.p, n_, Vb64, n_, nullptr);
        for (int t = 0; t < nc && locked_total < k_target; ++t) {
          const int j = cand[(std::size_t)t];
          double *v = Vb64 + (std::size_t)t * n_;
          if (prepare(v, nullptr))
            try_accept(v, j, avo[(std::size_t)j]);
          else
            best_unacc = std::min(best_unacc, avo[(std::size_t)j]);
        }
      }
    }
    S_.locked = locked_total;
    S_.iters = its;
    S_.best_unaccepted_residual = std::isfinite(best_unacc) ? best_unacc : 0.0;
    if (P_.verbose && P_.progress_every > 0 && (its == 1 || its % P_.progress_every == 0))
      std::fprintf(stderr, "[LOBPCG-HIP] it=%lld locked=%lld/%lld best=%.3e thr=%.3e\n",
                  (long long)its, (long long)locked_total, (long long)k_target,
                  S_.best_unaccepted_residual, accept_thr);
    if (locked_total >= k_target)
      break;
    const int ini = P_.soft_locking ? (int)nconv : 0;

    if (nconv >= bs - guards) {  // hard lock restart (:624-651)
      const int remain = B - (int)nconv;
      for (int j = 0; j < remain; ++j) {
        dcopy(X.col(j), X.col((int)nconv + j), cb);
        dcopy(AX.col(j), AX.col((int)nconv + j), cb);
        theta[(std::size_t)j] = theta[(std::size_t)(nconv + j)];
        dcopy(Pb.col(j), Pb.col((int)nconv + j), cb);
      }
      for (int j = remain; j < B; ++j)
        random_col(X.col(j), bp_);
      if (!orthonormalize(X, 0, B, P_.gs_tol, P_.gs_reorth, scratch))
        throw std::runtime_error("LOBPCG: restart orthogonalization failed");
      apply_normal(X, B, AX);
      dense::Mat S1, G1;
      gram(X, AX, B, S1, G1);
      TinyGEP t1 = solve_tiny_gep(G1, S1, bp_ == Prec::F32);
      std::fill(theta.begin(), theta.end(), 0.0);
      const int lim = std::min<int>((int)t1.theta.size(), B);
      for (int j = 0; j < lim; ++j)
        theta[(std::size_t)j] = std::max(0.0, t1.theta[(std::size_t)j]);
      continue;
    }

    const int na = B - ini;
    // W = M^{-1} R on the active slice (:654-659) when the caller gives M_inv:
    // its FP64 device views, through FP64 copies for FP32 blocks
    if (P_.M_inv) {
      double *r64 = v64_.as<double>(), *z64 = v64b_.as<double>();
      for (int j = ini; j < B; ++j) {
        to64(W.col(j), bp_, r64);
        c_.sync();
        P_.M_inv->apply(c_, r64, z64, n_);
        convert(c_, Prec::F64, z64, bp_, W.col(j), n_);
      }
    }
    // W = R (or M^{-1} R) on the active slice, orthonormalised in place (:653-680):
    // on a failure W is replaced by random columns anyway, so the partially
    // processed state orthonormalize leaves does not matter
    if (!orthonormalize(W, ini, na, P_.gs_tol, P_.gs_reorth, scratch)) {
      for (int j = ini; j < B; ++j)
        random_col(W.col(j), bp_);
      for (int j = ini; j < B; ++j)
        project_out(W.col(j), bp_);
    }
    if (its > 1)  // (:683-691): P keeps whatever state orthonormalize leaves
      (void)orthonormalize(Pb, ini, na, P_.gs_tol, P_.gs_reorth, scratch);
    // Z = [X_act | W_act | P_act] (:693-704)
    const int np0 = (its > 1) ? na : 0;
    dcopy(Z.col(0), X.col(ini), (std::size_t)na * cb);
    dcopy(Z.col(na), W.col(ini), (std::size_t)na * cb);
    if (np0)
      dcopy(Z.col(2 * na), Pb.col(ini), (std::size_t)na * cb);
    const int nv = 2 * na + np0;
    if (nv == 0)
      continue;
    if (P_.max_tiny_dim > 0 && nv > P_.max_tiny_dim)
      throw std::runtime_error("LOBPCG: tiny problem too large (nv cap)");
    apply_normal(Z, nv, AZ);
    dense::Mat S, G;
    gram(Z, AZ, nv, S, G);
    auto th = Clock::now();
    TinyGEP sol;
    int np_ = np0;
    if (!try_tiny_gep(G, S, bp_ == Prec::F32, sol)) {
      if (bp_ != Prec::F32 || np_ == 0)
        throw std::runtime_error("LOBPCG tiny GEP: S not SPD (after jitter)");
      // FP32 blocks: P has become dependent on [X W] to working precision;
      // restart this step without it (leading 2na x 2na Gram blocks)
      np_ = 0;
      dense::Mat S2(2 * na, 2 * na), G2(2 * na, 2 * na);
      for (int j = 0; j < 2 * na; ++j)
        for (int i = 0; i < 2 * na; ++i) {
          S2(i, j) = S(i, j);
          G2(i, j) = G(i, j);
        }
      sol = solve_tiny_gep(G2, S2, true);
      ++S_.p_drops;
    }
    S_.t_host += since(th);
    S_.t_gep += since(th);
    const int take = std::min<int>(na, (int)sol.theta.size());
    // P_new = [W P] Y[na:, :take]; X_new = X Y[:na, :take] + P_new (:718-759)
    combine(Z, na, na + np_, sol.Y, na, take, Xn, 0);
    dcopy(Pb.col(ini), Xn.col(0), (std::size_t)take * cb);
    combine(Z, 0, na, sol.Y, 0, take, Xn, 0);
    for (int j = 0; j < take; ++j)
      axpby(c_, bp_, n_, 1.0, Pb.col(ini + j), 1.0, Xn.col(j));
    dcopy(X.col(ini), Xn.col(0), (std::size_t)take * cb);
    if (!orthonormalize(X, ini, take, P_.gs_tol, P_.gs_reorth, scratch))
      throw std::runtime_error("LOBPCG: X slice orthonormalization failed");
    full_rr();
  }
  S_.converged_to_target = (S_.locked >= k_target);
  if (P_.verbose)
    std::fprintf(stderr,
                 "[LOBPCG-HIP] done: it=%lld locked=%lld random columns=%lld (%.3f s) "
                 "tiny GEPs %.3f s host total %.3f s P drops %lld deferred refinements %lld "
                 "Cholesky-QR blocks %lld (MGS fallbacks %lld)\n",
                 (long long)its, (long long)S_.locked, (long long)S_.random_cols, S_.t_rng,
                 S_.t_gep, S_.t_host, (long long)S_.p_drops, (long long)S_.deferred,
                 (long long)S_.cholqr_blocks, (long long)S_.cholqr_fallbacks);
  if (locked_.count > 0) {  // final polish (:806-831)
    const double desired = accept_thr;
    const double target = std::max(1e-14, std::min(1e-10, 0.01 * desired));
    if (bp_ == Prec::F32)
      projector_polish(target);
    block_rr_polish(target, 80);
    for (int i = 0; i < (int)locked_.count; ++i) {
      A_.mv(c_, Prec::F64, locked_.col(i), tm64_.p);
      ++S_.spmv;
      if (nrm(Prec::F64, tm64_.p, m_) <= target)
        continue;
      (void)two_vec_sharpen(i, target, 120);
    }
    mgs_locked(1e-14);
  }
  c_.sync();
  return std::move(locked_);
}

}  // namespace

DeviceBasis lobpcg(Context &ctx, const DeviceOp &op, RefStream &rng, const LobpcgParams &P,
                   LobpcgStats &S) {
  LobpcgDriver d(ctx, op, rng, P, S);
  return d.run();
}

}  // namespace kh
