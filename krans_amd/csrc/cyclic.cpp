// Device-resident cyclic Lanczos (Golub-Kahan) null-space driver.
//
// Follows find_right_nullspace (src/Solvers/CyclicLanczosSolver.cpp:291-616)
// step for step: the same thresholds, restart/stagnation/fallback logic,
// counters and order of random draws. What differs is where and how the
// vector work runs:
//  * all vectors live in HBM; only the scalars the control flow branches on
//    (beta_j, candidate norms and residuals, CG stop state) come back;
//  * reorthogonalisation is classical Gram-Schmidt applied twice (CGS2, one
//    V^T w then one w -= V h per pass) instead of modified Gram-Schmidt;
//  * the candidates of a restart are processed in windows: the best-ranked
//    one alone, then the rest as one batch whose projector polishes run as a
//    multi-column CG (each column keeps its own scalar recurrence). The
//    host then replays the reference's sequential accept/stall/best logic
//    over the window; candidates after an acceptance are redone against the
//    enlarged basis, so every candidate sees exactly the basis it would see
//    in the sequential loop.
//  * with P.krylov == F32 the Golub-Kahan vectors, products and
//    reorthogonalisation run in FP32 (FP64 inner products); candidates,
//    the locked basis and the polish residuals stay FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>

#include "cyclic_driver.hpp"
#include "dense.hpp"

namespace kh {
namespace {

using Clock = std::chrono::steady_clock;
inline double since(Clock::time_point t0) {
  return std::chrono::duration<double>(Clock::now() - t0).count();
}

inline std::size_t esize(Prec p) { return p == Prec::F64 ? 8 : 4; }

// Leading dimension (elements) for a block of columns of length len: the
// length rounded up to 256 bytes plus 17 x 256 bytes.
inline int64_t pad_ld(int64_t len, std::size_t e) {
  const int64_t q = 256 / (int64_t)e;
  return (len + q - 1) / q * q + 17 * q;
}

// Read n device doubles into out (through the pinned slab mirror).
void readback(Context &ctx, const double *d, int n, double *out) {
  if (n <= 0)
    return;
  if (n > Context::kScalars)
    throw std::length_error("readback too large");
  KH_CHECK(hipMemcpyAsync(ctx.hscal, d, sizeof(double) * (std::size_t)n,
                          hipMemcpyDeviceToHost, ctx.stream));
  ctx.sync();
  std::copy(ctx.hscal, ctx.hscal + n, out);
}

void upload_doubles(Context &ctx, const double *h, int n, double *d) {
  if (n <= 0)
    return;
  if (n > Context::kScalars)
    throw std::length_error("upload too large");
  ctx.upload(d, h, sizeof(double) * (std::size_t)n);  // queued, no wait
}

// Upload a host vector of FP64 (any length) via pageable memory.
void upload_vec(Context &ctx, const double *h, int64_t n, double *d) {
  KH_CHECK(hipMemcpyAsync(d, h, sizeof(double) * (std::size_t)n, hipMemcpyHostToDevice,
                          ctx.stream));
  KH_CHECK(hipStreamSynchronize(ctx.stream));
}


// Automatic graph replay of the Golub-Kahan cycle up to this m + n. Zero:
// on MI355X replay was slower than direct launches at every size measured
// (n = 8K to 1M; the cycle waits on dependent-dispatch latency, not on host
// launch cost), so the automatic setting launches directly.
constexpr int64_t kGraphAutoLen = 0;

// Stall window of the FP32 projector CG of the mixed-precision polish: a
// column whose residual norm has not halved within this many iterations has
// reached the FP32 rounding level and goes to FP64.
constexpr int64_t kStall32 = 64;
// FP32 polish batches in a row whose every column was solved again in FP64
// after which the solve polishes in FP64 directly
constexpr int kPolish32Misses = 4;


}  // namespace

void CyclicDriver::alloc() {
  kp_ = P_.krylov;
  const int64_t ncv = std::max<int64_t>(std::max<int64_t>(2, P_.ncv), k_target_ + 2);
  ycap_ = ncv / 2 + 2;
  xcap_ = ncv / 2 + 2;
  basis_.len = n_;
  basis_.cap = k_target_;
  basis_.count = 0;
  basis_.buf = DBuf((std::size_t)std::max<int64_t>(1, k_target_) * n_ * 8);
  // Column strides are padded off powers of two (by 17 x 256 bytes) so that
  // the same row of consecutive columns does not map to one HBM channel.
  if (!ext_) {
    ldy_ = pad_ld(n_, esize(kp_));
    ldx_ = pad_ld(m_, esize(kp_));
    VB_ = DBuf((std::size_t)(ycap_ + std::max<int64_t>(1, k_target_)) * ldy_ * esize(kp_));
    X_ = DBuf((std::size_t)xcap_ * ldx_ * esize(kp_));
    wx_ = DBuf((std::size_t)m_ * esize(kp_));
    wy_ = DBuf((std::size_t)n_ * esize(kp_));
    vbp_ = VB_.p, xp_ = X_.p, wxp_ = wx_.p, wyp_ = wy_.p;
  }
  y0_ = DBuf((std::size_t)n_ * 8);
  best_v_ = DBuf((std::size_t)n_ * 8);
  tmp_m_ = DBuf((std::size_t)m_ * 8);
  tmp_n_ = DBuf((std::size_t)n_ * 8);
  const std::size_t hsz = (std::size_t)(k_target_ + ycap_ + xcap_ + 8);
  // layout: bnorm[ncv+8] | pass2[ncv+8] | h[4*hsz] | wsc[4096] | C[ycap*(ncv+2)] | H[...]
  const std::size_t n_b = 2 * ((std::size_t)ncv + 8), n_h = 4 * hsz, n_w = 4096,
                    n_c = (std::size_t)ycap_ * (std::size_t)(ncv + 2);
  const std::size_t n_H = (std::size_t)std::max<int64_t>(1, k_target_) * 512;
  small_ = DBuf((n_b + n_h + n_w + n_c + n_H) * 8);
  bnorm_ = small_.as<double>();
  pass2_ = bnorm_ + ncv + 8;
  h_ = bnorm_ + n_b;
  wsc_ = h_ + n_h;
  C_ = wsc_ + n_w;
  H_ = C_ + n_c;
  hbuf_.resize((std::size_t)std::max(n_, m_));
}

void CyclicDriver::ensure_window(int w) {
  if (w <= wcap_)
    return;
  wcap_ = std::max(w, 8);
  const std::size_t cn = (std::size_t)wcap_ * n_ * 8, cm = (std::size_t)wcap_ * m_ * 8;
  Wv_ = DBuf(cn);
  AV_ = DBuf(cm);
  Vc_ = DBuf(cn);
  AVc_ = DBuf(cm);
  Vs_ = DBuf(cn);
  Rs_ = DBuf(cm);
  Vp_ = DBuf(cn);
  AVp_ = DBuf(cm);
  Tn_ = DBuf(cn);
  Wm64_ = DBuf(cm);
  if (P_.polish == Prec::F32)
    Wm_ = DBuf((std::size_t)wcap_ * m_ * 4 * 2);  // FP32 w and rhs
  if (kp_ == Prec::F32) {
    Wv32_ = DBuf((std::size_t)wcap_ * n_ * 4);
    AV32_ = DBuf((std::size_t)wcap_ * m_ * 4);
  }
}

void CyclicDriver::random_vec(double *d) {
  auto t0 = Clock::now();
  rng_.fill(hbuf_.data(), (std::size_t)n_);
  upload_vec(c_, hbuf_.data(), n_, d);
  S_.t.rng += since(t0);
}

double CyclicDriver::norm_dev(Prec p, const void *x, int64_t len, double *slot) {
  dot(c_, p, x, p, x, len, slot, true);
  double v;
  readback(c_, slot, 1, &v);
  return v;
}

// :324-330 gs_against_basis (2 passes), as CGS2 on the locked basis.
void CyclicDriver::gs_basis_vec(double *v) {
  const int nb = (int)basis_.count;
  if (nb == 0)
    return;
  for (int pass = 0; pass < 2; ++pass) {
    gemv_t(c_, Prec::F64, basis_.col(0), n_, nb, Prec::F64, v, n_, h_);
    gemv_n_sub(c_, Prec::F64, basis_.col(0), n_, nb, h_, Prec::F64, v, n_);
  }
}

void CyclicDriver::gs_basis_block(double *W, int q) {
  const int nb = (int)basis_.count;
  if (nb == 0 || q == 0)
    return;
  if ((std::size_t)nb * q > (std::size_t)std::max<int64_t>(1, k_target_) * 512)
    throw std::length_error("candidate window too large for the coefficient buffer");
  for (int pass = 0; pass < 2; ++pass) {
    gemm_tn(c_, Prec::F64, basis_.col(0), n_, nb, Prec::F64, W, n_, q, n_, H_);
    gemm_nn_sub(c_, Prec::F64, basis_.col(0), n_, nb, H_, q, Prec::F64, W, n_, n_);
  }
}

// One classical Gram-Schmidt pass of w against the groups: every coefficient
// is computed from the same w, then all projections are removed.
void CyclicDriver::cgs_pass(void *w, Prec pw, int64_t len, const std::vector<Group> &groups,
                            double *nrm_out) {
  std::size_t off = 0;
  std::vector<std::size_t> offs;
  for (const auto &g : groups) {
    offs.push_back(off);
    if (g.count > 0)
      gemv_t(c_, g.p, g.V, g.ld, g.count, pw, w, len, h_ + off);
    off += (std::size_t)g.count;
  }
  int last = -1;
  for (int i = (int)groups.size() - 1; i >= 0; --i)
    if (groups[(std::size_t)i].count > 0) {
      last = i;
      break;
    }
  for (std::size_t i = 0; i < groups.size(); ++i) {
    const auto &g = groups[i];
    if (g.count <= 0)
      continue;
    gemv_n_sub(c_, g.p, g.V, g.ld, g.count, h_ + offs[i], pw, w, len,
               ((int)i == last) ? nrm_out : nullptr);
  }
  if (last < 0 && nrm_out)
    dot(c_, pw, w, pw, w, len, nrm_out, true);
}

// :115-127
double CyclicDriver::estimate_fro() {
  const std::size_t samples = (std::size_t)std::max(1, P_.frob_samples);
  double sum2 = 0.0;
  for (std::size_t s = 0; s < samples; ++s) {
    random_vec(tmp_n_.as<double>());
    A_.mv(c_, Prec::F64, tmp_n_.as<double>(), tmp_m_.as<double>());
    ++S_.spmv;
    const double nn = norm_dev(Prec::F64, tmp_m_.as<double>(), m_, wsc_);
    sum2 += nn * nn;
  }
  return std::sqrt(sum2 / (double)samples);
}

void CyclicDriver::push_basis(const double *v) {
  const int64_t j = basis_.count;
  KH_CHECK(hipMemcpyAsync(basis_.col(j), v, (std::size_t)n_ * 8, hipMemcpyDeviceToDevice,
                          c_.stream));
  convert(c_, Prec::F64, v, kp_,
          static_cast<char *>(vbp_) + (std::size_t)(ycap_ + j) * ldy_ * esize(kp_), n_);
  basis_.count = j + 1;
}

void CyclicDriver::upload_mask(const std::vector<int> &mask) {
  if (mask_.bytes < mask.size() * sizeof(int))
    mask_ = DBuf(std::max<std::size_t>(mask.size(), 64) * sizeof(int));
  c_.upload(mask_.p, mask.data(), mask.size() * sizeof(int));
}

void CyclicDriver::cg_solve(int q, const void *rhs, void *w, Prec pc, const double *diag,
                            int64_t maxit, double rel_tol, std::vector<CgState> &st,
                            int64_t stall) {
  cg_solve_ops(c_, A_, cg_, q, rhs, w, pc, diag, maxit, rel_tol, P_.cg_breakdown_tol, st,
               stall);
  for (int c = 0; c < q; ++c) {
    const CgState &s = st[(std::size_t)c];
    // an iteration that stops on breakdown (reason 2) or a non-finite r.z
    // (3) has applied A A^T before stopping (CyclicLanczosSolver.cpp:79-95)
    const int64_t prods = s.k + ((s.reason == 2 || s.reason == 3) ? 1 : 0);
    S_.cg_iters += s.k;
    S_.spmv += prods;
    S_.spmtv += prods;
  }
}

// cg_normal_eq_prec (CyclicLanczosSolver.cpp:48-113) with the caller's
// preconditioner P_.precond (or none), for one column in FP64: the same
// recurrence, stop tests and returned (iterations, residual) as the batched
// kernels, with the scalars read back every iteration and the callback
// given the device residual r and the device vector z = M^{-1} r.
CgState CyclicDriver::cg_callback(const double *rhs, double *w, int64_t maxit,
                                  double rel_tol) {
  cg_.ensure(1, m_, n_, Prec::F64);
  double *r = cg_.r.as<double>(), *z = cg_.z.as<double>(), *p = cg_.pv.as<double>();
  double *Ap = cg_.Ap.as<double>(), *tn = cg_.tmp_n.as<double>();
  double *slot = wsc_ + 3072;
  auto dotp = [&](const double *a, const double *b) {
    dot(c_, Prec::F64, a, Prec::F64, b, m_, slot);
    double v = 0.0;
    readback(c_, slot, 1, &v);
    return v;
  };
  auto precondition = [&]() {
    if (P_.precond) {
      c_.sync();
      P_.precond->apply(c_, r, z, m_);
    } else {
      copy(c_, Prec::F64, m_, r, z);
    }
  };
  CgState S{};
  S.done = 1;
  axpby(c_, Prec::F64, m_, 0.0, nullptr, 0.0, w);  // w = 0
  copy(c_, Prec::F64, m_, rhs, r);
  const double rhsn = std::sqrt(std::max(0.0, dotp(r, r)));
  S.rhsn = S.rnorm = rhsn;
  if (!(rhsn > 0.0) || !std::isfinite(rhsn)) {  // zero or non-finite rhs: w = 0
    S.reason = 1;
    return S;
  }
  precondition();
  copy(c_, Prec::F64, m_, z, p);
  double rz = dotp(r, z);
  if (!(rz > 0.0) || !std::isfinite(rz)) {  // preconditioner not SPD on r
    S.reason = 1;
    return S;
  }
  double rnorm = std::sqrt(rz);
  const double tol = std::max(rel_tol * rnorm, 1e-30);
  S.rnorm = rnorm;
  S.tol = tol;
  if (rnorm <= tol)
    return S;
  int64_t k = 0;
  int reason = 4;  // stopped at maxit unless another test ends the loop
  for (; k < maxit; ++k) {
    A_.mtv(c_, Prec::F64, p, tn);
    A_.mv(c_, Prec::F64, tn, Ap);
    ++S_.spmv;
    ++S_.spmtv;
    const double pAp = dotp(p, Ap);
    if (!std::isfinite(pAp) || std::fabs(pAp) <= P_.cg_breakdown_tol) {
      reason = 2;
      break;
    }
    const double alpha = rz / pAp;
    axpby(c_, Prec::F64, m_, alpha, p, 1.0, w);
    axpby(c_, Prec::F64, m_, -alpha, Ap, 1.0, r);
    precondition();
    const double rz_new = dotp(r, z);
    if (!std::isfinite(rz_new)) {
      reason = 3;
      break;
    }
    rnorm = std::sqrt(std::max(0.0, rz_new));
    if (rnorm <= tol) {
      ++k;
      reason = 0;
      break;
    }
    const double beta = rz_new / std::max(1e-300, rz);
    rz = rz_new;
    axpby(c_, Prec::F64, m_, 1.0, z, beta, p);  // p = beta p + z
  }
  S.k = (int32_t)k;
  S.rnorm = rnorm;
  S.reason = reason;
  S_.cg_iters += k;
  return S;
}

void CyclicDriver::polish_cg(int q, const double *rhs, double *w, Prec pc, const double *diag,
                             std::vector<CgState> &st) {
  if (P_.precond) {  // the caller's preconditioner: FP64, column by column
    st.resize((std::size_t)q);
    for (int a = 0; a < q; ++a)
      st[(std::size_t)a] = cg_callback(rhs + (int64_t)a * m_, w + (int64_t)a * m_,
                                       P_.cg_maxit, P_.cg_rel_tol);
    return;
  }
  if (pc == Prec::F64) {
    cg_solve(q, rhs, w, Prec::F64, diag, P_.cg_maxit, P_.cg_rel_tol, st);
    return;
  }
  // Where the FP32 recurrence keeps failing (every column of the last
  // kPolish32Misses batches was solved again in FP64: residuals on the
  // smallest singular values of an ill-conditioned A), the FP32 attempt
  // only costs time: a failed column ends with the FP64 solve below from
  // the same right-hand side, bitwise what this direct FP64 solve gives.
  // The rest of the solve polishes in FP64 directly (the reference's
  // polish), and its columns count as escalated.
  if (polish32_off_) {
    cg_solve(q, rhs, w, Prec::F64, diag, P_.cg_maxit, P_.cg_rel_tol, st);
    S_.escalated += q;
    return;
  }
  // FP32 CG on the rounded FP64 residual; the FP64 residual of the updated
  // v decides whether the pass is kept
  float *W32 = Wm_.as<float>();
  float *R32 = W32 + (int64_t)wcap_ * m_;
  convert(c_, Prec::F64, rhs, Prec::F32, R32, (int64_t)q * m_);
  cg_solve(q, R32, W32, Prec::F32, diag, P_.cg_maxit, std::max(P_.cg_rel_tol, 1e-6), st,
           kStall32);
  convert(c_, Prec::F32, W32, Prec::F64, w, (int64_t)q * m_);
  // A column whose FP32 CG did not converge (its residual stalled at the
  // FP32 rounding level, reached cg_maxit or broke down) is solved again in
  // FP64 from its FP64 right-hand side, as the FP64 polish would: when the
  // residual of A v lies on the smallest singular values, FP32 rounding of
  // A A^T w can keep the recurrence from ever reaching its tolerance, and w
  // then carries error along the null space of A^T large enough to spoil
  // A^T w. Its state (iterations, stop reason) replaces the FP32 one.
  std::vector<int> redo;
  for (int a = 0; a < q; ++a)
    if (st[(std::size_t)a].reason >= 2)
      redo.push_back(a);
  polish32_miss_ = (q > 0 && (int)redo.size() == q) ? polish32_miss_ + 1 : 0;
  if (polish32_miss_ >= kPolish32Misses)
    polish32_off_ = true;
  if (redo.empty())
    return;
  const int qr = (int)redo.size();
  double *Rr = AVp_.as<double>();       // gathered FP64 right-hand sides
  double *Wr = Wm_.as<double>();        // FP64 solutions (the FP32 block is spent)
  for (int t = 0; t < qr; ++t)
    KH_CHECK(hipMemcpyAsync(Rr + (int64_t)t * m_, rhs + (int64_t)redo[(std::size_t)t] * m_,
                            (std::size_t)m_ * 8, hipMemcpyDeviceToDevice, c_.stream));
  std::vector<CgState> st64;
  cg_solve(qr, Rr, Wr, Prec::F64, diag, P_.cg_maxit, P_.cg_rel_tol, st64);
  for (int t = 0; t < qr; ++t) {
    const int a = redo[(std::size_t)t];
    KH_CHECK(hipMemcpyAsync(w + (int64_t)a * m_, Wr + (int64_t)t * m_, (std::size_t)m_ * 8,
                            hipMemcpyDeviceToDevice, c_.stream));
    st[(std::size_t)a] = st64[(std::size_t)t];
  }
  S_.escalated += qr;
}

// Batched projector polish of the window columns `cols` (indices into the
// window block Wv/AV): per column the pass loop of
// CyclicLanczosSolver.cpp:509-542. Slot i holds the current v of cols[i] in
// Vc(:,i) and A v in AVc(:,i).
void CyclicDriver::polish_batch(const std::vector<int> &cols, std::vector<ColOut> &out) {
  const int q = (int)cols.size();
  double *Wv = Wv_.as<double>(), *AV = AV_.as<double>();
  double *Vc = Vc_.as<double>(), *AVc = AVc_.as<double>();
  double *Vs = Vs_.as<double>(), *Rs = Rs_.as<double>();
  double *Vp = Vp_.as<double>(), *AVp = AVp_.as<double>();
  double *Tn = Tn_.as<double>(), *Wm64 = Wm64_.as<double>();
  auto dcopy = [&](double *dst, const double *src, int64_t len) {
    KH_CHECK(hipMemcpyAsync(dst, src, (std::size_t)len * 8, hipMemcpyDeviceToDevice,
                            c_.stream));
  };
  for (int i = 0; i < q; ++i) {
    const int c = cols[(std::size_t)i];
    dcopy(Vc + (int64_t)i * n_, Wv + (int64_t)c * n_, n_);
    dcopy(AVc + (int64_t)i * m_, AV + (int64_t)c * m_, m_);
  }
  std::vector<double> res(q);
  std::vector<int> active(q, 1);
  for (int i = 0; i < q; ++i)
    res[(std::size_t)i] = out[(std::size_t)cols[(std::size_t)i]].res_raw;
  const double *diag = P_.jacobi_AAt ? A_.diag_AAt() : nullptr;
  if (P_.jacobi_AAt && !diag)
    throw std::invalid_argument("Jacobi preconditioner needs a CSR operator");
  std::vector<CgState> st;
  for (int pass = 0; pass < std::max(1, P_.projector_passes); ++pass) {
    std::vector<int> act;
    for (int i = 0; i < q; ++i)
      if (active[(std::size_t)i] && res[(std::size_t)i] > accept_thr_)
        act.push_back(i);
    const int qa = (int)act.size();
    if (qa == 0)
      break;
    const bool dense_act = (qa == q);
    double *Vw = Vc, *Rw = AVc;
    if (!dense_act) {
      Vw = Vs;
      Rw = Rs;
      for (int a = 0; a < qa; ++a) {
        const int i = act[(std::size_t)a];
        dcopy(Vw + (int64_t)a * n_, Vc + (int64_t)i * n_, n_);
        dcopy(Rw + (int64_t)a * m_, AVc + (int64_t)i * m_, m_);
      }
    }
    polish_cg(qa, Rw, Wm64, P_.polish, diag, st);
    for (int a = 0; a < qa; ++a) {
      ColOut &o = out[(std::size_t)cols[(std::size_t)act[(std::size_t)a]]];
      ++o.passes;
      if (st[(std::size_t)a].k >= P_.cg_maxit) {  // :522-526, counted per pass
        ++o.hits;
        o.hit = true;
      }
    }
    // v' = v - A^T w; GS vs basis; normalise; ||A v'||
    block_product(c_, A_, Prec::F64, qa, Wm64, m_, Tn, n_, true, nullptr, rin_, rout_);
    S_.spmtv += qa;
    sub_cols(c_, Prec::F64, n_, qa, Vw, n_, Tn, n_, Vp, n_, nullptr);
    gs_basis_block(Vp, qa);
    // ||v'||, A v' and ||A v'|| in one pass and one read-back; v' and A v'
    // are then normalised in place
    block_product(c_, A_, Prec::F64, qa, Vp, n_, AVp, m_, false, nullptr, rin_, rout_);
    col_norms(c_, Prec::F64, Vp, n_, qa, n_, wsc_);
    col_norms(c_, Prec::F64, AVp, m_, qa, m_, wsc_ + 64);
    std::vector<double> nr((std::size_t)(64 + qa));
    readback(c_, wsc_, 64 + qa, nr.data());
    std::vector<double> sc(qa, 0.0), rp(qa, 0.0);
    std::vector<int> ok(qa, 0);
    for (int a = 0; a < qa; ++a) {
      const double v = nr[(std::size_t)a];
      if (std::isfinite(v) && v >= 1e-14) {
        ok[(std::size_t)a] = 1;
        sc[(std::size_t)a] = 1.0 / v;
        rp[(std::size_t)a] = nr[(std::size_t)(64 + a)] / v;
      }
    }
    upload_doubles(c_, sc.data(), qa, wsc_ + 1024);
    scale_cols(c_, Prec::F64, n_, qa, Vp, n_, wsc_ + 1024, true);
    scale_cols(c_, Prec::F64, m_, qa, AVp, m_, wsc_ + 1024, true);
    for (int a = 0; a < qa; ++a)
      S_.spmv += ok[(std::size_t)a];
    for (int a = 0; a < qa; ++a) {
      const int i = act[(std::size_t)a];
      if (!ok[(std::size_t)a] || !(rp[(std::size_t)a] < res[(std::size_t)i])) {
        active[(std::size_t)i] = 0;
        continue;
      }
      dcopy(Vc + (int64_t)i * n_, Vp + (int64_t)a * n_, n_);
      dcopy(AVc + (int64_t)i * m_, AVp + (int64_t)a * m_, m_);
      res[(std::size_t)i] = rp[(std::size_t)a];
    }
    if (P_.verbose)
      for (int a = 0; a < qa; ++a)
        std::fprintf(stderr, "[Cyclic-HIP] polish pass=%d cg_its=%d reason=%d ||A v'||=%.3e\n",
                     pass, st[(std::size_t)a].k, st[(std::size_t)a].reason, rp[(std::size_t)a]);
  }
  for (int i = 0; i < q; ++i)
    out[(std::size_t)cols[(std::size_t)i]].res = res[(std::size_t)i];
}

// :140-214
bool CyclicDriver::fallback_one() {
  double *v = best_v_.as<double>();
  double *tm = tmp_m_.as<double>(), *tn = tmp_n_.as<double>();
  const int nb0 = (int)basis_.count;
  auto gs1 = [&]() {  // single pass vs basis
    if (basis_.count == 0)
      return;
    gemv_t(c_, Prec::F64, basis_.col(0), n_, (int)basis_.count, Prec::F64, v, n_, h_);
    gemv_n_sub(c_, Prec::F64, basis_.col(0), n_, (int)basis_.count, h_, Prec::F64, v, n_);
  };
  (void)nb0;
  random_vec(v);
  gs1();
  double nv = norm_dev(Prec::F64, v, n_, wsc_);
  if (nv < 1e-14)
    return false;
  axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, v);
  A_.mv(c_, Prec::F64, v, tm);
  ++S_.spmv;
  ensure_window(1);
  double *w = Wm64_.as<double>();
  std::vector<CgState> st;
  cg_solve(1, tm, w, Prec::F64, nullptr, 500, 1e-12, st);
  const int its = st[0].k;
  const double rnorm = st[0].rnorm;
  const double ntm = norm_dev(Prec::F64, tm, m_, wsc_);
  if (its > 0 || rnorm < ntm) {
    A_.mtv(c_, Prec::F64, w, tn);
    ++S_.spmtv;
    axpby(c_, Prec::F64, n_, -1.0, tn, 1.0, v);
  }
  gs1();
  nv = norm_dev(Prec::F64, v, n_, wsc_);
  if (nv >= 1e-14)
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, v);
  A_.mv(c_, Prec::F64, v, tm);
  ++S_.spmv;
  double res = norm_dev(Prec::F64, tm, m_, wsc_);
  if (P_.verbose)
    std::fprintf(stderr, "[Cyclic-HIP] fallback: cg_its=%d rnorm=%.3e ||A v||=%.3e thr=%.3e\n",
                 its, rnorm, res, accept_thr_);
  if (res <= accept_thr_) {
    push_basis(v);
    return true;
  }
  const double fn = estimate_fro();
  const double L2 = std::max(1e-30, fn * fn);
  const double tau = 0.9 / L2;
  for (int it = 0; it < P_.grad_max_iters && res > accept_thr_; ++it) {
    A_.mv(c_, Prec::F64, v, tm);
    A_.mtv(c_, Prec::F64, tm, tn);
    S_.spmv++;
    S_.spmtv++;
    axpby(c_, Prec::F64, n_, -tau, tn, 1.0, v);
    gs1();
    nv = norm_dev(Prec::F64, v, n_, wsc_);
    if (nv < 1e-14)
      return false;
    axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nv, v);
    A_.mv(c_, Prec::F64, v, tm);
    ++S_.spmv;
    res = norm_dev(Prec::F64, tm, m_, wsc_);
  }
  if (P_.verbose)
    std::fprintf(stderr, "[Cyclic-HIP] fallback gradient: ||A v||=%.3e\n", res);
  if (res <= accept_thr_) {
    push_basis(v);
    return true;
  }
  return false;
}

DeviceBasis CyclicDriver::run() {
  if (prepare())
    return std::move(basis_);
  while (cycle_begin()) {
    const CycleQueued q = queue_cycle();
    const int st = cycle_end(q.queued, q.nmv, q.nmtv);
    if (st == kDone)
      return finish();
    if (st == kStop)
      break;
  }
  return after_loop();
}

// The phases of the restart loop of find_right_nullspace
// (CyclicLanczosSolver.cpp:291-616). prepare: the norm estimate, thresholds
// and start vector (:306-342); 1 when there is nothing to compute.
int CyclicDriver::prepare() {
  basis_.len = n_;
  if (n_ == 0 || m_ == 0)
    return 1;
  k_target_ = (P_.nullity_known >= 0) ? P_.nullity_known : P_.k;
  if (k_target_ == 0) {
    S_.converged_to_target = true;
    return 1;
  }
  if (P_.rank_known >= 0 && n_ <= P_.rank_known)
    return 1;
  ncv_ = std::max<int64_t>(std::max<int64_t>(2, P_.ncv), k_target_ + 2);
  alloc();

  S_.frob_norm_est = estimate_fro();
  accept_thr_ = std::max(P_.tol_abs, P_.tol_rel * std::max(1.0, S_.frob_norm_est));
  prescreen_thr_ = std::max(0.0, P_.prescreen_factor) * std::max(1.0, S_.frob_norm_est);
  // unit vectors have ||A v|| <= ||A||_2 <= ||A||_F: a screen below a tenth
  // of the norm can reject candidates
  screen_ = kp_ == Prec::F32 && P_.prescreen_factor < 0.1 && P_.screen >= 0;
  graphs_ = !ext_ && A_.capturable() && c_.trace_mask == 0 &&
            (P_.graphs > 0 || (P_.graphs == 0 && m_ + n_ <= kGraphAutoLen));
  select2_ = P_.reorth == 2 &&
             (P_.second_pass == 2 || (P_.second_pass == 0 && kp_ == Prec::F32));
  dgks_cancel_ = 1.0 - P_.second_pass_eta * P_.second_pass_eta;

  double *y0 = y0_.as<double>();
  if (P_.y0_host)
    upload_vec(c_, P_.y0_host, n_, y0);
  else
    random_vec(y0);
  bh_.assign(2 * ((std::size_t)ncv_ + 8), 0.0);
  restart_ = 0;
  restarts_since_accept_ = 0;
  return 0;
}

void CyclicDriver::random_right() {  // :332-336
  double *y0 = y0_.as<double>();
  random_vec(y0);
  gs_basis_vec(y0);
}

void *CyclicDriver::Ycol(int64_t j) const {
  return static_cast<char *>(vbp_) + (std::size_t)(ycap_ - 1 - j) * ldy_ * esize(kp_);
}

void *CyclicDriver::Xcol(int64_t j) const {
  return static_cast<char *>(xp_) + (std::size_t)j * ldx_ * esize(kp_);
}

// The start vector z0 = [0; y0] of the next cycle in Y's first column
// (:360-372); false when the restart loop ends.
bool CyclicDriver::cycle_begin() {
  if (restart_ >= P_.max_restarts)
    return false;
  S_.restarts = restart_;
  alpha_.clear();
  beta_.clear();
  tcyc_ = Clock::now();
  const Prec kp = kp_;
  double *y0 = y0_.as<double>();
  // ---- start vector z0 = [0; y0] orthogonal to the locked vectors
  gs_basis_vec(y0);
  double ny = norm_dev(Prec::F64, y0, n_, wsc_);
  if (!std::isfinite(ny) || ny < 1e-14) {
    random_right();
    ny = norm_dev(Prec::F64, y0, n_, wsc_);
    if (!(ny >= 1e-14))
      return false;  // the basis already spans R^n
  }
  axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / ny, y0);
  convert(c_, Prec::F64, y0, kp, Ycol(0), n_);
  return true;
}

// The Golub-Kahan build of one cycle (:374-421) on this solve's own storage.
CyclicDriver::CycleQueued CyclicDriver::queue_cycle(int64_t j0, int64_t force2) {
  const Prec kp = kp_;
  char *Xb = static_cast<char *>(xp_);
  const int64_t ncv = ncv_;
  // ---- Golub-Kahan build (:387-421), queued without host round trips:
  // beta_j stays on the device (bnorm_[j]) where the next step's -beta z
  // term and the w / beta store read it. The reference's stop test
  // (non-finite or < 1e-14) is replayed on the host afterwards; steps
  // queued after a stop only write columns that are then ignored.
  const int nb = (int)basis_.count;
  // step j starts from 1 + j / 2 y-vectors and (j + 1) / 2 x-vectors
  int64_t nyc = 1 + j0 / 2, nxc = (j0 + 1) / 2, nmv = 0, nmtv = 0;
  // The queued cycle depends on nothing but the basis size: same kernels,
  // shapes and buffers, with beta_j read on the device. Where the operator
  // only enqueues kernels it is captured once per basis size and replayed.
  auto build = [&]() {
    stepkind_.resize((std::size_t)ncv, 0);
    std::fill(stepkind_.begin() + j0, stepkind_.end(), (char)0);
    for (int64_t j = j0; j < ncv; ++j) {
      const bool from_y = nyc > nxc;
      void *w = from_y ? wxp_ : wyp_;
      const int64_t len = from_y ? m_ : n_;
      if (from_y) {
        A_.mv(c_, kp, Ycol(nyc - 1), w, j == 0 ? nullptr : Xcol(nxc - 1),
              j == 0 ? nullptr : bnorm_ + (j - 1));
        ++nmv;
      } else {
        A_.mtv(c_, kp, Xcol(nxc - 1), w, Ycol(nyc - 1), bnorm_ + (j - 1));
        ++nmtv;
      }
      const int passes = P_.reorth;
      std::vector<Group> groups;
      if (from_y)
        groups.push_back({Xb, ldx_, (int)nxc, kp});
      else  // [Y_{nyc-1} .. Y_0 | b_0 .. b_{nb-1}], contiguous
        groups.push_back({Ycol(nyc - 1), ldy_, (int)nyc + nb, kp});
      const Group &g0 = groups.front();
      bool stored = false;  // w / beta_j already written to its basis column
      if (passes == 2 && g0.count > 0 && select2_ && g0.count < 512 && j != force2) {
        // Selective second pass, speculative: one CGS pass (h1 = V^T w with
        // ||w||^2 from the same read of w; w -= V h1 with ||w||), and the
        // DGKS test of the pass (its second pass is due when it cancelled
        // half of ||w||^2 or more) recorded in pass2_[j]. The host reads
        // the tests back with beta_j and, should a step need its second
        // pass, queues the cycle again from that step with both passes
        // there: the cycle ends as if each step had decided on the device,
        // and the common step costs two reads of V and five launches.
        gemv_t(c_, g0.p, g0.V, g0.ld, g0.count, kp, w, len, h_, /*wsq=*/true);
        if (nyc + nxc < ncv) {
          gemv_n_sub_normalise(c_, g0.p, g0.V, g0.ld, g0.count, h_, kp, w, len, bnorm_ + j,
                               from_y ? Xcol(nxc) : Ycol(nyc), h_, pass2_ + j, dgks_cancel_);
          stored = true;
        } else {
          gemv_n_sub(c_, g0.p, g0.V, g0.ld, g0.count, h_, kp, w, len, bnorm_ + j, h_,
                     pass2_ + j, dgks_cancel_);
        }
        stepkind_[(std::size_t)j] = 3;
      } else if (passes == 2 && g0.count > 0) {
        // CGS2 with three reads of V: h1 = V^T w; [w -= V h1, h2 = V^T w]
        // fused; w -= V h2 with ||w||. The arithmetic of two cgs_pass calls,
        // except that h2's partial sums are grouped per block of the fused
        // kernel.
        double *h2 = h_ + g0.count;
        gemv_t(c_, g0.p, g0.V, g0.ld, g0.count, kp, w, len, h_);
        gemv_n_sub_t(c_, g0.p, g0.V, g0.ld, g0.count, h_, kp, w, len, h2);
        if (nyc + nxc < ncv) {  // beta_j and the stored w / beta_j in one launch
          gemv_n_sub_normalise(c_, g0.p, g0.V, g0.ld, g0.count, h2, kp, w, len, bnorm_ + j,
                               from_y ? Xcol(nxc) : Ycol(nyc));
          stored = true;
        } else {
          gemv_n_sub(c_, g0.p, g0.V, g0.ld, g0.count, h2, kp, w, len, bnorm_ + j);
        }
        stepkind_[(std::size_t)j] = 1;
      } else {
        for (int pass = 0; pass < passes; ++pass)
          cgs_pass(w, kp, len, groups, pass == passes - 1 ? bnorm_ + j : nullptr);
      }
      if (passes <= 0)
        dot(c_, kp, w, kp, w, len, bnorm_ + j, true);
      if (nyc + nxc >= ncv)
        break;
      if (!stored)
        scale_copy_dev(c_, kp, len, bnorm_ + j, w, from_y ? Xcol(nxc) : Ycol(nyc));
      if (from_y)
        ++nxc;
      else
        ++nyc;
    }
  };
  if (graphs_ && j0 == 0 && force2 < 0) {
    if (!gk_graph_.holds(nb)) {
      gk_graph_.begin(c_);
      try {
        build();
      } catch (...) {
        gk_graph_.end(c_, nb, false);
        throw;
      }
      gk_graph_.end(c_, nb, true);
      ++S_.graph_captures;
      gk_nyc_ = nyc, gk_nxc_ = nxc, gk_nmv_ = nmv, gk_nmtv_ = nmtv;
      gk_kinds_ = stepkind_;
    } else {
      nyc = gk_nyc_, nxc = gk_nxc_, nmv = gk_nmv_, nmtv = gk_nmtv_;
      stepkind_ = gk_kinds_;
    }
    gk_graph_.launch(c_);
    ++S_.graph_replays;
  } else {
    build();
  }
  return {nyc + nxc, nmv, nmtv};  // steps queued (the last one stores nothing)
}

// Everything after the build of a cycle (:423-585): the replay of the beta_j
// stop test, the projected problem, the candidates with their polish, the
// stagnation guard and the restart vector.
int CyclicDriver::cycle_end(int64_t queued, int64_t nmv, int64_t nmtv) {
  S_.spmv += nmv;
  S_.spmtv += nmtv;
  auto t0 = tcyc_;
  const Prec kp = kp_;
  const int64_t ncv = ncv_;
  double *y0 = y0_.as<double>();
  // beta_j, and the DGKS tests of speculative single passes: the first step
  // up to the stop whose test asks for the second pass is queued again with
  // both passes there (and the steps after it), until none does
  const auto kind_of = [&](int64_t j) {
    return j < (int64_t)stepkind_.size() ? stepkind_[(std::size_t)j] : (char)0;
  };
  for (;;) {
    bool spec = false;
    for (int64_t j = 0; j < queued; ++j)
      spec = spec || kind_of(j) == 3;
    readback(c_, bnorm_, (int)(spec ? ncv + 8 + queued : queued), bh_.data());
    int64_t redo = -1;
    for (int64_t j = 0; spec && j < queued; ++j) {
      if (kind_of(j) == 3 && bh_[(std::size_t)(ncv + 8 + j)] > 0.5) {
        redo = j;
        break;
      }
      const double bj = bh_[(std::size_t)j];
      if (!std::isfinite(bj) || bj < 1e-14 || j + 1 >= ncv)
        break;  // the stop step: later steps are ignored
    }
    if (redo < 0)
      break;
    ++S_.gk_requeued;
    const CycleQueued q = queue_cycle(redo, redo);
    queued = q.queued;
    S_.spmv += q.nmv;
    S_.spmtv += q.nmtv;
  }
  for (int64_t j = 0; j < queued; ++j) {
    const char kind = kind_of(j);
    if (kind == 0)
      continue;
    ++S_.cgs_pairs;
    if (kind == 1)
      ++S_.second_passes;
  }
  double beta_last = 0.0;
  int64_t local_steps = 0;
  for (int64_t j = 0; j < queued; ++j) {
    const double bj = bh_[(std::size_t)j];
    alpha_.push_back(0.0);
    beta_last = bj;
    if (!std::isfinite(bj) || bj < 1e-14 || j + 1 >= ncv)
      break;
    beta_.push_back(bj);
    ++local_steps;
  }
  S_.krylov_steps_total += local_steps;
  S_.t.gk += since(t0);
  const int64_t l = (int64_t)alpha_.size();
  if (l == 0 || !std::isfinite(beta_last))
    return kStop;

  // ---- tiny problem & interior selection (:428-453)
  t0 = Clock::now();
  std::vector<double> ev;
  dense::Mat V;
  std::vector<int> ord;
  bool finite = true;
  if (P_.use_harmonic) {
    dense::Mat Tb((int)l + 1, (int)l);
    for (int64_t i = 0; i < l; ++i) {
      Tb((int)i, (int)i) = alpha_[(std::size_t)i] - P_.target_tau;
      if (i + 1 < l) {
        Tb((int)i, (int)i + 1) = beta_[(std::size_t)i];
        Tb((int)i + 1, (int)i) = beta_[(std::size_t)i];
      }
    }
    Tb((int)l, (int)l - 1) = beta_last;
    // with a zero diagonal (target_tau = 0) T-bar splits into its even- and
    // odd-column blocks; candidates read only even rows, so only the even
    // block's vectors are formed (bitwise the same as the full sweeps)
    dense::svd_right(Tb, ev, V, true, true);
    ord.resize(ev.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  } else {
    dense::Mat T((int)l, (int)l);
    for (int64_t i = 0; i < l; ++i) {
      T((int)i, (int)i) = alpha_[(std::size_t)i];
      if (i + 1 < l) {
        T((int)i, (int)i + 1) = beta_[(std::size_t)i];
        T((int)i + 1, (int)i) = beta_[(std::size_t)i];
      }
    }
    finite = dense::sym_eig(T, ev, V);
    ord.resize(ev.size());
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return std::abs(ev[a] - P_.target_tau) < std::abs(ev[b] - P_.target_tau);
    });
  }
  double smax = 0.0;  // largest singular value of the projected problem
  for (double x : ev) {
    finite = finite && std::isfinite(x);
    smax = std::max(smax, std::fabs(x));
  }
  for (double x : V.a)
    finite = finite && std::isfinite(x);
  S_.t.tiny += since(t0);
  if (!finite) {
    if (P_.verbose)
      std::fprintf(stderr, "[Cyclic-HIP] non-finite projected problem, random restart\n");
    random_right();
    ++restart_;
    return kContinue;
  }

  // ---- candidates (:455-567), processed in windows
  t0 = Clock::now();
  const int64_t test_cap = std::min<int64_t>(l, std::max<int64_t>(k_target_, 8));
  const int ny_used = (int)((l + 1) / 2);
  std::vector<int> cl;
  for (int jbest : ord) {
    double ce = 0.0;
    for (int64_t i = 0; i < l; i += 2)
      ce += V((int)i, jbest) * V((int)i, jbest);
    if (ce >= 1e-20)
      cl.push_back(jbest);
  }
  // coefficient column of candidate jb (ny_used entries); Y_{ny_used-1} ..
  // Y_0 are contiguous (reversed), so the coefficients are stored in
  // reverse order
  auto coeffs = [&](int jb, double *dst) {
    for (int64_t i = 0; i < l; i += 2)
      dst[(std::size_t)(ny_used - 1 - i / 2)] = V((int)i, jb);
  };
  // Mixed precision: every candidate of the restart (at most 64) is screened
  // in FP32 once, before any of them is tested: w = Y c, ||A w|| / ||w||,
  // one read-back. The candidates are orthogonal to the locked basis to
  // FP32 rounding (the basis is part of every y-side reorthogonalisation),
  // and removing basis components from w can only shrink ||w||, so a
  // candidate whose FP32 residual clears the prescreen threshold by more
  // than the FP32 error margin is rejected as the FP64 test would reject
  // it, against the basis of any later point of the restart (its FP32
  // vector becomes the best-so-far restart candidate when it is the best).
  // Every other candidate (possible null vectors) is evaluated in FP64
  // exactly as without the screen. Screening switches off for the solve
  // when it keeps failing to reject candidates (thresholds above the
  // spectrum).
  //
  // With refined extraction and a zero target (the defaults) the screen
  // needs no vectors at all: T-bar then has a zero diagonal, the candidates
  // are the singular vectors s of its even-column block, u = Z s = [0; Y c]
  // and C u = Z_{l+1} T-bar s = [A Y c; 0], so ||A v|| / ||v|| for v = Y c
  // is the singular value itself (X and Y orthonormal; in FP32 to a few
  // units of 6e-8 ||A||_2 per step, far inside the margin). The candidate
  // vector of the best screened candidate is formed only if it is still the
  // best at the end of the restart.
  const int ncl = (int)std::min<std::size_t>(cl.size(), 64);
  std::vector<char> rej32((std::size_t)ncl, 0);
  std::vector<double> nv32((std::size_t)ncl, 0.0), r32((std::size_t)ncl, 0.0);
  float *W32 = nullptr;
  const bool projected = screen_ && P_.screen == 0 && P_.use_harmonic && P_.target_tau == 0.0;
  if (projected && ncl > 0) {
    const double margin = 1e-5 * std::max(1.0, 2.0 * smax);
    for (int c = 0; c < ncl; ++c) {
      const double r = ev[(std::size_t)cl[(std::size_t)c]];
      nv32[(std::size_t)c] = 1.0;
      r32[(std::size_t)c] = r;
      if (std::isfinite(r) && r > prescreen_thr_ + margin)
        rej32[(std::size_t)c] = 1;
    }
  } else if (screen_ && ncl > 0) {
    ensure_window(ncl);
    W32 = Wv32_.as<float>();
    float *AV32 = AV32_.as<float>();
    std::vector<double> Ch((std::size_t)ny_used * ncl, 0.0);
    for (int c = 0; c < ncl; ++c)
      coeffs(cl[(std::size_t)c], Ch.data() + (std::size_t)c * ny_used);
    upload_doubles(c_, Ch.data(), (int)Ch.size(), C_);
    gemm_nn(c_, kp, Ycol(ny_used - 1), ldy_, ny_used, C_, ncl, Prec::F32, W32, n_, n_);
    block_product(c_, A_, Prec::F32, ncl, W32, n_, AV32, m_, false, nullptr, rin_, rout_);
    col_norms(c_, Prec::F32, W32, n_, ncl, n_, wsc_);
    col_norms(c_, Prec::F32, AV32, m_, ncl, m_, wsc_ + 64);
    std::vector<double> nr((std::size_t)(64 + ncl));
    readback(c_, wsc_, 64 + ncl, nr.data());
    // FP32 error of ||A v|| for a unit v: a few units of 6e-8 times
    // ||A||_2, estimated (from below) by the largest singular value of
    // the projected problem; the margin allows 1e-5 of twice that
    const double margin = 1e-5 * std::max(1.0, 2.0 * smax);
    int kept = 0;
    for (int c = 0; c < ncl; ++c) {
      const double nv = nr[(std::size_t)c], r = nr[(std::size_t)(64 + c)] / nv;
      nv32[(std::size_t)c] = nv;
      r32[(std::size_t)c] = r;
      if (std::isfinite(nv) && nv >= 1e-5 && std::isfinite(r) && r > prescreen_thr_ + margin)
        rej32[(std::size_t)c] = 1;
      else
        ++kept;
      if (P_.verbose)
        std::fprintf(stderr, "[Cyclic-HIP] screen cand=%d sigma=%.9e fp32 ||Aw||/||w||=%.9e ||w||=%.6e%s\n",
                     c, ev[(std::size_t)cl[(std::size_t)c]], r, nv, rej32[(std::size_t)c] ? "" : " kept");
    }
    if (ncl >= 4) {
      screen_miss_ = kept * 4 > ncl * 3 ? screen_miss_ + 1 : 0;
      if (screen_miss_ >= 3)
        screen_ = false;
    }
  }
  double best_res = std::numeric_limits<double>::infinity();
  bool have_best_v = false, accepted_cycle = false, polish_stalled = false;
  int best_proj = -1;  // projected screen: candidate (index into cl) still to form as best_v
  int64_t tested = 0;
  std::size_t pos = 0;
  bool single_next = true;
  bool done_all = false;
  while (pos < cl.size() && tested < test_cap) {
    int wn = single_next ? 1 : (int)std::min<int64_t>((int64_t)(cl.size() - pos), test_cap - tested);
    wn = std::min(wn, 64);
    ensure_window(wn);
    double *Wv = Wv_.as<double>(), *AV = AV_.as<double>();
    std::vector<ColOut> out((std::size_t)wn);
    std::vector<int> sub;             // window columns evaluated in FP64
    std::vector<char> screened((std::size_t)wn, 0);
    for (int c = 0; c < wn; ++c) {
      const std::size_t g = pos + (std::size_t)c;
      if (g < rej32.size() && rej32[g]) {
        screened[(std::size_t)c] = 1;
        ColOut &o = out[(std::size_t)c];
        o.valid = true;
        o.nv = nv32[g];
        o.res_raw = o.res = r32[g];
        ++S_.screened;
      } else {
        sub.push_back(c);
      }
    }
    if (!sub.empty()) {  // FP64 coefficients of the columns in `sub`
      std::vector<double> Cs((std::size_t)ny_used * sub.size(), 0.0);
      for (std::size_t t = 0; t < sub.size(); ++t)
        coeffs(cl[pos + (std::size_t)sub[t]], Cs.data() + t * ny_used);
      upload_doubles(c_, Cs.data(), (int)Cs.size(), C_);
    }
    // FP64 evaluation of the columns in `sub` (window column sub[t] sits in
    // column t of Wv / AV)
    const int ws = (int)sub.size();
    std::vector<int> t64((std::size_t)wn, -1);
    for (int t = 0; t < ws; ++t)
      t64[(std::size_t)sub[(std::size_t)t]] = t;
    slot_of_.assign((std::size_t)std::max(ws, 1), -1);
    if (ws > 0) {
      // w = Y c, Gram-Schmidt against the basis; A w, ||w|| and ||A w|| in
      // one pass and one read-back (v = w / ||w||, so ||A v|| = ||A w|| /
      // ||w||); v and A v are then scaled in place for the polish
      gemm_nn(c_, kp, Ycol(ny_used - 1), ldy_, ny_used, C_, ws, Prec::F64, Wv, n_, n_);
      gs_basis_block(Wv, ws);
      block_product(c_, A_, Prec::F64, ws, Wv, n_, AV, m_, false, nullptr, rin_, rout_);
      col_norms(c_, Prec::F64, Wv, n_, ws, n_, wsc_);
      col_norms(c_, Prec::F64, AV, m_, ws, m_, wsc_ + 64);
      std::vector<double> nr((std::size_t)(64 + ws));
      readback(c_, wsc_, 64 + ws, nr.data());
      std::vector<ColOut> o64((std::size_t)ws);
      std::vector<double> sc(ws, 0.0);
      for (int t = 0; t < ws; ++t) {
        const double nv = nr[(std::size_t)t];
        o64[(std::size_t)t].nv = nv;
        if (std::isfinite(nv) && nv >= 1e-8) {
          o64[(std::size_t)t].valid = true;
          sc[(std::size_t)t] = 1.0 / nv;
        }
      }
      upload_doubles(c_, sc.data(), ws, wsc_ + 1024);
      scale_cols(c_, Prec::F64, n_, ws, Wv, n_, wsc_ + 1024, true);
      scale_cols(c_, Prec::F64, m_, ws, AV, m_, wsc_ + 1024, true);
      std::vector<int> pol;
      for (int t = 0; t < ws; ++t) {
        ColOut &o = o64[(std::size_t)t];
        if (!o.valid)
          continue;
        ++S_.spmv;
        o.res_raw = o.res = nr[(std::size_t)(64 + t)] / o.nv;
        if (std::isfinite(o.res_raw) && !(o.res_raw > prescreen_thr_) && !polish_stalled &&
            o.res_raw > accept_thr_)
          pol.push_back(t);
      }
      for (std::size_t i = 0; i < pol.size(); ++i)
        slot_of_[(std::size_t)pol[i]] = (int)i;
      if (!pol.empty()) {
        auto tp = Clock::now();
        polish_batch(pol, o64);
        S_.t.polish += since(tp);
      }
      for (int t = 0; t < ws; ++t)
        out[(std::size_t)sub[(std::size_t)t]] = o64[(std::size_t)t];
    }
    // replay the sequential accept / stall / best logic over the window
    std::size_t consumed = (std::size_t)wn;
    bool cut = false;
    for (int c = 0; c < wn; ++c) {
      ColOut &o = out[(std::size_t)c];
      if (!o.valid)
        continue;
      ++tested;
      if (!std::isfinite(o.res_raw))
        continue;
      if (o.res_raw > prescreen_thr_ || polish_stalled) {
        if (o.res_raw < best_res) {
          best_res = o.res_raw;
          best_proj = -1;
          if (screened[(std::size_t)c] && projected) {
            best_proj = (int)(pos + (std::size_t)c);
          } else if (screened[(std::size_t)c]) {  // the unnormalised FP32 screen vector
            const double sc = 1.0 / o.nv;
            convert(c_, Prec::F32, W32 + (int64_t)(pos + (std::size_t)c) * n_, Prec::F64,
                    best_v_.p, n_);
            upload_doubles(c_, &sc, 1, wsc_ + 1024);
            scale_cols(c_, Prec::F64, n_, 1, best_v_.p, n_, wsc_ + 1024, true);
          } else
            KH_CHECK(hipMemcpyAsync(best_v_.p, Wv + (int64_t)t64[(std::size_t)c] * n_,
                                    (std::size_t)n_ * 8, hipMemcpyDeviceToDevice, c_.stream));
          have_best_v = true;
        }
        continue;
      }
      const int t = t64[(std::size_t)c];
      const int slot = slot_of_[(std::size_t)t];
      const double *vfin = slot >= 0 ? Vc_.as<double>() + (int64_t)slot * n_ : Wv + (int64_t)t * n_;
      S_.cg_solves += o.passes;
      S_.cg_maxit_hits += o.hits;
      if (P_.verbose)
        std::fprintf(stderr, "[Cyclic-HIP] restart=%lld raw=%.4e polished=%.4e thr=%.4e\n",
                    (long long)restart_, o.res_raw, o.res, accept_thr_);
      if (o.res <= accept_thr_) {
        push_basis(vfin);
        accepted_cycle = true;
        if (basis_.count >= k_target_) {
          S_.converged_to_target = true;
          done_all = true;
          break;
        }
        consumed = (std::size_t)c + 1;
        cut = true;
        break;
      }
      polish_stalled = polish_stalled || o.hit;
      if (o.res < best_res) {
        best_res = o.res;
        best_proj = -1;
        KH_CHECK(hipMemcpyAsync(best_v_.p, vfin, (std::size_t)n_ * 8,
                                hipMemcpyDeviceToDevice, c_.stream));
        have_best_v = true;
      }
    }
    if (done_all)
      break;
    pos += consumed;
    single_next = cut;
  }
  if (best_proj >= 0 && !done_all) {  // v = Y c of the best screened candidate, FP64
    std::vector<double> Cb((std::size_t)ny_used, 0.0);
    coeffs(cl[(std::size_t)best_proj], Cb.data());
    upload_doubles(c_, Cb.data(), ny_used, C_);
    gemm_nn(c_, kp, Ycol(ny_used - 1), ldy_, ny_used, C_, 1, Prec::F64, best_v_.p, n_, n_);
  }
  S_.t.cand += since(t0);
  if (done_all)
    return kDone;
  // ---- stagnation guard (:569-581)
  if (!accepted_cycle) {
    ++restarts_since_accept_;
    if (P_.stagnation_restarts > 0 && restarts_since_accept_ >= P_.stagnation_restarts) {
      if (P_.verbose)
        std::fprintf(stderr, "[Cyclic-HIP] stopping due to stagnation\n");
      return kStop;
    }
  } else {
    restarts_since_accept_ = 0;
  }
  // ---- restart vector (:584)
  if (have_best_v)
    KH_CHECK(hipMemcpyAsync(y0, best_v_.p, (std::size_t)n_ * 8, hipMemcpyDeviceToDevice,
                            c_.stream));
  else
    random_right();
  ++restart_;
  return kContinue;
}

// :587-602, then the mixed-precision refinement and the final wait.
DeviceBasis CyclicDriver::after_loop() {
  // ---- fallback (:587-602)
  if (basis_.count < k_target_) {
    auto t0 = Clock::now();
    int failures = 0;
    while (basis_.count < k_target_ && failures < 8) {
      if (fallback_one())
        failures = 0;
      else
        ++failures;
    }
    S_.converged_to_target = basis_.count >= k_target_;
    S_.t.fallback += since(t0);
  }
  return finish();
}

DeviceBasis CyclicDriver::finish() {
  if (P_.polish == Prec::F32) {
    auto t0 = Clock::now();
    refine_basis();
    S_.t.polish += since(t0);
  }
  c_.sync();
  return std::move(basis_);
}

// Mixed precision only: the FP32 projector polish accepts a vector once its
// FP64 residual is below the threshold, which can leave it just below it.
// One batched FP64 projector pass v <- v - A^T (A A^T)^{-1} A v (FP64 CG at
// cg_rel_tol, the Jacobi preconditioner when set) refines every locked vector
// whose residual is above the final-polish target of LOBPCGSolver.cpp:806-811,
// max(1e-14, min(1e-10, 0.01 thr)); a refined vector replaces the locked one
// only if its normalised residual dropped (the pass rule of
// CyclicLanczosSolver.cpp:527-539), and the basis is then orthonormalised
// again in place (CGS2, column by column). A basis already below the target
// costs one block product and is left bitwise unchanged.
void CyclicDriver::refine_basis() {
  const int k = (int)basis_.count;
  if (k == 0)
    return;
  const double target = std::max(1e-14, std::min(1e-10, 0.01 * accept_thr_));
  constexpr int kChunk = 64;  // window blocks and scalar slots hold 64 columns
  ensure_window(std::min(k, kChunk));
  double *AV = AV_.as<double>(), *Vs = Vs_.as<double>(), *Rs = Rs_.as<double>();
  double *Vp = Vp_.as<double>(), *AVp = AVp_.as<double>(), *Tn = Tn_.as<double>();
  double *Wm64 = Wm64_.as<double>();
  auto dcopy = [&](double *dst, const double *src, int64_t len) {
    KH_CHECK(hipMemcpyAsync(dst, src, (std::size_t)len * 8, hipMemcpyDeviceToDevice,
                            c_.stream));
  };
  const double *diag = P_.jacobi_AAt ? A_.diag_AAt() : nullptr;
  int replaced = 0;
  for (int j0 = 0; j0 < k; j0 += kChunk) {
    const int kc = std::min(kChunk, k - j0);
    block_product(c_, A_, Prec::F64, kc, basis_.col(j0), n_, AV, m_, false, nullptr, rin_,
                  rout_);
    S_.spmv += kc;
    col_norms(c_, Prec::F64, AV, m_, kc, m_, wsc_);
    std::vector<double> res((std::size_t)kc);
    readback(c_, wsc_, kc, res.data());
    std::vector<int> sel;  // chunk columns above the target
    for (int j = 0; j < kc; ++j)
      if (std::isfinite(res[(std::size_t)j]) && res[(std::size_t)j] > target)
        sel.push_back(j);
    if (sel.empty())
      continue;
    const int q = (int)sel.size();
    for (int a = 0; a < q; ++a) {
      dcopy(Vs + (int64_t)a * n_, basis_.col(j0 + sel[(std::size_t)a]), n_);
      dcopy(Rs + (int64_t)a * m_, AV + (int64_t)sel[(std::size_t)a] * m_, m_);
    }
    std::vector<CgState> st;
    polish_cg(q, Rs, Wm64, Prec::F64, diag, st);
    block_product(c_, A_, Prec::F64, q, Wm64, m_, Tn, n_, true, nullptr, rin_, rout_);
    S_.spmtv += q;
    sub_cols(c_, Prec::F64, n_, q, Vs, n_, Tn, n_, Vp, n_, nullptr);
    col_norms(c_, Prec::F64, Vp, n_, q, n_, wsc_);
    std::vector<double> np((std::size_t)q), sc((std::size_t)q, 0.0);
    readback(c_, wsc_, q, np.data());
    std::vector<int> ok((std::size_t)q, 0);
    // a vector that loses half its length to the projection was not close
    // to the null space: keep the locked one
    for (int a = 0; a < q; ++a)
      if (std::isfinite(np[(std::size_t)a]) && np[(std::size_t)a] >= 0.5) {
        ok[(std::size_t)a] = 1;
        sc[(std::size_t)a] = 1.0 / np[(std::size_t)a];
      }
    upload_doubles(c_, sc.data(), q, wsc_ + 1024);
    scale_cols(c_, Prec::F64, n_, q, Vp, n_, wsc_ + 1024, true);
    upload_mask(ok);
    block_product(c_, A_, Prec::F64, q, Vp, n_, AVp, m_, false, mask_.as<int>(), rin_, rout_);
    S_.spmv += q;
    col_norms(c_, Prec::F64, AVp, m_, q, m_, wsc_ + 2048);
    std::vector<double> rp((std::size_t)q);
    readback(c_, wsc_ + 2048, q, rp.data());
    for (int a = 0; a < q; ++a) {
      const int j = sel[(std::size_t)a];
      if (!ok[(std::size_t)a] || !(rp[(std::size_t)a] < res[(std::size_t)j]))
        continue;
      dcopy(basis_.col(j0 + j), Vp + (int64_t)a * n_, n_);
      ++replaced;
    }
  }
  S_.refined += replaced;
  if (replaced == 0)
    return;
  // orthonormalise again: column j against columns 0..j-1, two passes
  for (int j = 0; j < k; ++j) {
    double *u = basis_.col(j);
    if (j > 0)
      for (int pass = 0; pass < 2; ++pass) {
        gemv_t(c_, Prec::F64, basis_.col(0), n_, j, Prec::F64, u, n_, h_);
        gemv_n_sub(c_, Prec::F64, basis_.col(0), n_, j, h_, Prec::F64, u, n_);
      }
    const double nu = norm_dev(Prec::F64, u, n_, wsc_);
    if (std::isfinite(nu) && nu > 0.0)
      axpby(c_, Prec::F64, n_, 0.0, nullptr, 1.0 / nu, u);
  }
}

void block_product(Context &c, const DeviceOp &A, Prec p, int q, const void *X, int64_t ldx,
                   void *Y, int64_t ldy, bool trans, const int *mask, DBuf &rin, DBuf &rout) {
  if (q <= 0)
    return;
  const int64_t nin = trans ? A.rows() : A.cols(), nout = trans ? A.cols() : A.rows();
  auto by_cols = [&] {
    if (trans)
      A.mtm(c, p, q, X, ldx, Y, ldy, mask);
    else
      A.mm(c, p, q, X, ldx, Y, ldy, mask);
  };
  if (q < 2 || !A.has_rows()) {
    by_cols();
    return;
  }
  const std::size_t e = p == Prec::F64 ? 8 : 4;
  if (rin.bytes < (std::size_t)nin * q * e)
    rin = DBuf((std::size_t)nin * q * e);
  if (rout.bytes < (std::size_t)nout * q * e)
    rout = DBuf((std::size_t)nout * q * e);
  auto by_rows = [&] {
    cols_to_rows(c, p, nin, q, X, ldx, rin.p);
    A.mm_rows(c, p, q, rin.p, rout.p, trans);
    rows_to_cols(c, p, nout, q, rout.p, Y, ldy);
  };
  std::atomic<signed char> *cache = A.block_path(p, trans, q);
  signed char path = cache ? cache->load(std::memory_order_relaxed) : 2;
  if (path == 0) {  // first product of this kind: time both (second run of each)
    auto timed = [&](auto &&fn) {
      fn();
      hipEvent_t e0, e1;
      KH_CHECK(hipEventCreate(&e0));
      KH_CHECK(hipEventCreate(&e1));
      KH_CHECK(hipEventRecord(e0, c.stream));
      fn();
      KH_CHECK(hipEventRecord(e1, c.stream));
      KH_CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      KH_CHECK(hipEventElapsedTime(&ms, e0, e1));
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      return ms;
    };
    const float t_cols = timed(by_cols), t_rows = timed(by_rows);
    path = t_rows < t_cols ? 2 : 1;
    cache->store(path, std::memory_order_relaxed);
    return;  // Y holds the (identical) product of the last run
  }
  if (path == 2)
    by_rows();
  else
    by_cols();
}

void cg_solve_ops(Context &c, const DeviceOp &A, CgBatch &B, int q, const void *rhs, void *w,
                  Prec p, const double *diag, int64_t maxit, double rel_tol,
                  double breakdown_tol, std::vector<CgState> &st, int64_t stall) {
  const int64_t m = A.rows(), n = A.cols();
  B.ensure(q, m, n, p);
  cg_init(c, B, rhs, w, diag, rel_tol, maxit);
  st.resize((std::size_t)q);
  int64_t iters_done = 0;
  int chunk = 4;
  while (true) {
    KH_CHECK(hipMemcpyAsync(st.data(), B.st.p, sizeof(CgState) * (std::size_t)q,
                            hipMemcpyDeviceToHost, c.stream));
    KH_CHECK(hipStreamSynchronize(c.stream));
    int running = 0;
    for (int k = 0; k < q; ++k)
      running += st[(std::size_t)k].done == 0;
    if (running == 0 || iters_done >= maxit)
      break;
    const int it = (int)std::min<int64_t>(chunk, maxit - iters_done);
    const int *act = B.act.as<int>();
    // two or more columns: A (A^T p) on row-major copies, one coalesced
    // q-wide row per gather (inactive columns are computed too and ignored)
    const bool rows = q >= 2 && A.has_rows();
    if (rows) {
      const std::size_t need = (std::size_t)q * (std::size_t)m * (p == Prec::F64 ? 8 : 4);
      if (B.rows_m.bytes < need)
        B.rows_m = DBuf(need);
    }
    for (int k = 0; k < it; ++k) {
      if (rows) {
        cols_to_rows(c, p, m, q, B.pv.p, m, B.rows_m.p);
        A.mm_rows(c, p, q, B.rows_m.p, B.tmp_n.p, true);
        A.mm_rows(c, p, q, B.tmp_n.p, B.rows_m.p, false);
        rows_to_cols(c, p, m, q, B.rows_m.p, B.Ap.p, m);
      } else {
        A.mtm(c, p, q, B.pv.p, m, B.tmp_n.p, n, act);
        A.mm(c, p, q, B.tmp_n.p, n, B.Ap.p, m, act);
      }
      cg_after_products(c, B, w, diag, maxit, breakdown_tol, stall);
    }
    iters_done += it;
    chunk = std::min(chunk * 2, 16);
  }
}

double cyclic_footprint(int64_t m, int64_t n, const CyclicParams &P) {
  const int64_t k = std::max<int64_t>(1, P.nullity_known >= 0 ? P.nullity_known : P.k);
  const int64_t ncv = std::max<int64_t>(std::max<int64_t>(2, P.ncv), k + 2);
  const double ke = P.krylov == Prec::F64 ? 8.0 : 4.0;
  const double cap = (double)(ncv / 2 + 2);
  const double ldy = (double)pad_ld(n, (std::size_t)ke), ldx = (double)pad_ld(m, (std::size_t)ke);
  double b = 8.0 * k * n;                              // locked basis
  b += (cap + k) * ldy * ke + cap * ldx * ke;          // [Y | basis], X
  b += ke * (m + n) + 8.0 * (2 * n + m + n);           // w_x, w_y, y0, best_v, tmp
  const double w = (double)std::min<int64_t>(64, std::max<int64_t>(8, std::min(ncv, std::max<int64_t>(k, 8))));
  b += w * 8.0 * (5.0 * n + 5.0 * m);                  // candidate windows
  b += w * 8.0 * (5.0 * m + n);                        // CG state (+ row-major block)
  b += w * 8.0 * (m + n);                              // row-major block products
  b += (P.polish == Prec::F32 ? w * 8.0 * m : 0.0);
  b += (P.krylov == Prec::F32 ? w * 4.0 * (n + m) : 0.0);  // FP32 window screen
  return b + (double)(64 << 20);                       // scalars, partials, slack
}

DeviceBasis cyclic_right(Context &ctx, const DeviceOp &op, RefStream &rng,
                         const CyclicParams &P, CyclicStats &S) {
  CyclicDriver d(ctx, op, rng, P, S);
  return d.run();
}

}  // namespace kh
