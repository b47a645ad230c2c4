// Matrix-free device operators (see operators.hpp).
#include "operators.hpp"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace kh {

StencilOp::StencilOp(Context &ctx, int64_t n_, int64_t N_, double c_) : n(n_), N(N_), c(c_) {
  if (n <= 0 || N < 0 || N >= n)
    throw std::invalid_argument("shifted difference: need n > 0 and 0 <= N < n");
  const int64_t m = n - N;
  // the CSR of row i is (i, -c), (i + N, 1): diag(A A^T) and ||A||_F summed
  // in that order, as csr_create does for the materialised matrix
  double rs = 0.0;
  rs += (-c) * (-c);
  rs += 1.0 * 1.0;
  std::vector<double> h((std::size_t)m, rs);
  double fro2 = 0.0;
  for (int64_t i = 0; i < m; ++i)
    fro2 += rs;
  fro_ = std::sqrt(fro2);
  diag = DBuf((std::size_t)m * 8);
  KH_CHECK(hipMemcpyAsync(diag.p, h.data(), (std::size_t)m * 8, hipMemcpyHostToDevice,
                          ctx.stream));
  ctx.sync();
}

void StencilOp::mv(Context &ctx, Prec p, const void *x, void *y, const void *z,
                   const double *beta) const {
  sd_apply(ctx, n, N, c, false, p, x, y, z, beta);
}

void StencilOp::mtv(Context &ctx, Prec p, const void *x, void *y, const void *z,
                    const double *beta) const {
  sd_apply(ctx, n, N, c, true, p, x, y, z, beta);
}

void StencilOp::mm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y,
                   int64_t ldy, const int *act) const {
  sd_block(ctx, n, N, c, false, p, s, x, ldx, y, ldy, act, false);
}

void StencilOp::mtm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y,
                    int64_t ldy, const int *act) const {
  sd_block(ctx, n, N, c, true, p, s, x, ldx, y, ldy, act, false);
}

void StencilOp::mm_rows(Context &ctx, Prec p, int s, const void *X, void *Y, bool trans) const {
  sd_block(ctx, n, N, c, trans, p, s, X, s, Y, s, nullptr, true);
}

void FunctorOp::call(Context &ctx, Prec p, bool trans, int s, const void *x, int64_t ldx,
                     void *y, int64_t ldy) const {
  ctx.sync();  // the callback sees finished inputs
  const int rc = fn(user, (void *)ctx.stream, p == Prec::F64 ? 0 : 1, trans ? 1 : 0, s, x, ldx,
                    y, ldy);
  if (rc != 0)
    throw std::runtime_error("user operator returned " + std::to_string(rc));
}

void FunctorOp::mv(Context &ctx, Prec p, const void *x, void *y, const void *z,
                   const double *beta) const {
  call(ctx, p, false, 1, x, n, y, m);
  if (z)
    sub_beta(ctx, p, m, beta, z, y);
}

void FunctorOp::mtv(Context &ctx, Prec p, const void *x, void *y, const void *z,
                    const double *beta) const {
  call(ctx, p, true, 1, x, m, y, n);
  if (z)
    sub_beta(ctx, p, n, beta, z, y);
}

void FunctorOp::block(Context &ctx, Prec p, bool trans, int s, const void *x, int64_t ldx,
                      void *y, int64_t ldy, const int *act) const {
  if (s <= 0)
    return;
  if (!act) {
    call(ctx, p, trans, s, x, ldx, y, ldy);
    return;
  }
  // inactive columns keep their contents: apply into scratch, copy the rest
  const int64_t out = trans ? n : m;
  const std::size_t need = (std::size_t)out * s * (p == Prec::F64 ? 8 : 4);
  if (scratch.bytes < need)
    scratch = DBuf(need);
  call(ctx, p, trans, s, x, ldx, scratch.p, out);
  copy_active(ctx, p, out, s, scratch.p, out, y, ldy, act);
}

void FunctorOp::mm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y,
                   int64_t ldy, const int *act) const {
  block(ctx, p, false, s, x, ldx, y, ldy, act);
}

void FunctorOp::mtm(Context &ctx, Prec p, int s, const void *x, int64_t ldx, void *y,
                    int64_t ldy, const int *act) const {
  block(ctx, p, true, s, x, ldx, y, ldy, act);
}

}  // namespace kh
