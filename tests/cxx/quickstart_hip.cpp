// The quick-start program of the reference (examples/quickstart.cpp) written
// against the unchanged KraNS API, with the HIP adapters in place of the
// Std/Eigen ones: the shifted-difference operator (n=1000, N=3, c=0.5) as a
// device CSR, as the matrix-free HipShiftedDifferenceOperator (the
// reference's StdMatrixAdapter lambdas) and as a HipOperatorAdapter whose
// callback applies the CSR; a seed-42 factory shared by both solvers, and the
// same checks (vector count, max ||A u||, distance to the exact kernel). The
// three operands give the same products, so their bases must agree bitwise.
// Exit code 0 on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "krans/Solvers/CyclicLanczosSolver.hpp"
#include "krans/Solvers/LOBPCGSolver.hpp"
#include "krans/adapters/hip/HipAdapters.hpp"

int main() {
  const std::size_t n = 1000, N = 3;
  const double c = 0.5;
  const std::size_t m = n - N;
  std::vector<int32_t> rp(m + 1), col(2 * m);
  std::vector<double> val(2 * m);
  for (std::size_t i = 0; i < m; ++i) {  // row i: -c at i, +1 at i+N
    rp[i] = (int32_t)(2 * i);
    col[2 * i] = (int32_t)i;
    val[2 * i] = -c;
    col[2 * i + 1] = (int32_t)(i + N);
    val[2 * i + 1] = 1.0;
  }
  rp[m] = (int32_t)(2 * m);

  krans::hip::HipContext ctx(0);
  krans::hip::HipCsrMatrixAdapter A(ctx.get(), m, n, rp.data(), col.data(), val.data());
  krans::hip::HipShiftedDifferenceOperator S(ctx.get(), n, N, c);
  // a user operator: each column through the CSR's own product entry point
  krans::hip::HipOperatorAdapter F(
      ctx.get(), m, n,
      [&](void *, krans::hip::Precision p, bool t, int s, const void *x, int64_t ldx, void *y,
          int64_t ldy) {
        const int prec = p == krans::hip::Precision::FP64 ? KH_PREC_FP64 : KH_PREC_MIXED;
        const std::size_t e = p == krans::hip::Precision::FP64 ? 8 : 4;
        for (int j = 0; j < s; ++j)
          krans::hip::check(kh_spmm_dev(A.handle(), t ? 1 : 0, prec, 1,
                                        static_cast<const char *>(x) + (std::size_t)j * ldx * e,
                                        static_cast<char *>(y) + (std::size_t)j * ldy * e));
      });

  int failures = 0;
  auto report = [&](const char *name, const auto &U) {
    double max_res = 0.0, max_dist = 0.0;
    krans::hip::HipVector r(ctx.get(), m);
    for (const auto &u : U) {
      A.multiply(*u, r);
      max_res = std::max(max_res, r.norm());
      auto d = static_cast<const krans::hip::HipVector &>(*u).to_host();
      for (std::size_t j = 0; j < N; ++j) {
        double dot = 0.0, vv = 0.0;
        for (std::size_t i = j, k = 0; i < n; i += N, ++k) {
          dot += d[i] * std::pow(c, (double)k);
          vv += std::pow(c, 2.0 * k);
        }
        for (std::size_t i = j, k = 0; i < n; i += N, ++k)
          d[i] -= dot / vv * std::pow(c, (double)k);
      }
      double dist = 0.0;
      for (double e : d)
        dist += e * e;
      max_dist = std::max(max_dist, std::sqrt(dist));
    }
    std::printf("%-8s %-9zu %-13.1e %-20.1e\n", name, U.size(), max_res, max_dist);
    if (U.size() != N || max_dist > 1e-8)
      ++failures;
  };

  std::printf("solver   vectors   max ||A u||   max dist(u, ker A)\n");
  std::vector<std::vector<double>> first;  // bases of the CSR run
  const krans::MatrixInterface *ops[3] = {&A, &S, &F};
  const char *names[3] = {"csr", "stencil", "callback"};
  for (int o = 0; o < 3; ++o) {
    krans::hip::HipVectorFactory vf(ctx.get(), n, /*seed=*/42);
    std::printf("-- operand: %s\n", names[o]);
    std::vector<std::vector<double>> got;
    {
      krans::solvers::CyclicLanczosOptions opt;
      opt.k = N;
      auto U = krans::solvers::find_nullspace_cyclic_lanczos(*ops[o], vf, opt);
      report("cyclic", U);
      for (const auto &u : U)
        got.push_back(static_cast<const krans::hip::HipVector &>(*u).to_host());
    }
    {
      krans::solvers::LOBPCGOptions opt;
      opt.nev = N;
      auto U = krans::solvers::find_nullspace_lobpcg(*ops[o], vf, opt);
      report("lobpcg", U);
      for (const auto &u : U)
        got.push_back(static_cast<const krans::hip::HipVector &>(*u).to_host());
    }
    if (o == 0) {
      first = got;
    } else if (got != first) {
      std::printf("%s: basis differs from the CSR run\n", names[o]);
      ++failures;
    }
  }
  krans::hip::HipVectorFactory vf(ctx.get(), n, /*seed=*/42);
  // the reference's error convention: a foreign backend is rejected
  try {
    struct Other final : krans::VectorFactory {
      std::unique_ptr<krans::VectorInterface> create_zero() const override { return nullptr; }
      std::unique_ptr<krans::VectorInterface> create_random() const override { return nullptr; }
      std::unique_ptr<krans::VectorInterface> create_zero_of_size(std::size_t) const override {
        return nullptr;
      }
      std::unique_ptr<krans::VectorInterface> create_random_of_size(std::size_t) const override {
        return nullptr;
      }
    } other;
    krans::solvers::CyclicLanczosOptions opt;
    (void)krans::solvers::find_nullspace_cyclic_lanczos(A, other, opt);
    ++failures;
  } catch (const std::invalid_argument &) {
  }
  return failures == 0 ? 0 : 1;
}
