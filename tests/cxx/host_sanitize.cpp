// Host-side code of the library under CPU sanitizers (built twice by
// tests/test_cpu_sanitizers.py: AddressSanitizer + UndefinedBehaviorSanitizer,
// and ThreadSanitizer). It drives, without a GPU:
//  * the reference-compatible random stream (kh::RefStream): its prefetch
//    thread and conversion workers, several streams at once from several
//    threads, small rings that wrap, requests of every size, destruction
//    while the producer runs; every value checked bit for bit against
//    std::mt19937 + std::normal_distribution<double>;
//  * the projected-problem kernels (kh::dense): the split SVD, whose two
//    parity blocks are swept on two threads, against the unsplit sweep
//    (bitwise), the symmetric eigensolver and Cholesky;
//  * the matrix loaders (kh::load_matrix_market, kh::load_ija,
//    kh::csr_from_triplets) on well-formed and malformed input;
//  * with -DWITH_ORACLE (ASan/UBSan build only) the oracle's cyclic-Lanczos
//    and LOBPCG solvers on small problems.
// Exits non-zero on the first mismatch; the sanitizers abort on a finding.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dense.hpp"
#include "ingest.hpp"
#include "refstream.hpp"
#ifdef WITH_ORACLE
#include "oracle.hpp"
#endif

namespace {

int failures = 0;
void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++failures;
  }
}

bool same_bits(double a, double b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// One stream, requested in the given chunk sizes, against the standard library.
bool stream_matches(unsigned seed, std::size_t prefetch, const std::vector<std::size_t> &chunks) {
  kh::RefStream rs(seed, prefetch);
  std::mt19937 gen(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  std::vector<double> buf;
  for (std::size_t c : chunks) {
    buf.assign(c, 0.0);
    rs.fill(buf.data(), c);
    for (std::size_t i = 0; i < c; ++i)
      if (!same_bits(buf[i], dist(gen)))
        return false;
  }
  return true;
}

void test_refstream() {
  // chunk patterns: single values, odd sizes, sizes around the ring
  std::vector<std::size_t> a{1, 1, 2, 3, 1000, 65536, 7, 200000, 1};
  std::vector<std::size_t> b{300000, 5, 17, 65535, 65537, 1};
  check(stream_matches(42, std::size_t(1) << 16, a), "refstream seed 42, small ring");
  check(stream_matches(7, std::size_t(1) << 22, b), "refstream seed 7, default ring");
  // several streams at once on several threads
  std::vector<std::thread> th;
  std::vector<int> ok(6, 0);
  for (int t = 0; t < 6; ++t)
    th.emplace_back([&, t] {
      ok[(std::size_t)t] = stream_matches(100u + (unsigned)t,
                                          (std::size_t)1 << (14 + t % 3), t % 2 ? a : b);
    });
  for (auto &x : th)
    x.join();
  for (int t = 0; t < 6; ++t)
    check(ok[(std::size_t)t] != 0, "refstream concurrent streams");
  // destroyed while the producer is still filling its ring
  for (int r = 0; r < 20; ++r) {
    kh::RefStream rs(1000u + (unsigned)r, std::size_t(1) << 20);
    double v[3];
    if (r % 2)
      rs.fill(v, 3);
  }
}

// (l+1) x l matrix with zeros where i + j is even: the refined Golub-Kahan
// matrix's pattern, which svd_right splits into two blocks swept on two threads.
kh::dense::Mat gk_pattern(int l, unsigned seed) {
  std::mt19937 g(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  kh::dense::Mat A(l + 1, l);
  for (int j = 0; j < l; ++j)
    for (int i = 0; i <= l; ++i)
      if ((i + j) % 2)
        A(i, j) = u(g);
  return A;
}

void test_dense() {
  for (int l : {1, 2, 5, 16, 33, 66}) {
    const kh::dense::Mat A = gk_pattern(l, 11u + (unsigned)l);
    std::vector<double> s1, s2;
    kh::dense::Mat V1, V2;
    kh::dense::svd_right(A, s1, V1, true);
    kh::dense::svd_right(A, s2, V2, false);
    bool same = s1.size() == s2.size() && V1.a.size() == V2.a.size();
    for (std::size_t i = 0; same && i < s1.size(); ++i)
      same = same_bits(s1[i], s2[i]);
    for (std::size_t i = 0; same && i < V1.a.size(); ++i)
      same = same_bits(V1.a[i], V2.a[i]);
    check(same, "split SVD bitwise equal to the unsplit sweep");
    std::vector<double> s3;
    kh::dense::Mat V3;
    kh::dense::svd_right(A, s3, V3, true, true);
    check(s3.size() == s1.size(), "even-only SVD size");
  }
  for (int n : {1, 3, 10, 40}) {
    std::mt19937 g(5u + (unsigned)n);
    std::normal_distribution<double> d;
    kh::dense::Mat B(n, n), S(n, n);
    for (auto &x : B.a)
      x = d(g);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double t = 0.0;
        for (int k = 0; k < n; ++k)
          t += B(k, i) * B(k, j);
        S(i, j) = t + (i == j ? 1.0 : 0.0);
      }
    std::vector<double> w;
    kh::dense::Mat Q, L;
    check(kh::dense::sym_eig(S, w, Q), "sym_eig converges");
    check(kh::dense::cholesky(S, L), "cholesky of an SPD matrix");
    (void)kh::dense::lower_inverse(L);
    kh::dense::Mat N(n, n);
    N(0, 0) = -1.0;
    check(!kh::dense::cholesky(N, L), "cholesky rejects a negative pivot");
  }
}

std::string write_tmp(const std::string &name, const std::string &text) {
  const char *dir = std::getenv("TMPDIR");
  std::string p = std::string(dir ? dir : "/tmp") + "/" + name;
  std::ofstream f(p);
  f << text;
  return p;
}

template <class F> bool throws(F &&f) {
  try {
    f();
  } catch (const std::exception &) {
    return true;
  }
  return false;
}

void test_ingest(const std::string &golden) {
  const auto mm = write_tmp("kh_san_ok.mtx", "%%MatrixMarket matrix coordinate real general\n"
                                             "% comment\n3 4 5\n1 1 1.5\n3 4 -2\n2 2 1\n"
                                             "2 2 3\n9 9 1\n");
  const kh::HostCsr A = kh::load_matrix_market(mm);
  check(A.m == 3 && A.n == 4 && A.rowptr.size() == 4, "matrix market sizes");
  check(A.col.size() == 3, "matrix market duplicates summed, out-of-range skipped");
  check(throws([] { kh::load_matrix_market("/nonexistent/kh.mtx"); }), "missing file throws");
  const auto bad = write_tmp("kh_san_bad.mtx", "%%MatrixMarket\n% no size line\n");
  check(throws([&] { kh::load_matrix_market(bad); }), "matrix market without sizes throws");
  const auto trunc = write_tmp("kh_san_trunc.ija", "4 4 3\n0 0 1\n1 x 2\n");
  (void)kh::load_ija(trunc, -1, -1);  // stops at the first line that does not parse
  const kh::HostCsr G = kh::load_ija(golden, -1, -1);
  check(G.m > 0 && G.n > 0 && (int64_t)G.rowptr.size() == G.m + 1, "golden ija loads");
  const int32_t I[] = {0, 1, 1, 0}, J[] = {1, 0, 0, 1};
  const double V[] = {1.0, 2.0, 3.0, 4.0};
  const kh::HostCsr T = kh::csr_from_triplets(2, 2, I, J, V, 4);
  check(T.col.size() == 2 && T.val[0] == 5.0 && T.val[1] == 5.0, "triplets summed");
  const int32_t Ib[] = {2};
  check(throws([&] { kh::csr_from_triplets(2, 2, Ib, J, V, 1); }), "out-of-range triplet throws");
}

#ifdef WITH_ORACLE
void test_oracle() {
  // oriented incidence of two disjoint 4-cycles: null space = 2 indicators
  std::vector<int32_t> rp{0}, ci;
  std::vector<double> va;
  for (int g = 0; g < 2; ++g)
    for (int i = 0; i < 4; ++i) {
      int a = 4 * g + i, b = 4 * g + (i + 1) % 4;
      if (a > b)
        std::swap(a, b);
      ci.push_back(a);
      ci.push_back(b);
      va.push_back(-1.0);
      va.push_back(1.0);
      rp.push_back((int32_t)ci.size());
    }
  orc::CsrOp A(8, 8, rp.data(), ci.data(), va.data());
  orc::Factory f(42);
  orc::CyclicOpts o;
  o.k = 2;
  o.tol_rel = 1e-12;
  o.tol_abs = 0.0;
  const auto r = orc::cyclic_lanczos(A, f, o);
  check(r.right_null.size() == 2, "oracle cyclic finds both indicators");
  orc::ShiftDiffOp S(200, 3, 0.5);
  orc::Factory f2(7);
  orc::LobpcgOpts lo;
  lo.nev = 3;
  lo.tol_rel = 1e-12;
  lo.tol_abs = 0.0;
  const auto q = orc::lobpcg(S, f2, lo);
  check(q.right_null.size() == 3, "oracle LOBPCG finds the shifted-difference kernel");
}
#endif

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: host_sanitize <golden.ija>\n");
    return 2;
  }
  test_refstream();
  test_dense();
  test_ingest(argv[1]);
#ifdef WITH_ORACLE
  test_oracle();
#endif
  std::printf("host_sanitize: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}
