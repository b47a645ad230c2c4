// Matrix ingestion (see ingest.hpp).
//
// Canonical CSR from triplets: two stable counting sorts (by column, then by
// row) put the entries in (row, column) order while keeping the input order
// among repeats of one (row, column); repeats are then summed left to right,
// which is the order Eigen's setFromTriplets sums them (its row-major staging
// appends in input order and collapses into the first occurrence). Explicit
// zeros are kept. Parsers read the whole file and accept a number only if
// the entire token is one (decimal integer for indices, decimal float for
// values); the first entry that does not parse ends the entry list, like the
// stream extraction loop of the reference loader.
#include "ingest.hpp"

#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace kh {

namespace {

std::string read_file(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in)
    throw std::runtime_error("Cannot open file: " + path);
  std::ostringstream ss;
  ss << in.rdbuf();
  return ss.str();
}

inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}

inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

// [+-]?[0-9]+ fitting in `int`
bool parse_int(const char *b, const char *e, int64_t &out) {
  const char *p = b;
  if (p < e && (*p == '+' || *p == '-'))
    ++p;
  if (p == e)
    return false;
  for (const char *q = p; q < e; ++q)
    if (!is_digit(*q))
      return false;
  std::string t(b, e);
  errno = 0;
  const long long v = std::strtoll(t.c_str(), nullptr, 10);
  if (errno == ERANGE || v < INT_MIN || v > INT_MAX)
    return false;
  out = v;
  return true;
}

// [+-]? (digits [. digits?] | . digits) ([eE] [+-]? digits)?
bool parse_real(const char *b, const char *e, double &out) {
  const char *p = b;
  if (p < e && (*p == '+' || *p == '-'))
    ++p;
  int mant = 0;
  while (p < e && is_digit(*p)) {
    ++p;
    ++mant;
  }
  if (p < e && *p == '.') {
    ++p;
    while (p < e && is_digit(*p)) {
      ++p;
      ++mant;
    }
  }
  if (mant == 0)
    return false;
  if (p < e && (*p == 'e' || *p == 'E')) {
    ++p;
    if (p < e && (*p == '+' || *p == '-'))
      ++p;
    int ed = 0;
    while (p < e && is_digit(*p)) {
      ++p;
      ++ed;
    }
    if (ed == 0)
      return false;
  }
  if (p != e)
    return false;
  std::string t(b, e);
  out = std::strtod(t.c_str(), nullptr);  // correctly rounded, as the stream's num_get
  return true;
}

struct Tokens {
  const char *p, *end;
  bool next(const char *&b, const char *&e) {
    while (p < end && is_space(*p))
      ++p;
    if (p == end)
      return false;
    b = p;
    while (p < end && !is_space(*p))
      ++p;
    e = p;
    return true;
  }
};

}  // namespace

HostCsr csr_from_triplets(int64_t m, int64_t n, const int32_t *I, const int32_t *J,
                          const double *V, int64_t count) {
  if (m < 0 || n < 0 || m > INT_MAX || n > INT_MAX)
    throw std::invalid_argument("triplets: dimensions must be in [0, 2^31)");
  if (count < 0 || (count > 0 && (!I || !J || !V)))
    throw std::invalid_argument("triplets: bad entry arrays");
  if (count > INT_MAX)
    throw std::invalid_argument("triplets: more than 2^31-1 entries");
  for (int64_t t = 0; t < count; ++t)
    if (I[t] < 0 || I[t] >= m || J[t] < 0 || J[t] >= n)
      throw std::invalid_argument("triplets: entry " + std::to_string(t) + " (" +
                                  std::to_string(I[t]) + ", " + std::to_string(J[t]) +
                                  ") outside " + std::to_string(m) + " x " + std::to_string(n));
  // pass 1: stable counting sort by column
  std::vector<int64_t> cp((std::size_t)n + 1, 0);
  for (int64_t t = 0; t < count; ++t)
    ++cp[(std::size_t)J[t] + 1];
  for (int64_t j = 0; j < n; ++j)
    cp[(std::size_t)j + 1] += cp[(std::size_t)j];
  std::vector<int32_t> by_col((std::size_t)count);
  for (int64_t t = 0; t < count; ++t)
    by_col[(std::size_t)cp[(std::size_t)J[t]]++] = (int32_t)t;
  // pass 2: stable counting sort by row of the column-sorted sequence
  std::vector<int64_t> rp((std::size_t)m + 1, 0);
  for (int64_t t = 0; t < count; ++t)
    ++rp[(std::size_t)I[t] + 1];
  for (int64_t i = 0; i < m; ++i)
    rp[(std::size_t)i + 1] += rp[(std::size_t)i];
  std::vector<int32_t> order((std::size_t)count);
  {
    std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
    for (int64_t u = 0; u < count; ++u) {
      const int32_t t = by_col[(std::size_t)u];
      order[(std::size_t)pos[(std::size_t)I[t]]++] = t;
    }
  }
  // merge repeats, summing in input order
  HostCsr H;
  H.m = m;
  H.n = n;
  H.entries_read = count;
  H.rowptr.assign((std::size_t)m + 1, 0);
  H.col.reserve((std::size_t)count);
  H.val.reserve((std::size_t)count);
  for (int64_t i = 0; i < m; ++i) {
    int32_t last = -1;
    for (int64_t u = rp[(std::size_t)i]; u < rp[(std::size_t)i + 1]; ++u) {
      const int32_t t = order[(std::size_t)u];
      if (J[t] == last) {
        H.val.back() = H.val.back() + V[t];
      } else {
        H.col.push_back(J[t]);
        H.val.push_back(V[t]);
        last = J[t];
      }
    }
    H.rowptr[(std::size_t)i + 1] = (int32_t)H.col.size();
  }
  return H;
}

HostCsr load_ija(const std::string &path, int64_t rows, int64_t cols) {
  const std::string text = read_file(path);
  Tokens tk{text.data(), text.data() + text.size()};
  const char *b, *e;
  int64_t hdr[3];
  for (int h = 0; h < 3; ++h)
    if (!tk.next(b, e) || !parse_int(b, e, hdr[h]))
      throw std::runtime_error("Invalid IJA file header: " + path);
  if (rows < 0)
    rows = hdr[0];
  if (cols < 0)
    cols = hdr[1];
  if (rows < 0 || cols < 0)
    throw std::invalid_argument("IJA file " + path + ": negative dimensions");
  std::vector<int32_t> I, J;
  std::vector<double> V;
  if (hdr[2] > 0) {
    I.reserve((std::size_t)hdr[2]);
    J.reserve((std::size_t)hdr[2]);
    V.reserve((std::size_t)hdr[2]);
  }
  while (true) {
    int64_t i, j;
    double v;
    if (!tk.next(b, e) || !parse_int(b, e, i))
      break;
    if (!tk.next(b, e) || !parse_int(b, e, j))
      break;
    if (!tk.next(b, e) || !parse_real(b, e, v))
      break;
    I.push_back((int32_t)i);
    J.push_back((int32_t)j);
    V.push_back(v);
  }
  if ((int64_t)I.size() != hdr[2])
    std::fprintf(stderr, "Warning: Expected %lld entries but read %lld\n", (long long)hdr[2],
                 (long long)I.size());
  HostCsr H = csr_from_triplets(rows, cols, I.data(), J.data(), V.data(), (int64_t)I.size());
  H.declared_nnz = hdr[2];
  return H;
}

HostCsr load_matrix_market(const std::string &path) {
  const std::string text = read_file(path);
  const char *p = text.data(), *end = text.data() + text.size();
  bool sizes = false;
  int64_t M = -1, N = -1, NNZ = -1;
  std::vector<int32_t> I, J;
  std::vector<double> V;
  while (p < end) {
    const char *ls = p;
    while (p < end && *p != '\n')
      ++p;
    const char *le = p;
    if (p < end)
      ++p;
    if (ls < le && *ls == '%')
      continue;
    Tokens tk{ls, le};
    const char *b, *e;
    if (!sizes) {
      int64_t v[3] = {-1, -1, 0};
      int got = 0;
      while (got < 3 && tk.next(b, e)) {
        // sizes are Eigen Index values: parse as integers of any width
        const char *q = b;
        if (q < e && (*q == '+' || *q == '-'))
          ++q;
        bool ok = q < e;
        for (const char *r = q; r < e; ++r)
          ok = ok && is_digit(*r);
        if (!ok)
          break;
        v[got++] = std::strtoll(std::string(b, e).c_str(), nullptr, 10);
      }
      if (got >= 2 && v[0] > 0 && v[1] > 0) {
        M = v[0];
        N = v[1];
        NNZ = got == 3 ? v[2] : 0;
        sizes = true;
        if (NNZ > 0) {
          I.reserve((std::size_t)NNZ);
          J.reserve((std::size_t)NNZ);
          V.reserve((std::size_t)NNZ);
        }
      }
      continue;
    }
    // "i j value", 1-based; a missing or malformed value reads as 0 (a
    // failed stream extraction stores 0), a missing index rejects the line
    int64_t i = 0, j = 0;
    double v = 0.0;
    bool ok_i = tk.next(b, e) && parse_int(b, e, i);
    bool ok_j = ok_i && tk.next(b, e) && parse_int(b, e, j);
    if (ok_j && !(tk.next(b, e) && parse_real(b, e, v)))
      v = 0.0;
    i = ok_i ? i - 1 : -1;
    j = ok_j ? j - 1 : -1;
    if (i >= 0 && j >= 0 && i < M && j < N) {
      I.push_back((int32_t)i);
      J.push_back((int32_t)j);
      V.push_back(v);
    } else {
      std::fprintf(stderr, "Invalid read: %lld,%lld\n", (long long)i, (long long)j);
    }
  }
  if (!sizes)
    throw std::runtime_error("Failed to load Matrix Market file: " + path);
  if ((int64_t)I.size() != NNZ)
    std::fprintf(stderr, "%lld!=%lld\n", (long long)I.size(), (long long)NNZ);
  HostCsr H = csr_from_triplets(M, N, I.data(), J.data(), V.data(), (int64_t)I.size());
  H.declared_nnz = NNZ;
  return H;
}

}  // namespace kh
