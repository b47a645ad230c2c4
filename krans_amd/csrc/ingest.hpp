// Matrix ingestion: IJA triplet files, Matrix Market files and raw triplets
// to the canonical CSR that kh_matrix_create_csr takes (columns strictly
// ascending in each row, duplicates summed in input order, explicit zeros
// kept). Host-side; the device CSC twin is built by csr_create.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kh {

struct HostCsr {
  int64_t m = 0, n = 0;
  std::vector<int32_t> rowptr, col;
  std::vector<double> val;
  int64_t declared_nnz = -1;  // entry count the file header announces (-1: none)
  int64_t entries_read = 0;   // triplets taken from the input (before summing duplicates)
};

// setFromTriplets semantics (used by MatrixLoader.cpp:53-54): every (i, j)
// must lie in [0, m) x [0, n) (std::invalid_argument otherwise).
HostCsr csr_from_triplets(int64_t m, int64_t n, const int32_t *I, const int32_t *J,
                          const double *V, int64_t count);

// load_ija_triplet_matrix (MatrixLoader.cpp:21-56): header "rows cols nnz",
// then 0-based "i j value" triplets until the first one that does not parse;
// rows/cols >= 0 override the header's dimensions.
HostCsr load_ija(const std::string &path, int64_t rows, int64_t cols);

// load_matrix_market (MatrixLoader.cpp:12-18), i.e. Eigen 3.4's loadMarket:
// '%' lines skipped, the first other line with two positive sizes gives
// "M N NNZ", then 1-based "i j value" lines; out-of-range entries are skipped
// with a warning; the header's symmetry field is not applied.
HostCsr load_matrix_market(const std::string &path);

}  // namespace kh
