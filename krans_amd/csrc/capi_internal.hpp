// Definitions of the opaque C-ABI handles, shared by capi.cpp and cxx_api.cpp.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>

#include "operators.hpp"
#include "solver.hpp"

struct kh_context_s {
  std::unique_ptr<kh::Context> ctx;
};

// A matrix operand: a device CSR (with its CSC twin) or a matrix-free
// device operator (stencil or user callback).
struct kh_matrix_s {
  kh_context_s *owner = nullptr;
  std::unique_ptr<kh::DeviceCsr> A;  // CSR operands
  std::unique_ptr<kh::DeviceOp> op;  // matrix-free operands (A is null)
  int64_t nnz_equiv = 0;             // stored entries, or those a CSR would hold

  int64_t rows() const { return A ? A->m : op->rows(); }
  int64_t cols() const { return A ? A->n : op->cols(); }
  double fro() const { return A ? A->fro : op->fro(); }
  const kh::DeviceCsr &csr(const char *who) const {
    if (!A)
      throw std::invalid_argument(std::string(who) + ": needs a CSR matrix, not a matrix-free "
                                                     "operator");
    return *A;
  }
  // op(A) = A or A^T as a solver operand
  struct View {
    std::unique_ptr<kh::DeviceOp> own;
    const kh::DeviceOp *p = nullptr;
    const kh::DeviceOp &operator*() const { return *p; }
  };
  View view(bool transposed) const {
    View v;
    if (A)
      v.own = std::make_unique<kh::CsrOp>(*A, transposed);
    else if (transposed)
      v.own = std::make_unique<kh::TransposedOpView>(*op);
    if (v.own)
      v.p = v.own.get();
    else
      v.p = op.get();
    return v;
  }
};

struct kh_rng_s {
  kh::RefStream s;
  explicit kh_rng_s(unsigned seed) : s(seed) {}
};
