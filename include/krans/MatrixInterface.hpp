// Abstract matrix interface of the KraNS plugin API, kept unchanged so that
// code written against include/krans/MatrixInterface.hpp of the reference
// compiles against this engine (reference: include/krans/MatrixInterface.hpp:10-24).
#pragma once
#include <cstddef>

namespace krans {

class VectorInterface;

class MatrixInterface {
public:
  virtual ~MatrixInterface() = default;
  /** result <- A * x */
  virtual void multiply(const VectorInterface &x, VectorInterface &result) const = 0;
  /** result <- A^T * x */
  virtual void multiply_transpose(const VectorInterface &x,
                                  VectorInterface &result) const = 0;
  virtual std::size_t rows() const = 0;
  virtual std::size_t cols() const = 0;
};

}  // namespace krans
