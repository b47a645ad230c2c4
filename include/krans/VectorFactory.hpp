// Abstract vector factory of the KraNS plugin API, kept unchanged
// (reference: include/krans/VectorFactory.hpp:9-23).
#pragma once
#include <cstddef>
#include <memory>

#include "VectorInterface.hpp"

namespace krans {

class VectorFactory {
public:
  virtual ~VectorFactory() = default;
  virtual std::unique_ptr<VectorInterface> create_zero() const = 0;
  virtual std::unique_ptr<VectorInterface> create_random() const = 0;
  virtual std::unique_ptr<VectorInterface> create_zero_of_size(std::size_t n) const = 0;
  virtual std::unique_ptr<VectorInterface> create_random_of_size(std::size_t n) const = 0;
};

}  // namespace krans
