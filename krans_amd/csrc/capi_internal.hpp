// Definitions of the opaque C-ABI handles, shared by capi.cpp and cxx_api.cpp.
#pragma once
#include <memory>

#include "solver.hpp"

struct kh_context_s {
  std::unique_ptr<kh::Context> ctx;
};
struct kh_matrix_s {
  kh_context_s *owner = nullptr;
  std::unique_ptr<kh::DeviceCsr> A;
};
struct kh_rng_s {
  kh::RefStream s;
  explicit kh_rng_s(unsigned seed) : s(seed) {}
};
