// LOBPCG null-space solver on A^T A: options, result and entry points of the
// reference with the same names, fields and defaults (reference:
// include/krans/Solvers/LOBPCGSolver.hpp:68-149). Defined by
// libkrans_solvers_hip.so as forwards to krans::hip::find_nullspace_lobpcg[_ex]
// (HIP operands run the device-resident driver; M_inv receives HipVector
// views of device columns).
#pragma once
#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

#include "krans/MatrixInterface.hpp"
#include "krans/VectorFactory.hpp"
#include "krans/VectorInterface.hpp"

namespace krans::solvers {

using NormalPreconditioner =
    std::function<void(const VectorInterface &r_n, VectorInterface &y_n)>;

struct LOBPCGOptions {
  std::size_t nev = 1;
  std::size_t block_size = 0;
  std::size_t max_iters = 1000;
  double tol_abs = 1e-12;
  double tol_rel = 1e-10;
  bool soft_locking = true;
  double restart = 0.9;
  double gs_tol = 1e-14;
  bool gs_reorth = true;
  NormalPreconditioner M_inv = nullptr;
  std::vector<std::unique_ptr<VectorInterface>> initial_block;
  int frob_samples = 4;
  std::size_t max_tiny_dim = 0;
  int nullity_known = -1;
  int rank_known = -1;
  bool verbose = false;
  std::size_t progress_every = 100;
};

struct LOBPCGResult {
  std::vector<std::unique_ptr<VectorInterface>> right_null;
  std::vector<double> theta;
  std::vector<double> residuals_Av;
  std::size_t iters = 0;
  std::size_t locked = 0;
  double frob_norm_est = 0.0;
  double accept_threshold = 0.0;
  double best_unaccepted_residual = 0.0;
  std::size_t block_size = 0;
  bool converged_to_target = false;
};

LOBPCGResult find_nullspace_lobpcg_ex(const MatrixInterface &A, const VectorFactory &vf,
                                      const LOBPCGOptions &opt);

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_lobpcg(const MatrixInterface &A, const VectorFactory &vf,
                      const LOBPCGOptions &opt);

}  // namespace krans::solvers
