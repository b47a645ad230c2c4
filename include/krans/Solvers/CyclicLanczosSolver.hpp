// Cyclic Lanczos (Golub-Kahan) null-space solver: the options and result
// structs and entry points of the reference, with the same names, fields and
// defaults (reference: include/krans/Solvers/CyclicLanczosSolver.hpp:73-178).
// These krans::solvers entry points are defined by libkrans_solvers_hip.so,
// which forwards every call to krans::hip::find_nullspace_cyclic_lanczos[_ex]
// (libkrans_hip.so, krans/adapters/hip/HipSolvers.hpp): a HipVectorFactory
// with any MatrixInterface runs the device-resident driver; other factories
// raise std::invalid_argument ("backend mismatch"). Beside the reference's
// own library the forwarding is instead a dispatch in its entry points.
#pragma once
#include <cstddef>
#include <functional>
#include <memory>
#include <vector>

#include "krans/MatrixInterface.hpp"
#include "krans/VectorFactory.hpp"
#include "krans/VectorInterface.hpp"

namespace krans::solvers {

using AAtPreconditioner =
    std::function<void(const VectorInterface &r_m, VectorInterface &y_m)>;

enum class ReorthMode { None = 0, Single = 1, Double = 2 };

struct CyclicLanczosOptions {
  std::size_t k = 1;
  std::size_t ncv = 48;
  std::size_t max_restarts = 80;
  double tol_abs = 1e-12;
  double tol_rel = 1e-10;
  ReorthMode reorth = ReorthMode::Double;
  bool use_harmonic = true;
  double target_tau = 0.0;
  int nullity_known = -1;
  int rank_known = -1;
  std::vector<std::unique_ptr<VectorInterface>> initial_right_space;
  /** On the HIP backend the callback receives HipVector views of device
   * columns; krans::hip::JacobiAAtPreconditioner on the solve's own matrix
   * runs inside the CG kernels. */
  AAtPreconditioner precond_AAt = nullptr;
  std::size_t cg_maxit = 80;
  double cg_rel_tol = 1e-10;
  double cg_breakdown_tol = 1e-30;
  int projector_passes = 1;
  int grad_max_iters = 3000;
  int frob_samples = 4;
  double prescreen_factor = 0.8;
  std::size_t stagnation_restarts = 30;
  bool return_left_space = false;
  bool verbose = false;
};

struct CyclicLanczosResult {
  std::vector<std::unique_ptr<VectorInterface>> right_null;
  std::vector<std::unique_ptr<VectorInterface>> left_null;
  std::size_t restarts = 0;
  std::size_t krylov_steps_total = 0;
  double frob_norm_est = 0.0;
  bool converged_to_target = false;
  std::size_t cg_solves = 0;
  std::size_t cg_maxit_hits = 0;
};

CyclicLanczosResult find_nullspace_cyclic_lanczos_ex(const MatrixInterface &A,
                                                     const VectorFactory &vf,
                                                     const CyclicLanczosOptions &opt);

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_cyclic_lanczos(const MatrixInterface &A, const VectorFactory &vf,
                              const CyclicLanczosOptions &opt);

}  // namespace krans::solvers
