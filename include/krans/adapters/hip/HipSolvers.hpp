// The HIP backend's solver entry points under their own names, so that
// libkrans_hip.so links beside the reference's library, whose
// krans::solvers::find_nullspace_* (CyclicLanczosSolver.cpp:663-668,
// LOBPCGSolver.cpp:837-842) keep their symbols; a dispatch on the factory
// (is_hip_factory) in those entry points, or libkrans_solvers_hip.so where
// the reference's library is absent, routes HIP calls here.
//
// Operands: a HipVectorFactory (its context, seeded stream and precision)
// with any MatrixInterface. HipMatrix operands (CSR, stencil, device
// callback) on the factory's context run directly on the device; any other
// MatrixInterface (decorators such as a counting wrapper, user operators) is
// applied column by column through HipVector views. Preconditioner
// callbacks (precond_AAt, M_inv) receive HipVector views; the polish CG that
// calls one runs in FP64 with the reference's recurrence
// (CyclicLanczosSolver.cpp:48-113), one column at a time.
#pragma once
#include "krans/Solvers/CyclicLanczosSolver.hpp"
#include "krans/Solvers/LOBPCGSolver.hpp"
#include "krans/adapters/hip/HipAdapters.hpp"

namespace krans::hip {

inline bool is_hip_factory(const VectorFactory &vf) {
  return dynamic_cast<const HipVectorFactory *>(&vf) != nullptr;
}

solvers::CyclicLanczosResult find_nullspace_cyclic_lanczos_ex(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::CyclicLanczosOptions &opt);
std::vector<std::unique_ptr<VectorInterface>> find_nullspace_cyclic_lanczos(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::CyclicLanczosOptions &opt);
solvers::LOBPCGResult find_nullspace_lobpcg_ex(const MatrixInterface &A, const VectorFactory &vf,
                                               const solvers::LOBPCGOptions &opt);
std::vector<std::unique_ptr<VectorInterface>> find_nullspace_lobpcg(
    const MatrixInterface &A, const VectorFactory &vf, const solvers::LOBPCGOptions &opt);

}  // namespace krans::hip
