// The reference's entry points, krans::solvers::find_nullspace_cyclic_lanczos
// [_ex] and find_nullspace_lobpcg[_ex] (CyclicLanczosSolver.hpp:170-178,
// LOBPCGSolver.hpp:142-149), for programs that use this engine without the
// reference's library: each forwards to the HIP backend (libkrans_hip.so).
// Built as libkrans_solvers_hip.so; beside the reference's own library, which
// defines these symbols, a dispatch in its entry points takes its place.
#include "krans/adapters/hip/HipSolvers.hpp"

namespace krans::solvers {

CyclicLanczosResult find_nullspace_cyclic_lanczos_ex(const MatrixInterface &A,
                                                     const VectorFactory &vf,
                                                     const CyclicLanczosOptions &opt) {
  return krans::hip::find_nullspace_cyclic_lanczos_ex(A, vf, opt);
}

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_cyclic_lanczos(const MatrixInterface &A, const VectorFactory &vf,
                              const CyclicLanczosOptions &opt) {
  return krans::hip::find_nullspace_cyclic_lanczos(A, vf, opt);
}

LOBPCGResult find_nullspace_lobpcg_ex(const MatrixInterface &A, const VectorFactory &vf,
                                      const LOBPCGOptions &opt) {
  return krans::hip::find_nullspace_lobpcg_ex(A, vf, opt);
}

std::vector<std::unique_ptr<VectorInterface>>
find_nullspace_lobpcg(const MatrixInterface &A, const VectorFactory &vf,
                      const LOBPCGOptions &opt) {
  return krans::hip::find_nullspace_lobpcg(A, vf, opt);
}

}  // namespace krans::solvers
