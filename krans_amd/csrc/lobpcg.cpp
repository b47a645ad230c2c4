// Device-resident LOBPCG driver (placeholder until the block kernels land).
#include "solver.hpp"

namespace kh {
DeviceBasis lobpcg(Context &, const DeviceOp &, RefStream &, const LobpcgParams &,
                   LobpcgStats &) {
  throw std::runtime_error("LOBPCG device driver not built yet");
}
}  // namespace kh
