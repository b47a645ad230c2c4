// Device scratch of the entry point that allocates nothing
// (kh_cgs2_dev_noalloc): a file-scope device array holds the reduction
// partials and the second-pass coefficients of its kernels, so a call makes
// no hipMalloc. Each GPU the library runs on holds its own copy of the array
// (32 MiB) until the process ends.
//
// In host code the array's name is only a host-side stand-in: its device
// address on each GPU is looked up once per device with hipGetSymbolAddress
// (noalloc_prepare, called by kh_context_create with that device current)
// and only that address is handed to kernels.
#include <mutex>
#include <stdexcept>
#include <string>

#include "engine.hpp"

namespace kh {
constexpr std::size_t kNoAllocScratchDoubles = std::size_t(1) << 22;  // 32 MiB
}

static __device__ __align__(256) double scratch[kh::kNoAllocScratchDoubles];

namespace kh {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_prepare(int device) {
  if (device < 0 || device >= kMaxDevices)
    throw std::invalid_argument("device index out of range for the no-alloc scratch");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[device])
    return;
  int cur = -1;
  KH_CHECK(hipGetDevice(&cur));
  if (cur != device)
    KH_CHECK(hipSetDevice(device));
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != device)
    KH_CHECK(hipSetDevice(cur));
  KH_CHECK(e);
  if (!p)
    throw DeviceError("hipGetSymbolAddress returned a null scratch address");
  g_ptr[device] = static_cast<double *>(p);
}

double *noalloc_scratch(int device, std::size_t *cap_doubles) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0 || device >= kMaxDevices || !g_ptr[device])
    throw std::runtime_error("no-alloc scratch not prepared on device " +
                             std::to_string(device) + " (create a context on it first)");
  *cap_doubles = kNoAllocScratchDoubles;
  return g_ptr[device];
}

}  // namespace kh
