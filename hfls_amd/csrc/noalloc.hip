// Scratch memory of the allocation-free entry point (hfls_project_noalloc): a
// file-scope device array, one copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in symbol); it looks the device address up once per device with
// hipGetSymbolAddress, with that device current, and hands out that address.
#include <hip/hip_runtime.h>

#include <array>
#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.hpp"

namespace hfls {

static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
std::mutex scratch_mu;
std::array<double*, 64> scratch_addr{};  // per device ordinal
}  // namespace

double* noalloc_scratch(int device, int64_t doubles) {
  if (doubles < 0 || doubles > NOALLOC_SCRATCH_DOUBLES)
    throw std::runtime_error("the allocation-free entry point needs " + std::to_string(doubles) +
                             " doubles of scratch, more than its " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                             " (at most 16384 subdomains)");
  if (device < 0 || device >= (int)scratch_addr.size()) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress of the scratch array failed: ") + hipGetErrorString(e));
    scratch_addr[device] = static_cast<double*>(p);
  }
  return scratch_addr[device];
}

}  // namespace hfls
