// noalloc.hip -- device scratch of the no-allocation entry point
// (xnyz_compress_batch_device_noalloc): one file-scope array per GPU, sized
// for its limits, so a call carves every device buffer of the pipeline out of
// it instead of allocating.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is ever handed to kernels or
// copies.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "kernels.hpp"

namespace xnyz {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

void* noalloc_scratch(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices)
    throw Error(-1, "device index out of range for the no-allocation scratch");
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    int cur = 0;
    check_hip(hipGetDevice(&cur), "hipGetDevice");
    check_hip(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    const hipError_t e2 = hipSetDevice(cur);
    check_hip(e, "hipGetSymbolAddress(scratch)");
    check_hip(e2, "hipSetDevice");
    if (!p) throw Error(-7, "hipGetSymbolAddress returned no address for the scratch array");
    addr[device] = p;
  }
  return addr[device];
}

}  // namespace xnyz
