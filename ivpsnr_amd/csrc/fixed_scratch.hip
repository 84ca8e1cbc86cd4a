// Fixed device scratch of the allocation-free frame entry point (qmiv_frame_noalloc).
//
// One statically sized block per device, in the code object's zero-filled data: it is created
// when the library's code object is loaded onto a device and lives until the process ends. Host
// code never uses the array's name as an address (on the host it names a stand-in, not device
// memory): fixed_scratch_setup resolves the device address once per device, with that device
// current, and every later use goes through that address.
#include <mutex>

#include "qmiv_internal.h"

namespace {
constexpr size_t kFixedScratchBytes = (size_t)64 << 20;
constexpr int kMaxFixedDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

namespace qmiv {

size_t fixed_scratch_bytes() { return kFixedScratchBytes; }

hipError_t fixed_scratch_setup(int device, void** ptr) {
  static std::mutex mu;
  static void* table[kMaxFixedDevices] = {};
  *ptr = nullptr;
  if (device < 0 || device >= kMaxFixedDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(mu);
  if (!table[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    table[device] = p;
  }
  *ptr = table[device];
  return hipSuccess;
}

}  // namespace qmiv
