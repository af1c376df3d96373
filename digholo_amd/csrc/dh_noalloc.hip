// Device scratch of the allocation-free batch entry point
// (digHoloProcessBatchDeviceNoAlloc).  The staged kernels' intermediates -- the
// row-spectrum band, the Fourier window and the per-field peaks -- live in one
// file-scope device array of fixed size instead of buffers the library
// allocates, so that entry point calls no allocator.  Each GPU holds its own
// copy of the array for the life of the process; every call on a GPU shares it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dh_internal.h"

namespace dh {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};
}  // namespace

// Device address of this GPU's scratch array.  The array's name is only a
// host-side stand-in; its device address is looked up once per device, with
// that device current, and kept.
hipError_t noalloc_scratch(void** ptr) {
  *ptr = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (!g_scratchPtr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratchPtr[dev] = p;
  }
  *ptr = g_scratchPtr[dev];
  return hipSuccess;
}

}  // namespace dh
