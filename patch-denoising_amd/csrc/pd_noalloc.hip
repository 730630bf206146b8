// pdamd_denoise_noalloc: pdamd_denoise for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its workspace is a file-scope device array of fixed size
// that the runtime places once per GPU when the library's code object is
// loaded there and keeps until the process ends.
//
// The array's name is only a host-side stand-in: its device address is taken
// with hipGetSymbolAddress, once per device with that device current, and
// only that address is handed to kernels and copies.  Every call on one GPU
// shares its array: calls that overlap on one GPU (different streams or
// threads) race, and each call re-initialises whatever must start at zero.
#include <hip/hip_runtime.h>

#include <mutex>

#include "patch_denoise_amd.h"

namespace pdamd_detail {
void set_error(const char* msg);
int denoise_capped(pdamd_problem* P, void* ws, size_t cap, void* stream);
}  // namespace pdamd_detail

namespace {

constexpr size_t kScratchBytes = size_t(256) << 20;  // 256 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch array (looked up on first use).
int scratch_address(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    pdamd_detail::set_error("pdamd_denoise_noalloc: no usable current HIP device");
    return PDAMD_EHIP;
  }
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pdamd_detail::set_error("pdamd_denoise_noalloc: hipGetSymbolAddress of the scratch failed");
      return PDAMD_EHIP;
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return PDAMD_OK;
}

}  // namespace

extern "C" {

int pdamd_noalloc_scratch_bytes(size_t* bytes) {
  if (!bytes) {
    pdamd_detail::set_error("null output pointer");
    return PDAMD_EINVAL;
  }
  *bytes = kScratchBytes;
  return PDAMD_OK;
}

int pdamd_denoise_noalloc(pdamd_problem* prob, void* stream) {
  if (!prob) {
    pdamd_detail::set_error("null problem");
    return PDAMD_EINVAL;
  }
  if (prob->stage_ms || prob->stage_launches) {
    pdamd_detail::set_error("pdamd_denoise_noalloc does not collect stage timings");
    return PDAMD_EINVAL;
  }
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return pdamd_detail::denoise_capped(prob, ws, kScratchBytes, stream);
}

}  // extern "C"
