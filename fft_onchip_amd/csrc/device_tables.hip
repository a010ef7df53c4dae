// The library's constant device tables as module globals: W_4096^j for
// j = 0..4095 in FP64 (double2) and rounded to FP32 (float2).  Every device
// gets its own copy when the code object is loaded, so initialising a device
// only copies the host-built values in: no entry point of the library ever
// allocates device memory for its tables.
#include "fftoc_internal.hpp"
#include "../../include/fftoc_device.hpp"

namespace fftoc {
namespace {
__device__ __align__(256) double2 g_tw64[kTwN];
__device__ __align__(256) float2 g_tw32[kTwN];
}  // namespace

// Device addresses of the current device's tables (hipGetSymbolAddress, with
// the device current: each device's copy lives at its own address).
int device_table_addresses(void **tw64, void **tw32) {
  if (hipGetSymbolAddress(tw64, HIP_SYMBOL(g_tw64)) != hipSuccess ||
      hipGetSymbolAddress(tw32, HIP_SYMBOL(g_tw32)) != hipSuccess)
    return FFTOC_ERROR_DEVICE;
  return FFTOC_SUCCESS;
}

}  // namespace fftoc
