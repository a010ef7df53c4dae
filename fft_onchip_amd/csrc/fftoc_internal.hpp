// Host-side interfaces between the C ABI (api.cpp) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "../../include/fftoc.h"

namespace fftoc {

struct LaunchArgs {
  const void *in;
  void *out;
  int64_t batch;
  const void *twiddles;  // device table W_4096^j, j = 0..4095 (double2, or
                         // float2 for complex64 data)
  const void *twiddles32;  // the same table rounded to float2
  hipStream_t stream;
  bool inverse;
  double scale;          // 1.0, or 1/N with FFTOC_NORMALIZE
  int tile;              // signals per workgroup override (0 = default)
  uint32_t precision;    // FFTOC_PREC_* bits
  bool c64;              // complex64 data (FFTOC_DATA_C64)
};

constexpr int kMaxDevices = 64;

// Workgroups of KERNEL (THREADS threads each) resident on the whole current
// device: the grid of a persistent launch.  Queried once per (kernel,
// device); concurrent first calls may both query, and store the same value.
template <auto KERNEL, int THREADS>
int64_t resident_workgroups() {
  static std::atomic<int64_t> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, KERNEL, THREADS, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  if (dev < 0 ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const int64_t n = (int64_t)per_cu * cus;
  if (dev >= 0 && dev < kMaxDevices) cache[dev].store(n, std::memory_order_relaxed);
  return n;
}

// mfma_prec_kernels.hip: FP32 / split-BF16 stage GEMMs (n = 64, 256, 4096)
int launch_tensor_precision(int n, const LaunchArgs &a);

// Per-device tables shared with the plan / serving code (api.cpp).
int device_tables(const void **tw64, const void **tw32);
// device_tables.hip: the current device's copies of the module-global tables.
int device_table_addresses(void **tw64, void **tw32);
// batch_kernels.hip: locates the allocation-free batch entry point's scratch
// array on device `dev` (which must be current); called once per device by
// the device set-up.
int batch_noalloc_setup(int dev);
// Plans and servers hold the tables while they exist: fftoc_shutdown refuses
// to free them until every holder is destroyed.
void hold_tables();
void drop_tables();

// A kernel variant for one transform length: name, the entry-point families
// it may serve, and its launcher.
enum FamilyBits : unsigned {
  FAM_RADIX8 = 1u,   // fftoc_legacy8_fft  (radix-8 passes only)
  FAM_RADIX16 = 2u,  // fftoc_legacy16_fft (radix-16 passes only)
  FAM_TENSOR = 4u,   // fftoc_tensor_fft   (FP64 MFMA stages)
  FAM_ANY = 8u,      // fftoc_fft          (every variant)
};

struct Variant {
  const char *name;
  unsigned families;
  int (*launch)(const LaunchArgs &);
  int tile;  // signals per workgroup (per step for persistent forms): FFTOC_TILE
};

// radix_kernels.hip / mfma_kernels.hip: variants available for length n,
// preferred (default) variant of each family first.  c64: variants for
// complex64 data (FP32 arithmetic).
int radix_variants(int n, bool c64, const Variant **out);
int tensor_variants(int n, bool c64, const Variant **out);

}  // namespace fftoc
