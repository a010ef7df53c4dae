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
  int per_cu;            // at most this many workgroups per CU (0 = the kernel's occupancy)
  int *occupancy;        // non-null: the launcher stores its kernel's resident
                         // workgroups per CU here and launches nothing
};

constexpr int kMaxDevices = 64;
constexpr int kMaxPerCu = 63;  // FFTOC_WGS_PER_CU's range

// Resident workgroups per CU of KERNEL (THREADS threads each, no dynamic
// LDS) and the CU count of the current device.  Queried once per (kernel,
// device); concurrent first calls may both query, and store the same value.
struct Occupancy {
  int per_cu;
  int cus;
};
template <auto KERNEL, int THREADS>
Occupancy occupancy() {
  static std::atomic<int64_t> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[dev].load(std::memory_order_relaxed);
    if (v > 0) return Occupancy{(int)(v >> 32), (int)(v & 0xffffffff)};
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, KERNEL, THREADS, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  if (dev < 0 ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  if (dev >= 0 && dev < kMaxDevices)
    cache[dev].store(((int64_t)per_cu << 32) | cus, std::memory_order_relaxed);
  return Occupancy{per_cu, cus};
}

// Workgroups of KERNEL resident on the whole current device, at most
// `per_cu` per CU when per_cu > 0: the grid of a persistent launch.
template <auto KERNEL, int THREADS>
int64_t resident_workgroups(int per_cu = 0) {
  const Occupancy o = occupancy<KERNEL, THREADS>();
  const int k = (per_cu > 0 && per_cu < o.per_cu) ? per_cu : o.per_cu;
  return (int64_t)k * o.cus;
}

// Dynamic LDS (bytes, unused by the kernel) that leaves room for at most
// `per_cu` workgroups of KERNEL per CU: the smallest multiple of 256 bytes
// at which the runtime's occupancy calculation allows per_cu or fewer.  0
// when per_cu is 0 or not below the kernel's own occupancy.  This is how a
// one-workgroup-per-tile launch, whose grid covers the whole batch, is held
// to fewer resident workgroups.  Cached per (kernel, device, per_cu).
template <auto KERNEL, int THREADS>
size_t occupancy_lds(int per_cu) {
  if (per_cu <= 0 || per_cu > kMaxPerCu) return 0;
  const Occupancy o = occupancy<KERNEL, THREADS>();
  if (per_cu >= o.per_cu) return 0;
  static std::atomic<int64_t> cache[kMaxDevices][kMaxPerCu + 1];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[dev][per_cu].load(std::memory_order_relaxed);
    if (v > 0) return (size_t)(v - 1);
  }
  int lds = 0;
  if (dev < 0 ||
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
          hipSuccess ||
      lds <= 0)
    lds = 160 * 1024;
  hipFuncAttributes attr{};
  const size_t stat = hipFuncGetAttributes(&attr, reinterpret_cast<const void *>(KERNEL)) ==
                              hipSuccess
                          ? attr.sharedSizeBytes
                          : 0;
  // occupancy is non-increasing in the dynamic LDS: bisect in 256-byte steps
  size_t lo = 0, hi = ((size_t)lds > stat ? (size_t)lds - stat : 0) / 256;
  while (lo < hi) {
    const size_t mid = (lo + hi) / 2;
    int k = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, KERNEL, THREADS, mid * 256) !=
            hipSuccess ||
        k <= per_cu)
      hi = mid;
    else
      lo = mid + 1;
  }
  const size_t bytes = lo * 256;
  if (dev >= 0 && dev < kMaxDevices)
    cache[dev][per_cu].store((int64_t)bytes + 1, std::memory_order_relaxed);
  return bytes;
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
