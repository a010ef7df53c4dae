// Host-side interfaces between the C ABI (api.cpp) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fftoc.h"

namespace fftoc {

struct LaunchArgs {
  const void *in;
  void *out;
  int64_t batch;
  const void *twiddles;  // device table W_4096^j, j = 0..4095 (double2)
  hipStream_t stream;
  bool inverse;
  double scale;          // 1.0, or 1/N with FFTOC_NORMALIZE
  int tile;              // signals per workgroup override (0 = default)
  uint32_t precision;    // FFTOC_PREC_* bits
};

// radix_kernels.hip
int launch_radix8(int n, const LaunchArgs &a);
int launch_radix16(int n, const LaunchArgs &a);
int launch_radix_mixed(int n, const LaunchArgs &a);

// mfma_kernels.hip
int launch_tensor(int n, const LaunchArgs &a);

}  // namespace fftoc
