// Interfaces of the mixed-size batch kernel (batch_kernels.hip) used by the
// plan and serving host code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fftoc_internal.hpp"

namespace fftoc {

// One job as the device sees it.  Items item_begin .. item_begin + items - 1
// of the launch belong to this job; item i covers tiles
// [(i - item_begin) * tiles_per_item, ... + tiles_per_item) of ntiles.
struct JobDev {
  const void *in;
  void *out;
  int64_t batch;
  int64_t ntiles;
  int64_t item_begin;
  int32_t tiles_per_item;
  int32_t routine;
  int32_t inverse;
  int32_t c64;
  double scale;
};

int batch_routine_count();
int batch_routine_n(int id);
const char *batch_routine_name(int id);
int batch_routine_signals(int id);
int batch_resident_blocks();
int launch_batch(const JobDev *d_jobs, int njobs, int64_t total_items, unsigned long long *d_ctr,
                 const void *tw64, const void *tw32, int grid, hipStream_t stream);

}  // namespace fftoc
