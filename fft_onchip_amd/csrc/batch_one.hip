// Kernels specialized to one routine and data type with the direction chosen
// per job at run time (batch_one.hpp), and the host side of all three
// direction forms (the others in batch_one_fwd.hip, batch_one_inv.hip).
#define FFTOC_ONE_DIR 2
#include "batch_one.hpp"

#include <algorithm>

namespace fftoc {
namespace {
const void *const kOneKernels[][2] = {FFTOC_ROUTINES(FFTOC_ONE)};
static_assert(sizeof(kOneKernels) / sizeof(kOneKernels[0]) == kNumRoutines, "one kernel per routine");
}  // namespace

const void *batch_one_kernel_fwd(int routine, bool c64);
const void *batch_one_kernel_inv(int routine, bool c64);

constexpr int kOnePerCu = 2;

const void *one_kernel(int routine, bool c64, int dir) {
  if (routine < 0 || routine >= kNumRoutines) return nullptr;
  if (dir == 0) return batch_one_kernel_fwd(routine, c64);
  if (dir == 1) return batch_one_kernel_inv(routine, c64);
  return kOneKernels[routine][c64 ? 1 : 0];
}

int batch_one_occupancy(int routine, bool c64, int dir, int *cus) {
  const void *k = one_kernel(routine, c64, dir);
  if (!k || dir < 0 || dir > 2) return 0;
  static std::atomic<int64_t> cache[kNumRoutines][2][3][kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  const int t = c64 ? 1 : 0;
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[routine][t][dir][dev].load(std::memory_order_relaxed);
    if (v > 0) {
      *cus = (int)(v & 0xffffffff);
      return (int)(v >> 32);
    }
  }
  int per_cu = 0, n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kRoutines[routine].threads, 0) !=
          hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  if (dev < 0 ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      n <= 0)
    n = 256;
  if (dev >= 0 && dev < kMaxDevices)
    cache[routine][t][dir][dev].store(((int64_t)per_cu << 32) | n, std::memory_order_relaxed);
  *cus = n;
  return per_cu;
}

int64_t batch_one_resident(int routine, bool c64, int dir, int per_cu) {
  int cus = 0;
  const int occ = batch_one_occupancy(routine, c64, dir, &cus);
  if (occ <= 0) return 0;
  // default: at most two workgroups per CU, as the mixed-size kernel: more
  // (the light routines allow up to eight) measured slower on every length,
  // the claims and the launch's tail spread over more workgroups
  const int want = per_cu > 0 ? per_cu : kOnePerCu;
  return (int64_t)std::min(want, occ) * cus;
}

int launch_batch_one(int routine, bool c64, int dir, const JobDev *d_jobs, int njobs,
                     int64_t total_items, unsigned long long *d_ctr, const void *tw, int grid,
                     hipStream_t stream) {
  const void *k = one_kernel(routine, c64, dir);
  if (!k || grid <= 0) return FFTOC_ERROR_INVALID_BATCH;
  void *args[] = {&d_jobs, &njobs, &total_items, &d_ctr, &tw};
  if (hipLaunchKernel(k, dim3(grid),
                      dim3(kRoutines[routine].threads), args, 0, stream) != hipSuccess)
    return FFTOC_ERROR_LAUNCH;
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace fftoc
