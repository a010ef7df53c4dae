// Kernels specialized to one routine and data type (forward direction):
// batch_one.hpp.
#define FFTOC_ONE_DIR 0
#include "batch_one.hpp"

namespace fftoc {
namespace {
const void *const kOneKernels[][2] = {FFTOC_ROUTINES(FFTOC_ONE)};
static_assert(sizeof(kOneKernels) / sizeof(kOneKernels[0]) == kNumRoutines, "one kernel per routine");
}  // namespace

const void *batch_one_kernel_fwd(int routine, bool c64) {
  return (routine >= 0 && routine < kNumRoutines) ? kOneKernels[routine][c64 ? 1 : 0] : nullptr;
}

}  // namespace fftoc
