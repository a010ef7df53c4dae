// Mixed-size batched FFT: one persistent kernel processes many jobs of
// different length, direction and data type (complex128 or complex64).
//
// Work items are chunks of tiles of one job; workgroups claim items in order
// from a device counter (dynamic load balance over uneven lengths) and run the
// job's tile routine.  Every routine is a Stockham tile on exactly 256
// threads with at most 4096 complex values in LDS, so one workgroup shape and
// one LDS allocation serve every length.  The last workgroup to finish resets
// the counters, so the launch can be replayed from a hipGraph without any
// memset node.
#include "batch_internal.hpp"
#include "stockham.hpp"

namespace fftoc {
namespace {

constexpr int kBatchThreads = 256;
constexpr int kLdsElems = lds_elems(4096);

// Routine table: id -> (n, split, signals per tile, staged).  Two candidates
// per length; the plan autotuner picks between them.
#define FFTOC_ROUTINES(X)                              \
  X(0, 8, "4x2/staged/S128", (Split<4, 2>), 128, true)  \
  X(1, 8, "8/staged/S256", (Split<8>), 256, true)        \
  X(2, 16, "4x4/staged/S64", (Split<4, 4>), 64, true)    \
  X(3, 16, "16/staged/S256", (Split<16>), 256, true)     \
  X(4, 32, "8x4/staged/S64", (Split<8, 4>), 64, true)    \
  X(5, 32, "4x4x2/staged/S32", (Split<4, 4, 2>), 32, true) \
  X(6, 64, "8x8/tile/S32", (Split<8, 8>), 32, false)     \
  X(7, 64, "4x4x4/tile/S16", (Split<4, 4, 4>), 16, false) \
  X(8, 128, "16x8/tile/S32", (Split<16, 8>), 32, false)  \
  X(9, 128, "8x16/tile/S32", (Split<8, 16>), 32, false)  \
  X(10, 256, "16x16/tile/S16", (Split<16, 16>), 16, false) \
  X(11, 256, "8x8x4/tile/S8", (Split<8, 8, 4>), 8, false) \
  X(12, 512, "8x8x8/tile/S4", (Split<8, 8, 8>), 4, false) \
  X(13, 512, "16x16x2/tile/S8", (Split<16, 16, 2>), 8, false) \
  X(14, 1024, "16x16x4/tile/S4", (Split<16, 16, 4>), 4, false) \
  X(15, 1024, "8x8x16/tile/S4", (Split<8, 8, 16>), 4, false) \
  X(16, 2048, "16x16x8/tile/S2", (Split<16, 16, 8>), 2, false) \
  X(17, 2048, "8x16x16/tile/S2", (Split<8, 16, 16>), 2, false) \
  X(18, 4096, "16x16x16/tile/S1", (Split<16, 16, 16>), 1, false)

#define FFTOC_UNPAREN(...) __VA_ARGS__

#define FFTOC_CHECK_THREADS(ID, NN, NAME, SPLIT, S, STAGED)                              \
  static_assert((FFTOC_UNPAREN SPLIT::N / FFTOC_UNPAREN SPLIT::RM) * S == kBatchThreads, \
                "tile routine " #ID " must use 256 threads");                            \
  static_assert(FFTOC_UNPAREN SPLIT::N == NN, "tile routine " #ID " length");            \
  static_assert(S * NN <= 4096, "tile routine " #ID " LDS");
FFTOC_ROUTINES(FFTOC_CHECK_THREADS)
#undef FFTOC_CHECK_THREADS

struct RoutineInfo {
  int n;
  const char *name;
  int signals_per_tile;
};

#define FFTOC_INFO(ID, NN, NAME, SPLIT, S, STAGED) {NN, NAME, S},
constexpr RoutineInfo kRoutines[] = {FFTOC_ROUTINES(FFTOC_INFO)};
#undef FFTOC_INFO
constexpr int kNumRoutines = sizeof(kRoutines) / sizeof(kRoutines[0]);

// Each tile routine is a separate (not inlined) function so that its
// registers are allocated on their own; the kernel needs the maximum over the
// routines, not their union.
template <class K, class CT>
__device__ __noinline__ void tile_call(const CT *in, CT *out, int64_t batch, int64_t tile, CT *lds,
                                       const CT *__restrict__ tw, real_t<CT> scale, bool inv) {
  K::run_tile(in, out, batch, tile, lds, tw, scale, inv);
}

template <class CT>
__device__ __forceinline__ void run_routine(int id, const JobDev &J, int64_t tile, CT *lds,
                                            const CT *__restrict__ tw) {
  const CT *in = static_cast<const CT *>(J.in);
  CT *out = static_cast<CT *>(J.out);
  const real_t<CT> scale = (real_t<CT>)J.scale;
  const bool inv = J.inverse != 0;
  switch (id) {
#define FFTOC_CASE(ID, NN, NAME, SPLIT, S, STAGED)                                          \
  case ID:                                                                                \
    tile_call<StockhamKernel<FFTOC_UNPAREN SPLIT, S, 2, STAGED, CT>, CT>(in, out, J.batch, \
                                                                         tile, lds, tw,    \
                                                                         scale, inv);      \
    break;
    FFTOC_ROUTINES(FFTOC_CASE)
#undef FFTOC_CASE
    default:
      break;
  }
}

__global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_kernel(const JobDev *__restrict__ jobs,
                                                              int njobs, int64_t total_items,
                                                              unsigned long long *ctr,
                                                              const c2 *__restrict__ tw64,
                                                              const f2 *__restrict__ tw32) {
  __shared__ c2 lds[kLdsElems];
  __shared__ long long s_item;
  int jcur = 0;
  for (;;) {
    if (threadIdx.x == 0) s_item = (long long)atomicAdd(&ctr[0], 1ull);
    __syncthreads();
    const int64_t item = s_item;
    __syncthreads();  // s_item is read by all before thread 0 claims again
    if (item >= total_items) break;
    while (jcur + 1 < njobs && jobs[jcur + 1].item_begin <= item) ++jcur;
    const JobDev &J = jobs[jcur];
    const int64_t t0 = (item - J.item_begin) * J.tiles_per_item;
    const int64_t t1 = (t0 + J.tiles_per_item < J.ntiles) ? t0 + J.tiles_per_item : J.ntiles;
    for (int64_t t = t0; t < t1; ++t) {
      if (J.c64)
        run_routine<f2>(J.routine, J, t, reinterpret_cast<f2 *>(lds), tw32);
      else
        run_routine<c2>(J.routine, J, t, lds, tw64);
      __syncthreads();  // LDS is reused by the next tile
    }
  }
  // The last workgroup out resets the counters for the next launch / replay.
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned long long done = atomicAdd(&ctr[1], 1ull);
    if (done == (unsigned long long)gridDim.x - 1) {
      atomicExch(&ctr[0], 0ull);
      atomicExch(&ctr[1], 0ull);
    }
  }
}

}  // namespace

int batch_routine_count() { return kNumRoutines; }
int batch_routine_n(int id) { return (id >= 0 && id < kNumRoutines) ? kRoutines[id].n : 0; }
const char *batch_routine_name(int id) {
  return (id >= 0 && id < kNumRoutines) ? kRoutines[id].name : nullptr;
}
int batch_routine_signals(int id) {
  return (id >= 0 && id < kNumRoutines) ? kRoutines[id].signals_per_tile : 0;
}

int batch_resident_blocks() {
  static int cached = 0;
  if (cached == 0) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel, kBatchThreads, 0) !=
            hipSuccess || per_cu <= 0)
      per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0)
      cus = 256;
    cached = per_cu * cus;
  }
  return cached;
}

int launch_batch(const JobDev *d_jobs, int njobs, int64_t total_items, unsigned long long *d_ctr,
                 const void *tw64, const void *tw32, int grid, hipStream_t stream) {
  if (grid <= 0) return FFTOC_ERROR_INVALID_BATCH;
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBatchThreads), 0, stream, d_jobs, njobs,
                     total_items, d_ctr, static_cast<const c2 *>(tw64),
                     static_cast<const f2 *>(tw32));
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace fftoc
