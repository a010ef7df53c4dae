// Batched FFT kernels specialized to one tile routine and data type.  A plan
// whose jobs all run the same routine on the same data type launches one of
// these instead of the mixed-size kernel: the same work items, counters and
// dynamic claiming, with the routine's stream inlined into the kernel.  There
// is then no routine call, so none of the callee-saved registers a called
// routine stores to scratch and reloads (up to 112 per lane per call, some
// 50 us per launch of the mixed kernel at N = 4096), and the kernel's
// registers and LDS are those of its routine alone.
#include "batch_routines.hpp"

namespace fftoc {
namespace {

// LDS elements (of the data type) a routine's tile uses.
constexpr int routine_lds(TForm f, int n, int s, bool c128) {
  if (f == TForm::Lanes) return 1;
  if ((f == TForm::TensorK || f == TForm::TensorH) && c128) return 4096;
  return lds_elems(s * n);
}

template <class SP, int S, TForm F, class CT, int THREADS = kRoutineThreads<SP, S, F, CT>>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2))) void
batch_one_kernel(const JobDev *__restrict__ jobs, int njobs, int64_t total_items,
                 unsigned long long *ctr, const CT *__restrict__ tw) {
  using K = typename RoutineKernel<SP, S, F, CT>::type;
  __shared__ CT lds[routine_lds(F, SP::N, S, sizeof(CT) == 16)];
  __shared__ long long s_item;
  // the claims as in the mixed-size kernel: item b first, then from the
  // counter, ahead of the running item outside the launch's tail
  Cursor cur{(int64_t)blockIdx.x, 0, claim_ahead(blockIdx.x, total_items) ? 1 : 0};
  if (cur.ahead && threadIdx.x == 0) cur.claim = claim_item((global_ptr<unsigned long long>)ctr);
  int jcur = 0;
  for (;;) {
    const int64_t item = (int64_t)uniform_u64((uint64_t)cur.item);
    if (item >= total_items) break;
    while (jcur + 1 < njobs && jobs[jcur + 1].item_begin <= item) ++jcur;
    const JobDev &J = jobs[jcur];
    const int64_t ib = (int64_t)uniform_u64((uint64_t)J.item_begin);
    const int64_t tpi = __builtin_amdgcn_readfirstlane(J.tiles_per_item);
    const int64_t ntiles = (int64_t)uniform_u64((uint64_t)J.ntiles);
    const int64_t ie = ib + (ntiles + tpi - 1) / tpi;
    cur.item = item;
    cur = stream_routine<K, F, CT>(
        (global_ptr<const CT>)J.in, (global_ptr<CT>)J.out, J.batch, cur, ib, ie, total_items,
        ntiles, (int)tpi, (global_ptr<unsigned long long>)ctr, (lds_ptr<long long>)&s_item,
        (lds_ptr<CT>)lds, (global_ptr<const CT>)tw, (real_t<CT>)J.scale, J.inverse != 0);
    __syncthreads();  // LDS is reused by the next job's tiles
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

#define FFTOC_ONE(ID, NN, NAME, SPLIT, S, FORM)                                   \
  {reinterpret_cast<const void *>(&batch_one_kernel<FFTOC_UNPAREN SPLIT, S, FORM, c2>), \
   reinterpret_cast<const void *>(&batch_one_kernel<FFTOC_UNPAREN SPLIT, S, FORM, f2>)},
const void *const kOneKernels[][2] = {FFTOC_ROUTINES(FFTOC_ONE)};
#undef FFTOC_ONE
static_assert(sizeof(kOneKernels) / sizeof(kOneKernels[0]) == kNumRoutines, "one kernel per routine");

}  // namespace

constexpr int kOnePerCu = 2;

int64_t batch_one_resident(int routine, bool c64) {
  if (routine < 0 || routine >= kNumRoutines) return 0;
  static std::atomic<int64_t> cache[kNumRoutines][2][kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = -1;
  const int t = c64 ? 1 : 0;
  if (dev >= 0 && dev < kMaxDevices) {
    const int64_t v = cache[routine][t][dev].load(std::memory_order_relaxed);
    if (v > 0) return v;
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kOneKernels[routine][t],
                                                   kRoutines[routine].threads, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  // at most two workgroups per CU, as the mixed-size kernel: more (the light
  // routines allow up to eight) measured slower on every length, the claims
  // and the launch's tail spread over more workgroups
  if (per_cu > kOnePerCu) per_cu = kOnePerCu;
  if (dev < 0 ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const int64_t n = (int64_t)per_cu * cus;
  if (dev >= 0 && dev < kMaxDevices) cache[routine][t][dev].store(n, std::memory_order_relaxed);
  return n;
}

int launch_batch_one(int routine, bool c64, const JobDev *d_jobs, int njobs, int64_t total_items,
                     unsigned long long *d_ctr, const void *tw, int grid, hipStream_t stream) {
  if (routine < 0 || routine >= kNumRoutines || grid <= 0) return FFTOC_ERROR_INVALID_BATCH;
  void *args[] = {&d_jobs, &njobs, &total_items, &d_ctr, &tw};
  if (hipLaunchKernel(kOneKernels[routine][c64 ? 1 : 0], dim3(grid),
                      dim3(kRoutines[routine].threads), args, 0, stream) != hipSuccess)
    return FFTOC_ERROR_LAUNCH;
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace fftoc
