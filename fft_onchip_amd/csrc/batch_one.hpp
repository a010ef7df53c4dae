// Batched FFT kernels specialized to one tile routine, data type and
// direction (batch_one*.hip).  A plan whose jobs all run the same routine on
// the same data type launches one of these instead of the mixed-size kernel:
// the same work items, counters and dynamic claiming, with the routine's
// stream inlined into the kernel.  There is then no routine call, so none of
// the callee-saved registers a called routine stores to scratch and reloads
// (up to 112 per lane per call, some 50 us per launch of the mixed kernel at
// N = 4096); the kernel's registers and LDS are those of its routine alone;
// and when the jobs share their direction too it is fixed at compile time
// (DIR 0 forward, 1 inverse; 2: each job's own, chosen at run time).
#pragma once
#include "batch_routines.hpp"

namespace fftoc {
namespace {

// LDS elements (of the data type) a routine's tile uses.
constexpr int routine_lds(TForm f, int n, int s, bool c128) {
  if (f == TForm::Lanes) return 1;
  if ((f == TForm::TensorK || f == TForm::TensorH || f == TForm::TensorK1) && c128) return 4096;
  return lds_elems(s * n);
}

template <class SP, int S, TForm F, class CT, int DIR, int THREADS = kRoutineThreads<SP, S, F, CT>>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(2))) void
batch_one_kernel(const JobDev *__restrict__ jobs, int njobs, int64_t total_items,
                 unsigned long long *ctr, const CT *__restrict__ tw) {
  using K = typename RoutineKernel<SP, S, F, CT, DIR>::type;
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


// The kernels of every routine for one direction, [routine][complex64].
#define FFTOC_ONE(ID, NN, NAME, SPLIT, S, FORM)                                        \
  {reinterpret_cast<const void *>(&batch_one_kernel<FFTOC_UNPAREN SPLIT, S, FORM, c2, FFTOC_ONE_DIR>), \
   reinterpret_cast<const void *>(&batch_one_kernel<FFTOC_UNPAREN SPLIT, S, FORM, f2, FFTOC_ONE_DIR>)},

}  // namespace
}  // namespace fftoc
