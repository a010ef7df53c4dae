// Mixed-size batched FFT: one persistent kernel processes many jobs of
// different length, direction and data type (complex128 or complex64).
//
// Work items are chunks of tiles of one job; workgroups claim items in order
// from a device counter (dynamic load balance over uneven lengths) and run the
// job's tile routine.  Every routine is a tile on exactly 256 threads with at
// most 4096 complex values in LDS (Stockham tiles; for N = 8 ... 64 also the
// lane-shuffle tiles of lane_fft.hpp, which use none), so one workgroup shape
// and one LDS allocation serve every length.  The last workgroup to finish resets
// the counters, so the launch can be replayed from a hipGraph without any
// memset node.
#include "batch_routines.hpp"

namespace fftoc {
namespace {

__global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(2))) void batch_kernel(const JobDev *__restrict__ jobs,
                                                              int njobs, int64_t total_items,
                                                              unsigned long long *ctr,
                                                              const c2 *__restrict__ tw64,
                                                              const f2 *__restrict__ tw32) {
  __shared__ c2 lds[kLdsElems];
  __shared__ long long s_item;
  // Workgroup b starts with item b and claims the items after the grid's
  // first ones from the counter (claim_item), one item ahead of the one
  // running (claim_ahead): the start waits for no atomic.
  Cursor cur{(int64_t)blockIdx.x, 0, claim_ahead(blockIdx.x, total_items) ? 1 : 0};
  if (cur.ahead && threadIdx.x == 0) cur.claim = claim_item((global_ptr<unsigned long long>)ctr);
  int jcur = 0;
  for (;;) {
    // readfirstlane: the value is the same in every lane, and the compiler
    // then knows it, so the loop exit below is a uniform branch (a loop it
    // took to exit per lane would run this loop's barriers on split lanes)
    const int64_t item = (int64_t)uniform_u64((uint64_t)cur.item);
    if (item >= total_items) break;
    while (jcur + 1 < njobs && jobs[jcur + 1].item_begin <= item) ++jcur;
    const JobDev &J = jobs[jcur];
    const int64_t ib = (int64_t)uniform_u64((uint64_t)J.item_begin);
    const int64_t tpi = __builtin_amdgcn_readfirstlane(J.tiles_per_item);
    const int64_t ntiles = (int64_t)uniform_u64((uint64_t)J.ntiles);
    const int64_t ie = ib + (ntiles + tpi - 1) / tpi;
    const int routine = __builtin_amdgcn_readfirstlane(J.routine);
    const bool c64 = __builtin_amdgcn_readfirstlane(J.c64) != 0;
    cur.item = item;
    if (c64)
      cur = run_routine<f2>(routine, J, cur, ib, ie, total_items, ctr, &s_item,
                            reinterpret_cast<f2 *>(lds), tw32);
    else
      cur = run_routine<c2>(routine, J, cur, ib, ie, total_items, ctr, &s_item, lds, tw64);
    __syncthreads();  // LDS is reused by the next routine
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

// ---- device-resident serving loop -------------------------------------------

// Polling reads are relaxed (they bypass non-coherent cache lines but
// invalidate nothing); a poll that finds new work is followed by one acquire
// fence.  An acquire load in the loop itself would invalidate the L2 of the
// wave's XCD on every iteration.
__device__ __forceinline__ uint64_t ld_sys_relaxed(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_sys32_relaxed(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// Copies descriptor `c` from host memory in one round trip (all words loaded
// together); false if the slot does not hold request c.  The caller has
// acquired `published` > c, and the host rewrites a descriptor slot only after
// the request it held has completed, which needs this copy first: the words
// are stable, and the sequence check only confirms the slot.
__device__ bool read_host_desc(const ReqDesc *slot, uint64_t c, ReqDesc &d) {
  const uint64_t *src = reinterpret_cast<const uint64_t *>(slot);
  uint64_t *dst = reinterpret_cast<uint64_t *>(&d);
#pragma unroll
  for (int w = 0; w < (int)(sizeof(ReqDesc) / 8); ++w)
    dst[w] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  return d.seq == c;
}

// Device descriptor table entry of request c.  The dispatcher writes it before
// releasing the request's work items, and the caller has acquired a published
// item of request c, so the entry is complete; seq confirms the slot.
__device__ bool read_dtab(const ReqDesc *e, uint64_t c, JobDev &J) {
  uint64_t w[sizeof(ReqDesc) / 8];
  const uint64_t *src = reinterpret_cast<const uint64_t *>(e);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(ReqDesc) / 8); ++k)
    w[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  ReqDesc d;
  __builtin_memcpy(&d, w, sizeof(d));
  if (d.seq != c) return false;
  J.in = d.in;
  J.out = d.out;
  J.batch = d.batch;
  J.ntiles = d.ntiles;
  J.item_begin = d.nitems;  // holds nitems for served requests
  J.routine = d.routine;
  J.tiles_per_item = d.tiles_per_item;
  J.inverse = d.flags & 1;
  J.c64 = (d.flags >> 1) & 1;
  J.scale = d.scale;
  return true;
}

enum ServeState : int { kWork = 0, kIdle = 1, kExit = 2, kIdleNoTicket = 3 };

// Wave 0 of workgroup 0: moves newly published requests from the host ring
// into the device table, up to 64 per round (one per lane, so one PCIe round
// trip per round), and expands them into the work-item ring.  Entries follow
// the sequence lock seq = invalid, fields, seq = id, each step ordered by a
// release fence; the work items are published (wpub) after their
// descriptors.  Leaves when the host sets `stop` or after
// idle_timeout_ticks without requests, then raises dev_stop so the workers
// leave once idle.
__device__ void dispatcher(const ServerDev &sd) {
  const int lane = threadIdx.x;  // 0..63
  uint64_t dpub = __hip_atomic_load(sd.dev_pub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t wp = __hip_atomic_load(sd.wpub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t idle_since = wall_clock64();
  const uint64_t ns = (uint64_t)sd.nslots;
  for (;;) {
    uint64_t pub = ld_sys_relaxed(&sd.ring->published);
    pub = uniform_u64(pub);
    if (pub > dpub) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope: the host's descriptors
      const uint64_t cnt = (pub - dpub) < 64 ? (pub - dpub) : 64;
      const uint64_t c = dpub + (uint64_t)lane;
      ReqDesc d;
      const bool ok = ((uint64_t)lane < cnt) && read_host_desc(&sd.descs[c % ns], c, d);
      const unsigned long long mask = __ballot(ok);
      const int k = (~mask == 0ull) ? 64 : __builtin_ctzll(~mask);  // leading lanes ready
      if (k > 0) {
        ReqDesc *e = &sd.dtab[c % ns];
        uint64_t *dst = reinterpret_cast<uint64_t *>(e);
        const uint64_t *src = reinterpret_cast<const uint64_t *>(&d);
        // the entry (fields, then seq) and the work items become visible to
        // the workers together, through the release before wpub below
        if (lane < k) {
#pragma unroll
          for (int w = 1; w < (int)(sizeof(ReqDesc) / 8); ++w)
            __hip_atomic_store(dst + w, src[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(&e->seq, c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // work items of these k requests: exclusive prefix sum of nitems
        uint64_t ni = (lane < k) ? (uint64_t)d.nitems : 0;
        uint64_t incl = ni;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint64_t v = __shfl_up(incl, off);
          if (lane >= off) incl += v;
        }
        const uint64_t total = __shfl(incl, 63);
        const uint64_t first = wp + incl - ni;
        // the items of each request in turn, written by all lanes (a request
        // of thousands of items takes tens of store instructions, not thousands)
        for (int r = 0; r < k; ++r) {
          const uint64_t cr = dpub + (uint64_t)r;
          const uint64_t nr = __shfl((unsigned long long)ni, r);
          const uint64_t fr = __shfl((unsigned long long)first, r);
          for (uint64_t i = (uint64_t)lane; i < nr; i += 64)
            __hip_atomic_store(&sd.witems[(fr + i) & (kWorkRing - 1)],
                               (unsigned long long)((cr << kItemBits) | i), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        release_agent();
        wp += total;
        dpub += (uint64_t)k;
        if (lane == 0) {
          __hip_atomic_store(sd.wpub, (unsigned long long)wp, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(sd.dev_pub, (unsigned long long)dpub, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        idle_since = wall_clock64();
        continue;
      }
    }
    // a flag only: nothing is read on its strength
    const unsigned stop = __builtin_amdgcn_readfirstlane(ld_sys32_relaxed(&sd.ring->stop));
    const uint64_t now = wall_clock64();
    const bool timeout = __builtin_amdgcn_readfirstlane((unsigned)(now - idle_since > sd.idle_timeout_ticks));
    if (stop || timeout) break;
    __builtin_amdgcn_s_sleep(2);
  }
  if (lane == 0) __hip_atomic_store(sd.dev_stop, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Identifies the CU a wave runs on from the hardware ID registers: XCC_ID
// (hwreg 20) bits 2:0, HW_ID (hwreg 4) SE bits 15:13, SH bit 12, CU bits 11:8.
// s_getreg's operand is id | offset << 6 | (size - 1) << 11.
__device__ inline unsigned cu_key() {
  const unsigned hw = __builtin_amdgcn_s_getreg(4 | (31 << 11));
  const unsigned xcc = __builtin_amdgcn_s_getreg(20 | (31 << 11)) & 7u;
  return (xcc << 8) | (((hw >> 13) & 7u) << 5) | (((hw >> 12) & 1u) << 4) | ((hw >> 8) & 15u);
}

// The first worker seated on a CU is its primary and takes work tickets in
// order, holding one while it waits.  Any other worker on the CU is secondary:
// it takes a ticket only while published items wait unclaimed beyond the
// tickets idle primaries hold, so at low load items run alone on their CUs and
// few workers poll the queue, and under a backlog every slot works.  A worker
// claims its next ticket only when it has finished an item: a ticket claimed
// ahead would hold an item back from an idle worker until the current item is
// done.  (Claiming single published items by compare-and-swap instead starves
// the primaries' atomics.)
__device__ inline bool claim_ticket(const ServerDev &sd, bool secondary, uint64_t &ticket) {
  if (secondary) {
    const uint64_t p = __hip_atomic_load(sd.wpub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t t = __hip_atomic_load(sd.wnext, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((int64_t)(p - t) <= (int64_t)sd.claim_backlog) return false;
  }
  ticket = atomicAdd(sd.wnext, 1ull);
  return true;
}

// Persistent serving kernel.  Workgroup 0 is the dispatcher; each other
// workgroup takes work tickets in order (one atomic), waits until its ticket
// is published, reads the work item (request, item), runs the request's tile
// routine and counts the finished item; the workgroup that finishes a
// request's last item publishes its completion record in host memory.
__global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(2)))
void serve_kernel(ServerDev sd) {
  if (blockIdx.x == 0) {
    if (threadIdx.x < 64) {
      dispatcher(sd);
      if (threadIdx.x == 0)
        __hip_atomic_store(&sd.live[0], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    return;
  }
  __shared__ c2 lds[kLdsElems];
  __shared__ JobDev s_job;
  __shared__ long long s_item;
  __shared__ unsigned long long s_req;
  __shared__ int s_state;
  // Scheduling state lives in thread 0; its decision reaches the workgroup
  // through s_state, so every wave takes the same branch.
  __shared__ unsigned s_seat;
  if (threadIdx.x == 0) s_seat = atomicAdd(&sd.cu_seats[cu_key() & (kCuKeys - 1)], 1u);
  __syncthreads();
  const bool secondary = s_seat != 0;
  bool has_ticket = false;
  uint64_t ticket = 0;
  bool have = false;       // s_job holds request cached_c
  uint64_t cached_c = 0;
  unsigned pending = 0;    // finished items of s_req not yet counted
  uint64_t idle_since = wall_clock64();
  for (;;) {
    if (threadIdx.x == 0) {
      int state = kIdle;
      if (!has_ticket) has_ticket = claim_ticket(sd, secondary, ticket);
      if (has_ticket &&
          ticket < __hip_atomic_load(sd.wpub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        acquire_agent();  // the item and its request's table entry
        const uint64_t e = __hip_atomic_load(&sd.witems[ticket & (kWorkRing - 1)], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t c = e >> kItemBits;
        bool ready = have && cached_c == c;
        if (!ready) {
          JobDev J;
          ready = read_dtab(&sd.dtab[c % (uint64_t)sd.nslots], c, J);
          if (ready) {
            s_job = J;
            have = true;
            cached_c = c;
          }
        }
        if (ready) {
          s_item = (long long)(e & ((1ull << kItemBits) - 1));
          s_req = c;
          has_ticket = false;
          state = kWork;
        }
      }
      if (state == kIdle) {
        // A worker holding a ticket leaves only with the dispatcher (dev_stop):
        // the item of that ticket may be published at any later time while the
        // dispatcher runs, and a ticket given up here would never be run (the
        // host relaunches only once the dispatcher has left too).  The own
        // timer is a backstop for workers without a ticket.
        if (__hip_atomic_load(sd.dev_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            (!has_ticket && wall_clock64() - idle_since > sd.idle_timeout_ticks + 100000000ull))
          state = kExit;
        else if (!has_ticket)
          state = kIdleNoTicket;
      } else {
        idle_since = wall_clock64();
      }
      s_state = state;
    }
    __syncthreads();
    const int state = __builtin_amdgcn_readfirstlane(s_state);  // uniform branches below
    if (state == kExit) break;
    if (state != kWork) {
      if (state == kIdleNoTicket)  // a secondary looks for unclaimed items less often
        __builtin_amdgcn_s_sleep(64);
      else
        __builtin_amdgcn_s_sleep(4);
      __syncthreads();  // s_state has been read before thread 0 rewrites it
      continue;
    }
    const JobDev J = s_job;
    // the item and the routine, read from LDS, made uniform: the tile
    // stream holds barriers, so its branches must be the same in every lane
    const int64_t item = (int64_t)uniform_u64((uint64_t)s_item);
    const int routine = __builtin_amdgcn_readfirstlane(J.routine);
    const bool c64 = __builtin_amdgcn_readfirstlane(J.c64) != 0;
    // the one item (tiles item * tiles_per_item on, request-relative), pipelined
    const Cursor one{item, 0, 0};
    if (c64)
      run_routine<f2>(routine, J, one, 0, item + 1, item + 1, nullptr, &s_item,
                      reinterpret_cast<f2 *>(lds), static_cast<const f2 *>(sd.tw32));
    else
      run_routine<c2>(routine, J, one, 0, item + 1, item + 1, nullptr, &s_item, lds,
                      static_cast<const c2 *>(sd.tw64));
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t r = s_req;
      ++pending;
      // A request with more items than there are workers: the next ticket is
      // claimed now, and if its item is published and of the same request,
      // this item is counted with that one (one release and one count for
      // both: the release writes back this XCD's L2); the worker runs that
      // item next, so the count is only delayed, never left.  (Smaller
      // requests are counted at once: the claim's round trip would delay
      // their completion.)
      bool defer = false;
      if (J.item_begin > (int64_t)gridDim.x - 1 && !has_ticket)
        has_ticket = claim_ticket(sd, secondary, ticket);
      if (J.item_begin > (int64_t)gridDim.x - 1 && has_ticket &&
          ticket < __hip_atomic_load(sd.wpub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        acquire_agent();  // the item word
        const uint64_t e = __hip_atomic_load(&sd.witems[ticket & (kWorkRing - 1)], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        defer = (e >> kItemBits) == r;
      }
      if (!defer) {
        // the counted items' outputs reach memory before they are counted
        // (release only: the worker that counts the last item takes the
        // acquire side from the system fence below)
        release_agent();
        unsigned *done = &sd.done[r % (uint64_t)sd.nslots];
        const unsigned fin = atomicAdd(done, pending);
        if ((int64_t)fin + pending == J.item_begin) {
          // every item is counted: the slot's counter starts from zero for the
          // next request in it, which the host submits only after reading the
          // completion record written below
          __hip_atomic_store(done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          Completion *cp = &sd.comps[r % (uint64_t)sd.nslots];
          __hip_atomic_store(&cp->gpu_time, wall_clock64(), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_SYSTEM);
          __threadfence_system();
          // ordered after the outputs, the counter reset and gpu_time by the
          // system fence above
          __hip_atomic_store(&cp->seq, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        pending = 0;
      }
    }
    __syncthreads();
  }
  // Lets the host see, by reading its own memory, that this workgroup left.
  if (threadIdx.x == 0)
    __hip_atomic_store(&sd.live[blockIdx.x], 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Before each (re)launch: tickets handed out beyond the published items
// belonged to workgroups that left while idle; start over from wpub.
__global__ void serve_reset_kernel(ServerDev sd) {
  for (int i = threadIdx.x; i < kCuKeys; i += blockDim.x) sd.cu_seats[i] = 0u;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const unsigned long long p =
        __hip_atomic_load(sd.wpub, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sd.wnext, p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(sd.dev_stop, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace

int server_grid(int max_per_cu) {
  // every resident slot (two workgroups per CU on gfx950: LDS and VGPRs),
  // the dispatcher workgroup included
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, serve_kernel, kBatchThreads, 0) !=
          hipSuccess || per_cu <= 0)
    per_cu = 1;
  if (max_per_cu > 0 && per_cu > max_per_cu) per_cu = max_per_cu;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  return per_cu * cus;
}

int launch_server(const ServerDev &sd, int grid, hipStream_t stream) {
  hipLaunchKernelGGL(serve_reset_kernel, dim3(1), dim3(64), 0, stream, sd);
  hipLaunchKernelGGL(serve_kernel, dim3(grid), dim3(kBatchThreads), 0, stream, sd);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

int batch_routine_count() { return kNumRoutines; }
// Untuned choices per length (index log2 n - 3) and data type, from single
// 1 GiB jobs on MI355X.  kPlan*: plans, whose jobs sharing one routine run the
// kernel specialized to it (lane-shuffle tiles up to N = 32, then mostly the
// plain tile loops, which gain most from a compile-time direction).  All are
// kBatchThreads routines, so a job's untuned result is the same alone and in
// a mix (the mixed-size kernel); the wide routines are there for set_routine
// and the autotuner.  kServe*: the serving kernel, one item per routine call
// with the direction chosen at run time, where the pipelined streams run best.
constexpr int kPlanC128[10] = {20, 21, 22, 7, 19, 10, 13, 15, 16, 18};
constexpr int kPlanC64[10] = {20, 21, 22, 25, 26, 10, 13, 15, 33, 35};
constexpr int kServeC128[10] = {20, 21, 22, 25, 26, 10, 30, 32, 39, 37};
constexpr int kServeC64[10] = {20, 21, 22, 25, 26, 29, 30, 31, 33, 35};
int length_index(int n) {
  for (int k = 0; k < 10; ++k)
    if (n == (8 << k)) return k;
  return -1;
}

int batch_routine_default(int n, bool c64) {
  const int k = length_index(n);
  return k < 0 ? -1 : c64 ? kPlanC64[k] : kPlanC128[k];
}
int batch_routine_server(int n, bool c64) {
  const int k = length_index(n);
  return k < 0 ? -1 : c64 ? kServeC64[k] : kServeC128[k];
}
int batch_routine_mixed(int id, bool c64) {
  if (id < 0 || id >= kNumRoutines || kRoutines[id].threads == kBatchThreads) return id;
  return batch_routine_default(kRoutines[id].n, c64);
}
int batch_routine_threads(int id) {
  return (id >= 0 && id < kNumRoutines) ? kRoutines[id].threads : 0;
}
int batch_routine_n(int id) { return (id >= 0 && id < kNumRoutines) ? kRoutines[id].n : 0; }
const char *batch_routine_name(int id) {
  return (id >= 0 && id < kNumRoutines) ? kRoutines[id].name : nullptr;
}
int batch_routine_signals(int id) {
  return (id >= 0 && id < kNumRoutines) ? kRoutines[id].signals_per_tile : 0;
}

int batch_resident_blocks(int per_cu) {
  return (int)resident_workgroups<batch_kernel, kBatchThreads>(per_cu);
}
int batch_occupancy() { return occupancy<batch_kernel, kBatchThreads>().per_cu; }

// ---- allocation-free batches --------------------------------------------------
// Scratch of fftoc_batch_fft_noalloc: the job descriptors (JobDev[32]) and the
// two work counters, 4 KiB per device as a module global.  Host code never
// passes `scratch` itself to a kernel (on the host it names a stand-in, not
// device memory): it uses the address hipGetSymbolAddress returned for the
// current device at set-up.  Every call on a device shares its array, so calls
// that overlap on one device race.
constexpr size_t kNoAllocScratchBytes = 4096;
constexpr size_t kNoAllocCtrOffset = kNoAllocMaxJobs * sizeof(JobDev);
static_assert(kNoAllocCtrOffset % 8 == 0 &&
                  kNoAllocCtrOffset + 2 * sizeof(unsigned long long) <= kNoAllocScratchBytes,
              "scratch layout");
static_assert(sizeof(JobDev) % 8 == 0, "descriptors are copied as 8-byte words");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

int batch_noalloc_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return FFTOC_ERROR_DEVICE;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
    return FFTOC_ERROR_DEVICE;
  g_scratch_ptr[dev] = ptr;
  return FFTOC_SUCCESS;
}

namespace {
// Copies the descriptors from the kernel argument into the scratch array and
// clears the work counters (a kernel, not a memset: the launch sequence stays
// capturable into a graph).
__global__ __launch_bounds__(64) void batch_noalloc_setup_kernel(NoAllocJobs jobs, int njobs,
                                                                 unsigned char *sp) {
  const uint64_t *src = reinterpret_cast<const uint64_t *>(&jobs.j[0]);
  uint64_t *dst = reinterpret_cast<uint64_t *>(sp);
  const int words = njobs * (int)(sizeof(JobDev) / 8);
  for (int w = threadIdx.x; w < words; w += 64) dst[w] = src[w];
  if (threadIdx.x < 2)
    reinterpret_cast<unsigned long long *>(sp + kNoAllocCtrOffset)[threadIdx.x] = 0ull;
}
}  // namespace

int launch_batch_noalloc(const NoAllocJobs &jobs, int njobs, int64_t total_items,
                         const void *tw64, const void *tw32, int grid, hipStream_t stream) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return FFTOC_ERROR_DEVICE;
  unsigned char *sp = static_cast<unsigned char *>(g_scratch_ptr[dev]);
  if (!sp) return FFTOC_ERROR_DEVICE;  // the device was not set up
  if (njobs < 1 || njobs > kNoAllocMaxJobs || grid <= 0) return FFTOC_ERROR_INVALID_BATCH;
  hipLaunchKernelGGL(batch_noalloc_setup_kernel, dim3(1), dim3(64), 0, stream, jobs, njobs, sp);
  hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBatchThreads), 0, stream,
                     reinterpret_cast<const JobDev *>(sp), njobs, total_items,
                     reinterpret_cast<unsigned long long *>(sp + kNoAllocCtrOffset),
                     static_cast<const c2 *>(tw64), static_cast<const f2 *>(tw32));
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
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
