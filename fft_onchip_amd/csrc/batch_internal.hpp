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

// ---- serving loop -----------------------------------------------------------
// Request descriptor in fine-grained (coherent) pinned host memory, written by
// the host under a sequence lock: seq = kSeqInvalid, fields, seq = request id,
// then `published` advanced past it; the device reads it only after acquiring
// that count.
constexpr uint64_t kSeqInvalid = ~0ull;
struct alignas(64) ReqDesc {
  uint64_t seq;
  const void *in;
  void *out;
  int64_t batch;
  int64_t ntiles;
  int64_t nitems;
  int32_t n;
  int32_t routine;
  int32_t tiles_per_item;
  int32_t flags;  // bit 0 inverse, bit 1 complex64
  double scale;
};
struct alignas(16) Completion {
  uint64_t seq;       // request id once complete
  uint64_t gpu_time;  // wall_clock64() at completion (100 MHz)
};
struct HostRing {  // pinned, coherent, mapped into the device address space
  uint64_t published;  // requests 0 .. published-1 have descriptors
  uint32_t stop;       // host asks the loop to exit
  uint32_t pad;
};
struct ServerDev {
  const HostRing *ring;    // device view of the host ring header
  uint32_t *live;          // [grid] in host memory: 1 while that workgroup runs
  const ReqDesc *descs;    // [nslots] host memory, read by the dispatcher only
  ReqDesc *dtab;           // [nslots] device copy of the descriptors
  unsigned long long *dev_pub;   // requests copied into dtab (dispatcher state)
  unsigned long long *witems;    // [kWorkRing] work items: request << 24 | item
  unsigned long long *wpub;      // work items published
  unsigned long long *wnext;     // next work ticket
  unsigned *dev_stop;      // the dispatcher has left: workers leave when idle
  Completion *comps;       // [nslots]
  unsigned *done;          // [nslots] items finished of the slot's request; the
                           // worker finishing the last one resets it
  int nslots;
  unsigned *cu_seats;      // [kCuKeys] workers seated per CU in this launch
  int claim_backlog;       // unassigned items above which secondary workers join
  uint64_t idle_timeout_ticks;  // wall_clock64 ticks without work before exit
  const void *tw64;
  const void *tw32;
};
constexpr uint64_t kWorkRing = 1ull << 22;  // work-item ring entries
constexpr int kItemBits = 24;                // items per request < 2^24
constexpr uint64_t kMaxRequestId = 1ull << (64 - kItemBits);  // ids fit a work item
constexpr int kCuKeys = 2048;                // (XCC, SE, SH, CU) of a workgroup
int launch_server(const ServerDev &sd, int grid, hipStream_t stream);
int server_grid(int max_per_cu);

// ---- allocation-free batches (fftoc_batch_fft_noalloc) ----------------------
// The descriptors travel as a kernel argument into a fixed per-device scratch
// array of the library, which also holds the work counters.
constexpr int kNoAllocMaxJobs = 32;
struct NoAllocJobs {
  JobDev j[kNoAllocMaxJobs];
};
int launch_batch_noalloc(const NoAllocJobs &jobs, int njobs, int64_t total_items,
                         const void *tw64, const void *tw32, int grid, hipStream_t stream);

int batch_routine_count();
int batch_routine_default(int n, bool c64);  // untuned routine of a length / data type
int batch_routine_server(int n, bool c64);   // the serving kernel's routine
// the routine of a mixed-size kernel (or the server) for routine id: id
// itself, or for a wide routine (512 threads, one-routine kernels only) the
// kBatchThreads routine of its length
int batch_routine_mixed(int id, bool c64);
int batch_routine_threads(int id);
int batch_routine_n(int id);
const char *batch_routine_name(int id);
int batch_routine_signals(int id);
// workgroups of the mixed-size kernel resident on the device, at most
// per_cu per CU when per_cu > 0; batch_occupancy: its workgroups per CU
int batch_resident_blocks(int per_cu = 0);
int batch_occupancy();
int launch_batch(const JobDev *d_jobs, int njobs, int64_t total_items, unsigned long long *d_ctr,
                 const void *tw64, const void *tw32, int grid, hipStream_t stream);
// batch_one*.hip: the kernel specialized to one routine and data type (every
// job of the launch runs that routine on that data type) and direction (dir
// 0: every job forward, 1: every job inverse, 2: each job its own)
// batch_one_resident: the grid, per_cu workgroups per CU (0: two) or the
// occupancy if lower; batch_one_occupancy: workgroups per CU and the CUs
int64_t batch_one_resident(int routine, bool c64, int dir, int per_cu = 0);
int batch_one_occupancy(int routine, bool c64, int dir, int *cus);
int launch_batch_one(int routine, bool c64, int dir, const JobDev *d_jobs, int njobs,
                     int64_t total_items, unsigned long long *d_ctr, const void *tw, int grid,
                     hipStream_t stream);

}  // namespace fftoc
