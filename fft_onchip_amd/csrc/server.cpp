// Continuous-batching FFT server: a device-resident request loop
// (serve_kernel in batch_kernels.hip) fed through a request ring in coherent
// pinned host memory.  The host admits requests while a ring slot and enough
// of the HBM budget (the output pool) are free; requests that arrive while
// others run are picked up by idle workgroups at once, so their work overlaps
// on the device.  Completions are posted by the device into the ring.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "batch_internal.hpp"

namespace {

// Pool allocations are whole 256-byte granules.
constexpr size_t granule(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// First-fit allocator over [0, size) with coalescing; 256-byte granules (the
// size is a multiple of the granule, so anything that fits the capacity can
// be placed once the pool is empty).
class PoolAllocator {
 public:
  void reset(size_t size) {
    free_.clear();
    free_[0] = size;
    size_ = size;
  }
  bool alloc(size_t bytes, size_t *off) {
    bytes = granule(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t start = it->first + bytes;
        free_.erase(it);
        if (rest) free_[start] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = granule(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t capacity() const { return size_; }
  size_t used() const { return used_; }

 private:
  std::map<size_t, size_t> free_;
  size_t size_ = 0, used_ = 0;
};

struct Slot {
  bool busy = false;
  uint64_t seq = 0;
  size_t out_off = 0, out_bytes = 0;
  int64_t items = 0;
};

// The device loop exits after this long without requests (so it never
// outlives a vanished host, nor holds its hardware queue for long); submit /
// poll relaunch it when it is needed again.
constexpr double kTicksPerSecond = 1e8;  // wall_clock64 runs at 100 MHz
constexpr uint64_t kIdleTimeoutTicks = (uint64_t)(0.020 * kTicksPerSecond);  // 20 ms

// Output bytes in flight above which the device counts as busy (about four
// 256 KiB items per worker).
constexpr size_t kDeepBacklogBytes = 256ull << 20;
// Under such a backlog a work item holds about backlog / (4 x workgroups) of
// input, within these bounds: enough items to keep every workgroup busy,
// few enough that per-item claims and fences stay cheap (saturated C5 mix:
// 2.6 TB/s with 256 KiB items, 3.5 TB/s with 4 MiB items).
constexpr int64_t kMinDeepItemBytes = 256 << 10;
constexpr int64_t kMaxDeepItemBytes = 4 << 20;

}  // namespace

struct fftoc_server_t {
  int device = 0;
  std::mutex m;
  PoolAllocator pool;
  char *d_pool = nullptr;
  int nslots = 0;
  std::vector<Slot> slots;
  std::map<uint64_t, int> slot_of;  // request id -> slot
  fftoc::HostRing *ring = nullptr;  // pinned host memory
  fftoc::ReqDesc *descs = nullptr;
  fftoc::Completion *comps = nullptr;
  uint32_t *live = nullptr;         // pinned: per-workgroup "running" flags
  double last_check = 0;            // time of the last ensure_running from poll
  int grid = 0;
  int spread_workers = 1;           // a request at low load is cut into this many items
  fftoc::ServerDev sd{};
  uint64_t next_seq = 0;
  int64_t inflight_items = 0;       // bounded by the device work-item ring
  hipStream_t stream = nullptr;
  bool running = false;
  int routine[13][2];
  const void *tw64 = nullptr, *tw32 = nullptr;
  bool holds_tables = false;
};

namespace fftoc {
namespace {

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int lg(int n) { return 31 - __builtin_clz((unsigned)n); }

void free_server(fftoc_server_t *s) {
  if (s->d_pool) (void)hipFree(s->d_pool);
  if (s->sd.done) (void)hipFree(s->sd.done);
  if (s->sd.dtab) (void)hipFree(s->sd.dtab);
  if (s->sd.dev_pub) (void)hipFree(s->sd.dev_pub);
  if (s->sd.dev_stop) (void)hipFree(s->sd.dev_stop);
  if (s->sd.witems) (void)hipFree(s->sd.witems);
  if (s->sd.wpub) (void)hipFree(s->sd.wpub);
  if (s->sd.wnext) (void)hipFree(s->sd.wnext);
  if (s->sd.cu_seats) (void)hipFree(s->sd.cu_seats);
  if (s->ring) (void)hipHostFree(s->ring);
  if (s->descs) (void)hipHostFree(s->descs);
  if (s->comps) (void)hipHostFree(s->comps);
  if (s->live) (void)hipHostFree(s->live);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->holds_tables) drop_tables();
  delete s;
}

template <class T>
int host_alloc(T **h, const T **d_const, T **d, size_t count) {
  void *p = nullptr;
  if (hipHostMalloc(&p, count * sizeof(T), hipHostMallocCoherent | hipHostMallocMapped) !=
      hipSuccess)
    return FFTOC_ERROR_ALLOC;
  std::memset(p, 0, count * sizeof(T));
  void *dp = nullptr;
  if (hipHostGetDevicePointer(&dp, p, 0) != hipSuccess) {
    (void)hipHostFree(p);
    return FFTOC_ERROR_DEVICE;
  }
  *h = static_cast<T *>(p);
  if (d_const) *d_const = static_cast<const T *>(dp);
  if (d) *d = static_cast<T *>(dp);
  return FFTOC_SUCCESS;
}

// Launches the device loop.  The live flags are raised before the launch and
// each workgroup lowers its own flag as it exits, so "all flags clear" means
// the loop is gone (a relaunch on the same stream starts after it ends).
int start_loop(fftoc_server_t *s) {
  for (int i = 0; i < s->grid; ++i) __atomic_store_n(&s->live[i], 1u, __ATOMIC_RELAXED);
  std::atomic_thread_fence(std::memory_order_seq_cst);
  // ordered on the loop's stream after the previous launch; launch_server
  // resets the ticket counter and dev_stop first
  return launch_server(s->sd, s->grid, s->stream);
}

// Relaunches the device loop if it has exited after its idle timeout.  Called
// with the server mutex held.  Workgroup 0's flag is read first; the full
// scan runs only once it is clear (the loop is leaving or gone).
int ensure_running(fftoc_server_t *s) {
  if (__atomic_load_n(&s->live[0], __ATOMIC_ACQUIRE)) return FFTOC_SUCCESS;
  for (int i = 1; i < s->grid; ++i)
    if (__atomic_load_n(&s->live[i], __ATOMIC_ACQUIRE)) return FFTOC_SUCCESS;
  return start_loop(s);
}

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace fftoc

extern "C" {

int fftoc_server_create(size_t hbm_budget_bytes, int max_inflight, fftoc_server *out) {
  using namespace fftoc;
  if (!out) return FFTOC_ERROR_NULL_POINTER;
  *out = nullptr;
  if (hbm_budget_bytes < 4096 || max_inflight <= 0 || max_inflight > (1 << 16))
    return FFTOC_ERROR_INVALID_BATCH;
  const void *tw64 = nullptr, *tw32 = nullptr;
  int rc = device_tables(&tw64, &tw32);
  if (rc != FFTOC_SUCCESS) return rc;
  fftoc_server_t *s = new (std::nothrow) fftoc_server_t;
  if (!s) return FFTOC_ERROR_ALLOC;
  (void)hipGetDevice(&s->device);
  s->tw64 = tw64;
  s->tw32 = tw32;
  for (int k = 0; k < 13; ++k) {
    s->routine[k][0] = batch_routine_server(1 << k, false);
    s->routine[k][1] = batch_routine_server(1 << k, true);
  }
  s->nslots = max_inflight;
  s->slots.assign(max_inflight, Slot{});
  s->sd.nslots = max_inflight;
  s->sd.idle_timeout_ticks = kIdleTimeoutTicks;
  s->sd.tw64 = tw64;
  s->sd.tw32 = tw32;
  rc = FFTOC_SUCCESS;
  if (hipMalloc(&s->d_pool, hbm_budget_bytes) != hipSuccess ||
      hipMalloc(&s->sd.done, (size_t)max_inflight * sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&s->sd.dtab, (size_t)max_inflight * sizeof(ReqDesc)) != hipSuccess ||
      hipMalloc(&s->sd.dev_pub, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&s->sd.dev_stop, sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&s->sd.witems, kWorkRing * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&s->sd.wpub, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&s->sd.wnext, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&s->sd.cu_seats, kCuKeys * sizeof(unsigned)) != hipSuccess)
    rc = FFTOC_ERROR_ALLOC;
  if (rc == FFTOC_SUCCESS &&
      (hipMemset(s->sd.done, 0, (size_t)max_inflight * sizeof(unsigned)) != hipSuccess ||
       hipMemset(s->sd.dev_pub, 0, sizeof(unsigned long long)) != hipSuccess ||
       hipMemset(s->sd.wpub, 0, sizeof(unsigned long long)) != hipSuccess ||
       hipMemset(s->sd.dtab, 0xff, (size_t)max_inflight * sizeof(ReqDesc)) != hipSuccess))
    rc = FFTOC_ERROR_DEVICE;
  HostRing *ring_d = nullptr;
  {
    // FFTOC_SERVER_WORKERS_PER_CU caps the workers per CU (default: as many as fit)
    const char *e = getenv("FFTOC_SERVER_WORKERS_PER_CU");
    s->grid = server_grid(e ? atoi(e) : 0);
  }
  {
    // one item per CU: finer cuts only add per-item claims and dispatcher work
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      cus = 0;
    s->spread_workers = std::max(1, std::min(cus, s->grid - 1));
    // secondaries join as soon as any published item is left beyond the
    // tickets idle primaries hold (a margin of 64 items measured worse)
    s->sd.claim_backlog = 0;
  }
  if (rc == FFTOC_SUCCESS) rc = host_alloc<uint32_t>(&s->live, nullptr, &s->sd.live, s->grid);
  if (rc == FFTOC_SUCCESS) rc = host_alloc<HostRing>(&s->ring, nullptr, &ring_d, 1);
  if (rc == FFTOC_SUCCESS) rc = host_alloc<ReqDesc>(&s->descs, &s->sd.descs, nullptr, max_inflight);
  if (rc == FFTOC_SUCCESS) rc = host_alloc<Completion>(&s->comps, nullptr, &s->sd.comps, max_inflight);
  if (rc != FFTOC_SUCCESS) {
    free_server(s);
    return rc;
  }
  s->sd.ring = ring_d;
  for (int i = 0; i < max_inflight; ++i) {
    s->descs[i].seq = kSeqInvalid;
    s->comps[i].seq = kSeqInvalid;
  }
  s->pool.reset(hbm_budget_bytes & ~size_t(255));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    free_server(s);
    return FFTOC_ERROR_DEVICE;
  }
  std::atomic_thread_fence(std::memory_order_seq_cst);
  rc = start_loop(s);
  if (rc != FFTOC_SUCCESS) {
    free_server(s);
    return rc;
  }
  s->running = true;
  hold_tables();
  s->holds_tables = true;
  *out = s;
  return FFTOC_SUCCESS;
}

int fftoc_server_submit(fftoc_server s, const void *in, int64_t batch, int n, uint32_t flags,
                        uint64_t *request_id) {
  using namespace fftoc;
  if (!s || !s->running) return FFTOC_ERROR_INVALID_HANDLE;
  if (!request_id) return FFTOC_ERROR_NULL_POINTER;
  const uint32_t known = FFTOC_INVERSE | FFTOC_NORMALIZE | FFTOC_DATA_C64;
  if (flags & FFTOC_PREC_MASK) return FFTOC_ERROR_UNSUPPORTED;
  if (flags & ~known) return FFTOC_ERROR_INVALID_FLAGS;
  if (!pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  if (batch <= 0) return FFTOC_ERROR_INVALID_BATCH;
  if (!in) return FFTOC_ERROR_NULL_POINTER;
  const bool c64 = (flags & FFTOC_DATA_C64) != 0;
  const size_t es = c64 ? 8 : 16;
  if (reinterpret_cast<uintptr_t>(in) & (es - 1)) return FFTOC_ERROR_MISALIGNED;
  const size_t bytes = (size_t)batch * (size_t)n * es;
  if (bytes / es / (size_t)n != (size_t)batch) return FFTOC_ERROR_INVALID_BATCH;
  const int id = s->routine[lg(n)][c64 ? 1 : 0];
  const int S = batch_routine_signals(id);
  const int64_t ntiles = (batch + S - 1) / S;
  std::lock_guard<std::mutex> lock(s->m);
  // Work-item size follows the load: with little work queued a request is
  // spread over the workers (latency); with a deep queue items grow with the
  // backlog (throughput: fewer per-item claims and fences).
  const int64_t tile_bytes = (int64_t)S * n * (int64_t)es;
  const int64_t used = (int64_t)s->pool.used();
  const int64_t item_bytes =
      std::clamp<int64_t>(used / (4 * (int64_t)s->grid), kMinDeepItemBytes, kMaxDeepItemBytes);
  const int64_t big = std::max<int64_t>(1, item_bytes / tile_bytes);
  const bool deep = used > (int64_t)kDeepBacklogBytes;
  const int64_t spread = std::max<int64_t>(1, ntiles / (int64_t)s->spread_workers);
  const int32_t tpi = (int32_t)(deep ? big : std::min(big, spread));
  const int64_t nitems = (ntiles + tpi - 1) / tpi;
  constexpr int64_t kRingItems = (int64_t)kWorkRing / 2;  // keep half the ring as margin
  if (granule(bytes) > s->pool.capacity() || nitems > kRingItems || nitems >= (1 << kItemBits))
    return FFTOC_ERROR_BUDGET;
  if (s->next_seq >= kMaxRequestId) return FFTOC_ERROR_QUEUE_FULL;  // 2^40 per server
  const int slot = (int)(s->next_seq % (uint64_t)s->nslots);
  if (s->slots[slot].busy) return FFTOC_ERROR_QUEUE_FULL;  // ring full
  if (s->inflight_items + nitems > kRingItems) return FFTOC_ERROR_QUEUE_FULL;
  size_t off = 0;
  if (!s->pool.alloc(bytes, &off)) return FFTOC_ERROR_QUEUE_FULL;  // budget in use
  const uint64_t seq = s->next_seq++;
  Slot &sl = s->slots[slot];
  sl.busy = true;
  sl.seq = seq;
  sl.out_off = off;
  sl.out_bytes = bytes;
  sl.items = nitems;
  s->inflight_items += nitems;
  s->slot_of[seq] = slot;

  ReqDesc &d = s->descs[slot];
  volatile uint64_t *seqp = &d.seq;
  *seqp = kSeqInvalid;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  d.in = in;
  d.out = s->d_pool + off;
  d.batch = batch;
  d.ntiles = ntiles;
  d.nitems = nitems;
  d.n = n;
  d.routine = id;
  d.tiles_per_item = tpi;
  d.flags = ((flags & FFTOC_INVERSE) ? 1 : 0) | (c64 ? 2 : 0);
  d.scale = (flags & FFTOC_NORMALIZE) ? 1.0 / (double)n : 1.0;
  s->comps[slot].seq = kSeqInvalid;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  *seqp = seq;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  __atomic_store_n(&s->ring->published, seq + 1, __ATOMIC_RELEASE);
  *request_id = seq;
  return ensure_running(s);
}

// *done = 1 and *out = the result (device pointer, valid until release) once
// the request has completed; *done = 0 while it runs.
int fftoc_server_poll(fftoc_server s, uint64_t request_id, int *done, void **out) {
  if (!s) return FFTOC_ERROR_INVALID_HANDLE;
  if (!done) return FFTOC_ERROR_NULL_POINTER;
  std::lock_guard<std::mutex> lock(s->m);
  auto it = s->slot_of.find(request_id);
  if (it == s->slot_of.end()) return FFTOC_ERROR_INVALID_HANDLE;
  const int slot = it->second;
  const uint64_t seq = __atomic_load_n(&s->comps[slot].seq, __ATOMIC_ACQUIRE);
  *done = (seq == request_id) ? 1 : 0;
  if (out) *out = *done ? (void *)(s->d_pool + s->slots[slot].out_off) : nullptr;
  if (*done) return FFTOC_SUCCESS;
  // a pending request needs a running loop; check at most once per millisecond
  const double t = fftoc::now_s();
  if (t - s->last_check < 1e-3) return FFTOC_SUCCESS;
  s->last_check = t;
  return fftoc::ensure_running(s);
}

// Returns a completed request's output memory to the budget and its slot to
// the ring.
int fftoc_server_release(fftoc_server s, uint64_t request_id) {
  if (!s) return FFTOC_ERROR_INVALID_HANDLE;
  std::lock_guard<std::mutex> lock(s->m);
  auto it = s->slot_of.find(request_id);
  if (it == s->slot_of.end()) return FFTOC_ERROR_INVALID_HANDLE;
  const int slot = it->second;
  if (__atomic_load_n(&s->comps[slot].seq, __ATOMIC_ACQUIRE) != request_id)
    return FFTOC_ERROR_UNSUPPORTED;  // still running
  Slot &sl = s->slots[slot];
  s->pool.release(sl.out_off, sl.out_bytes);
  s->inflight_items -= sl.items;
  sl.busy = false;
  s->slot_of.erase(it);
  return FFTOC_SUCCESS;
}

// Device time (100 MHz ticks) at which a completed request finished.
int fftoc_server_gpu_time(fftoc_server s, uint64_t request_id, uint64_t *ticks) {
  if (!s || !ticks) return FFTOC_ERROR_NULL_POINTER;
  std::lock_guard<std::mutex> lock(s->m);
  auto it = s->slot_of.find(request_id);
  if (it == s->slot_of.end()) return FFTOC_ERROR_INVALID_HANDLE;
  const fftoc::Completion &c = s->comps[it->second];
  if (__atomic_load_n(&c.seq, __ATOMIC_ACQUIRE) != request_id) return FFTOC_ERROR_UNSUPPORTED;
  *ticks = __atomic_load_n(&c.gpu_time, __ATOMIC_RELAXED);
  return FFTOC_SUCCESS;
}

// Replays a request trace against the server: request i arrives at
// arrival_s[i] seconds after the call starts; requests are admitted in
// arrival order as slots and budget free up, completions are collected and
// released.  latency_s[i] = completion seen by the host - arrival (NaN for a
// request that was not served).  Admission that stays refused while none of
// the trace's requests is in flight (a slot or the budget held by requests
// submitted outside the trace and not released) ends the replay with
// FFTOC_ERROR_QUEUE_FULL after kReplayStallSeconds; on any early exit the
// requests the replay admitted are waited for and released first.
int fftoc_server_replay(fftoc_server s, const fftoc_job *reqs, int count, const double *arrival_s,
                        double *latency_s, double *span_s) {
  using namespace fftoc;
  constexpr double kReplayStallSeconds = 1.0;
  constexpr double kDrainSeconds = 30.0;
  if (!s || !reqs || !arrival_s || !latency_s) return FFTOC_ERROR_NULL_POINTER;
  if (count < 0) return FFTOC_ERROR_INVALID_BATCH;
  for (int i = 0; i < count; ++i) {
    if (reqs[i].batch <= 0 || !reqs[i].in) return FFTOC_ERROR_INVALID_BATCH;
    if (granule((size_t)reqs[i].batch * (size_t)reqs[i].n *
                ((reqs[i].flags & FFTOC_DATA_C64) ? 8 : 16)) > s->pool.capacity())
      return FFTOC_ERROR_BUDGET;
  }
  for (int i = 0; i < count; ++i) latency_s[i] = std::nan("");
  std::deque<int> waiting;
  struct Live {
    uint64_t id;
    int slot;
    int index;
  };
  std::vector<Live> inflight;
  int next = 0, completed = 0, rc = FFTOC_SUCCESS;
  double last_check = 0;
  const double t0 = now_s();
  double progress = t0;  // last admission or completion
  // One pass over the completion records under one lock: releases what has
  // completed and records its latency.
  auto collect = [&]() -> int {
    std::lock_guard<std::mutex> lock(s->m);
    int got = 0;
    for (size_t k = 0; k < inflight.size();) {
      const Live &lv = inflight[k];
      if (__atomic_load_n(&s->comps[lv.slot].seq, __ATOMIC_ACQUIRE) == lv.id) {
        latency_s[lv.index] = (now_s() - t0) - arrival_s[lv.index];
        Slot &sl = s->slots[lv.slot];
        s->pool.release(sl.out_off, sl.out_bytes);
        s->inflight_items -= sl.items;
        sl.busy = false;
        s->slot_of.erase(lv.id);
        inflight[k] = inflight.back();
        inflight.pop_back();
        ++got;
      } else {
        ++k;
      }
    }
    int erc = FFTOC_SUCCESS;
    if (!inflight.empty()) {
      const double t = now_s();
      if (t - last_check > 1e-3) {
        last_check = t;
        erc = ensure_running(s);
      }
    }
    completed += got;
    if (got) progress = now_s();
    return erc;
  };
  while (completed < count && rc == FFTOC_SUCCESS) {
    const double now = now_s() - t0;
    while (next < count && arrival_s[next] <= now) waiting.push_back(next++);
    while (!waiting.empty()) {
      const fftoc_job &r = reqs[waiting.front()];
      uint64_t id = 0;
      const int src = fftoc_server_submit(s, r.in, r.batch, r.n, r.flags, &id);
      if (src == FFTOC_ERROR_QUEUE_FULL) {
        // nothing of the trace in flight: only a release from outside can
        // free the slot or the budget
        if (inflight.empty() && now_s() - progress > kReplayStallSeconds)
          rc = FFTOC_ERROR_QUEUE_FULL;
        break;
      }
      if (src != FFTOC_SUCCESS) {
        rc = src;
        break;
      }
      inflight.push_back({id, (int)(id % (uint64_t)s->nslots), waiting.front()});
      waiting.pop_front();
      progress = now_s();
    }
    const int crc = collect();
    if (rc == FFTOC_SUCCESS) rc = crc;
  }
  // an early exit leaves nothing of the trace behind: its admitted requests
  // finish on the device and their slots and budget are returned
  const double drain_end = now_s() + kDrainSeconds;
  while (!inflight.empty() && now_s() < drain_end) (void)collect();
  if (!inflight.empty() && rc == FFTOC_SUCCESS) rc = FFTOC_ERROR_DEVICE;
  if (span_s) *span_s = now_s() - t0;
  return rc;
}

int fftoc_server_stats(fftoc_server s, size_t *budget_used, int *inflight) {
  if (!s) return FFTOC_ERROR_INVALID_HANDLE;
  std::lock_guard<std::mutex> lock(s->m);
  if (budget_used) *budget_used = s->pool.used();
  if (inflight) *inflight = (int)s->slot_of.size();
  return FFTOC_SUCCESS;
}

int fftoc_server_set_idle_timeout(fftoc_server s, double seconds) {
  if (!s) return FFTOC_ERROR_INVALID_HANDLE;
  if (!(seconds >= 1e-4 && seconds <= 60.0)) return FFTOC_ERROR_INVALID_BATCH;
  std::lock_guard<std::mutex> lock(s->m);
  s->sd.idle_timeout_ticks = (uint64_t)(seconds * kTicksPerSecond);
  return FFTOC_SUCCESS;
}

// Stops the loop and frees the server.  The wait for the loop to leave is
// bounded: the stop flag ends the dispatcher at its next poll and the workers
// at their next idle check (an item in progress finishes first).  A loop
// still running after kDestroyWaitSeconds is left alone with the memory it
// uses (freeing it under a running kernel would fault the device), and
// FFTOC_ERROR_DEVICE is returned.
int fftoc_server_destroy(fftoc_server s) {
  constexpr double kDestroyWaitSeconds = 10.0;
  if (!s) return FFTOC_ERROR_INVALID_HANDLE;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(s->device);
  int rc = FFTOC_SUCCESS;
  if (s->running) {
    __atomic_store_n(&s->ring->stop, 1u, __ATOMIC_RELEASE);
    const double end = fftoc::now_s() + kDestroyWaitSeconds;
    hipError_t q = hipStreamQuery(s->stream);
    while (q == hipErrorNotReady && fftoc::now_s() < end) {
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      q = hipStreamQuery(s->stream);
    }
    s->running = false;
    if (q == hipErrorNotReady) {
      (void)hipSetDevice(cur);
      return FFTOC_ERROR_DEVICE;  // the loop's memory stays allocated
    }
    if (q != hipSuccess) rc = FFTOC_ERROR_DEVICE;
  }
  fftoc::free_server(s);
  (void)hipSetDevice(cur);
  return rc;
}

}  // extern "C"
