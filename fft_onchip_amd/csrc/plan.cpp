// Mixed-size batch plans: many FFT jobs of different length, direction and
// data type executed by ONE persistent kernel launch (batch_kernels.hip),
// optionally replayed from a captured hipGraph, with the tile routine of each
// (length, data type) picked by an autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <vector>

#include "batch_internal.hpp"

struct fftoc_plan_t {
  int device = 0;
  std::vector<fftoc_job> jobs;          // as given by the caller
  std::vector<int> order;               // device order (heaviest items first)
  std::vector<fftoc::JobDev> host;      // descriptors in device order
  fftoc::JobDev *d_jobs = nullptr;
  unsigned long long *d_ctr = nullptr;  // [0] next item, [1] finished workgroups
  int64_t total_items = 0;
  int grid = 0;
  int one_routine = -1;                 // the routine of every job, if they share
  bool one_c64 = false;                 // one (routine, data type); else -1
  int one_dir = 2;                      // 0 / 1: every job forward / inverse
  int per_cu = 0;                       // workgroups per CU (0: the default)
  int routine[13][2];                   // [log2 n][complex64] -> routine id
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *tw64 = nullptr, *tw32 = nullptr;
  bool holds_tables = false;
};

namespace fftoc {
namespace {

constexpr int kMaxPlanPerCu = 16;  // largest workgroups-per-CU cap the autotuner tries

bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int lg(int n) { return 31 - __builtin_clz((unsigned)n); }

int validate_job(const fftoc_job &j) {
  const uint32_t known = FFTOC_INVERSE | FFTOC_NORMALIZE | FFTOC_DATA_C64;
  if (j.flags & FFTOC_PREC_MASK) return FFTOC_ERROR_UNSUPPORTED;
  if (j.flags & ~known) return FFTOC_ERROR_INVALID_FLAGS;
  if (!pow2(j.n) || j.n < 8 || j.n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  if (j.batch < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (j.batch == 0) return FFTOC_SUCCESS;
  if (!j.in || !j.out) return FFTOC_ERROR_NULL_POINTER;
  const uint64_t es = (j.flags & FFTOC_DATA_C64) ? 8u : 16u;
  const uintptr_t i0 = reinterpret_cast<uintptr_t>(j.in), o0 = reinterpret_cast<uintptr_t>(j.out);
  if ((i0 | o0) & (es - 1)) return FFTOC_ERROR_MISALIGNED;
  const uint64_t bytes = (uint64_t)j.batch * (uint64_t)j.n * es;
  if (bytes / es / (uint64_t)j.n != (uint64_t)j.batch) return FFTOC_ERROR_INVALID_BATCH;
  if (i0 != o0 && i0 < o0 + bytes && o0 < i0 + bytes) return FFTOC_ERROR_OVERLAP;
  return FFTOC_SUCCESS;
}

// Every job's output range must be disjoint from every other job's input and
// output range (a job may be in place): the jobs run concurrently.
int check_cross_overlap(const fftoc_job *jobs, int njobs) {
  struct Range {
    uintptr_t lo, hi;
    int job;
    bool written;
  };
  std::vector<Range> r;
  r.reserve(2 * (size_t)njobs);
  for (int i = 0; i < njobs; ++i) {
    const fftoc_job &j = jobs[i];
    if (j.batch == 0) continue;
    const uintptr_t bytes = (uintptr_t)j.batch * (uintptr_t)j.n *
                            ((j.flags & FFTOC_DATA_C64) ? 8u : 16u);
    const uintptr_t in = reinterpret_cast<uintptr_t>(j.in);
    const uintptr_t o = reinterpret_cast<uintptr_t>(j.out);
    r.push_back({o, o + bytes, i, true});
    if (in != o) r.push_back({in, in + bytes, i, false});
  }
  std::sort(r.begin(), r.end(), [](const Range &a, const Range &b) { return a.lo < b.lo; });
  // sweep: compare each range with the earlier ones still open
  std::vector<Range> open;
  for (const Range &c : r) {
    open.erase(std::remove_if(open.begin(), open.end(),
                              [&](const Range &o) { return o.hi <= c.lo; }),
               open.end());
    for (const Range &o : open)
      if (o.job != c.job && (o.written || c.written)) return FFTOC_ERROR_OVERLAP;
    open.push_back(c);
  }
  return FFTOC_SUCCESS;
}


// Descriptor of one job whose work items start at item_begin: about 64 KiB
// of input per item.
JobDev make_jobdev(const fftoc_job &j, int routine, int64_t item_begin) {
  const bool c64 = (j.flags & FFTOC_DATA_C64) != 0;
  JobDev d{};
  d.in = j.in;
  d.out = j.out;
  d.batch = j.batch;
  d.routine = routine;
  const int S = batch_routine_signals(d.routine);
  d.ntiles = (j.batch + S - 1) / S;
  const int64_t tile_bytes = (int64_t)S * j.n * (c64 ? 8 : 16);
  d.tiles_per_item = (int32_t)std::max<int64_t>(1, 65536 / tile_bytes);
  d.item_begin = item_begin;
  d.inverse = (j.flags & FFTOC_INVERSE) ? 1 : 0;
  d.c64 = c64 ? 1 : 0;
  d.scale = (j.flags & FFTOC_NORMALIZE) ? 1.0 / (double)j.n : 1.0;
  return d;
}

int64_t job_items(const JobDev &d) { return (d.ntiles + d.tiles_per_item - 1) / d.tiles_per_item; }

// Device order of the jobs: heaviest items first (longer transforms, then
// bigger batches); stable.
bool heavier(const fftoc_job &a, const fftoc_job &b) {
  if (a.n != b.n) return a.n > b.n;
  return a.batch > b.batch;
}

// Lays the jobs out as work items and uploads the descriptors.
int layout(fftoc_plan_t *p) {
  // Jobs sharing one routine and data type run the kernel specialized to it;
  // otherwise the mixed-size kernel runs, with kBatchThreads routines only.
  p->one_routine = -1;
  p->one_dir = 2;
  bool first = true;
  for (int idx : p->order) {
    const fftoc_job &j = p->jobs[idx];
    if (j.batch == 0) continue;
    const bool c64 = (j.flags & FFTOC_DATA_C64) != 0;
    const int r = p->routine[lg(j.n)][c64 ? 1 : 0];
    const int dir = (j.flags & FFTOC_INVERSE) ? 1 : 0;
    if (first) {
      p->one_routine = r;
      p->one_c64 = c64;
      p->one_dir = dir;
      first = false;
    } else if (r != p->one_routine || c64 != p->one_c64) {
      p->one_routine = -1;
      break;
    } else if (dir != p->one_dir) {
      p->one_dir = 2;
    }
  }
  p->host.clear();
  p->total_items = 0;
  for (int idx : p->order) {
    const fftoc_job &j = p->jobs[idx];
    if (j.batch == 0) continue;
    const bool c64 = (j.flags & FFTOC_DATA_C64) != 0;
    int r = p->routine[lg(j.n)][c64 ? 1 : 0];
    if (p->one_routine < 0) r = batch_routine_mixed(r, c64);
    const JobDev d = make_jobdev(j, r, p->total_items);
    p->total_items += job_items(d);
    p->host.push_back(d);
  }
  if (p->d_jobs) {
    (void)hipFree(p->d_jobs);
    p->d_jobs = nullptr;
  }
  if (!p->host.empty()) {
    const size_t bytes = p->host.size() * sizeof(JobDev);
    if (hipMalloc(&p->d_jobs, bytes) != hipSuccess) return FFTOC_ERROR_ALLOC;
    if (hipMemcpy(p->d_jobs, p->host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
      return FFTOC_ERROR_DEVICE;
  }
  const int64_t resident =
      p->one_routine >= 0 ? batch_one_resident(p->one_routine, p->one_c64, p->one_dir, p->per_cu)
                          : batch_resident_blocks(p->per_cu);
  p->grid = (int)std::max<int64_t>(1, std::min<int64_t>(resident, p->total_items));
  if (p->exec) {
    (void)hipGraphExecDestroy(p->exec);
    p->exec = nullptr;
  }
  if (p->graph) {
    (void)hipGraphDestroy(p->graph);
    p->graph = nullptr;
  }
  return FFTOC_SUCCESS;
}

int launch(fftoc_plan_t *p, hipStream_t s) {
  if (p->total_items == 0) return FFTOC_SUCCESS;
  if (p->one_routine >= 0)
    return launch_batch_one(p->one_routine, p->one_c64, p->one_dir, p->d_jobs,
                            (int)p->host.size(), p->total_items, p->d_ctr,
                            p->one_c64 ? p->tw32 : p->tw64, p->grid, s);
  return launch_batch(p->d_jobs, (int)p->host.size(), p->total_items, p->d_ctr, p->tw64, p->tw32,
                      p->grid, s);
}

int check_device(fftoc_plan_t *p) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return FFTOC_ERROR_DEVICE;
  return dev == p->device ? FFTOC_SUCCESS : FFTOC_ERROR_DEVICE;
}

}  // namespace
}  // namespace fftoc

extern "C" {

int fftoc_plan_create(const fftoc_job *jobs, int njobs, fftoc_plan *out) {
  using namespace fftoc;
  if (!out) return FFTOC_ERROR_NULL_POINTER;
  *out = nullptr;
  if (njobs < 0 || (njobs > 0 && !jobs)) return FFTOC_ERROR_NULL_POINTER;
  for (int i = 0; i < njobs; ++i) {
    const int rc = validate_job(jobs[i]);
    if (rc != FFTOC_SUCCESS) return rc;
  }
  int rc = check_cross_overlap(jobs, njobs);
  if (rc != FFTOC_SUCCESS) return rc;
  const void *tw64 = nullptr, *tw32 = nullptr;
  rc = device_tables(&tw64, &tw32);
  if (rc != FFTOC_SUCCESS) return rc;
  fftoc_plan_t *p = new (std::nothrow) fftoc_plan_t;
  if (!p) return FFTOC_ERROR_ALLOC;
  (void)hipGetDevice(&p->device);
  p->tw64 = tw64;
  p->tw32 = tw32;
  p->jobs.assign(jobs, jobs + njobs);
  p->order.resize(njobs);
  for (int i = 0; i < njobs; ++i) p->order[i] = i;
  // heaviest items first: longer transforms, then bigger batches
  std::stable_sort(p->order.begin(), p->order.end(),
                   [&](int a, int b) { return heavier(p->jobs[a], p->jobs[b]); });
  for (int k = 0; k < 13; ++k) {
    p->routine[k][0] = batch_routine_default(1 << k, false);
    p->routine[k][1] = batch_routine_default(1 << k, true);
  }
  if (hipMalloc(&p->d_ctr, 2 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(p->d_ctr, 0, 2 * sizeof(unsigned long long)) != hipSuccess) {
    fftoc_plan_destroy(p);
    return FFTOC_ERROR_ALLOC;
  }
  rc = layout(p);
  if (rc != FFTOC_SUCCESS) {
    fftoc_plan_destroy(p);
    return rc;
  }
  hold_tables();
  p->holds_tables = true;
  *out = p;
  return FFTOC_SUCCESS;
}

int fftoc_plan_execute(fftoc_plan p, void *stream) {
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  const int rc = fftoc::check_device(p);
  if (rc != FFTOC_SUCCESS) return rc;
  return fftoc::launch(p, static_cast<hipStream_t>(stream));
}

int fftoc_plan_execute_graph(fftoc_plan p, void *stream) {
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  int rc = fftoc::check_device(p);
  if (rc != FFTOC_SUCCESS) return rc;
  if (p->total_items == 0) return FFTOC_SUCCESS;
  if (!p->exec) {
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return FFTOC_ERROR_DEVICE;
    rc = FFTOC_SUCCESS;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      rc = FFTOC_ERROR_DEVICE;
    } else {
      const int lrc = fftoc::launch(p, cs);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(cs, &g);
      if (lrc != FFTOC_SUCCESS || e != hipSuccess || !g) {
        rc = lrc != FFTOC_SUCCESS ? lrc : FFTOC_ERROR_DEVICE;
        if (g) (void)hipGraphDestroy(g);
      } else if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        rc = FFTOC_ERROR_DEVICE;
      } else {
        p->graph = g;
      }
    }
    (void)hipStreamDestroy(cs);
    if (rc != FFTOC_SUCCESS) return rc;
  }
  return hipGraphLaunch(p->exec, static_cast<hipStream_t>(stream)) == hipSuccess
             ? FFTOC_SUCCESS
             : FFTOC_ERROR_LAUNCH;
}

int fftoc_plan_autotune(fftoc_plan p, int reps, void *stream_p) {
  using namespace fftoc;
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  if (reps <= 0) return FFTOC_ERROR_INVALID_BATCH;
  int rc = check_device(p);
  if (rc != FFTOC_SUCCESS) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream_p);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return FFTOC_ERROR_DEVICE;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return FFTOC_ERROR_DEVICE;
  }
  // Mean over `reps` back-to-back launches after a warm-up launch: plans run
  // in loops, and a launch timed alone after a synchronisation starts from an
  // idle GPU, which favours kernels with more workgroups.
  auto time_plan = [&](double *us) -> int {
    if ((rc = launch(p, s)) != FFTOC_SUCCESS) return rc;  // warm-up
    float ms = 0;
    if (hipEventRecord(e0, s) != hipSuccess) return FFTOC_ERROR_DEVICE;
    for (int r = 0; r < reps; ++r)
      if ((rc = launch(p, s)) != FFTOC_SUCCESS) return rc;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return FFTOC_ERROR_DEVICE;
    *us = 1000.0 * ms / reps;
    return FFTOC_SUCCESS;
  };
  // coordinate descent over the (length, data type) classes present
  std::map<std::pair<int, int>, bool> classes;
  for (const fftoc_job &j : p->jobs)
    if (j.batch > 0) classes[{lg(j.n), (j.flags & FFTOC_DATA_C64) ? 1 : 0}] = true;
  rc = FFTOC_SUCCESS;
  for (const auto &kv : classes) {
    const int k = kv.first.first, c = kv.first.second;
    int best = p->routine[k][c];
    double best_us = 1e300;
    for (int id = 0; id < batch_routine_count() && rc == FFTOC_SUCCESS; ++id) {
      if (batch_routine_n(id) != (1 << k)) continue;
      p->routine[k][c] = id;
      if ((rc = layout(p)) != FFTOC_SUCCESS) break;
      double us = 0;
      if ((rc = time_plan(&us)) != FFTOC_SUCCESS) break;
      if (us < best_us) {
        best_us = us;
        best = id;
      }
    }
    p->routine[k][c] = best;
    if (rc != FFTOC_SUCCESS) break;
  }
  // then the workgroups per CU of the chosen kernel: the default, then every
  // cap up to its occupancy; a cap must win by 1 % (timing noise)
  if (rc == FFTOC_SUCCESS) {
    const int keep = p->per_cu;
    p->per_cu = 0;
    rc = layout(p);
    int occ = 0, cus = 0;
    if (rc == FFTOC_SUCCESS)
      occ = p->one_routine >= 0 ? batch_one_occupancy(p->one_routine, p->one_c64, p->one_dir, &cus)
                                : batch_occupancy();
    double best_us = 1e300;
    int best = 0;
    for (int cap = 0; cap <= std::min(occ, kMaxPlanPerCu) && rc == FFTOC_SUCCESS; ++cap) {
      p->per_cu = cap;
      if ((rc = layout(p)) != FFTOC_SUCCESS) break;
      double us = 0;
      if ((rc = time_plan(&us)) != FFTOC_SUCCESS) break;
      if (cap == 0 || us < best_us * 0.99) {
        best_us = us;
        best = cap;
      }
    }
    p->per_cu = rc == FFTOC_SUCCESS ? best : keep;
  }
  const int lrc = layout(p);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc != FFTOC_SUCCESS ? rc : lrc;
}

int fftoc_plan_workgroups_per_cu(fftoc_plan p) { return p ? p->per_cu : -1; }

int fftoc_plan_set_workgroups_per_cu(fftoc_plan p, int per_cu) {
  using namespace fftoc;
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  if (per_cu < 0 || per_cu > kMaxPerCu) return FFTOC_ERROR_INVALID_FLAGS;
  const int rc = check_device(p);
  if (rc != FFTOC_SUCCESS) return rc;
  p->per_cu = per_cu;
  return layout(p);
}

int fftoc_plan_routine(fftoc_plan p, int n, uint32_t flags, const char **name) {
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  if (!fftoc::pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  const int id = p->routine[fftoc::lg(n)][(flags & FFTOC_DATA_C64) ? 1 : 0];
  if (name) *name = fftoc::batch_routine_name(id);
  return FFTOC_SUCCESS;
}

int fftoc_plan_set_routine(fftoc_plan p, int n, uint32_t flags, const char *name) {
  using namespace fftoc;
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  if (!name) return FFTOC_ERROR_NULL_POINTER;
  if (!pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  const int rc = check_device(p);
  if (rc != FFTOC_SUCCESS) return rc;
  for (int id = 0; id < batch_routine_count(); ++id) {
    if (batch_routine_n(id) == n && std::strcmp(batch_routine_name(id), name) == 0) {
      p->routine[lg(n)][(flags & FFTOC_DATA_C64) ? 1 : 0] = id;
      return layout(p);
    }
  }
  return FFTOC_ERROR_UNSUPPORTED;
}

int fftoc_batch_routine_name(int n, int index, const char **name) {
  using namespace fftoc;
  if (!name) return FFTOC_ERROR_NULL_POINTER;
  if (!pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  int k = 0;
  for (int id = 0; id < batch_routine_count(); ++id) {
    if (batch_routine_n(id) != n) continue;
    if (k++ == index) {
      *name = batch_routine_name(id);
      return FFTOC_SUCCESS;
    }
  }
  return FFTOC_ERROR_INVALID_HANDLE;
}

int fftoc_plan_destroy(fftoc_plan p) {
  if (!p) return FFTOC_ERROR_INVALID_HANDLE;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(p->device);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->d_jobs) (void)hipFree(p->d_jobs);
  if (p->d_ctr) (void)hipFree(p->d_ctr);
  (void)hipSetDevice(cur);
  if (p->holds_tables) fftoc::drop_tables();
  delete p;
  return FFTOC_SUCCESS;
}

// One launch over up to kNoAllocMaxJobs jobs with no allocation: the same
// validation, job order, tile routines (the untuned defaults) and work items
// as a plan of these jobs, with the descriptors and the work counter in the
// library's fixed per-device scratch array instead of plan memory.  Uses no
// heap either: the descriptors are built on the stack and reach the device as
// a kernel argument.
int fftoc_batch_fft_noalloc(const fftoc_job *jobs, int njobs, void *stream) {
  using namespace fftoc;
  if (njobs < 0 || (njobs > 0 && !jobs)) return FFTOC_ERROR_NULL_POINTER;
  if (njobs > kNoAllocMaxJobs) return FFTOC_ERROR_BUDGET;
  for (int i = 0; i < njobs; ++i) {
    const int rc = validate_job(jobs[i]);
    if (rc != FFTOC_SUCCESS) return rc;
  }
  // every job's output disjoint from every other job's input and output
  auto range = [](const void *p, const fftoc_job &j, uintptr_t *lo, uintptr_t *hi) {
    *lo = reinterpret_cast<uintptr_t>(p);
    *hi = *lo + (uintptr_t)j.batch * (uintptr_t)j.n * ((j.flags & FFTOC_DATA_C64) ? 8u : 16u);
  };
  auto meets = [](uintptr_t a0, uintptr_t a1, uintptr_t b0, uintptr_t b1) {
    return a0 < b1 && b0 < a1;
  };
  for (int a = 0; a < njobs; ++a) {
    if (jobs[a].batch == 0) continue;
    uintptr_t ao0, ao1;
    range(jobs[a].out, jobs[a], &ao0, &ao1);
    for (int b = 0; b < njobs; ++b) {
      if (b == a || jobs[b].batch == 0) continue;
      uintptr_t bi0, bi1, bo0, bo1;
      range(jobs[b].in, jobs[b], &bi0, &bi1);
      range(jobs[b].out, jobs[b], &bo0, &bo1);
      if (meets(ao0, ao1, bi0, bi1) || meets(ao0, ao1, bo0, bo1)) return FFTOC_ERROR_OVERLAP;
    }
  }
  int order[kNoAllocMaxJobs];
  int live = 0;
  for (int i = 0; i < njobs; ++i) {
    if (jobs[i].batch == 0) continue;
    int k = live++;
    while (k > 0 && heavier(jobs[i], jobs[order[k - 1]])) {  // stable insertion
      order[k] = order[k - 1];
      --k;
    }
    order[k] = i;
  }
  if (live == 0) return FFTOC_SUCCESS;
  const void *tw64 = nullptr, *tw32 = nullptr;
  const int rc = device_tables(&tw64, &tw32);  // sets the device up on first use
  if (rc != FFTOC_SUCCESS) return rc;
  NoAllocJobs desc;
  int64_t total = 0;
  for (int k = 0; k < live; ++k) {
    const fftoc_job &j = jobs[order[k]];
    const bool c64 = (j.flags & FFTOC_DATA_C64) != 0;
    desc.j[k] = make_jobdev(j, batch_routine_mixed(batch_routine_default(j.n, c64), c64), total);
    total += job_items(desc.j[k]);
  }
  const int grid = (int)std::min<int64_t>(batch_resident_blocks(), total);
  return launch_batch_noalloc(desc, live, total, tw64, tw32, grid,
                              static_cast<hipStream_t>(stream));
}

}  // extern "C"
