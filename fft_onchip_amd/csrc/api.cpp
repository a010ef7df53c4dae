// C ABI of libfftoc.so: argument checking, per-device FP64 twiddle tables and
// dispatch to the gfx950 kernels.  Declarations and semantics: include/fftoc.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fftoc_internal.hpp"

namespace fftoc {
namespace {

struct DeviceState {
  bool ready = false;
  int status = FFTOC_SUCCESS;
  void *twiddles = nullptr;    // kTwN x double2 (module global of device_tables.hip)
  void *twiddles32 = nullptr;  // kTwN x float2, rounded from the FP64 table
};

std::mutex g_mutex;
DeviceState g_dev[kMaxDevices];

// Host copies of the tables, built once (static storage: initialising a
// device allocates nothing, on the host or on the device).
constexpr int kTw = 4096;
double g_tw_host[2 * kTw];
float g_tw32_host[2 * kTw];
bool g_tw_built = false;

// W_4096^j = exp(-2 pi i j / 4096) to FP64.  The first octant (j <= 512) is
// evaluated with long-double cosl/sinl; every other entry is an exact
// symmetry of it (swap / negate), so all 4096 values are the correctly
// rounded FP64 values of long-double results.  Called with g_mutex held.
void build_twiddles() {
  if (g_tw_built) return;
  const int n = kTw;
  long double c[kTw / 8 + 1], s[kTw / 8 + 1];
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int j = 0; j <= n / 8; ++j) {
    const long double ang = two_pi * (long double)j / (long double)n;
    c[j] = cosl(ang);
    s[j] = sinl(ang);
  }
  c[n / 8] = s[n / 8] = 0.707106781186547524400844362104849039L;
  for (int j = 0; j < n; ++j) {
    const int q = j / (n / 4);        // quadrant
    const int r = j - q * (n / 4);    // 0 .. n/4-1
    long double cr, sr;               // cos, sin of 2 pi r / n
    if (r <= n / 8) {
      cr = c[r];
      sr = s[r];
    } else {
      cr = s[n / 4 - r];
      sr = c[n / 4 - r];
    }
    long double cv, sv;  // cos, sin of 2 pi j / n
    switch (q) {
      case 0: cv = cr; sv = sr; break;
      case 1: cv = -sr; sv = cr; break;
      case 2: cv = -cr; sv = -sr; break;
      default: cv = sr; sv = -cr; break;
    }
    g_tw_host[2 * j] = (double)cv;
    g_tw_host[2 * j + 1] = (double)(-sv);  // forward root: exp(-i angle)
  }
  for (int i = 0; i < 2 * n; ++i) g_tw32_host[i] = (float)g_tw_host[i];
  g_tw_built = true;
}

// Sets up the current device once: checks it is a gfx950, copies the tables
// into its module globals and locates the allocation-free batch entry point's
// scratch array on it.  Allocates nothing.
int init_current_device(int *dev_out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return FFTOC_ERROR_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return FFTOC_ERROR_DEVICE;
  *dev_out = dev;
  std::lock_guard<std::mutex> lock(g_mutex);
  DeviceState &st = g_dev[dev];
  if (st.ready) return st.status;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return FFTOC_ERROR_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    // kernels are built for gfx950 only
    return FFTOC_ERROR_DEVICE;
  }
  build_twiddles();
  void *d = nullptr, *d32 = nullptr;
  if (device_table_addresses(&d, &d32) != FFTOC_SUCCESS) return FFTOC_ERROR_DEVICE;
  if (hipMemcpy(d, g_tw_host, sizeof g_tw_host, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d32, g_tw32_host, sizeof g_tw32_host, hipMemcpyHostToDevice) != hipSuccess)
    return FFTOC_ERROR_DEVICE;
  const int rc = batch_noalloc_setup(dev);
  if (rc != FFTOC_SUCCESS) return rc;
  st.twiddles = d;
  st.twiddles32 = d32;
  st.status = FFTOC_SUCCESS;
  st.ready = true;
  return FFTOC_SUCCESS;
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
int log2i(int n) { return 31 - __builtin_clz((unsigned)n); }

enum class Family { Tensor, Radix8, Radix16, Best };

unsigned family_bit(Family f) {
  switch (f) {
    case Family::Tensor: return FAM_TENSOR;
    case Family::Radix8: return FAM_RADIX8;
    case Family::Radix16: return FAM_RADIX16;
    default: return FAM_ANY;
  }
}
int family_index(Family f) { return (int)f; }

// ---- variant registry and tuning table --------------------------------------
// Variants of length n = 2^k (k = 3..12): tensor variants first, then radix.
constexpr int kBuckets = 32;  // batch bucket = floor(log2(batch)), 0..31
int batch_bucket(int64_t batch) {
  if (batch <= 1) return 0;
  const int b = 63 - __builtin_clzll((unsigned long long)batch);
  return b < kBuckets ? b : kBuckets - 1;
}

struct SizeRegistry {
  std::vector<Variant> v;
  std::atomic<int> choice[4][kBuckets];  // [Family][batch bucket] -> index into v
  std::atomic<int> per_cu[4][kBuckets];  // workgroups per CU of that choice (0: its occupancy)
  std::vector<double> tuned_us;  // last autotune time per variant (<0: not run),
                                 // best over the workgroups-per-CU caps tried
  std::vector<int> tuned_per_cu;  // the cap of that time (0: the occupancy);
                                  // both guarded by g_tune_mutex
};

SizeRegistry g_reg[13][2];  // [log2 n][complex64 data]
std::mutex g_tune_mutex;
std::once_flag g_reg_once;

// Default choice of fftoc_fft per length and batch bucket -- variant and
// workgroups-per-CU cap (0: the kernel's occupancy) -- from MI355X
// measurements of the variants at 2 MiB - 1 GiB batches, buffers rotating
// past the 256 MB last-level cache (replaced by fftoc_autotune /
// fftoc_tuning_load).  The caps pay only on large batches: fewer resident
// workgroups per CU (or one persistent workgroup looping over tiles) gave
// +1.5-5 % from 128 MiB up, and lost 10-20 % at 32 MiB.
struct Default {
  const char *name;
  int per_cu;
};
Default default_best(int n, int bucket) {
  switch (n) {
    case 64:  // up to 2^16 signals the persistent form hides the launch tail
      return {bucket <= 16 ? "radix:8x8/persistent/S32" : "radix:8x8/tile/S16", 0};
    // N = 8, 16, 32 from 2^18 / 2^17 / 2^16 signals: one-wave workgroups of
    // 1-2 KiB, whose loads and stores stay in the compact, in-order address
    // window the dispatcher gives small workgroups (+2-4 % over 256 threads)
    case 8: return {bucket >= 18 ? "radix:8/lanes/S8/64 threads" : "radix:8/lanes/S128", 0};
    case 16: return {bucket >= 17 ? "radix:16/lanes/S4/64 threads" : "radix:16/lanes/S64", 0};
    case 32: return {bucket >= 16 ? "radix:32/lanes/S4/64 threads" : "radix:32/lanes/S32", 0};
    case 128: return {"radix:16x8/tile/S8", bucket >= 16 ? 6 : 0};
    case 256: return {"radix:16x16/tile/S4", bucket >= 15 ? 6 : 0};
    case 512:  // two-wave pipelined workgroups, two per CU (one wave per SIMD)
      return bucket >= 13 ? Default{"radix:8x8x8/pipelined/S2", 2}
                          : Default{"radix:8x8x8/tile/S1", 0};
    case 1024:
      return bucket >= 13 ? Default{"radix:8x8x8x2/pipelined/S1", 3}
                          : Default{"radix:8x8x16/tile/S2", 0};
    case 2048: return {"radix:8x16x16/tile/S1", 0};
    default: return {"radix:8x8x8x8/pipelined/S1", 0};
  }
}

// complex64 data: the pipelined form (with a cap of workgroups per CU) where
// it measured ahead of the list's first variant; elsewhere that first one.
Default default_c64(int n, int bucket) {
  switch (n) {
    case 64: return {"radix:4x4x4/pipelined/S16", 3};
    case 128: return bucket <= 18 ? Default{"radix:8x4x4/pipelined/S16", 4} : Default{nullptr, 0};
    case 512: return {"radix:8x8x8/pipelined/S4", 2};
    default: return {nullptr, 0};
  }
}

// Default choice of fftoc_tensor_fft where it depends on the batch size (else
// the first tensor variant of the length).  N = 128 from 2^15 signals and
// N = 256, 512 from 2^14: persistent kernels with one-wave workgroups held
// to four per CU (one wave per SIMD), one unit of 4-8 KiB per iteration:
// N = 128 0.68-0.78 of HBM against 0.64-0.78, N = 256 0.75-0.80 against
// 0.69-0.74, N = 512 0.68-0.75 against 0.65-0.69 for four-wave workgroups at
// their occupancy; three or five per CU measured lower (N = 512: 0.63-0.68 /
// 0.58).
Default default_tensor(int n, int bucket) {
  if (n == 64 && bucket <= 16) return {"tensor:16x4/8 per wave/persistent", 0};
  if (n == 8 && bucket >= 18) return {"tensor:diag(F8,F8)/32 per wave/LDS-staged/1 wave", 0};
  if (n == 128 && bucket >= 15) return {"tensor:16x8/2 per wave/persistent/1 wave", 4};
  if (n == 256 && bucket >= 14) return {"tensor:16x16/1 per wave/persistent/1 wave", 4};
  if (n == 512 && bucket >= 14) return {"tensor:16x2x16/1 per wave/persistent/1 wave", 4};
  return {nullptr, 0};
}

void build_registry() {
  for (int k = 3; k <= 12; ++k) for (int c64 = 0; c64 < 2; ++c64) {
    const int n = 1 << k;
    SizeRegistry &r = g_reg[k][c64];
    const Variant *t = nullptr, *x = nullptr;
    const int nt = tensor_variants(n, c64 != 0, &t), nx = radix_variants(n, c64 != 0, &x);
    for (int i = 0; i < nt; ++i) r.v.push_back(t[i]);
    for (int i = 0; i < nx; ++i) r.v.push_back(x[i]);
    r.tuned_us.assign(r.v.size(), -1.0);
    r.tuned_per_cu.assign(r.v.size(), 0);
    const Family fams[4] = {Family::Tensor, Family::Radix8, Family::Radix16, Family::Best};
    for (Family f : fams) {
      int first = -1;
      for (size_t i = 0; i < r.v.size(); ++i) {
        if (r.v[i].families & family_bit(f)) {
          first = (int)i;
          break;
        }
      }
      for (int b = 0; b < kBuckets; ++b) {
        int pick = first, cap = 0;
        // fftoc_fft's default; the radix families take it too where it is
        // one of their variants
        Default want{nullptr, 0};
        if (!c64 && f != Family::Tensor) want = default_best(n, b);
        if (c64 && f != Family::Tensor) want = default_c64(n, b);
        if (!c64 && f == Family::Tensor) want = default_tensor(n, b);
        if (want.name) {
          for (size_t i = 0; i < r.v.size(); ++i)
            if (std::strcmp(r.v[i].name, want.name) == 0 && (r.v[i].families & family_bit(f))) {
              pick = (int)i;
              cap = want.per_cu;
            }
        }
        r.choice[family_index(f)][b].store(pick);
        r.per_cu[family_index(f)][b].store(cap);
      }
    }
  }
}

SizeRegistry *registry(int n, bool c64 = false) {
  if (!is_pow2(n) || n < 8 || n > 4096) return nullptr;
  std::call_once(g_reg_once, build_registry);
  return &g_reg[log2i(n)][c64 ? 1 : 0];
}

int check_and_fill(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                   void *stream, LaunchArgs &a) {
  const uint32_t known = FFTOC_INVERSE | FFTOC_NORMALIZE | FFTOC_PREC_MASK | FFTOC_DATA_C64 |
                         FFTOC_TILE_MASK | FFTOC_WGS_PER_CU_MASK;
  if (flags & ~known) return FFTOC_ERROR_INVALID_FLAGS;
  const uint32_t prec = flags & FFTOC_PREC_MASK;
  if (prec != FFTOC_PREC_FP64 && prec != FFTOC_PREC_FP32 && prec != FFTOC_PREC_BF16X3)
    return FFTOC_ERROR_INVALID_FLAGS;
  if (!is_pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  if (batch < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (batch > 0) {
    if (!in || !out) return FFTOC_ERROR_NULL_POINTER;
    const uint64_t esize = (flags & FFTOC_DATA_C64) ? 8u : 16u;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & (esize - 1u))
      return FFTOC_ERROR_MISALIGNED;
    const uint64_t bytes = (uint64_t)batch * (uint64_t)n * esize;
    if (bytes / esize / (uint64_t)n != (uint64_t)batch) return FFTOC_ERROR_INVALID_BATCH;
    const uintptr_t i0 = reinterpret_cast<uintptr_t>(in), o0 = reinterpret_cast<uintptr_t>(out);
    if (i0 != o0 && i0 < o0 + bytes && o0 < i0 + bytes) return FFTOC_ERROR_OVERLAP;
  }
  a.in = in;
  a.out = out;
  a.batch = batch;
  a.stream = static_cast<hipStream_t>(stream);
  a.inverse = (flags & FFTOC_INVERSE) != 0;
  a.scale = (flags & FFTOC_NORMALIZE) ? 1.0 / (double)n : 1.0;
  a.tile = (int)((flags & FFTOC_TILE_MASK) >> FFTOC_TILE_SHIFT);
  a.precision = prec;
  a.twiddles = nullptr;
  a.c64 = (flags & FFTOC_DATA_C64) != 0;
  a.per_cu = (int)((flags & FFTOC_WGS_PER_CU_MASK) >> FFTOC_WGS_PER_CU_SHIFT);
  a.occupancy = nullptr;
  return FFTOC_SUCCESS;
}

int run(Family fam, const void *in, void *out, int64_t batch, int n, uint32_t flags,
        void *stream, int variant = -1) {
  LaunchArgs a;
  int rc = check_and_fill(in, out, batch, n, flags, stream, a);
  if (rc != FFTOC_SUCCESS) return rc;
  // family-specific size / precision rules (checked before touching the device)
  switch (fam) {
    case Family::Radix8:
      if (n != 8 && n != 64 && n != 512 && n != 4096) return FFTOC_ERROR_INVALID_SIZE;
      if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
      break;
    case Family::Radix16:
      if (n != 16 && n != 256 && n != 4096) return FFTOC_ERROR_INVALID_SIZE;
      if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
      break;
    case Family::Tensor:
      if (n == 1024 || n == 2048) return FFTOC_ERROR_INVALID_SIZE;
      if (a.precision != FFTOC_PREC_FP64 && (a.c64 || (n != 64 && n != 256 && n != 4096)))
        return FFTOC_ERROR_UNSUPPORTED;
      break;
    case Family::Best:
      if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
      break;
  }
  SizeRegistry *reg = registry(n, a.c64);
  if (!reg) return FFTOC_ERROR_INVALID_SIZE;
  int pick = variant >= 0 ? variant : reg->choice[family_index(fam)][batch_bucket(batch)].load();
  if (pick < 0 || pick >= (int)reg->v.size())
    return a.c64 ? FFTOC_ERROR_UNSUPPORTED : FFTOC_ERROR_INVALID_SIZE;
  const int requested_per_cu = a.per_cu;  // FFTOC_WGS_PER_CU of the call, 0 if none
  // the tuned workgroups-per-CU cap belongs to the tuned choice
  if (variant < 0 && a.per_cu == 0 && a.tile == 0)
    a.per_cu = reg->per_cu[family_index(fam)][batch_bucket(batch)].load();
  if (a.tile > 0 && reg->v[pick].tile != a.tile) {
    // FFTOC_TILE(s): the family's variant with s signals per workgroup (the
    // tuned choice when it has that tile, else the first listed)
    if (variant >= 0) return FFTOC_ERROR_INVALID_FLAGS;  // contradicts the named variant
    pick = -1;
    for (size_t i = 0; i < reg->v.size() && pick < 0; ++i)
      if ((reg->v[i].families & family_bit(fam)) && reg->v[i].tile == a.tile) pick = (int)i;
    if (pick < 0) return FFTOC_ERROR_UNSUPPORTED;
  }
  if (batch == 0) return FFTOC_SUCCESS;
  int dev = 0;
  rc = init_current_device(&dev);
  if (rc != FFTOC_SUCCESS) return rc;
  a.twiddles = a.c64 ? g_dev[dev].twiddles32 : g_dev[dev].twiddles;
  a.twiddles32 = g_dev[dev].twiddles32;
  if (fam == Family::Tensor && a.precision != FFTOC_PREC_FP64) {
    // the split-precision kernels are not FP64 variants: only the call's own
    // cap applies (their defaults are in launch_tensor_precision)
    a.per_cu = requested_per_cu;
    a.twiddles = g_dev[dev].twiddles;
    return launch_tensor_precision(n, a);
  }
  return reg->v[pick].launch(a);
}

// Times every variant of length n and data type on zero-filled scratch
// buffers (FFT timing does not depend on the data), at its occupancy and at
// each smaller workgroups-per-CU cap, and makes the fastest (variant, cap)
// of each family the choice of that family's entry point.
constexpr int kMaxSweptPerCu = 16;
constexpr double kCapMargin = 0.01;
int autotune(int n, bool c64, int64_t batch, int reps, void *stream_p) {
  SizeRegistry *reg = registry(n, c64);
  if (!reg) return FFTOC_ERROR_INVALID_SIZE;
  if (batch <= 0 || reps <= 0) return FFTOC_ERROR_INVALID_BATCH;
  int dev = 0;
  int rc = init_current_device(&dev);
  if (rc != FFTOC_SUCCESS) return rc;
  hipStream_t stream = static_cast<hipStream_t>(stream_p);
  const size_t bytes = (size_t)batch * (size_t)n * (c64 ? 8u : 16u);
  void *in = nullptr, *out = nullptr;
  if (hipMalloc(&in, bytes) != hipSuccess) return FFTOC_ERROR_ALLOC;
  if (hipMalloc(&out, bytes) != hipSuccess) {
    (void)hipFree(in);
    return FFTOC_ERROR_ALLOC;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = FFTOC_SUCCESS;
  if (hipMemsetAsync(in, 0, bytes, stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = FFTOC_ERROR_DEVICE;
  std::vector<double> us(reg->v.size(), -1.0);
  std::vector<int> caps(reg->v.size(), 0);
  for (size_t i = 0; i < reg->v.size() && rc == FFTOC_SUCCESS; ++i) {
    LaunchArgs a{in,     out,   batch, c64 ? g_dev[dev].twiddles32 : g_dev[dev].twiddles,
                 g_dev[dev].twiddles32, stream, false, 1.0, 0, FFTOC_PREC_FP64, c64, 0, nullptr};
    // workgroups per CU: the kernel's occupancy (0), then every smaller cap
    // up to kMaxSweptPerCu (the reference's blocks-per-SM sweep)
    int occ = 0;
    a.occupancy = &occ;
    if (reg->v[i].launch(a) != FFTOC_SUCCESS) occ = 0;
    a.occupancy = nullptr;
    const int last = std::min(occ - 1, kMaxSweptPerCu);
    for (int cap = 0; cap <= last || cap == 0; ++cap) {
      a.per_cu = cap;
      if (reg->v[i].launch(a) != FFTOC_SUCCESS) {  // warm-up
        if (cap == 0) break;
        continue;
      }
      std::vector<float> t;
      for (int r = 0; r < reps; ++r) {
        float ms = 0.f;
        if (hipEventRecord(e0, stream) != hipSuccess || reg->v[i].launch(a) != FFTOC_SUCCESS ||
            hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = FFTOC_ERROR_DEVICE;
          break;
        }
        t.push_back(ms);
      }
      if (rc != FFTOC_SUCCESS || t.empty()) break;
      std::sort(t.begin(), t.end());
      const double m = 1000.0 * t[t.size() / 2];
      // a cap must win by more than the timing noise to replace the occupancy
      if (us[i] < 0 || m < us[i] * (1.0 - kCapMargin)) {
        us[i] = m;
        caps[i] = cap;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(stream);
  (void)hipFree(in);
  (void)hipFree(out);
  if (rc != FFTOC_SUCCESS) return rc;
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    reg->tuned_us = us;
    reg->tuned_per_cu = caps;
  }
  const Family fams[4] = {Family::Tensor, Family::Radix8, Family::Radix16, Family::Best};
  for (Family f : fams) {
    int best = -1;
    for (size_t i = 0; i < us.size(); ++i)
      if ((reg->v[i].families & family_bit(f)) && us[i] >= 0 && (best < 0 || us[i] < us[best]))
        best = (int)i;
    if (best >= 0) {
      // a call running concurrently may pair the old choice with the new
      // cap (or the reverse): that changes its speed, never its result
      reg->per_cu[family_index(f)][batch_bucket(batch)].store(caps[best]);
      reg->choice[family_index(f)][batch_bucket(batch)].store(best);
    }
  }
  return FFTOC_SUCCESS;
}

const char *kFamilyNames[4] = {"tensor", "legacy8", "legacy16", "fft"};

}  // namespace

std::atomic<int> g_table_holders{0};
void hold_tables() { g_table_holders.fetch_add(1); }
void drop_tables() { g_table_holders.fetch_sub(1); }

int device_tables(const void **tw64, const void **tw32) {
  int dev = 0;
  const int rc = init_current_device(&dev);
  if (rc != FFTOC_SUCCESS) return rc;
  *tw64 = g_dev[dev].twiddles;
  *tw32 = g_dev[dev].twiddles32;
  return FFTOC_SUCCESS;
}

}  // namespace fftoc

using fftoc::Family;

extern "C" {

int fftoc_init(void) {
  int dev = 0;
  return fftoc::init_current_device(&dev);
}

int fftoc_twiddle_table(const void **fp64, const void **fp32) {
  if (!fp64 && !fp32) return FFTOC_ERROR_NULL_POINTER;
  const void *a = nullptr, *b = nullptr;
  const int rc = fftoc::device_tables(&a, &b);
  if (rc != FFTOC_SUCCESS) return rc;
  if (fp64) *fp64 = a;
  if (fp32) *fp32 = b;
  return FFTOC_SUCCESS;
}

int fftoc_shutdown(void) {
  std::lock_guard<std::mutex> lock(fftoc::g_mutex);
  // a plan's or server's kernels read the tables
  if (fftoc::g_table_holders.load() > 0) return FFTOC_ERROR_UNSUPPORTED;
  // the tables are module globals that live as long as the library is
  // loaded; a device is set up again (tables copied in) on its next use
  for (int d = 0; d < fftoc::kMaxDevices; ++d) fftoc::g_dev[d] = fftoc::DeviceState{};
  return FFTOC_SUCCESS;
}

const char *fftoc_strerror(int status) {
  switch (status) {
    case FFTOC_SUCCESS: return "success";
    case FFTOC_ERROR_INVALID_SIZE: return "transform length not supported by this entry point";
    case FFTOC_ERROR_INVALID_BATCH: return "invalid batch count";
    case FFTOC_ERROR_NULL_POINTER: return "null buffer pointer";
    case FFTOC_ERROR_INVALID_FLAGS: return "invalid flags";
    case FFTOC_ERROR_OVERLAP: return "input and output buffers partially overlap";
    case FFTOC_ERROR_UNSUPPORTED: return "operation not supported by this path";
    case FFTOC_ERROR_DEVICE: return "HIP device error (is this a gfx950 GPU?)";
    case FFTOC_ERROR_LAUNCH: return "kernel launch failed";
    case FFTOC_ERROR_ALLOC: return "memory allocation failed";
    case FFTOC_ERROR_MISALIGNED: return "buffers must be 16-byte aligned";
    case FFTOC_ERROR_BUDGET: return "request exceeds the memory budget";
    case FFTOC_ERROR_QUEUE_FULL: return "no free request slot";
    case FFTOC_ERROR_INVALID_HANDLE: return "invalid handle";
    default: return "unknown status";
  }
}

int fftoc_version(void) { return 200; }

int fftoc_tensor_fft(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                     void *stream) {
  return fftoc::run(Family::Tensor, in, out, batch, n, flags, stream);
}

#define FFTOC_TENSOR_ALIAS(N)                                                               \
  int fftoc_tensor_fft_##N(const void *in, void *out, int64_t batch, uint32_t flags,        \
                           void *stream) {                                                  \
    return fftoc::run(Family::Tensor, in, out, batch, N, flags, stream);                    \
  }
FFTOC_TENSOR_ALIAS(8)
FFTOC_TENSOR_ALIAS(16)
FFTOC_TENSOR_ALIAS(64)
FFTOC_TENSOR_ALIAS(128)
FFTOC_TENSOR_ALIAS(256)
FFTOC_TENSOR_ALIAS(512)
FFTOC_TENSOR_ALIAS(4096)
#undef FFTOC_TENSOR_ALIAS

int fftoc_legacy8_fft(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                      void *stream) {
  return fftoc::run(Family::Radix8, in, out, batch, n, flags, stream);
}

int fftoc_legacy16_fft(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                       void *stream) {
  return fftoc::run(Family::Radix16, in, out, batch, n, flags, stream);
}

int fftoc_fft(const void *in, void *out, int64_t batch, int n, uint32_t flags, void *stream) {
  return fftoc::run(Family::Best, in, out, batch, n, flags, stream);
}

namespace {
bool dtype_c64(uint32_t flags) { return (flags & FFTOC_DATA_C64) != 0; }
const char *kDtypeNames[2] = {"c128", "c64"};
}  // namespace

int fftoc_variant_count(int n, uint32_t flags) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  return r ? (int)r->v.size() : 0;
}

const char *fftoc_variant_name(int n, uint32_t flags, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r || variant < 0 || variant >= (int)r->v.size()) return nullptr;
  return r->v[variant].name;
}

int fftoc_fft_variant(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                      int variant, void *stream) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r) return FFTOC_ERROR_INVALID_SIZE;
  if (variant < 0 || variant >= (int)r->v.size()) return FFTOC_ERROR_INVALID_HANDLE;
  const Family fam = (r->v[variant].families & fftoc::FAM_TENSOR) ? Family::Tensor : Family::Best;
  return fftoc::run(fam, in, out, batch, n, flags, stream, variant);
}

int fftoc_get_variant(int n, uint32_t flags, int64_t batch) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  return r ? r->choice[fftoc::family_index(Family::Best)][fftoc::batch_bucket(batch)].load() : -1;
}

int fftoc_set_variant(int n, uint32_t flags, int64_t batch, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r) return FFTOC_ERROR_INVALID_SIZE;
  if (variant < 0 || variant >= (int)r->v.size()) return FFTOC_ERROR_INVALID_HANDLE;
  for (int b = 0; b < fftoc::kBuckets; ++b)
    if (batch < 0 || b == fftoc::batch_bucket(batch)) {
      r->choice[fftoc::family_index(Family::Best)][b].store(variant);
      r->per_cu[fftoc::family_index(Family::Best)][b].store(0);
    }
  return FFTOC_SUCCESS;
}

int fftoc_get_workgroups_per_cu(int n, uint32_t flags, int64_t batch) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  return r ? r->per_cu[fftoc::family_index(Family::Best)][fftoc::batch_bucket(batch)].load() : -1;
}

int fftoc_set_workgroups_per_cu(int n, uint32_t flags, int64_t batch, int per_cu) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r) return FFTOC_ERROR_INVALID_SIZE;
  if (per_cu < 0 || per_cu > fftoc::kMaxPerCu) return FFTOC_ERROR_INVALID_FLAGS;
  for (int b = 0; b < fftoc::kBuckets; ++b)
    if (batch < 0 || b == fftoc::batch_bucket(batch))
      r->per_cu[fftoc::family_index(Family::Best)][b].store(per_cu);
  return FFTOC_SUCCESS;
}

int fftoc_variant_occupancy(int n, uint32_t flags, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r || variant < 0 || variant >= (int)r->v.size()) return -1;
  int dev = 0;
  if (fftoc::init_current_device(&dev) != FFTOC_SUCCESS) return -1;
  int occ = 0;
  fftoc::LaunchArgs a{};
  a.precision = FFTOC_PREC_FP64;
  a.c64 = dtype_c64(flags);
  a.scale = 1.0;
  a.occupancy = &occ;
  return r->v[variant].launch(a) == FFTOC_SUCCESS && occ > 0 ? occ : -1;
}

int fftoc_variant_per_cu(int n, uint32_t flags, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r || variant < 0 || variant >= (int)r->v.size()) return -1;
  std::lock_guard<std::mutex> lock(fftoc::g_tune_mutex);
  return r->tuned_us[variant] < 0 ? -1 : r->tuned_per_cu[variant];
}

int fftoc_autotune(int n, uint32_t flags, int64_t batch, int reps, void *stream) {
  return fftoc::autotune(n, dtype_c64(flags), batch, reps, stream);
}

double fftoc_variant_time_us(int n, uint32_t flags, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r || variant < 0 || variant >= (int)r->v.size()) return -1.0;
  std::lock_guard<std::mutex> lock(fftoc::g_tune_mutex);
  return r->tuned_us[variant];
}

int fftoc_variant_tile(int n, uint32_t flags, int variant) {
  fftoc::SizeRegistry *r = fftoc::registry(n, dtype_c64(flags));
  if (!r || variant < 0 || variant >= (int)r->v.size()) return -1;
  return r->v[variant].tile;
}

// Text format, one line per (entry point, length, data type, batch bucket):
//   <family> <n> <c128|c64> <floor(log2 batch)> [wgs=<k>] <variant name up to end of line>
// wgs=<k>: at most k workgroups per CU (absent: the kernel's occupancy).
int fftoc_tuning_save(const char *path) {
  if (!path) return FFTOC_ERROR_NULL_POINTER;
  FILE *f = std::fopen(path, "w");
  if (!f) return FFTOC_ERROR_INVALID_HANDLE;
  std::fprintf(f, "# fftoc tuning table: family n dtype log2(batch) [wgs=per-CU cap] variant\n");
  for (int k = 3; k <= 12; ++k) {
    for (int c64 = 0; c64 < 2; ++c64) {
      fftoc::SizeRegistry *r = fftoc::registry(1 << k, c64 != 0);
      for (int fi = 0; fi < 4; ++fi) {
        for (int b = 0; b < fftoc::kBuckets; ++b) {
          const int c = r->choice[fi][b].load();
          const int cap = r->per_cu[fi][b].load();
          if (c < 0) continue;
          if (cap > 0)
            std::fprintf(f, "%s %d %s %d wgs=%d %s\n", fftoc::kFamilyNames[fi], 1 << k,
                         kDtypeNames[c64], b, cap, r->v[c].name);
          else
            std::fprintf(f, "%s %d %s %d %s\n", fftoc::kFamilyNames[fi], 1 << k, kDtypeNames[c64],
                         b, r->v[c].name);
        }
      }
    }
  }
  const int err = std::ferror(f);
  std::fclose(f);
  return err ? FFTOC_ERROR_INVALID_HANDLE : FFTOC_SUCCESS;
}

// Lines that do not parse, or name an unknown family, length, data type or
// variant (or a variant of another family), are skipped and make the call
// return FFTOC_ERROR_INVALID_HANDLE; every valid line is applied.
int fftoc_tuning_load(const char *path) {
  if (!path) return FFTOC_ERROR_NULL_POINTER;
  FILE *f = std::fopen(path, "r");
  if (!f) return FFTOC_ERROR_INVALID_HANDLE;
  char line[512];
  int rc = FFTOC_SUCCESS;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char fam[32], dt[8];
    int n = 0, bucket = 0, used = 0;
    if (std::sscanf(line, "%31s %d %7s %d %n", fam, &n, dt, &bucket, &used) != 4 || bucket < 0 ||
        bucket >= fftoc::kBuckets) {
      rc = FFTOC_ERROR_INVALID_HANDLE;
      continue;
    }
    int c64 = -1;
    for (int i = 0; i < 2; ++i)
      if (std::strcmp(dt, kDtypeNames[i]) == 0) c64 = i;
    int cap = 0, cap_used = 0;
    if (std::strncmp(line + used, "wgs=", 4) == 0) {
      if (std::sscanf(line + used, "wgs=%d %n", &cap, &cap_used) != 1 || cap_used == 0 ||
          cap < 0 || cap > fftoc::kMaxPerCu) {
        rc = FFTOC_ERROR_INVALID_HANDLE;
        continue;
      }
      used += cap_used;
    }
    std::string name(line + used);
    while (!name.empty() && (name.back() == '\n' || name.back() == '\r')) name.pop_back();
    fftoc::SizeRegistry *r = c64 >= 0 ? fftoc::registry(n, c64 != 0) : nullptr;
    int fi = -1;
    for (int i = 0; i < 4; ++i)
      if (std::strcmp(fam, fftoc::kFamilyNames[i]) == 0) fi = i;
    if (!r || fi < 0) {
      rc = FFTOC_ERROR_INVALID_HANDLE;
      continue;
    }
    int match = -1;
    for (size_t i = 0; i < r->v.size(); ++i)
      if (name == r->v[i].name) match = (int)i;
    const unsigned need = fi == 0   ? fftoc::FAM_TENSOR
                          : fi == 1 ? fftoc::FAM_RADIX8
                          : fi == 2 ? fftoc::FAM_RADIX16
                                    : fftoc::FAM_ANY;
    if (match < 0 || !(r->v[match].families & need)) {
      rc = FFTOC_ERROR_INVALID_HANDLE;
      continue;
    }
    r->per_cu[fi][bucket].store(cap);
    r->choice[fi][bucket].store(match);
  }
  std::fclose(f);
  return rc;
}

}  // extern "C"
