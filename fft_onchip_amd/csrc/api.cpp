// C ABI of libfftoc.so: argument checking, per-device FP64 twiddle tables and
// dispatch to the gfx950 kernels.  Declarations and semantics: include/fftoc.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "fftoc_internal.hpp"

namespace fftoc {
namespace {

constexpr int kMaxDevices = 64;

struct DeviceState {
  bool ready = false;
  int status = FFTOC_SUCCESS;
  void *twiddles = nullptr;  // kTwN x double2
};

std::mutex g_mutex;
DeviceState g_dev[kMaxDevices];

// W_4096^j = exp(-2 pi i j / 4096) to FP64.  The first octant (j <= 512) is
// evaluated with long-double cosl/sinl; every other entry is an exact
// symmetry of it (swap / negate), so all 4096 values are the correctly
// rounded FP64 values of long-double results.
std::vector<double> build_twiddles() {
  const int n = 4096;
  std::vector<long double> c(n / 8 + 1), s(n / 8 + 1);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int j = 0; j <= n / 8; ++j) {
    const long double ang = two_pi * (long double)j / (long double)n;
    c[j] = cosl(ang);
    s[j] = sinl(ang);
  }
  c[n / 8] = s[n / 8] = 0.707106781186547524400844362104849039L;
  std::vector<double> t(2 * n);
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
    t[2 * j] = (double)cv;
    t[2 * j + 1] = (double)(-sv);  // forward root: exp(-i angle)
  }
  return t;
}

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
  std::vector<double> tw = build_twiddles();
  void *d = nullptr;
  if (hipMalloc(&d, tw.size() * sizeof(double)) != hipSuccess) return FFTOC_ERROR_ALLOC;
  if (hipMemcpy(d, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return FFTOC_ERROR_DEVICE;
  }
  st.twiddles = d;
  st.status = FFTOC_SUCCESS;
  st.ready = true;
  return FFTOC_SUCCESS;
}

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

enum class Family { Tensor, Radix8, Radix16, Best };

int check_and_fill(const void *in, void *out, int64_t batch, int n, uint32_t flags,
                   void *stream, LaunchArgs &a) {
  const uint32_t known = FFTOC_INVERSE | FFTOC_NORMALIZE | FFTOC_PREC_MASK | FFTOC_TILE_MASK;
  if (flags & ~known) return FFTOC_ERROR_INVALID_FLAGS;
  const uint32_t prec = flags & FFTOC_PREC_MASK;
  if (prec != FFTOC_PREC_FP64 && prec != FFTOC_PREC_FP32 && prec != FFTOC_PREC_BF16X3)
    return FFTOC_ERROR_INVALID_FLAGS;
  if (!is_pow2(n) || n < 8 || n > 4096) return FFTOC_ERROR_INVALID_SIZE;
  if (batch < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (batch > 0) {
    if (!in || !out) return FFTOC_ERROR_NULL_POINTER;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u)
      return FFTOC_ERROR_MISALIGNED;
    const uint64_t bytes = (uint64_t)batch * (uint64_t)n * 16u;
    if (bytes / 16u / (uint64_t)n != (uint64_t)batch) return FFTOC_ERROR_INVALID_BATCH;
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
  return FFTOC_SUCCESS;
}

int run(Family fam, const void *in, void *out, int64_t batch, int n, uint32_t flags,
        void *stream) {
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
      if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
      break;
    case Family::Best:
      if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
      break;
  }
  if (batch == 0) return FFTOC_SUCCESS;
  int dev = 0;
  rc = init_current_device(&dev);
  if (rc != FFTOC_SUCCESS) return rc;
  a.twiddles = g_dev[dev].twiddles;
  switch (fam) {
    case Family::Radix8: return launch_radix8(n, a);
    case Family::Radix16: return launch_radix16(n, a);
    case Family::Tensor: return launch_tensor(n, a);
    case Family::Best: return launch_radix_mixed(n, a);
  }
  return FFTOC_ERROR_UNSUPPORTED;
}

}  // namespace
}  // namespace fftoc

using fftoc::Family;

extern "C" {

int fftoc_init(void) {
  int dev = 0;
  return fftoc::init_current_device(&dev);
}

int fftoc_shutdown(void) {
  std::lock_guard<std::mutex> lock(fftoc::g_mutex);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (int d = 0; d < fftoc::kMaxDevices; ++d) {
    fftoc::DeviceState &st = fftoc::g_dev[d];
    if (st.twiddles) {
      (void)hipSetDevice(d);
      (void)hipFree(st.twiddles);
    }
    st = fftoc::DeviceState{};
  }
  (void)hipSetDevice(cur);
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

int fftoc_version(void) { return 100; }

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

}  // extern "C"
