// Example of fusing the transform into a caller's kernel with the in-LDS
// functors of include/fftoc_lds.hpp (the reference's functor-on-shared-memory
// form, includes/testing.cuh:39-43): each workgroup loads S signals into LDS
// applying an optional Hann window (pre-processing), transforms them in place
// with fftoc::LdsFft and stores them.  Built as libfftoc_fused_example.so; the
// GPU tests check it against the oracle.
#include <hip/hip_runtime.h>

#include "../../include/fftoc.h"
#include "../../include/fftoc_lds.hpp"

namespace {

using fftoc::LdsFft;

// Signals per workgroup: 256 threads where the length allows, <= 64 KiB LDS.
template <int N>
constexpr int kSignals = (256 / LdsFft<N, 1>::kThreads) * N * 16 <= 65536
                             ? (256 / LdsFft<N, 1>::kThreads > 0 ? 256 / LdsFft<N, 1>::kThreads : 1)
                             : 65536 / (N * 16);

template <int N>
using Fft = LdsFft<N, kSignals<N>>;

template <int N>
__global__ __launch_bounds__(Fft<N>::kThreads) void windowed_fft_kernel(
    const double2 *in, double2 *out, int64_t batch, const double2 *__restrict__ tw, int inverse,
    double scale, int window) {
  using F = Fft<N>;
  constexpr int S = kSignals<N>;
  __shared__ double2 buf[F::kElems];
  const int64_t s0 = (int64_t)blockIdx.x * S;
  for (int e = threadIdx.x; e < F::kElems; e += F::kThreads) {
    const int64_t s = s0 + e / N;
    double2 v = s < batch ? in[s0 * N + e] : make_double2(0.0, 0.0);
    if (window) {
      const double w = 0.5 - 0.5 * cospi(2.0 * (double)(e % N) / (double)N);
      v.x *= w;
      v.y *= w;
    }
    buf[e] = v;
  }
  __syncthreads();
  F(buf, tw)(inverse != 0, scale);
  for (int e = threadIdx.x; e < F::kElems; e += F::kThreads) {
    const int64_t s = s0 + e / N;
    if (s < batch) out[s0 * N + e] = buf[e];
  }
}

template <int N>
int launch(const void *in, void *out, int64_t batch, const void *tw, int inverse, double scale,
           int window, hipStream_t stream) {
  const int64_t grid = (batch + kSignals<N> - 1) / kSignals<N>;
  if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  hipLaunchKernelGGL(windowed_fft_kernel<N>, dim3((unsigned)grid), dim3(Fft<N>::kThreads), 0,
                     stream, static_cast<const double2 *>(in), static_cast<double2 *>(out), batch,
                     static_cast<const double2 *>(tw), inverse, scale, window);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace

extern "C" {

// Windowed (window != 0: Hann, w[e] = 0.5 - 0.5 cos(2 pi e / n)) forward or
// inverse FP64 FFT of `batch` signals of n points, each output times `scale`.
// `tw` is the table from fftoc_twiddle_table.  Out of place (in != out).
int fftoc_example_windowed_fft(const void *in, void *out, int64_t batch, int n, const void *tw,
                               int inverse, double scale, int window, void *stream) {
  if (batch < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (batch == 0) return FFTOC_SUCCESS;
  if (!in || !out || !tw) return FFTOC_ERROR_NULL_POINTER;
  if (in == out) return FFTOC_ERROR_OVERLAP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (n) {
    case 8: return launch<8>(in, out, batch, tw, inverse, scale, window, s);
    case 16: return launch<16>(in, out, batch, tw, inverse, scale, window, s);
    case 32: return launch<32>(in, out, batch, tw, inverse, scale, window, s);
    case 64: return launch<64>(in, out, batch, tw, inverse, scale, window, s);
    case 128: return launch<128>(in, out, batch, tw, inverse, scale, window, s);
    case 256: return launch<256>(in, out, batch, tw, inverse, scale, window, s);
    case 512: return launch<512>(in, out, batch, tw, inverse, scale, window, s);
    case 1024: return launch<1024>(in, out, batch, tw, inverse, scale, window, s);
    case 2048: return launch<2048>(in, out, batch, tw, inverse, scale, window, s);
    case 4096: return launch<4096>(in, out, batch, tw, inverse, scale, window, s);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

}  // extern "C"
