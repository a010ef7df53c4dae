// Example of fusing the transform into a caller's kernel with the in-LDS
// functors of include/fftoc_lds.hpp (the reference's functor-on-shared-memory
// form, includes/testing.cuh:39-43): each workgroup loads its signals into LDS
// applying an optional Hann window (pre-processing), transforms them in place
// with fftoc::LdsFft (Stockham radix passes) or fftoc::LdsTensorFft (matrix
// cores) and stores them.  A second entry point is the reference's own timing
// harness (includes/testing.cuh:16-57, 115-124): load once, run the functor R
// times on the LDS-resident signals, store once, so that the difference of two
// repeat counts is the on-chip transform time with HBM excluded.  Built as
// libfftoc_fused_example.so; the GPU tests check it against the oracle.
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

// Tensor functor per length: units of one wave (N <= 256) with several
// signals each, or one four-wave unit (N = 4096); 256 threads per workgroup.
// C: the contraction (fftoc::LdsContraction).  N = 8 has the Karatsuba
// contraction only; the one-matrix-stage form exists where a 256-point core
// runs (N = 256, 4096).
template <int N, int C> struct TensorCfg;
template <int C> struct TensorCfg<8, C> { using F = fftoc::LdsTensorFft<8, 4, 64, fftoc::kKaratsuba>; };
template <int C> struct TensorCfg<16, C> { using F = fftoc::LdsTensorFft<16, 4, 16, C>; };
template <int C> struct TensorCfg<64, C> { using F = fftoc::LdsTensorFft<64, 4, 8, C>; };
template <int C> struct TensorCfg<128, C> { using F = fftoc::LdsTensorFft<128, 4, 4, C>; };
template <int C> struct TensorCfg<256, C> { using F = fftoc::LdsTensorFft<256, 4, 2, C>; };
template <int C> struct TensorCfg<512, C> { using F = fftoc::LdsTensorFft<512, 4, 1, C>; };
template <int C> struct TensorCfg<4096, C> { using F = fftoc::LdsTensorFft<4096, 1, 1, C>; };
// One-wave shapes for batches of a few signals per CU (the reference's
// saturation sweep at small k): the fewest signals one unit takes, so that
// k FFTs per CU still spread over every CU.
template <int N, int C> struct TensorCfgSmall;
template <int C> struct TensorCfgSmall<8, C> { using F = fftoc::LdsTensorFft<8, 1, 32, fftoc::kKaratsuba>; };
template <int C> struct TensorCfgSmall<16, C> { using F = fftoc::LdsTensorFft<16, 1, 16, C>; };
template <int C> struct TensorCfgSmall<64, C> { using F = fftoc::LdsTensorFft<64, 1, 4, C>; };
template <int C> struct TensorCfgSmall<128, C> { using F = fftoc::LdsTensorFft<128, 1, 2, C>; };
template <int C> struct TensorCfgSmall<256, C> { using F = fftoc::LdsTensorFft<256, 1, 1, C>; };
template <int C> struct TensorCfgSmall<512, C> { using F = fftoc::LdsTensorFft<512, 1, 1, C>; };
template <int C> struct TensorCfgSmall<4096, C> { using F = fftoc::LdsTensorFft<4096, 1, 1, C>; };
template <int N>
using FftSmall = LdsFft<N, (kSignals<N> < 4 ? kSignals<N> : 4)>;

template <int N, int C>
constexpr bool kHasTensor =
    (N == 16 || N == 64 || N == 128 || N == 256 || N == 512 || N == 4096 || N == 8) &&
    (C == fftoc::kKaratsuba || (C == fftoc::kTwoProduct && N != 8) ||
     (C == fftoc::kKaratsubaOneStage && (N == 256 || N == 4096)));

// Threads and signals per workgroup of either functor family.
template <class F> struct Shape {
  static constexpr int threads = F::max_threads_per_block;
};
template <int N, int S, class CT> struct Shape<LdsFft<N, S, CT>> {
  static constexpr int threads = LdsFft<N, S, CT>::kThreads;
};

// Loads the workgroup's signals (s0 on) into LDS, optionally windowed.
template <int N, int ELEMS, int THREADS>
__device__ __forceinline__ void load_block(const double2 *in, double2 *buf, int64_t s0,
                                           int64_t batch, int window) {
  for (int e = threadIdx.x; e < ELEMS; e += THREADS) {
    const int64_t s = s0 + e / N;
    double2 v = s < batch ? in[s0 * N + e] : make_double2(0.0, 0.0);
    if (window) {
      const double w = 0.5 - 0.5 * cospi(2.0 * (double)(e % N) / (double)N);
      v.x *= w;
      v.y *= w;
    }
    buf[e] = v;
  }
}
template <int N, int ELEMS, int THREADS>
__device__ __forceinline__ void store_block(double2 *out, const double2 *buf, int64_t s0,
                                            int64_t batch) {
  for (int e = threadIdx.x; e < ELEMS; e += THREADS) {
    const int64_t s = s0 + e / N;
    if (s < batch) out[s0 * N + e] = buf[e];
  }
}

template <int N, class F>
__global__ __launch_bounds__(Shape<F>::threads) void windowed_tensor_kernel(
    const double2 *in, double2 *out, int64_t batch, const double2 *__restrict__ tw, int inverse,
    double scale, int window) {
  constexpr int S = F::kElems / N;
  __shared__ double2 buf[F::kElems];
  const int64_t s0 = (int64_t)blockIdx.x * S;
  load_block<N, F::kElems, Shape<F>::threads>(in, buf, s0, batch, window);
  __syncthreads();
  F f(buf, tw);
  f(inverse != 0, scale);
  __syncthreads();
  store_block<N, F::kElems, Shape<F>::threads>(out, buf, s0, batch);
}

template <int N, int C>
int launch_tensor(const void *in, void *out, int64_t batch, const void *tw, int inverse,
                  double scale, int window, hipStream_t stream) {
  using F = typename TensorCfg<N, C>::F;
  constexpr int S = F::kElems / N;
  const int64_t grid = (batch + S - 1) / S;
  if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  hipLaunchKernelGGL((windowed_tensor_kernel<N, F>), dim3((unsigned)grid), dim3(Shape<F>::threads),
                     0, stream, static_cast<const double2 *>(in), static_cast<double2 *>(out),
                     batch, static_cast<const double2 *>(tw), inverse, scale, window);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

template <int N>
int tensor(int contraction, const void *in, void *out, int64_t batch, const void *tw, int inverse,
           double scale, int window, hipStream_t s) {
  if constexpr (kHasTensor<N, fftoc::kKaratsuba>)
    if (contraction == fftoc::kKaratsuba)
      return launch_tensor<N, fftoc::kKaratsuba>(in, out, batch, tw, inverse, scale, window, s);
  if constexpr (kHasTensor<N, fftoc::kTwoProduct>)
    if (contraction == fftoc::kTwoProduct)
      return launch_tensor<N, fftoc::kTwoProduct>(in, out, batch, tw, inverse, scale, window, s);
  if constexpr (kHasTensor<N, fftoc::kKaratsubaOneStage>)
    if (contraction == fftoc::kKaratsubaOneStage)
      return launch_tensor<N, fftoc::kKaratsubaOneStage>(in, out, batch, tw, inverse, scale,
                                                         window, s);
  return FFTOC_ERROR_UNSUPPORTED;
}

// The reference's timing harness with either functor family: one load, R
// forward transforms of the LDS-resident signals (each output times `scale`,
// 1/sqrt(N) keeping the values' magnitude), one store, in place.
template <int N, class F>
__global__ __launch_bounds__(Shape<F>::threads) void onchip_kernel(double2 *data, int64_t batch,
                                                                   const double2 *__restrict__ tw,
                                                                   int repeats, double scale) {
  constexpr int S = F::kElems / N;
  __shared__ double2 buf[F::kElems];
  const int64_t s0 = (int64_t)blockIdx.x * S;
  load_block<N, F::kElems, Shape<F>::threads>(data, buf, s0, batch, 0);
  __syncthreads();
  F f(buf, tw);
  for (int r = 0; r < repeats; ++r) f(false, scale);
  __syncthreads();
  store_block<N, F::kElems, Shape<F>::threads>(data, buf, s0, batch);
}

template <int N, class F>
int launch_onchip(void *data, int64_t batch, const void *tw, int repeats, double scale,
                  hipStream_t stream) {
  constexpr int S = F::kElems / N;
  const int64_t grid = (batch + S - 1) / S;
  if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  hipLaunchKernelGGL((onchip_kernel<N, F>), dim3((unsigned)grid), dim3(Shape<F>::threads), 0,
                     stream, static_cast<double2 *>(data), batch,
                     static_cast<const double2 *>(tw), repeats, scale);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

template <int N, bool SMALL>
int onchip(int family, void *data, int64_t batch, const void *tw, int repeats, double scale,
           hipStream_t s) {
  using R = std::conditional_t<SMALL, FftSmall<N>, Fft<N>>;
  if (family == 0) return launch_onchip<N, R>(data, batch, tw, repeats, scale, s);
  if constexpr (kHasTensor<N, fftoc::kKaratsuba>) {
    using T = std::conditional_t<SMALL, TensorCfgSmall<N, fftoc::kKaratsuba>,
                                 TensorCfg<N, fftoc::kKaratsuba>>;
    if (family == 1) return launch_onchip<N, typename T::F>(data, batch, tw, repeats, scale, s);
  }
  if constexpr (kHasTensor<N, fftoc::kTwoProduct>) {
    using T = std::conditional_t<SMALL, TensorCfgSmall<N, fftoc::kTwoProduct>,
                                 TensorCfg<N, fftoc::kTwoProduct>>;
    if (family == 2) return launch_onchip<N, typename T::F>(data, batch, tw, repeats, scale, s);
  }
  if constexpr (kHasTensor<N, fftoc::kKaratsubaOneStage>) {
    using T = std::conditional_t<SMALL, TensorCfgSmall<N, fftoc::kKaratsubaOneStage>,
                                 TensorCfg<N, fftoc::kKaratsubaOneStage>>;
    if (family == 3) return launch_onchip<N, typename T::F>(data, batch, tw, repeats, scale, s);
  }
  return FFTOC_ERROR_UNSUPPORTED;
}

template <bool SMALL>
int onchip_n(int n, int family, void *data, int64_t batch, const void *tw, int repeats,
             double scale, hipStream_t s) {
  switch (n) {
    case 8: return onchip<8, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 16: return onchip<16, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 32: return onchip<32, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 64: return onchip<64, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 128: return onchip<128, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 256: return onchip<256, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 512: return onchip<512, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 1024: return onchip<1024, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 2048: return onchip<2048, SMALL>(family, data, batch, tw, repeats, scale, s);
    case 4096: return onchip<4096, SMALL>(family, data, batch, tw, repeats, scale, s);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

}  // namespace

extern "C" {

int fftoc_example_onchip_fft_shape(void *data, int64_t batch, int n, const void *tw, int family,
                                   int shape, int repeats, double scale, void *stream);

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

// The matrix-core functors (fftoc::LdsTensorFft) in the same example: n in
// {8, 16, 64, 128, 256, 512, 4096}; contraction 0 = Karatsuba, 1 = the
// two-real-product form (not at n = 8), 2 = Karatsuba on one matrix stage of
// the 256-point core (n = 256, 4096).
int fftoc_example_windowed_tensor_fft(const void *in, void *out, int64_t batch, int n,
                                      const void *tw, int inverse, double scale, int window,
                                      int contraction, void *stream) {
  if (batch < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (batch == 0) return FFTOC_SUCCESS;
  if (!in || !out || !tw) return FFTOC_ERROR_NULL_POINTER;
  if (in == out) return FFTOC_ERROR_OVERLAP;
  if (contraction < 0 || contraction > 2) return FFTOC_ERROR_INVALID_FLAGS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (n) {
    case 8: return tensor<8>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 16: return tensor<16>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 64: return tensor<64>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 128: return tensor<128>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 256: return tensor<256>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 512: return tensor<512>(contraction, in, out, batch, tw, inverse, scale, window, s);
    case 4096: return tensor<4096>(contraction, in, out, batch, tw, inverse, scale, window, s);
    default: return FFTOC_ERROR_INVALID_SIZE;
  }
}

// On-chip timing harness: `repeats` forward transforms (outputs times `scale`)
// of the batch held in LDS between one load and one store, in place on data.
// family 0 = fftoc::LdsFft (any n), 1 = fftoc::LdsTensorFft with Karatsuba, 2 =
// with the two-product contraction, 3 = Karatsuba on one matrix stage (tensor
// n in {8, 16, 64, 128, 256, 512, 4096}, two-product not at 8, family 3 at
// n = 256 and 4096).
int fftoc_example_onchip_fft(void *data, int64_t batch, int n, const void *tw, int family,
                             int repeats, double scale, void *stream) {
  return fftoc_example_onchip_fft_shape(data, batch, n, tw, family, 0, repeats, scale, stream);
}

// The same harness with a workgroup shape: shape 0 = the shapes above (up to
// 256 threads and 64 KiB of signals per workgroup), 1 = one-wave workgroups
// holding the fewest signals a functor unit takes (radix: up to 4; tensor: 32
// at n = 8, 16 at 16, 4 at 64, 2 at 128, 1 from 256), for batches of a few
// signals per CU.
int fftoc_example_onchip_fft_shape(void *data, int64_t batch, int n, const void *tw, int family,
                                   int shape, int repeats, double scale, void *stream) {
  if (batch < 0 || repeats < 0) return FFTOC_ERROR_INVALID_BATCH;
  if (batch == 0) return FFTOC_SUCCESS;
  if (!data || !tw) return FFTOC_ERROR_NULL_POINTER;
  if (family < 0 || family > 3 || shape < 0 || shape > 1) return FFTOC_ERROR_INVALID_FLAGS;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return shape ? onchip_n<true>(n, family, data, batch, tw, repeats, scale, s)
               : onchip_n<false>(n, family, data, batch, tw, repeats, scale, s);
}

}  // extern "C"
