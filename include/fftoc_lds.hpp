// Device-callable in-LDS FFT functors (HIP, gfx950) for fusing the transform
// into a caller's own kernel: pre-processing writes the signals into LDS,
// the functor transforms them in place, post-processing reads them back.
//
// This is the operator form of the reference's FFT functors -- construct on a
// shared-memory pointer, call operator() (fft::tensor_fft_N / legacy_fft:
// e.g. includes/tensor_fft_64.cuh:12-19,43-45, includes/legacy8_fft.cuh:
// 11-19,35, as driven by testing::fft_tester, includes/testing.cuh:39-43) --
// on the library's Stockham radix kernels.
//
//   #include "fftoc_lds.hpp"
//   using F = fftoc::LdsFft<1024, 2>;          // 2 signals of 1024 points
//   __global__ __launch_bounds__(F::kThreads) void k(..., const double2 *tw) {
//     __shared__ double2 buf[F::kElems];       // natural order, signal-major
//     ... fill buf ...; __syncthreads();
//     F(buf, tw)();                            // forward, in place
//     ... use buf ...
//   }
//
// `tw` is the library's table (fftoc_twiddle_table in fftoc.h).  The workgroup
// must have exactly kThreads threads (1-D), all of which call operator();
// it synchronises the workgroup.  Element e of signal s is buf[s * N + e].
#pragma once

#include "fftoc_stockham.hpp"

namespace fftoc {

// Radix split per length: those of the tuned one-tile library kernels.
template <int N> struct LdsSplit;
template <> struct LdsSplit<8> { using type = Split<4, 2>; };
template <> struct LdsSplit<16> { using type = Split<4, 4>; };
template <> struct LdsSplit<32> { using type = Split<8, 4>; };
template <> struct LdsSplit<64> { using type = Split<8, 8>; };
template <> struct LdsSplit<128> { using type = Split<16, 8>; };
template <> struct LdsSplit<256> { using type = Split<16, 16>; };
template <> struct LdsSplit<512> { using type = Split<8, 8, 8>; };
template <> struct LdsSplit<1024> { using type = Split<8, 8, 16>; };
template <> struct LdsSplit<2048> { using type = Split<8, 16, 16>; };
template <> struct LdsSplit<4096> { using type = Split<16, 16, 16>; };

// SIGNALS transforms of N points (CT = double2, FP64; or float2, FP32
// arithmetic with the float2 table) held in LDS.
template <int N, int SIGNALS = 1, class CT = double2>
struct LdsFft {
  using Kernel = StockhamKernel<typename LdsSplit<N>::type, SIGNALS, 2, true, CT, false>;
  static constexpr int kThreads = Kernel::THREADS;  // workgroup size
  static constexpr int kElems = N * SIGNALS;       // LDS elements used

  CT *buf;
  const CT *tw;
  __device__ LdsFft(CT *smem, const CT *twiddles) : buf(smem), tw(twiddles) {}

  // Forward (inverse = false) or inverse DFT, each output multiplied by
  // `scale` (1/N gives the normalised inverse).
  __device__ void operator()(bool inverse = false, real_t<CT> scale = 1) const {
    Kernel::run_lds(buf, tw, scale, inverse);
  }
};

}  // namespace fftoc
