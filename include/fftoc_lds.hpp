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

#include <type_traits>

#include "fftoc_mfma.hpp"
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

// ---------------------------------------------------------------------------
// FP64 matrix-core (MFMA) in-LDS functors: the tensor path in the reference's
// own functor form -- template <CT, Size, UPB, FPU> struct tensor_fft_N with
// static threads / units_per_block / ffts_per_unit / max_threads_per_block,
// constructed on the block's shared memory (its per-thread DFT-matrix and
// twiddle values set up once there), operator() transforming each unit's
// ffts_per_unit signals in place (includes/tensor_fft_64.cuh:12-77,
// includes/tensor_fft_256.cuh:10-123, driven by includes/testing.cuh:39-43).
//
//   using F = fftoc::LdsTensorFft<256, 4, 2>;   // 4 units x 2 signals of 256
//   __global__ __launch_bounds__(F::max_threads_per_block) void k(..., const double2 *tw) {
//     __shared__ double2 buf[F::kElems];        // natural order, signal-major
//     ... fill buf ...; __syncthreads();
//     F f(buf, tw);                              // loads the lane constants
//     f();                                       // forward, in place
//     __syncthreads(); ... use buf ...
//   }
//
// Unit u (u = threadIdx.x / threads, 0 <= u < units_per_block) owns the
// signals at buf + u * N * ffts_per_unit.  A unit is one wave (N = 16, 64,
// 256) or four waves (N = 4096, 64 KiB of LDS per signal).  Every thread of
// the block calls operator(); units of one wave need no barrier, the N = 4096
// unit synchronises the block.  `tw` is the library's FP64 table
// (fftoc_twiddle_table).  CONTRACTION selects the DFT16 products:
// kKaratsuba (0, Karatsuba's three real products), kTwoProduct (1, or true:
// the two-real-product form, the library's "2-product" variants) or
// kKaratsubaOneStage (2: where a 256-point core runs, N = 256 and 4096, only
// its first DFT16 on the matrix cores with Karatsuba and its second as 4 x 4
// on the vector ALUs, the library's N = 4096 default; elsewhere the same as
// kKaratsuba).  operator()(inverse, scale): forward or inverse DFT, each
// output multiplied by scale.
enum LdsContraction : int { kKaratsuba = 0, kTwoProduct = 1, kKaratsubaOneStage = 2 };
template <int N, int UNITS = 1, int FPU = 1, int CONTRACTION = kKaratsuba>
struct LdsTensorFft;

namespace lds_detail {
__device__ __forceinline__ double2 ld_dir(const double2 *p, bool inverse) {
  const double2 v = *p;
  return inverse ? make_double2(v.y, v.x) : v;
}
__device__ __forceinline__ void st_dir(double2 *p, double2 v, bool inverse, double scale) {
  v = make_double2(v.x * scale, v.y * scale);
  *p = inverse ? make_double2(v.y, v.x) : v;
}
template <int C>
using Dft16 = std::conditional_t<C == kTwoProduct, TwoProductDft16, KaratsubaDft16>;
template <int C>
using Core256 = std::conditional_t<C == kTwoProduct, TwoProductCore,
                                   std::conditional_t<C == kKaratsubaOneStage, KaratsubaHalfCore,
                                                      KaratsubaCore>>;
}  // namespace lds_detail

// N = 16: one DFT16 per matrix row, 16 signals per matrix product.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<16, UNITS, FPU, CONTRACTION> {
  static_assert(FPU % 16 == 0, "N = 16: ffts_per_unit must be a multiple of 16");
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 16 * FPU * UNITS;
  double2 *buf;
  lds_detail::Dft16<CONTRACTION> d;
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 16 * FPU * (threadIdx.x / threads)) {
    d.load(tw, threadIdx.x & 63);
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    for (int t = 0; t < FPU / 16; ++t) {
      double2 *x = buf + 256 * t;
      c2 a[4], y[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) a[c] = lds_detail::ld_dir(x + 16 * j + 4 * c + g, inverse);
      d.run(a, y);
      wave_lds_sync();  // the wave's reads of the tile precede its writes
#pragma unroll
      for (int r = 0; r < 4; ++r) lds_detail::st_dir(x + 16 * (4 * r + g) + d.col, y[r], inverse, scale);
      wave_lds_sync();
    }
  }
};

// N = 64 = 16 x 4: four signals per matrix product (rows = (signal, n2)),
// the DFT16 over n1 on the matrix cores, twiddle W_64^{n2 k1} and DFT4 over n2
// in registers.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<64, UNITS, FPU, CONTRACTION> {
  static_assert(FPU % 4 == 0, "N = 64: ffts_per_unit must be a multiple of 4");
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 64 * FPU * UNITS;
  double2 *buf;
  lds_detail::Dft16<CONTRACTION> d;
  c2 w[4];  // W_64^{n2 k1} of register r (n2 = r, k1 = the lane's frequency)
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 64 * FPU * (threadIdx.x / threads)) {
    d.load(tw, threadIdx.x & 63);
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = tw[(r * d.col) * (kTwN / 64)];
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    for (int t = 0; t < FPU / 4; ++t) {
      double2 *x = buf + 256 * t;
      // row j <-> (signal j % 4, n2 = j / 4); K = n1 = 4c + g; n = 4 n1 + n2
      c2 a[4], y[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        a[c] = lds_detail::ld_dir(x + 64 * (j & 3) + 4 * (4 * c + g) + (j >> 2), inverse);
      d.run(a, y);  // y[r]: signal g, n2 = r, k1 = d.col
#pragma unroll
      for (int r = 1; r < 4; ++r) y[r] = cmul(y[r], w[r]);
      dft4(y[0], y[1], y[2], y[3]);
      wave_lds_sync();
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2)
        lds_detail::st_dir(x + 64 * g + d.col + 16 * k2, y[k2], inverse, scale);
      wave_lds_sync();
    }
  }
};

// N = 8: 32 signals per matrix product (the A operand holds two signals per
// row, the DFT matrix is diag(F8, F8)); Karatsuba contraction only.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<8, UNITS, FPU, CONTRACTION> {
  static_assert(CONTRACTION != kTwoProduct, "N = 8: the Karatsuba contraction only");
  static_assert(FPU % 32 == 0, "N = 8: ffts_per_unit must be a multiple of 32");
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 8 * FPU * UNITS;
  double2 *buf;
  Frag f;
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 8 * FPU * (threadIdx.x / threads)) {
    f = dft8x2_frag(tw, threadIdx.x & 63);
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    for (int t = 0; t < FPU / 32; ++t) {
      double2 *x = buf + 256 * t;
      // row j: signals j (K chunks 0, 1) and 16 + j (chunks 2, 3)
      c2 a[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        a[c] = lds_detail::ld_dir(x + 8 * ((c < 2 ? 0 : 16) + j) + 4 * (c & 1) + g, inverse);
      v4d yr, yi;
      kgemm(a, f, yr, yi);
      wave_lds_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r)
        lds_detail::st_dir(x + 8 * ((j < 8 ? 0 : 16) + 4 * r + g) + (j & 7),
                           make_double2(yr[r], yi[r]), inverse, scale);
      wave_lds_sync();
    }
  }
};

// N = 128 = 16 x 8: two signals per matrix product (rows = (signal, n2)),
// the DFT16 over n1 on the matrix cores, twiddle W_128^{n2 k1}, DFT8 over n2
// as a DFT4 in registers and one radix-2 step with the lane 32 apart.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<128, UNITS, FPU, CONTRACTION> {
  static_assert(FPU % 2 == 0, "N = 128: ffts_per_unit must be even");
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 128 * FPU * UNITS;
  double2 *buf;
  lds_detail::Dft16<CONTRACTION> d;
  c2 w[4];  // W_128^{n2 k1}, n2 = 2r + g / 2
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 128 * FPU * (threadIdx.x / threads)) {
    const int lane = threadIdx.x & 63;
    d.load(tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) w[r] = tw[((2 * r + (lane >> 5)) * d.col) * (kTwN / 128)];
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15, h = g >> 1;
    for (int t = 0; t < FPU / 2; ++t) {
      double2 *x = buf + 256 * t;
      // row j <-> (signal j % 2, n2 = j / 2); K = n1 = 4c + g; n = 8 n1 + n2
      c2 a[4], y[4];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        a[c] = lds_detail::ld_dir(x + 128 * (j & 1) + 8 * (4 * c + g) + (j >> 1), inverse);
      d.run(a, y);  // y[r]: signal g % 2, n2 = 2r + h, k1 = d.col
#pragma unroll
      for (int r = 0; r < 4; ++r) y[r] = cmul(y[r], w[r]);
      // DIT over n2 = 2m + h: E/O = DFT4 over m, X[k] = E[k] + W8^k O[k],
      // X[k + 4] = E[k] - W8^k O[k]
      dft4(y[0], y[1], y[2], y[3]);
      if (h == 1) {
        y[1] = mul_w8_1(y[1]);
        y[2] = mul_mi(y[2]);
        y[3] = mul_w8_3(y[3]);
      }
      c2 p[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[r].x = __shfl_xor(y[r].x, 32);
        p[r].y = __shfl_xor(y[r].y, 32);
      }
      wave_lds_sync();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const c2 v = (h == 0) ? cadd(y[k], p[k]) : csub(p[k], y[k]);
        lds_detail::st_dir(x + 128 * (g & 1) + d.col + 16 * (4 * h + k), v, inverse, scale);
      }
      wave_lds_sync();
    }
  }
};

// N = 512 = 16 x 2 x 16: one signal per pass of the wave.  n = 32 n1 + 16 h +
// n2'; a DFT16 over n1 for h = 0, 1, twiddle W_512^{n2 k1}, a radix-2 over h
// with W_32^{n2' k2a}, a DFT16 over n2' for k2a = 0, 1; X[k1 + 16 k2a + 32 k2b].
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<512, UNITS, FPU, CONTRACTION> {
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 512 * FPU * UNITS;
  using D = lds_detail::Dft16<CONTRACTION>;
  double2 *buf;
  D d;
  c2 w1[2][4];  // W_512^{n2 k1}, n2 = 16 h + 4 r + g, k1 = d.col
  c2 w2[4];     // W_32^{n2' k2a = 1}, n2' = 4 r + g
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 512 * FPU * (threadIdx.x / threads)) {
    const int lane = threadIdx.x & 63, g = lane >> 4;
    d.load(tw, lane);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int h = 0; h < 2; ++h) w1[h][r] = tw[((16 * h + 4 * r + g) * d.col) * (kTwN / 512)];
      w2[r] = tw[(4 * r + g) * (kTwN / 32)];
    }
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    for (int t = 0; t < FPU; ++t) {
      double2 *x = buf + 512 * t;
      c2 yh[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        c2 a[4];  // row j = n2' , K = n1 = 4c + g
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = lds_detail::ld_dir(x + 32 * (4 * c + g) + 16 * h + j, inverse);
        d.run(a, yh[h]);  // register r: n2' = 4r + g, k1 = d.col
#pragma unroll
        for (int r = 0; r < 4; ++r) yh[h][r] = cmul(yh[h][r], w1[h][r]);
      }
      c2 b[2][4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        b[0][r] = cadd(yh[0][r], yh[1][r]);
        b[1][r] = cmul(csub(yh[0][r], yh[1][r]), w2[r]);
      }
      wave_lds_sync();
#pragma unroll
      for (int k2a = 0; k2a < 2; ++k2a) {
        c2 z[4];  // rows = the lanes' k1 columns of stage 1, K = n2'
        d.run(b[k2a], z);  // register r: row 4r + g (k1 = freq(4r + g)), k2b = d.col
#pragma unroll
        for (int r = 0; r < 4; ++r)
          lds_detail::st_dir(x + D::freq(4 * r + g) + 16 * k2a + 32 * d.col, z[r], inverse, scale);
      }
      wave_lds_sync();
    }
  }
};

// N = 256 = 16 x 16: the two-stage core, one signal per pass of the wave.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<256, UNITS, FPU, CONTRACTION> {
  static constexpr int threads = 64, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 256 * FPU * UNITS;
  double2 *buf;
  lds_detail::Core256<CONTRACTION> core;
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 256 * FPU * (threadIdx.x / threads)) {
    core.load(tw, threadIdx.x & 63);
  }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int lane = threadIdx.x & 63;
    for (int t = 0; t < FPU; ++t) {
      double2 *x = buf + 256 * t;
      c2 a[4], z[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) a[c] = lds_detail::ld_dir(x + 64 * c + lane, inverse);
      core.run(a, z);
      wave_lds_sync();
#pragma unroll
      for (int r = 0; r < 4; ++r) lds_detail::st_dir(x + core.out_index(r), z[r], inverse, scale);
      wave_lds_sync();
    }
  }
};

// N = 4096 = 16 x 256: stage A (VALU, thread t = column m: a DFT16 over n1 in
// registers, twiddle W_4096^{m k1}, back into the column), stage B (MFMA: the
// 256-point core of each row k1, four rows per wave, results back into the
// row in core order), then a gather into natural order.
template <int UNITS, int FPU, int CONTRACTION>
struct LdsTensorFft<4096, UNITS, FPU, CONTRACTION> {
  static constexpr int threads = 256, units_per_block = UNITS, ffts_per_unit = FPU;
  static constexpr int max_threads_per_block = UNITS * threads;
  static constexpr int kElems = 4096 * FPU * UNITS;
  double2 *buf;
  lds_detail::Core256<CONTRACTION> core;
  TwBase<16, double2> ta;
  __device__ LdsTensorFft(double2 *smem, const double2 *tw)
      : buf(smem + 4096 * FPU * (threadIdx.x / threads)) {
    core.load(tw, threadIdx.x & 63);
    ta.fetch(tw, threadIdx.x % threads);
  }
  // position of X[k1 + 16 k2] inside row k1 (XOR of the low four bits: the
  // row's writes and the gather's reads spread over the LDS banks)
  __device__ static int rowpos(int k1, int k2) { return k1 * 256 + (k2 ^ (((k2 >> 4) ^ k1) & 15)); }
  __device__ void operator()(bool inverse = false, double scale = 1.0) const {
    const int t = threadIdx.x % threads, lane = t & 63, w = t >> 6;
    for (int sgl = 0; sgl < FPU; ++sgl) {
      double2 *x = buf + 4096 * sgl;
      {
        c2 v[16];
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) v[n1] = lds_detail::ld_dir(x + 256 * n1 + t, inverse);
        dft<16>(v);
        apply_twiddle_base<16>(v, ta);
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) x[256 * k1 + t] = v[k1];
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k1 = 4 * w + q;
        c2 a[4], z[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = x[256 * k1 + 64 * c + lane];
        core.run(a, z);
        wave_lds_sync();  // the wave alone reads and writes row k1
#pragma unroll
        for (int r = 0; r < 4; ++r) x[rowpos(k1, core.out_index(r))] = z[r];
      }
      __syncthreads();
      c2 o[16];
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        const int e = 256 * c + t;
        o[c] = x[rowpos(e & 15, e >> 4)];
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 16; ++c) lds_detail::st_dir(x + 256 * c + t, o[c], inverse, scale);
      __syncthreads();
    }
  }
};

}  // namespace fftoc
