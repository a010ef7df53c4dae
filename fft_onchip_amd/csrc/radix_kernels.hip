// Stockham autosort FFT kernels on the vector ALUs (FP64), one template for
// every radix split.  These are the gfx950 counterparts of the reference's
// register-radix kernels legacy_fft<CT,N> (radix 8, includes/legacy8_fft.cuh:
// 57-118) and legacy16_fft<CT,N> (radix 16, includes/legacy16_fft.cuh:
// 99-172), which use the same pass structure: pass p with radix R reads
// a[j + m*N/R], multiplies by W_{R*L}^{m*k} (k = j mod L, L = product of the
// earlier radices), runs a radix-R butterfly and writes
// a[(j - k)*R + k + q*L].
//
// MI355X mapping:
//  * T = N / Rmax lanes own one signal, each holding Rmax complex values in
//    VGPRs; a workgroup processes S signals (a contiguous tile of S*N*16 B).
//  * Pass 0 reads straight from HBM: for fixed m the T lanes of a signal
//    read T consecutive complex values (16 B per lane, dwordx4), so each
//    load instruction covers whole 64..256 B runs.  The last pass writes
//    out[j + q*L] with j < L, again T consecutive values per signal.  For
//    T < 4 (N = 8, 16, 32) the tile is instead staged through LDS with
//    fully contiguous 1 KiB-per-instruction copies.
//  * Streaming data uses non-temporal loads and stores.
//  * Persistent variant: a workgroup loops over tiles and issues the HBM
//    loads of its next tile right after the current tile's first pass has
//    gone to LDS.  Pipelined variant (StockhamKernel::run_pipelined_persistent):
//    the same, with each lane's base twiddles loaded once before the loop, so
//    the later passes issue no vector-memory load (the vector-memory counter
//    retires in order: a twiddle load issued after the prefetch would make
//    its pass wait for the prefetch), and the next tile's first pass runs
//    before the current tile's stores; the N = 4096 default.
//  * Between passes the data crosses lanes through LDS.  When a signal lives
//    inside one wave (T <= 64) no workgroup barrier is needed: the LDS
//    operations of a wave are executed in order.
//  * Twiddles W_4096^j come from a 64 KiB FP64 table built on the host in
//    long double (L2/L1 resident); fixed roots inside the butterflies are
//    exact-rounded constants.
#include "fftoc_internal.hpp"
#include "lane_fft.hpp"
#include "../../include/fftoc_stockham.hpp"

namespace fftoc {
namespace {

template <class SP, int S>
constexpr int kThreads = (SP::N / SP::RM) * S;

template <class SP, int S, bool INV, bool STAGED, class CT>
__global__ __launch_bounds__((kThreads<SP, S>)) void stockham_kernel(
    const CT *in, CT *out, int64_t batch, const CT *__restrict__ tw, real_t<CT> scale) {
  StockhamKernel<SP, S, INV, STAGED, CT>::run(in, out, batch, tw, scale);
}

template <class SP, int S, bool INV, class CT>
__global__ __launch_bounds__((kThreads<SP, S>)) __attribute__((amdgpu_waves_per_eu(2)))
void stockham_persistent_kernel(
    const CT *in, CT *out, int64_t batch, const CT *__restrict__ tw, real_t<CT> scale) {
  StockhamKernel<SP, S, INV, false, CT>::run_persistent(in, out, batch, tw, scale);
}

template <class SP, int S, bool INV, class CT>
__global__ __launch_bounds__((kThreads<SP, S>)) __attribute__((amdgpu_waves_per_eu(2)))
void stockham_pipelined_kernel(
    const CT *in, CT *out, int64_t batch, const CT *__restrict__ tw, real_t<CT> scale) {
  StockhamKernel<SP, S, INV, false, CT>::run_pipelined_persistent(in, out, batch, tw, scale);
}

// N = 8 ... 64 with the butterflies across lanes (lane_fft.hpp): each wave
// transforms C * 64 / N signals, C load instructions of 64 consecutive
// values; TH threads per workgroup.
constexpr int kLaneThreads = 256;
static_assert(LaneTile<16, 4, 0, c2>::THREADS == kLaneThreads);

template <int N, int C, bool INV, class CT, int TH>
__global__ __launch_bounds__(TH) void lane_fft_kernel(
    const CT *in, CT *out, int64_t batch, const CT *__restrict__ tw, real_t<CT> scale) {
  LaneTile<N, C, INV ? 1 : 0, CT, TH>::run_tile(in, out, batch, blockIdx.x, nullptr, tw, scale);
}

template <int N, int C, class CT, int TH = kLaneThreads>
int launch_lane(const LaunchArgs &a) {
  constexpr auto K = lane_fft_kernel<N, C, false, CT, TH>;
  if (a.occupancy) {
    *a.occupancy = occupancy<K, TH>().per_cu;
    return FFTOC_SUCCESS;
  }
  constexpr int64_t per_block = (int64_t)TH * C;  // elements
  const int64_t blocks = (a.batch * N + per_block - 1) / per_block;
  if (blocks > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  const CT *in = static_cast<const CT *>(a.in);
  CT *out = static_cast<CT *>(a.out);
  const CT *tw = static_cast<const CT *>(a.twiddles);
  const real_t<CT> scale = (real_t<CT>)a.scale;
  const size_t lds = occupancy_lds<K, TH>(a.per_cu);
  if (a.inverse)
    hipLaunchKernelGGL((lane_fft_kernel<N, C, true, CT, TH>), dim3((unsigned)blocks),
                       dim3(TH), lds, a.stream, in, out, a.batch, tw, scale);
  else
    hipLaunchKernelGGL((lane_fft_kernel<N, C, false, CT, TH>), dim3((unsigned)blocks),
                       dim3(TH), lds, a.stream, in, out, a.batch, tw, scale);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

#define FFTOC_LANE(NAME, FAM, N, C, CT) \
  Variant { NAME, FAM, &launch_lane<N, C, CT>, kLaneThreads * (C) / (N) }
#define FFTOC_LANE_T(NAME, FAM, N, C, CT, TH) \
  Variant { NAME, FAM, &launch_lane<N, C, CT, TH>, (TH) * (C) / (N) }

enum class Form { OneTile, Staged, Persistent, Pipelined };

// The forward kernel of a split form (occupancy queries; the inverse
// instantiation uses the same resources).
template <class SP, int S, Form F, class CT>
constexpr auto split_kernel() {
  if constexpr (F == Form::Persistent)
    return stockham_persistent_kernel<SP, S, false, CT>;
  else if constexpr (F == Form::Pipelined)
    return stockham_pipelined_kernel<SP, S, false, CT>;
  else
    return stockham_kernel<SP, S, false, F == Form::Staged, CT>;
}

template <class SP, int S, Form F, class CT = c2>
int launch_split(const LaunchArgs &a) {
  constexpr int threads = kThreads<SP, S>;
  const int64_t ntiles = (a.batch + S - 1) / S;
  if (ntiles > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  const CT *in = static_cast<const CT *>(a.in);
  CT *out = static_cast<CT *>(a.out);
  const CT *tw = static_cast<const CT *>(a.twiddles);
  const real_t<CT> scale = (real_t<CT>)a.scale;
  if (a.occupancy) {
    *a.occupancy = occupancy<split_kernel<SP, S, F, CT>(), threads>().per_cu;
    return FFTOC_SUCCESS;
  }
  if constexpr (F == Form::Persistent) {
    const int64_t resident =
        resident_workgroups<stockham_persistent_kernel<SP, S, false, CT>, threads>(a.per_cu);
    const unsigned grid = (unsigned)(ntiles < resident ? ntiles : resident);
    if (a.inverse)
      hipLaunchKernelGGL((stockham_persistent_kernel<SP, S, true, CT>), dim3(grid), dim3(threads), 0,
                         a.stream, in, out, a.batch, tw, scale);
    else
      hipLaunchKernelGGL((stockham_persistent_kernel<SP, S, false, CT>), dim3(grid), dim3(threads), 0,
                         a.stream, in, out, a.batch, tw, scale);
  } else if constexpr (F == Form::Pipelined) {
    const int64_t resident =
        resident_workgroups<stockham_pipelined_kernel<SP, S, false, CT>, threads>(a.per_cu);
    const unsigned grid = (unsigned)(ntiles < resident ? ntiles : resident);
    if (a.inverse)
      hipLaunchKernelGGL((stockham_pipelined_kernel<SP, S, true, CT>), dim3(grid), dim3(threads), 0,
                         a.stream, in, out, a.batch, tw, scale);
    else
      hipLaunchKernelGGL((stockham_pipelined_kernel<SP, S, false, CT>), dim3(grid), dim3(threads), 0,
                         a.stream, in, out, a.batch, tw, scale);
  } else {
    constexpr bool staged = (F == Form::Staged);
    const size_t lds = occupancy_lds<split_kernel<SP, S, F, CT>(), threads>(a.per_cu);
    if (a.inverse)
      hipLaunchKernelGGL((stockham_kernel<SP, S, true, staged, CT>), dim3((unsigned)ntiles),
                         dim3(threads), lds, a.stream, in, out, a.batch, tw, scale);
    else
      hipLaunchKernelGGL((stockham_kernel<SP, S, false, staged, CT>), dim3((unsigned)ntiles),
                         dim3(threads), lds, a.stream, in, out, a.batch, tw, scale);
  }
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

template <class SP, int S, Form F, class CT = c2>
constexpr int tile_of() {
  return S;
}

#define FFTOC_V(NAME, FAM, ...) \
  Variant { NAME, FAM, &launch_split<__VA_ARGS__>, tile_of<__VA_ARGS__>() }

constexpr unsigned R8 = FAM_RADIX8 | FAM_ANY;
constexpr unsigned R16 = FAM_RADIX16 | FAM_ANY;
constexpr unsigned MIX = FAM_ANY;

const Variant kV8[] = {
    FFTOC_LANE("radix:8/lanes/S128", R8, 8, 4, c2),
    FFTOC_LANE("radix:8/lanes/S256", R8, 8, 8, c2),
    FFTOC_V("radix:8/staged/S128", R8, Split<8>, 128, Form::Staged),
    FFTOC_V("radix:8/staged/S256", R8, Split<8>, 256, Form::Staged),
    FFTOC_V("radix:4x2/staged/S128", MIX, Split<4, 2>, 128, Form::Staged),
    FFTOC_LANE_T("radix:8/lanes/S8/64 threads", R8, 8, 1, c2, 64),
    FFTOC_LANE_T("radix:8/lanes/S16/64 threads", R8, 8, 2, c2, 64),
};
const Variant kV16[] = {
    FFTOC_LANE("radix:16/lanes/S64", R16, 16, 4, c2),
    FFTOC_LANE("radix:16/lanes/S128", R16, 16, 8, c2),
    FFTOC_V("radix:16/staged/S64", R16, Split<16>, 64, Form::Staged),
    FFTOC_V("radix:16/staged/S128", R16, Split<16>, 128, Form::Staged),
    FFTOC_V("radix:4x4/staged/S64", MIX, Split<4, 4>, 64, Form::Staged),
    FFTOC_LANE_T("radix:16/lanes/S4/64 threads", R16, 16, 1, c2, 64),
    FFTOC_LANE_T("radix:16/lanes/S8/64 threads", R16, 16, 2, c2, 64),
};
const Variant kV32[] = {
    FFTOC_V("radix:8x4/staged/S32", MIX, Split<8, 4>, 32, Form::Staged),
    FFTOC_V("radix:8x4/tile/S32", MIX, Split<8, 4>, 32, Form::OneTile),
    FFTOC_V("radix:8x4/tile/S64", MIX, Split<8, 4>, 64, Form::OneTile),
    FFTOC_V("radix:4x8/staged/S32", MIX, Split<4, 8>, 32, Form::Staged),
    FFTOC_V("radix:4x4x2/staged/S64", MIX, Split<4, 4, 2>, 64, Form::Staged),
    FFTOC_LANE("radix:32/lanes/S32", MIX, 32, 4, c2),
    FFTOC_LANE("radix:32/lanes/S64", MIX, 32, 8, c2),
    FFTOC_LANE_T("radix:32/lanes/S4/64 threads", MIX, 32, 2, c2, 64),
};
const Variant kV64[] = {
    FFTOC_V("radix:8x8/tile/S16", R8, Split<8, 8>, 16, Form::OneTile),
    FFTOC_V("radix:8x8/tile/S32", R8, Split<8, 8>, 32, Form::OneTile),
    FFTOC_V("radix:8x8/persistent/S32", R8, Split<8, 8>, 32, Form::Persistent),
    FFTOC_V("radix:16x4/tile/S16", MIX, Split<16, 4>, 16, Form::OneTile),
    FFTOC_LANE("radix:64/lanes/S16", MIX, 64, 4, c2),
    FFTOC_LANE("radix:64/lanes/S32", MIX, 64, 8, c2),
    FFTOC_V("radix:8x8/pipelined/S32", R8, Split<8, 8>, 32, Form::Pipelined),
};
const Variant kV128[] = {
    FFTOC_V("radix:16x8/tile/S16", MIX, Split<16, 8>, 16, Form::OneTile),
    FFTOC_V("radix:16x8/tile/S8", MIX, Split<16, 8>, 8, Form::OneTile),
    FFTOC_V("radix:16x8/persistent/S16", MIX, Split<16, 8>, 16, Form::Persistent),
    FFTOC_V("radix:8x16/tile/S16", MIX, Split<8, 16>, 16, Form::OneTile),
    FFTOC_V("radix:8x4x4/pipelined/S16", MIX, Split<8, 4, 4>, 16, Form::Pipelined),
};
const Variant kV256[] = {
    FFTOC_V("radix:16x16/tile/S4", R16, Split<16, 16>, 4, Form::OneTile),
    FFTOC_V("radix:16x16/tile/S8", R16, Split<16, 16>, 8, Form::OneTile),
    FFTOC_V("radix:16x16/persistent/S8", R16, Split<16, 16>, 8, Form::Persistent),
    FFTOC_V("radix:8x8x4/tile/S8", MIX, Split<8, 8, 4>, 8, Form::OneTile),
    FFTOC_V("radix:8x8x4/pipelined/S8", MIX, Split<8, 8, 4>, 8, Form::Pipelined),
};
const Variant kV512[] = {
    FFTOC_V("radix:8x8x8/tile/S1", R8, Split<8, 8, 8>, 1, Form::OneTile),
    FFTOC_V("radix:8x8x8/tile/S2", R8, Split<8, 8, 8>, 2, Form::OneTile),
    FFTOC_V("radix:8x8x8/tile/S4", R8, Split<8, 8, 8>, 4, Form::OneTile),
    FFTOC_V("radix:8x8x8/persistent/S4", R8, Split<8, 8, 8>, 4, Form::Persistent),
    FFTOC_V("radix:16x16x2/tile/S4", MIX, Split<16, 16, 2>, 4, Form::OneTile),
    FFTOC_V("radix:16x16x2/persistent/S4", MIX, Split<16, 16, 2>, 4, Form::Persistent),
    FFTOC_V("radix:16x16x2/tile/S2", MIX, Split<16, 16, 2>, 2, Form::OneTile),
    FFTOC_V("radix:8x8x8/pipelined/S4", R8, Split<8, 8, 8>, 4, Form::Pipelined),
    FFTOC_V("radix:8x8x8/pipelined/S2", R8, Split<8, 8, 8>, 2, Form::Pipelined),
};
const Variant kV1024[] = {
    FFTOC_V("radix:16x16x4/persistent/S2", MIX, Split<16, 16, 4>, 2, Form::Persistent),
    FFTOC_V("radix:16x16x4/tile/S2", MIX, Split<16, 16, 4>, 2, Form::OneTile),
    FFTOC_V("radix:16x16x4/tile/S1", MIX, Split<16, 16, 4>, 1, Form::OneTile),
    FFTOC_V("radix:8x8x16/tile/S2", MIX, Split<8, 8, 16>, 2, Form::OneTile),
    FFTOC_V("radix:4x16x16/tile/S2", MIX, Split<4, 16, 16>, 2, Form::OneTile),
    FFTOC_V("radix:8x8x8x2/pipelined/S1", MIX, Split<8, 8, 8, 2>, 1, Form::Pipelined),
};
const Variant kV2048[] = {
    FFTOC_V("radix:16x16x8/persistent/S1", MIX, Split<16, 16, 8>, 1, Form::Persistent),
    FFTOC_V("radix:16x16x8/tile/S1", MIX, Split<16, 16, 8>, 1, Form::OneTile),
    FFTOC_V("radix:16x8x16/persistent/S1", MIX, Split<16, 8, 16>, 1, Form::Persistent),
    FFTOC_V("radix:8x16x16/tile/S1", MIX, Split<8, 16, 16>, 1, Form::OneTile),
    FFTOC_V("radix:8x8x8x4/pipelined/S1", MIX, Split<8, 8, 8, 4>, 1, Form::Pipelined),
};
const Variant kV4096[] = {
    FFTOC_V("radix:16x16x16/tile/S1", R16, Split<16, 16, 16>, 1, Form::OneTile),
    FFTOC_V("radix:16x16x16/persistent/S1", R16, Split<16, 16, 16>, 1, Form::Persistent),
    FFTOC_V("radix:8x8x8x8/pipelined/S1", R8, Split<8, 8, 8, 8>, 1, Form::Pipelined),
    FFTOC_V("radix:8x8x8x8/tile/S1", R8, Split<8, 8, 8, 8>, 1, Form::OneTile),
    FFTOC_V("radix:8x8x8x8/persistent/S1", R8, Split<8, 8, 8, 8>, 1, Form::Persistent),
};

// complex64 data (FP32 arithmetic, float2 twiddles rounded from the FP64 table).
// The one-wave lane workgroups that lead for complex128 at N = 8, 16, 32 lose
// here (N = 8 at 2^22 signals: 0.59 of HBM against 0.80): one load
// instruction moves 512 B, too little work per wave.  One-wave 4x4x4 tiles
// at N = 64 (S4 / S8, one-shot or pipelined) stayed 1-10 % below the S16
// pipelined default.
#define FFTOC_V32(NAME, FAM, SPLIT_S_FORM) \
  Variant { NAME, FAM, &launch_split<FFTOC_UNPAREN SPLIT_S_FORM, f2>, \
            tile_of<FFTOC_UNPAREN SPLIT_S_FORM, f2>() }
#define FFTOC_UNPAREN(...) __VA_ARGS__
const Variant kW8[] = {FFTOC_LANE("radix:8/lanes/S128", R8, 8, 4, f2),
                       FFTOC_LANE("radix:8/lanes/S256", R8, 8, 8, f2),
                       FFTOC_V32("radix:4x2/staged/S128", MIX, (Split<4, 2>, 128, Form::Staged)),
                       FFTOC_V32("radix:8/staged/S128", R8, (Split<8>, 128, Form::Staged))};
const Variant kW16[] = {FFTOC_LANE("radix:16/lanes/S128", R16, 16, 8, f2),
                        FFTOC_LANE("radix:16/lanes/S64", R16, 16, 4, f2),
                        FFTOC_V32("radix:4x4/staged/S64", MIX, (Split<4, 4>, 64, Form::Staged)),
                        FFTOC_V32("radix:16/staged/S64", R16, (Split<16>, 64, Form::Staged))};
const Variant kW32[] = {FFTOC_LANE("radix:32/lanes/S64", MIX, 32, 8, f2),
                        FFTOC_LANE("radix:32/lanes/S32", MIX, 32, 4, f2),
                        FFTOC_V32("radix:8x4/staged/S64", MIX, (Split<8, 4>, 64, Form::Staged)),
                        FFTOC_V32("radix:8x4/staged/S32", MIX, (Split<8, 4>, 32, Form::Staged))};
// complex64: with 8-byte values the T lanes of a signal read T*8 bytes per
// instruction, so splits with a smaller largest radix (more lanes per signal)
// read longer runs: N=64 4x4x4 and N=128 8x4x4 (T = 16) stream at 5.8 TB/s
// against 4.9 / 4.3 for 8x8 / 16x8 (T = 8).  A complex64 signal is half the
// bytes of a complex128 one for the same per-signal work (twiddle fetches,
// index arithmetic, LDS exchanges), so the pipelined form -- each lane's base
// twiddles loaded once before its tile loop, the next tile's loads in flight
// during the current tile -- gains more here: +5-12 % at N = 64, 128, 512
// from 2^14 to 2^18 signals.
// Splits of radix 4 throughout (N = 256 4x4x4x4, N = 128 4x4x4x2) read
// longer runs still but pay a fourth LDS exchange, and measured 10-30 % below.
// At N = 4096 the one-signal radix-16 tile kernel stays ahead (0.74 of HBM at
// 2^15 signals) of every pipelined form (8x8x8x8 S1 0.72; S2 and radix-16
// pipelined 0.62-0.72) and of two signals per workgroup (0.72).
const Variant kW64[] = {FFTOC_V32("radix:4x4x4/tile/S16", MIX, (Split<4, 4, 4>, 16, Form::OneTile)),
                        FFTOC_V32("radix:8x8/staged/S32", R8, (Split<8, 8>, 32, Form::Staged)),
                        FFTOC_V32("radix:8x8/tile/S16", R8, (Split<8, 8>, 16, Form::OneTile)),
                        FFTOC_LANE("radix:64/lanes/S16", MIX, 64, 4, f2),
                        FFTOC_LANE("radix:64/lanes/S32", MIX, 64, 8, f2),
                        FFTOC_V32("radix:4x4x4/pipelined/S16", MIX, (Split<4, 4, 4>, 16, Form::Pipelined)),
                        FFTOC_V32("radix:4x4x4/pipelined/S32", MIX, (Split<4, 4, 4>, 32, Form::Pipelined))};
const Variant kW128[] = {FFTOC_V32("radix:8x4x4/tile/S16", MIX, (Split<8, 4, 4>, 16, Form::OneTile)),
                         FFTOC_V32("radix:16x8/tile/S8", MIX, (Split<16, 8>, 8, Form::OneTile)),
                         FFTOC_V32("radix:8x8x2/tile/S16", MIX, (Split<8, 8, 2>, 16, Form::OneTile)),
                         FFTOC_V32("radix:8x4x4/pipelined/S16", MIX, (Split<8, 4, 4>, 16, Form::Pipelined))};
const Variant kW256[] = {FFTOC_V32("radix:16x16/tile/S4", R16, (Split<16, 16>, 4, Form::OneTile))};
const Variant kW512[] = {FFTOC_V32("radix:8x8x8/tile/S2", R8, (Split<8, 8, 8>, 2, Form::OneTile)),
                         FFTOC_V32("radix:8x8x8/pipelined/S4", R8, (Split<8, 8, 8>, 4, Form::Pipelined))};
const Variant kW1024[] = {FFTOC_V32("radix:8x8x16/tile/S2", MIX, (Split<8, 8, 16>, 2, Form::OneTile)),
                          FFTOC_V32("radix:8x8x8x2/pipelined/S1", MIX, (Split<8, 8, 8, 2>, 1, Form::Pipelined))};
const Variant kW2048[] = {FFTOC_V32("radix:8x16x16/tile/S1", MIX, (Split<8, 16, 16>, 1, Form::OneTile))};
const Variant kW4096[] = {FFTOC_V32("radix:16x16x16/tile/S1", R16, (Split<16, 16, 16>, 1, Form::OneTile)),
                          FFTOC_V32("radix:8x8x8x8/tile/S1", R8, (Split<8, 8, 8, 8>, 1, Form::OneTile)),
                          FFTOC_V32("radix:8x8x8x8/pipelined/S1", R8, (Split<8, 8, 8, 8>, 1, Form::Pipelined))};
#undef FFTOC_V32
#undef FFTOC_LANE
#undef FFTOC_LANE_T
#undef FFTOC_UNPAREN
#undef FFTOC_V

template <size_t K>
int set_out(const Variant (&v)[K], const Variant **out) {
  *out = v;
  return (int)K;
}

}  // namespace

int radix_variants(int n, bool c64, const Variant **out) {
  if (c64) {
    switch (n) {
      case 8: return set_out(kW8, out);
      case 16: return set_out(kW16, out);
      case 32: return set_out(kW32, out);
      case 64: return set_out(kW64, out);
      case 128: return set_out(kW128, out);
      case 256: return set_out(kW256, out);
      case 512: return set_out(kW512, out);
      case 1024: return set_out(kW1024, out);
      case 2048: return set_out(kW2048, out);
      case 4096: return set_out(kW4096, out);
      default: *out = nullptr; return 0;
    }
  }
  switch (n) {
    case 8: return set_out(kV8, out);
    case 16: return set_out(kV16, out);
    case 32: return set_out(kV32, out);
    case 64: return set_out(kV64, out);
    case 128: return set_out(kV128, out);
    case 256: return set_out(kV256, out);
    case 512: return set_out(kV512, out);
    case 1024: return set_out(kV1024, out);
    case 2048: return set_out(kV2048, out);
    case 4096: return set_out(kV4096, out);
    default: *out = nullptr; return 0;
  }
}

}  // namespace fftoc
