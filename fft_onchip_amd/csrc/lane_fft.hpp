// Radix-8 / radix-16 butterflies across lanes with DPP (gfx950 data-parallel
// primitives) and the gfx950 lane-swap instructions, for the shortest
// lengths, N = 8 ... 64, where one signal is 128 B ... 1 KiB.
//
// Layout: a wave reads 64 consecutive complex values per load instruction
// (1 KiB of FP64 data, fully contiguous), so lane l holds element l mod N of
// signal l / N of that instruction's slice.  The signal's N-point DFT is
// computed in place across its N lanes as log2(N) radix-2
// decimation-in-frequency steps.  Step D (D = N/2, ..., 1) pairs lane p with
// lane p ^ D of the group:
//   lower lane (p & D == 0):  y = x_p + x_{p+D}
//   upper lane:               y = (x_{p-D} - x_p) * W_{2D}^{p mod D}
// The partner's value arrives by DPP for D = 1, 2 (quad_perm inside 4-lane
// quads) and D = 4, 8 (row_ror inside 16-lane rows), and by
// v_permlane16_swap / v_permlane32_swap for D = 16 / 32 (across rows / wave
// halves).  After the last step lane p of the group holds X[bitrev(p)]; the
// store writes it straight to that index, so each store instruction again
// covers the same contiguous 1 KiB (a permutation inside every N-lane
// group).  No LDS is used, so occupancy is bounded by VGPRs only.
//
// This is a radix-N pass of the reference's legacy kernels
// (includes/legacy16_fft.cuh:99-172 with N = 16 a single pass; radix 8 in
// includes/legacy8_fft.cuh:57-118), with the butterfly's data exchange done
// by lane shuffles instead of per-thread registers.
#pragma once

#include "../../include/fftoc_device.hpp"

namespace fftoc {

// Value of lane l ^ 16 (ROWS = true) or l ^ 32 (ROWS = false) of a 32-bit
// register.  permlane{16,32}_swap(x, x) returns (x with its odd rows / upper
// half replaced by the even rows / lower half, x with its even rows / lower
// half replaced by the odd rows / upper half): the partner is the second
// result on even rows / the lower half and the first one elsewhere.
template <bool ROWS>
__device__ __forceinline__ unsigned swap_u32(unsigned v, bool high) {
  if constexpr (ROWS) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return high ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return high ? r[0] : r[1];
  }
}
template <bool ROWS>
__device__ __forceinline__ float swap_r(float v, bool high) {
  return __builtin_bit_cast(float, swap_u32<ROWS>(__builtin_bit_cast(unsigned, v), high));
}
template <bool ROWS>
__device__ __forceinline__ double swap_r(double v, bool high) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = swap_u32<ROWS>((unsigned)u, high);
  const unsigned hi = swap_u32<ROWS>((unsigned)(u >> 32), high);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <bool ROWS, class CT>
__device__ __forceinline__ CT swap_c(CT v, bool high) {
  return mk<CT>(swap_r<ROWS>(v.x, high), swap_r<ROWS>(v.y, high));
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

template <int N>
__device__ __forceinline__ int bit_reverse(int p) {
  int r = 0;
#pragma unroll
  for (int b = 0; b < ilog2(N); ++b) r |= ((p >> b) & 1) << (ilog2(N) - 1 - b);
  return r;
}

// N-point DFT of values spread one per lane over aligned groups of N lanes
// (N = 8, 16, 32 or 64); every lane of the wave must call it (the exchanges
// read other lanes).
template <int N, class CT>
struct LaneFft {
  static_assert(N == 8 || N == 16 || N == 32 || N == 64, "N-lane groups inside one wave");
  using R = real_t<CT>;
  static constexpr int kSteps = ilog2(N);

  R sgn[kSteps];      // +1 on lower lanes, -1 on upper lanes of each step
  CT w[kSteps - 1];   // step twiddle (1 on lower lanes); the D = 1 step has none
  bool upper4;        // lane takes the D = 4 partner from l - 4 (else l + 4)
  bool odd_row;       // lane bit 4: the D = 16 partner is in the row below
  bool high_half;     // lane bit 5: the D = 32 partner is in the lower half

  // tw: W_4096^j, j = 0..4095; lane: lane index in the wave.
  __device__ LaneFft(const CT *__restrict__ tw, int lane) {
    const int p = lane & (N - 1);
#pragma unroll
    for (int s = 0; s < kSteps; ++s) {
      const int d = (N / 2) >> s;
      const bool up = (p & d) != 0;
      sgn[s] = up ? (R)-1 : (R)1;
      // lower lanes read tw[0] = 1 exactly
      if (s < kSteps - 1) w[s] = tw[up ? (p & (d - 1)) * (kTwN / (2 * d)) : 0];
    }
    upper4 = (p & 4) != 0;
    odd_row = (lane & 16) != 0;
    high_half = (lane & 32) != 0;
  }

  template <int D>
  __device__ __forceinline__ CT partner(CT x) const {
    if constexpr (D == 32) {
      return swap_c<false>(x, high_half);
    } else if constexpr (D == 16) {
      return swap_c<true>(x, odd_row);
    } else if constexpr (D == 8) {
      return dpp_c<kDppRowRor8>(x);
    } else if constexpr (D == 4) {
      const CT from_below = dpp_c<kDppRowRor4>(x);   // lane l - 4
      const CT from_above = dpp_c<kDppRowRor12>(x);  // lane l + 4
      return upper4 ? from_below : from_above;
    } else if constexpr (D == 2) {
      return dpp_c<kDppQuadXor2>(x);
    } else {
      return dpp_c<kDppQuadXor1>(x);
    }
  }

  template <int S>
  __device__ __forceinline__ CT step(CT x) const {
    constexpr int D = (N / 2) >> S;
    const CT q = partner<D>(x);
    // lower: x + q; upper: q - x (q is the lower partner's value)
    CT y = mk<CT>(fma(sgn[S], x.x, q.x), fma(sgn[S], x.y, q.y));
    if constexpr (D > 1) y = cmul(y, w[S]);
    return y;
  }

  // Returns X[bitrev(p)] on lane p of each group.
  __device__ __forceinline__ CT operator()(CT x) const {
    x = step<0>(x);
    x = step<1>(x);
    x = step<2>(x);
    if constexpr (kSteps > 3) x = step<3>(x);
    if constexpr (kSteps > 4) x = step<4>(x);
    if constexpr (kSteps > 5) x = step<5>(x);
    return x;
  }
};

// One tile of S = TH * C / N consecutive signals on a TH-thread workgroup
// (256 unless given): wave w transforms elements [tile * S * N + w * C * 64,
// + C * 64), C load instructions of 64 consecutive values.  INVM: 0 forward,
// 1 inverse, 2 the `inv` argument.  Same interface as
// StockhamKernel::run_tile (the LDS argument is unused).
template <int N, int C, int INVM, class CT, int TH = 256>
struct LaneTile {
  static constexpr int THREADS = TH;
  static_assert(TH % 64 == 0 && TH <= 1024, "whole waves");
  static constexpr int S = THREADS * C / N;
  static_assert(S * N == THREADS * C, "whole signals per wave");
  static_assert(C * 64 <= kTwN, "prefetch dummies inside the twiddle table");

  // Prefetch form: plain loads (past the end: the last element), fixed up
  // where the values are used, so no use is tied to the loads; issued whether
  // or not there is a next tile (`live` false: from `dummy`, an array of at
  // least 64 C elements that stays in cache, at distinct offsets), so the
  // loop's vector-memory operations are the same on every pass.
  __device__ __forceinline__ static void load_raw(CT (&v)[C], const CT *in, int64_t total,
                                                  int64_t base, int lane, bool live,
                                                  const CT *dummy) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int64_t e = base + c * 64 + lane;
      v[c] = ld_stream(live ? in + (e < total ? e : total - 1) : dummy + (c * 64 + lane));
    }
  }
  __device__ __forceinline__ static void prep(CT (&v)[C], int64_t total, int64_t base, int lane,
                                              bool sw) {
#pragma unroll
    for (int c = 0; c < C; ++c) pin_vgpr(v[c]);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (base + c * 64 + lane >= total) v[c] = mk<CT>(0, 0);
      if (sw) v[c] = cswap(v[c]);
    }
  }
  __device__ static int64_t base_of(int64_t tile) {
    return tile * (S * N) + (int64_t)(threadIdx.x >> 6) * (C * 64);
  }
  __device__ static void load(CT (&v)[C], const CT *in, int64_t total, int64_t base, int lane,
                              bool sw) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      // past the end: read the last element (no branch) and use zero
      const int64_t e = base + c * 64 + lane;
      v[c] = ld_stream(in + (e < total ? e : total - 1));
      if (e >= total) v[c] = mk<CT>(0, 0);
      if (sw) v[c] = cswap(v[c]);
    }
  }
  __device__ static void store(const CT (&v)[C], CT *out, int64_t total, int64_t base, int lane,
                               int to, real_t<CT> scale, bool sw) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int64_t e = base + c * 64 + lane;  // a whole N-lane group is valid or not
      CT r = v[c];
      if (scale != (real_t<CT>)1) r = cscale(r, scale);
      if (sw) r = cswap(r);
      if (e < total) st_stream(out + e + to, r);
    }
  }

  __device__ static void run_tile(const CT *in, CT *out, int64_t batch, int64_t tile, CT *,
                                  const CT *__restrict__ tw, real_t<CT> scale, bool inv = false) {
    const bool sw = INVM == 0 ? false : INVM == 1 ? true : inv;
    const int lane = threadIdx.x & 63;
    const int64_t total = batch * N;
    const int64_t base = base_of(tile);
    const LaneFft<N, CT> f(tw, lane);
    const int p = lane & (N - 1);
    const int to = bit_reverse<N>(p) - p;  // the group's output index minus its input index
    CT v[C];
    load(v, in, total, base, lane, sw);
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = f(v[c]);
    store(v, out, total, base, lane, to, scale, sw);
  }

  // Pipelined stream of tiles (the interface of StockhamKernel::run_stream):
  // `tile`, then the tiles next() returns until it returns a negative index;
  // the next tile's loads are in flight during this tile's butterflies and
  // stores.
  template <class Next>
  __device__ __forceinline__ static void run_stream(const CT *in, CT *out, int64_t batch,
                                                    int64_t tile, CT *, const CT *__restrict__ tw,
                                                    real_t<CT> scale, bool inv, Next &&next) {
    const bool sw = INVM == 0 ? false : INVM == 1 ? true : inv;
    const int lane = threadIdx.x & 63;
    const int64_t total = batch * N;
    const LaneFft<N, CT> f(tw, lane);
    const int p = lane & (N - 1);
    const int to = bit_reverse<N>(p) - p;
    // two buffers in turn: a copy of the prefetched values would wait for them
    CT a[C], b[C];
    load_raw(a, in, total, base_of(tile), lane, true, tw);
    int64_t nt = next();
    auto step = [&](CT (&cur)[C], CT (&pre)[C]) -> bool {
      const bool more = nt >= 0;  // uniform over the workgroup
      load_raw(pre, in, total, base_of(more ? nt : 0), lane, more, tw);
      prep(cur, total, base_of(tile), lane, sw);
#pragma unroll
      for (int c = 0; c < C; ++c) cur[c] = f(cur[c]);
      const int64_t nnt = more ? next() : -1;
      store(cur, out, total, base_of(tile), lane, to, scale, sw);
      tile = nt;
      nt = nnt;
      return more;
    };
    while (step(a, b) && step(b, a)) {
    }
  }
};

}  // namespace fftoc
