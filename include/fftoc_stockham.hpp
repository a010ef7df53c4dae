// Stockham autosort FFT device code (FP64, vector ALUs) shared by the
// per-length kernels (radix_kernels.hip), the mixed-size batch kernel and the
// serving loop.  Pass p with radix R reads a[j + m*N/R], multiplies by
// W_{R*L}^{m*k} (k = j mod L, L = product of the earlier radices), runs a
// radix-R butterfly and writes a[(j - k)*R + k + q*L] -- the pass structure of
// the reference's legacy_fft / legacy16_fft (includes/legacy8_fft.cuh:57-118,
// includes/legacy16_fft.cuh:99-172).  See radix_kernels.hip for the mapping
// onto the MI355X.
#pragma once

#include <utility>

#include "fftoc_device.hpp"

namespace fftoc {

template <int... Rs>
struct Split {
  static constexpr int P = sizeof...(Rs);
  static constexpr int R[P] = {Rs...};
  static constexpr int N = (Rs * ...);
  static constexpr int RM = [] {
    int m = 1;
    for (int r : {Rs...}) m = r > m ? r : m;
    return m;
  }();
  static constexpr int L(int p) {
    int l = 1;
    for (int i = 0; i < p; ++i) l *= R[i];
    return l;
  }
};

// LDS element index of tile element E = s*N + e: one pad slot after every 16
// elements keeps strided butterfly accesses on distinct banks.
__device__ __forceinline__ int lds_index(int E) { return E + (E >> 4); }
constexpr int lds_elems(int n) { return n + (n >> 4); }

// INVM: 0 = forward, 1 = inverse, 2 = direction chosen at run time by the
// `inv` argument (used by the mixed-size kernels).  PAD: LDS element E lives
// at lds_index(E) (the library's own buffers) or at E (a caller's buffer in
// natural order, see run_lds).
template <class SP, int S, int INVM, bool STAGED, class CT = c2, bool PAD = true>
struct StockhamKernel {
  using R_ = real_t<CT>;
  static constexpr int N = SP::N;
  static constexpr int RM = SP::RM;
  static constexpr int T = N / RM;
  static constexpr int THREADS = T * S;
  static constexpr bool WAVE_LOCAL = (T <= 64) && (64 % T == 0);

  __device__ static int at(int E) {
    if constexpr (PAD) return lds_index(E);
    else return E;
  }

  __device__ static bool swap_io(bool inv) {
    if constexpr (INVM == 0) return false;
    else if constexpr (INVM == 1) return true;
    else return inv;
  }

  __device__ static void sync() {
    if constexpr (WAVE_LOCAL) {
      wave_lds_sync();
    } else {
      __syncthreads();
    }
  }

  // Pass-0 inputs of a prefetch: plain loads from a signal that exists (the
  // caller clamps it), so no use of the values is tied to the loads; prep()
  // then zeroes a signal past the end and applies the direction, where the
  // values are used.  The loads are issued whether or not there is a next
  // tile (`live` false: from the same offsets of `dummy`, an array of at
  // least N elements that stays in cache -- the twiddle table -- rather than
  // all from one element, which every workgroup's last tile hit at once), so
  // every pass through the loop issues the same vector-memory operations and
  // the compiler's wait counts stay exact.
  __device__ __forceinline__ static void load_raw(CT (&reg)[RM], const CT *gin, int t, bool live,
                                                  const CT *dummy) {
    constexpr int R = SP::R[0];
    constexpr int NB = RM / R;
    const CT *base = live ? gin : dummy;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
      for (int m = 0; m < R; ++m) reg[i * R + m] = ld_stream(base + (t + i * T + m * (N / R)));
    }
  }
  __device__ __forceinline__ static void prep(CT (&reg)[RM], bool valid, bool inv) {
#pragma unroll
    for (int i = 0; i < RM; ++i) pin_vgpr(reg[i]);
#pragma unroll
    for (int i = 0; i < RM; ++i) {
      CT v = valid ? reg[i] : mk<CT>(0, 0);
      if (swap_io(inv)) v = cswap(v);
      reg[i] = v;
    }
  }

  // Pass-0 inputs straight from HBM (non-staged kernels).
  __device__ static void load_first(CT (&reg)[RM], const CT *gin, int t, bool valid, bool inv) {
    constexpr int R = SP::R[0];
    constexpr int NB = RM / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
      for (int m = 0; m < R; ++m) {
        CT v = valid ? ld_stream(gin + t + i * T + m * (N / R)) : mk<CT>(0, 0);
        if (swap_io(inv)) v = cswap(v);
        reg[i * R + m] = v;
      }
    }
  }

  // One Stockham pass.  IN_REGS: inputs already in reg (pass 0 of the
  // non-staged kernels); otherwise they are gathered from LDS.
  template <int p, bool IN_REGS>
  __device__ static void pass(CT (&reg)[RM], CT *lds, CT *gout, int t, int sbase, bool valid,
                              const CT *__restrict__ tw, R_ scale, bool inv) {
    constexpr int R = SP::R[p];
    constexpr int L = SP::L(p);
    constexpr int NB = RM / R;
    constexpr bool LAST = (p == SP::P - 1);

    if constexpr (!IN_REGS) {
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int j = t + i * T;
#pragma unroll
        for (int m = 0; m < R; ++m) reg[i * R + m] = lds[at(sbase + j + m * (N / R))];
      }
      // all lanes have read this pass's inputs before anyone overwrites them
      sync();
    }

    // ---- twiddle W_{R L}^{m k}, k = j mod L, and butterfly
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      if constexpr (L > 1) {
        const int k = j & (L - 1);
        constexpr int step = kTwN / (R * L);
        apply_twiddles<R>(&reg[i * R], tw, k * step);
      }
      dft<R>(&reg[i * R]);
    }

    // ---- scatter outputs
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      const int k = j & (L - 1);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int d = (j - k) * R + k + q * L;
        CT v = reg[i * R + q];
        if constexpr (LAST && !STAGED) {
          if (scale != 1.0) v = cscale(v, scale);
          if (swap_io(inv)) v = cswap(v);
          if (valid) st_stream(gout + d, v);
        } else {
          lds[at(sbase + d)] = v;
        }
      }
    }
    if constexpr (!LAST || STAGED) sync();
  }

  template <int... ps>
  __device__ static void later_passes(std::integer_sequence<int, ps...>, CT (&reg)[RM], CT *lds,
                                      CT *gout, int t, int sbase, bool valid,
                                      const CT *__restrict__ tw, R_ scale, bool inv) {
    (pass<ps + 1, false>(reg, lds, gout, t, sbase, valid, tw, scale, inv), ...);
  }

  // Tile `tile` (signals tile*S .. tile*S + S-1) on a workgroup of THREADS
  // threads with LDS buffer `lds` of lds_elems(S*N) elements.
  __device__ static void run_tile(const CT *in, CT *out, int64_t batch, int64_t tile, CT *lds,
                                  const CT *__restrict__ tw, R_ scale, bool inv = false) {
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int sl = tid / T;
    const int64_t tile0 = tile * S;
    const int64_t sig = tile0 + sl;
    const bool valid = sig < batch;
    CT reg[RM];

    if constexpr (STAGED) {
      // contiguous tile copy HBM -> LDS (16 B per lane per instruction)
      const int64_t nvalid = (batch - tile0) < S ? (batch - tile0) : S;
      const CT *src = in + tile0 * N;
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        CT v = (E < nvalid * N) ? ld_stream(src + E) : mk<CT>(0, 0);
        if (swap_io(inv)) v = cswap(v);
        lds[at(E)] = v;
      }
      __syncthreads();
      pass<0, false>(reg, lds, nullptr, t, sl * N, valid, tw, scale, inv);
    } else {
      load_first(reg, in + sig * N, t, valid, inv);
      pass<0, true>(reg, lds, out + sig * N, t, sl * N, valid, tw, scale, inv);
    }
    later_passes(std::make_integer_sequence<int, SP::P - 1>{}, reg, lds, out + sig * N, t, sl * N,
                 valid, tw, scale, inv);

    if constexpr (STAGED) {
      __syncthreads();
      const int64_t nvalid = (batch - tile0) < S ? (batch - tile0) : S;
      CT *dst = out + tile0 * N;
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        CT v = lds[at(E)];
        if (scale != 1.0) v = cscale(v, scale);
        if (swap_io(inv)) v = cswap(v);
        if (E < nvalid * N) st_stream(dst + E, v);
      }
    }
  }

  // In-LDS transform of the S signals of N elements at `lds` (index at(E)
  // for element E = s*N + e), in place, by a workgroup of exactly THREADS
  // threads, all of which call it.  Direction and scale are run-time
  // arguments; `tw` is the library's W_4096 table (fftoc_twiddle_table).
  __device__ static void run_lds(CT *lds, const CT *__restrict__ tw, R_ scale, bool inv) {
    static_assert(STAGED, "the in-LDS form ends every pass in LDS");
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int sl = tid / T;
    if (inv) {  // inverse = forward on swapped (re, im), swapped back
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        lds[at(E)] = cswap(lds[at(E)]);
      }
      __syncthreads();
    }
    CT reg[RM];
    pass<0, false>(reg, lds, nullptr, t, sl * N, true, tw, scale, false);
    later_passes(std::make_integer_sequence<int, SP::P - 1>{}, reg, lds, nullptr, t, sl * N, true,
                 tw, scale, false);
    __syncthreads();  // the fix-up below crosses signals and waves
    if (inv || scale != (R_)1) {
#pragma unroll
      for (int c = 0; c < (S * N) / THREADS; ++c) {
        const int E = tid + c * THREADS;
        CT v = lds[at(E)];
        if (scale != (R_)1) v = cscale(v, scale);
        if (inv) v = cswap(v);
        lds[at(E)] = v;
      }
      __syncthreads();
    }
  }

  // One tile per workgroup.
  __device__ static void run(const CT *in, CT *out, int64_t batch, const CT *__restrict__ tw,
                             R_ scale) {
    __shared__ CT lds[lds_elems(S * N)];
    run_tile(in, out, batch, blockIdx.x, lds, tw, scale);
  }

  // ---- pieces of a pass, for the pipelined persistent form below --------

  static constexpr int kK(int r) { return r == 2 ? 1 : r == 4 ? 2 : r == 8 ? 3 : 4; }
  // base twiddles of passes 1 .. P-1, per pass and group, in one array
  static constexpr int kTwOff(int p) {
    int o = 0;
    for (int q = 1; q < p; ++q) o += (RM / SP::R[q]) * kK(SP::R[q]);
    return o;
  }
  static constexpr int kTwCount = kTwOff(SP::P);

  template <int p>
  __device__ __forceinline__ static void fetch_pass_tw(CT *hw, const CT *__restrict__ tw, int t) {
    constexpr int R = SP::R[p];
    constexpr int L = SP::L(p);
    constexpr int NB = RM / R;
    constexpr int step = kTwN / (R * L);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      TwBase<R, CT> b;
      b.fetch(tw, ((t + i * T) & (L - 1)) * step);
#pragma unroll
      for (int w = 0; w < kK(R); ++w) hw[kTwOff(p) + i * kK(R) + w] = b.w[w];
    }
  }
  template <int... ps>
  __device__ __forceinline__ static void fetch_all_tw(std::integer_sequence<int, ps...>, CT *hw,
                                      const CT *__restrict__ tw, int t) {
    (fetch_pass_tw<ps + 1>(hw, tw, t), ...);
  }

  // LDS -> registers (the pass's inputs), then an ordering point
  template <int p>
  __device__ __forceinline__ static void gather(CT (&reg)[RM], const CT *lds, int t, int sbase) {
    constexpr int R = SP::R[p];
    constexpr int NB = RM / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
#pragma unroll
      for (int m = 0; m < R; ++m) reg[i * R + m] = lds[at(sbase + j + m * (N / R))];
    }
    sync();  // every lane has read before anyone overwrites
  }
  // twiddles (base values from hw; none in pass 0) and butterflies
  template <int p>
  __device__ __forceinline__ static void butterflies(CT (&reg)[RM], const CT *hw) {
    constexpr int R = SP::R[p];
    constexpr int NB = RM / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      if constexpr (p > 0) {
        TwBase<R, CT> b;
#pragma unroll
        for (int w = 0; w < kK(R); ++w) b.w[w] = hw[kTwOff(p) + i * kK(R) + w];
        apply_twiddle_base<R>(&reg[i * R], b);
      }
      dft<R>(&reg[i * R]);
    }
  }
  template <int p>
  __device__ __forceinline__ static void scatter_lds(const CT (&reg)[RM], CT *lds, int t, int sbase) {
    constexpr int R = SP::R[p];
    constexpr int L = SP::L(p);
    constexpr int NB = RM / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      const int k = j & (L - 1);
#pragma unroll
      for (int q = 0; q < R; ++q) lds[at(sbase + (j - k) * R + k + q * L)] = reg[i * R + q];
    }
    sync();
  }
  __device__ __forceinline__ static void scatter_out(const CT (&reg)[RM], CT *gout, int t, bool valid, R_ scale,
                                                     bool inv = false) {
    constexpr int p = SP::P - 1;
    constexpr int R = SP::R[p];
    constexpr int L = SP::L(p);
    constexpr int NB = RM / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int j = t + i * T;
      const int k = j & (L - 1);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        CT v = reg[i * R + q];
        if (scale != 1.0) v = cscale(v, scale);
        if (swap_io(inv)) v = cswap(v);
        if (valid) st_stream(gout + (j - k) * R + k + q * L, v);
      }
    }
  }
  template <int... ps>
  __device__ __forceinline__ static void middle_passes(std::integer_sequence<int, ps...>, CT (&reg)[RM], CT *lds,
                                       int t, int sbase, const CT *hw) {
    ((gather<ps + 1>(reg, lds, t, sbase), butterflies<ps + 1>(reg, hw),
      scatter_lds<ps + 1>(reg, lds, t, sbase)),
     ...);
  }

  // Pipelined stream of tiles: `tile` first, then the tiles next() returns
  // until it returns a negative index (next() is called by every thread at
  // the same point, once per tile, and returns the same value in all; it may
  // hold workgroup barriers).  The base twiddles are read once, before the
  // loop, so the passes after the first issue no vector-memory loads: the
  // next tile's loads, issued right after the current tile's first pass, then
  // overlap all of its later passes (the vector-memory counter retires in
  // order, so any load issued after them -- a twiddle read -- would wait for
  // them).  The next tile's first pass runs before the current tile's stores
  // are issued, so its wait for those loads does not wait for these stores;
  // next() for the tile after it is called there too, when no load is in
  // flight.  `lds` holds lds_elems(S * N) elements.
  template <class Next>
  __device__ __forceinline__ static void run_stream(const CT *in, CT *out, int64_t batch, int64_t tile,
                                                    CT *lds, const CT *__restrict__ tw, R_ scale,
                                                    bool inv, Next &&next) {
    static_assert(!STAGED && SP::P >= 2, "pipelined form needs an LDS pass");
    const int tid = threadIdx.x;
    const int t0 = tid % T, sl0 = tid / T, sbase0 = sl0 * N;
    const int t = t0, sl = sl0, sbase = sbase0;
    CT hw[kTwCount];
    fetch_all_tw(std::make_integer_sequence<int, SP::P - 1>{}, hw, tw, t);
    CT reg[RM];
    {
      const int64_t sig = tile * S + sl;
      load_raw(reg, in + (sig < batch ? sig : batch - 1) * N, t, true, tw);
      prep(reg, sig < batch, inv);
    }
    butterflies<0>(reg, hw);
    scatter_lds<0>(reg, lds, t, sbase);
    int64_t nt = next();
    for (;;) {
      // the thread's indices redefined on every pass: the compiler then
      // computes each LDS and HBM address where it is used instead of
      // hoisting them all out of the loop, where they stayed live and spilled
      int t = t0, sl = sl0, sbase = sbase0;
      asm volatile("" : "+v"(t), "+v"(sl), "+v"(sbase));
      const int64_t sig = tile * S + sl;
      const bool more = nt >= 0;  // uniform over the workgroup
      const int64_t nsig = nt * S + sl;
      CT nxt[RM];
      load_raw(nxt, in + (nsig < batch ? nsig : batch - 1) * N, t, more, tw);
      middle_passes(std::make_integer_sequence<int, SP::P - 2>{}, reg, lds, t, sbase, hw);
      gather<SP::P - 1>(reg, lds, t, sbase);  // LDS free for the next tile after this
      butterflies<SP::P - 1>(reg, hw);
      int64_t nnt = -1;
      if (more) {
        prep(nxt, nsig < batch, inv);
        butterflies<0>(nxt, hw);
        scatter_lds<0>(nxt, lds, t, sbase);
        nnt = next();
      }
      scatter_out(reg, out + sig * N, t, sig < batch, scale, inv);
      if (!more) break;
      tile = nt;
      nt = nnt;
    }
  }

  // Persistent, pipelined: tiles blockIdx.x, blockIdx.x + gridDim.x, ...
  __device__ __forceinline__ static void run_pipelined_persistent(const CT *in, CT *out, int64_t batch,
                                                  const CT *__restrict__ tw, R_ scale) {
    static_assert(INVM != 2, "direction fixed at compile time");
    __shared__ CT lds[lds_elems(S * N)];
    const int64_t ntiles = (batch + S - 1) / S;
    int64_t tile = blockIdx.x;
    if (tile >= ntiles) return;  // uniform over the workgroup
    run_stream(in, out, batch, tile, lds, tw, scale, false, [&]() -> int64_t {
      tile += gridDim.x;
      return tile < ntiles ? tile : -1;
    });
  }

  // Persistent workgroups: tiles blockIdx.x, blockIdx.x + gridDim.x, ...;
  // the next tile's HBM loads overlap the current tile's LDS passes.
  __device__ static void run_persistent(const CT *in, CT *out, int64_t batch,
                                        const CT *__restrict__ tw, R_ scale) {
    static_assert(!STAGED && SP::P >= 2 && INVM != 2, "persistent form needs an LDS pass");
    __shared__ CT lds[lds_elems(S * N)];
    const int tid = threadIdx.x;
    const int t = tid % T;
    const int sl = tid / T;
    const int64_t ntiles = (batch + S - 1) / S;
    int64_t tile = blockIdx.x;
    if (tile >= ntiles) return;  // uniform over the workgroup
    CT reg[RM];
    {
      const int64_t sig = tile * S + sl;
      load_first(reg, in + sig * N, t, sig < batch, false);
    }
    for (;;) {
      const int64_t sig = tile * S + sl;
      const bool valid = sig < batch;
      const int64_t next = tile + gridDim.x;
      // Re-read the twiddles every tile instead of letting the compiler hoist
      // all of them out of the loop (they would pin ~120 VGPRs and halve the
      // occupancy); the empty asm only hides the pointer's loop invariance.
      const CT *twl = tw;
      asm volatile("" : "+s"(twl));
      pass<0, true>(reg, lds, out + sig * N, t, sl * N, valid, twl, scale, false);
      CT nxt[RM];
      if (next < ntiles) {
        const int64_t nsig = next * S + sl;
        load_first(nxt, in + nsig * N, t, nsig < batch, false);
      }
      later_passes(std::make_integer_sequence<int, SP::P - 1>{}, reg, lds, out + sig * N, t,
                   sl * N, valid, twl, scale, false);
      if (next >= ntiles) break;
      tile = next;
#pragma unroll
      for (int i = 0; i < RM; ++i) reg[i] = nxt[i];
    }
  }
};


}  // namespace fftoc
