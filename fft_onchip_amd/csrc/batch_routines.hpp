// Tile routines of the mixed-size batched FFT kernels (batch_kernels.hip) and
// of the kernels specialized to one routine (batch_one.hip): the routine
// table, the forms of a tile, and the stream that runs a routine over the
// tiles of consecutive work items of one job.
#pragma once
#include "batch_internal.hpp"
#include "lane_fft.hpp"
#include "../../include/fftoc_stockham.hpp"
#include "../../include/fftoc_mfma.hpp"

namespace fftoc {
namespace {

constexpr int kBatchThreads = 256;
constexpr int kLdsElems = lds_elems(4096);

// Tile forms: a Stockham tile read from HBM directly (Tile) or staged
// through LDS (Staged), one tile after another; the Stockham tile pipelined
// over the routine's tiles (Stream: the next tile's loads in flight during
// this tile's passes); the lane-shuffle tile, pipelined (Lanes, N = 8 ... 64);
// the N = 4096 matrix-core tile, pipelined, with the Karatsuba (TensorK) or
// two-product (TensorH) DFT16 contraction, or Karatsuba on one of the
// 256-point core's two DFT16 stages (TensorK1, the library's tensor default)
// (complex128; complex64 jobs run the Stockham stream of the same split).
enum class TForm { Tile, Staged, Stream, Lanes, TensorK, TensorH, TensorK1 };

// N = 4096 = 16 x 256 on one 256-thread workgroup, as the library's tensor
// kernel: stage A is the DFT16 over n1 of column m = thread (VALU) times
// W_4096^{m k1}, written to LDS row k1; stage B is the 256-point core of each
// row on the matrix cores (four rows per wave); the outputs are gathered in
// natural order.  Pipelined over the routine's tiles: the next signal's loads
// go out when stage A has written the current one to LDS (into the same
// registers) and are in flight during stage B and the stores; the stream's
// next() comes after stage B, before the stores, when those loads have had a
// stage B's time to arrive.
template <class CORE, int INVM = 2>
struct TensorTile4096 {
  // INVM: 0 forward, 1 inverse, 2 the run-time `inv` argument
  __device__ static bool swap_io(bool inv) { return INVM == 2 ? inv : INVM == 1; }
  static constexpr int THREADS = 256;
  __device__ static int zpos(int k1, int m) { return k1 * 256 + (m ^ k1); }
  __device__ static int rowpos(int k1, int k2) { return k1 * 256 + (k2 ^ (((k2 >> 4) ^ k1) & 15)); }

  // stage A of the signal in v (column m = t): DFT16 over n1, times
  // W_4096^{m k1}, into LDS row k1
  __device__ __forceinline__ static void stage_a(c2 (&v)[16], c2 *lds, const TwBase<16, c2> &ta,
                                                 int t, bool inv) {
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
      pin_vgpr(v[n1]);
      if (swap_io(inv)) v[n1] = cswap(v[n1]);
    }
    dft<16>(v);
    apply_twiddle_base<16>(v, ta);
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) lds[zpos(k1, t)] = v[k1];
  }

  template <class Next>
  __device__ __forceinline__ static void run_stream(const c2 *in, c2 *out, int64_t batch,
                                                    int64_t tile, c2 *lds,
                                                    const c2 *__restrict__ tw, double scale,
                                                    bool inv, Next &&next) {
    (void)batch;  // one signal per tile: every tile handed out exists
    const int t0 = threadIdx.x, lane0 = t0 & 63, w0 = t0 >> 6;
    // table values before any signal load (the vector-memory counter retires
    // in order: a table load issued after them would wait for them)
    CORE core;
    core.load(tw, lane0);
    TwBase<16, c2> ta;
    ta.fetch(tw, t0);
    c2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = ld_stream(in + tile * 4096 + 256 * n1 + t0);
    stage_a(v, lds, ta, t0, inv);
    int64_t nt = next();
    // Each pass: the next signal's loads (into v, free once stage A has
    // run), stage B and the stores of this one, then the next signal's
    // stage A.  That stage A's wait for its loads follows this pass's stores
    // in straight-line code, so it waits for the loads alone; at the top of
    // the loop (reached from the prologue too) it waited for the stores.
    for (;;) {
      const bool more = nt >= 0;  // uniform over the workgroup
      // the thread's indices redefined on every pass, so its LDS and HBM
      // addresses are computed where they are used and not kept live across
      // the loop (hoisted, they took some sixty registers and spilled)
      int t = t0, lane = lane0, w = w0;
      asm volatile("" : "+v"(t), "+v"(lane), "+v"(w));
      {
        // the next signal (or, past the last tile, the twiddle table, which
        // stays in cache: the loop issues the same loads on every pass)
        const c2 *src = more ? in + nt * 4096 + t : tw + t;
        constexpr int stride = 256;
        // issued at raised wave priority: the other resident workgroup is
        // usually in its matrix stage and would otherwise delay these loads
        __builtin_amdgcn_s_setprio(3);
#pragma unroll
        for (int n1 = 0; n1 < 16; ++n1) v[n1] = ld_stream(src + stride * n1);
        __builtin_amdgcn_s_setprio(0);
      }
      __syncthreads();  // stage A's rows are in LDS
      // ---- stage B (MFMA): rows k1 = 4 w + q, results back into the same row
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k1 = 4 * w + q;
        c2 a[4], z[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = lds[zpos(k1, 64 * c + lane)];
        core.run(a, z);
#pragma unroll
        for (int r = 0; r < 4; ++r) lds[rowpos(k1, core.out_index(r))] = z[r];
      }
      __syncthreads();
      // the outputs in two halves, each read in full before its stores (one
      // LDS round trip per half, not one per store; all sixteen at once held
      // 32 more registers than the kernel has to spare)
      c2 *y = out + tile * 4096;
      int64_t nnt = -1;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        c2 o[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int e = 256 * (8 * h + c) + t;
          o[c] = lds[rowpos(e & 15, e >> 4)];
        }
        __builtin_amdgcn_sched_barrier(0);
        if (h == 1) {
          __syncthreads();  // LDS is rewritten by the next signal's stage A
          nnt = more ? next() : -1;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          c2 r = o[c];
          if (scale != 1.0) r = cscale(r, scale);
          if (swap_io(inv)) r = cswap(r);
          st_stream(y + 256 * (8 * h + c) + t, r);
        }
      }
      if (!more) break;
      stage_a(v, lds, ta, t, inv);
      tile = nt;
      nt = nnt;
    }
  }
};

// DIR: the direction, 0 forward, 1 inverse, 2 chosen at run time per job (a
// run-time choice costs some 6 % at N = 4096: the selects sit on every
// loaded and stored value); the mixed-size kernels run DIR = 2.
template <class SP, int S, TForm F, class CT, int DIR = 2>
struct RoutineKernel {
  using type = StockhamKernel<SP, S, DIR, F == TForm::Staged, CT>;
};
template <class SP, int S, int DIR>
struct RoutineKernel<SP, S, TForm::TensorK, c2, DIR> {
  using type = TensorTile4096<KaratsubaCore, DIR>;
};
template <class SP, int S, int DIR>
struct RoutineKernel<SP, S, TForm::TensorH, c2, DIR> {
  using type = TensorTile4096<TwoProductCore, DIR>;
};
template <class SP, int S, int DIR>
struct RoutineKernel<SP, S, TForm::TensorK1, c2, DIR> {
  using type = TensorTile4096<KaratsubaHalfCore, DIR>;
};
// which kernels a routine streams with run_stream (the others run run_tile
// per tile)
constexpr bool streams(TForm f) { return f != TForm::Tile && f != TForm::Staged; }
template <class SP, int S, class CT, int DIR>
struct RoutineKernel<SP, S, TForm::Lanes, CT, DIR> {
  using type = LaneTile<SP::N, S * SP::N / kBatchThreads, DIR, CT>;
};

// Routine table: id -> (n, split, signals per tile, form).  Two or three
// candidates per length; the plan autotuner picks among them, and
// batch_routine_default gives the untuned choice per data type.
#define FFTOC_ROUTINES(X)                              \
  X(0, 8, "4x2/staged/S128", (Split<4, 2>), 128, TForm::Staged)  \
  X(1, 8, "8/staged/S256", (Split<8>), 256, TForm::Staged)        \
  X(2, 16, "4x4/staged/S64", (Split<4, 4>), 64, TForm::Staged)    \
  X(3, 16, "16/staged/S256", (Split<16>), 256, TForm::Staged)     \
  X(4, 32, "8x4/staged/S64", (Split<8, 4>), 64, TForm::Staged)    \
  X(5, 32, "4x4x2/staged/S32", (Split<4, 4, 2>), 32, TForm::Staged) \
  X(6, 64, "8x8/tile/S32", (Split<8, 8>), 32, TForm::Tile)     \
  X(7, 64, "4x4x4/tile/S16", (Split<4, 4, 4>), 16, TForm::Tile) \
  X(8, 128, "16x8/tile/S32", (Split<16, 8>), 32, TForm::Tile)  \
  X(9, 128, "8x16/tile/S32", (Split<8, 16>), 32, TForm::Tile)  \
  X(10, 256, "16x16/tile/S16", (Split<16, 16>), 16, TForm::Tile) \
  X(11, 256, "8x8x4/tile/S8", (Split<8, 8, 4>), 8, TForm::Tile) \
  X(12, 512, "8x8x8/tile/S4", (Split<8, 8, 8>), 4, TForm::Tile) \
  X(13, 512, "16x16x2/tile/S8", (Split<16, 16, 2>), 8, TForm::Tile) \
  X(14, 1024, "16x16x4/tile/S4", (Split<16, 16, 4>), 4, TForm::Tile) \
  X(15, 1024, "8x8x16/tile/S4", (Split<8, 8, 16>), 4, TForm::Tile) \
  X(16, 2048, "16x16x8/tile/S2", (Split<16, 16, 8>), 2, TForm::Tile) \
  X(17, 2048, "8x16x16/tile/S2", (Split<8, 16, 16>), 2, TForm::Tile) \
  X(18, 4096, "16x16x16/tile/S1", (Split<16, 16, 16>), 1, TForm::Tile) \
  X(19, 128, "8x4x4/tile/S16", (Split<8, 4, 4>), 16, TForm::Tile) \
  X(20, 8, "8/lanes/S128", (Split<8>), 128, TForm::Lanes) \
  X(21, 16, "16/lanes/S64", (Split<16>), 64, TForm::Lanes) \
  X(22, 32, "32/lanes/S32", (Split<32>), 32, TForm::Lanes) \
  X(23, 64, "64/lanes/S16", (Split<64>), 16, TForm::Lanes) \
  X(24, 64, "8x8/stream/S32", (Split<8, 8>), 32, TForm::Stream) \
  X(25, 64, "4x4x4/stream/S16", (Split<4, 4, 4>), 16, TForm::Stream) \
  X(26, 128, "8x4x4/stream/S16", (Split<8, 4, 4>), 16, TForm::Stream) \
  X(27, 128, "16x8/stream/S32", (Split<16, 8>), 32, TForm::Stream) \
  X(28, 256, "8x8x4/stream/S8", (Split<8, 8, 4>), 8, TForm::Stream) \
  X(29, 256, "16x16/stream/S16", (Split<16, 16>), 16, TForm::Stream) \
  X(30, 512, "8x8x8/stream/S4", (Split<8, 8, 8>), 4, TForm::Stream) \
  X(31, 1024, "16x16x4/stream/S4", (Split<16, 16, 4>), 4, TForm::Stream) \
  X(32, 1024, "8x8x16/stream/S4", (Split<8, 8, 16>), 4, TForm::Stream) \
  X(33, 2048, "16x16x8/stream/S2", (Split<16, 16, 8>), 2, TForm::Stream) \
  X(34, 2048, "8x16x16/stream/S2", (Split<8, 16, 16>), 2, TForm::Stream) \
  X(35, 4096, "16x16x16/stream/S1", (Split<16, 16, 16>), 1, TForm::Stream) \
  X(36, 4096, "tensor:16x16x16/stream/S1", (Split<16, 16, 16>), 1, TForm::TensorK) \
  X(37, 4096, "tensor:16x16x16/stream/2-product/S1", (Split<16, 16, 16>), 1, TForm::TensorH) \
  X(38, 1024, "8x8x8x2/stream/S2", (Split<8, 8, 8, 2>), 2, TForm::Stream) \
  X(39, 2048, "8x8x8x4/stream/S1", (Split<8, 8, 8, 4>), 1, TForm::Stream) \
  X(40, 4096, "8x8x8x8/stream/S1 (512 threads)", (Split<8, 8, 8, 8>), 1, TForm::Stream) \
  X(41, 4096, "tensor:16x16x16/stream/one matrix stage/S1", (Split<16, 16, 16>), 1, TForm::TensorK1)

#define FFTOC_UNPAREN(...) __VA_ARGS__

// Routines run on workgroups of kBatchThreads threads, except wide routines
// (kWideThreads: the N = 4096 radix-8 stream, eight values per thread), which
// only the one-routine kernels run: mixed plans and the server use a
// kBatchThreads routine of the same length instead (batch_routine_mixed).
constexpr int kWideThreads = 512;
template <class SP, int S, TForm F, class CT>
constexpr int kRoutineThreads = RoutineKernel<SP, S, F, CT>::type::THREADS;

#define FFTOC_CHECK_THREADS(ID, NN, NAME, SPLIT, S, FORM)                                     \
  static_assert((kRoutineThreads<FFTOC_UNPAREN SPLIT, S, FORM, c2> == kBatchThreads ||         \
                 kRoutineThreads<FFTOC_UNPAREN SPLIT, S, FORM, c2> == kWideThreads) &&         \
                    kRoutineThreads<FFTOC_UNPAREN SPLIT, S, FORM, c2> ==                       \
                        kRoutineThreads<FFTOC_UNPAREN SPLIT, S, FORM, f2>,                     \
                "tile routine " #ID ": 256 or 512 threads");                                  \
  static_assert(FFTOC_UNPAREN SPLIT::N == NN, "tile routine " #ID " length");                 \
  static_assert(S * NN <= 4096, "tile routine " #ID " LDS");
FFTOC_ROUTINES(FFTOC_CHECK_THREADS)
#undef FFTOC_CHECK_THREADS

struct RoutineInfo {
  int n;
  const char *name;
  int signals_per_tile;
  int threads;
};

#define FFTOC_INFO(ID, NN, NAME, SPLIT, S, FORM) \
  {NN, NAME, S, kRoutineThreads<FFTOC_UNPAREN SPLIT, S, FORM, c2>},
constexpr RoutineInfo kRoutines[] = {FFTOC_ROUTINES(FFTOC_INFO)};
#undef FFTOC_INFO
constexpr int kNumRoutines = sizeof(kRoutines) / sizeof(kRoutines[0]);

// A routine's pointer parameters carry their address spaces (global, LDS):
// through generic pointers it would access HBM with flat instructions, whose
// LDS counter makes every LDS wait also wait for outstanding HBM traffic.
template <class T>
using global_ptr = __attribute__((address_space(1))) T *;
template <class T>
using lds_ptr = __attribute__((address_space(3))) T *;

// Where a workgroup stands in the item sequence: the item it runs (the same
// in every thread), whether thread 0 has claimed the next item ahead (the
// same in every thread) and, in thread 0, that claim.
struct Cursor {
  int64_t item;
  long long claim;
  int ahead;
};

// Claims are made one item ahead (the counter's round trip hidden behind the
// current item) except in the last two rounds of the grid: there an item
// claimed ahead waits behind its holder's current item while other
// workgroups find the counter exhausted, and the tail of the launch grows to
// two items.  Claims on one address serialize (some 12 ns each), so the
// first item of workgroup b is b and the counter hands out the items from
// the grid size on.
__device__ __forceinline__ bool claim_ahead(int64_t item, int64_t total) {
  return item + 2 * (int64_t)gridDim.x < total;
}
__device__ __forceinline__ long long claim_item(global_ptr<unsigned long long> ctr) {
  // a global (not flat) atomic: no LDS wait waits for it
  return (long long)gridDim.x +
         (long long)__hip_atomic_fetch_add(ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0's claimed item to every thread (two barriers: the slot is read by
// all before it is written again).
__device__ __forceinline__ int64_t publish_claim(long long claim, lds_ptr<long long> s_item) {
  if (threadIdx.x == 0) *s_item = claim;
  __syncthreads();
  const int64_t item = (int64_t)uniform_u64((uint64_t)*s_item);
  __syncthreads();
  return item;
}

// One routine call streams tiles: those of item cur.item of its job (items
// [item_begin, item_end) of tpi tiles each, relative tile (item - item_begin)
// * tpi on).  With a claim counter, at the end of an item the claimed item is
// published and the next one claimed; while the published item belongs to the
// same job the stream runs on into it, so the next tile's loads overlap this
// tile's passes across item boundaries too.  Returns the cursor at the first
// item outside the job.  Without a counter (ctr null) the stream ends with
// the item.
template <class K, TForm F, class CT>
__device__ __forceinline__ Cursor stream_routine(global_ptr<const CT> in, global_ptr<CT> out, int64_t batch,
                                           Cursor cur, int64_t item_begin, int64_t item_end,
                                           int64_t total, int64_t ntiles, int tpi,
                                           global_ptr<unsigned long long> ctr,
                                           lds_ptr<long long> s_item, lds_ptr<CT> lds,
                                           global_ptr<const CT> tw, real_t<CT> scale, bool inv) {
  // the stream's bounds drive its barriers: made uniform
  int64_t item = (int64_t)uniform_u64((uint64_t)cur.item);
  item_begin = (int64_t)uniform_u64((uint64_t)item_begin);
  item_end = (int64_t)uniform_u64((uint64_t)item_end);
  total = (int64_t)uniform_u64((uint64_t)total);
  ntiles = (int64_t)uniform_u64((uint64_t)ntiles);
  tpi = __builtin_amdgcn_readfirstlane(tpi);
  const bool claims = __builtin_amdgcn_readfirstlane(ctr != nullptr ? 1 : 0) != 0;
  long long claim = cur.claim;
  int64_t t = (item - item_begin) * tpi;
  int64_t t_end = (t + tpi < ntiles) ? t + tpi : ntiles;
  bool ahead = __builtin_amdgcn_readfirstlane(cur.ahead) != 0;
  auto next = [&]() -> int64_t {
    if (++t < t_end) return t;
    if (!claims) return -1;
    if (!ahead && threadIdx.x == 0) claim = claim_item(ctr);  // the tail: claimed now
    item = publish_claim(claim, s_item);
    ahead = claim_ahead(item, total);
    if (ahead && threadIdx.x == 0) claim = claim_item(ctr);
    if (item >= item_end) return -1;
    t = (item - item_begin) * tpi;
    t_end = (t + tpi < ntiles) ? t + tpi : ntiles;
    return t;
  };
  if constexpr (streams(F)) {
    K::run_stream((const CT *)in, (CT *)out, batch, t, (CT *)lds, (const CT *)tw, scale, inv, next);
  } else {
    for (int64_t tile = t; tile >= 0; tile = next()) {
      K::run_tile((const CT *)in, (CT *)out, batch, tile, (CT *)lds, (const CT *)tw, scale, inv);
      __syncthreads();  // LDS is reused by the next tile
    }
  }
  return Cursor{item, claim, ahead ? 1 : 0};
}

// The mixed-size kernels call each routine as a separate (not inlined)
// function, so that its registers are allocated on their own: the kernel
// needs the maximum over the routines, not their union.  The price is the
// callee-saved registers the routine stores to scratch and reloads on every
// call (up to 112 per lane), which the one-routine kernels do not pay.
template <class K, TForm F, class CT>
__device__ __noinline__ Cursor tile_stream(global_ptr<const CT> in, global_ptr<CT> out, int64_t batch,
                                           Cursor cur, int64_t item_begin, int64_t item_end,
                                           int64_t total, int64_t ntiles, int tpi,
                                           global_ptr<unsigned long long> ctr,
                                           lds_ptr<long long> s_item, lds_ptr<CT> lds,
                                           global_ptr<const CT> tw, real_t<CT> scale, bool inv) {
  return stream_routine<K, F, CT>(in, out, batch, cur, item_begin, item_end, total, ntiles, tpi, ctr,
                                  s_item, lds, tw, scale, inv);
}

// The call of the mixed-size kernels; a wide routine is never dispatched
// there (the host maps it to a kBatchThreads routine).
template <class K, TForm F, class CT>
__device__ __forceinline__ Cursor mixed_call(global_ptr<const CT> in, global_ptr<CT> out,
                                             int64_t batch, Cursor cur, int64_t item_begin,
                                             int64_t item_end, int64_t total, int64_t ntiles,
                                             int tpi, global_ptr<unsigned long long> ctr,
                                             lds_ptr<long long> s_item, lds_ptr<CT> lds,
                                             global_ptr<const CT> tw, real_t<CT> scale, bool inv) {
  if constexpr (K::THREADS == kBatchThreads)
    return tile_stream<K, F, CT>(in, out, batch, cur, item_begin, item_end, total, ntiles, tpi, ctr,
                                 s_item, lds, tw, scale, inv);
  else
    return Cursor{item_end, cur.claim, cur.ahead};
}

template <class CT>
__device__ __forceinline__ Cursor run_routine(int id, const JobDev &J, Cursor cur, int64_t item_begin,
                                              int64_t item_end, int64_t total,
                                              unsigned long long *ctr, long long *s_item, CT *lds,
                                              const CT *__restrict__ tw) {
  const CT *in = static_cast<const CT *>(J.in);
  CT *out = static_cast<CT *>(J.out);
  const real_t<CT> scale = (real_t<CT>)J.scale;
  const bool inv = J.inverse != 0;
  switch (id) {
#define FFTOC_CASE(ID, NN, NAME, SPLIT, S, FORM)                                                  \
  case ID:                                                                                      \
    return mixed_call<typename RoutineKernel<FFTOC_UNPAREN SPLIT, S, FORM, CT>::type, FORM, CT>( \
        (global_ptr<const CT>)in, (global_ptr<CT>)out, J.batch, cur, item_begin, item_end, total,  \
        J.ntiles, J.tiles_per_item, (global_ptr<unsigned long long>)ctr,                           \
        (lds_ptr<long long>)s_item, (lds_ptr<CT>)lds, (global_ptr<const CT>)tw, scale, inv);
    FFTOC_ROUTINES(FFTOC_CASE)
#undef FFTOC_CASE
    default:
      return Cursor{item_end, cur.claim, cur.ahead};
  }
}


}  // namespace
}  // namespace fftoc
