// FP64 MFMA ("tensor") FFT kernels for gfx950: DFT stages computed as dense
// 16x16 complex matrix products on v_mfma_f64_16x16x4_f64 with the
// three-real-product (Karatsuba) complex GEMM of the reference
// (includes/tensor_utils.cuh:25-36: s1 = Ar Br, s2 = Ai Bi,
// s3 = (Ar+Ai)(Br+Bi); Re = s1 - s2, Im = s3 - s1 - s2) and the
// DFT -> twiddle -> DFT structure of includes/c64_fft64.cuh:8-27.
//
// CDNA4 operand layout of v_mfma_f64_16x16x4_f64 (wave64, lane l, g = l/16,
// j = l%16), measured on MI355X:  A[i][k]: lane l holds A[j][g];
// B[k][col]: lane l holds B[g][j];  D[i][col]: lane l register r holds
// D[4r + g][j].
//
// Design points:
//  * The signal goes into the A operand (rows = independent sub-transforms,
//    K = the 16 inputs of one DFT16 in 4 K-chunks, k = 4c + g), the DFT
//    matrix into B.
//  * For two-stage transforms (256 = 16 x 16 and the sizes built on it) the
//    four D registers of stage 1 ARE the A operands of the four K-chunks of
//    stage 2 (register r = chunk r, row 4r + g = K index 4r + g): the data
//    transpose between the stages costs no data movement (the role of the
//    fragment reuse in includes/c64_fft64.cuh:24-26).
//  * Both stages use the same per-lane B fragment (F16 is symmetric), so a
//    wave keeps 12 doubles of DFT matrix (Br, Bi, Br+Bi per K-chunk).
//  * Twiddles and DFT-matrix entries come from the FP64 table W_4096^j built
//    on the host; only additions are done in-kernel (Br + Bi).
//  * HBM access: N = 8, 16, 32 stage each wave's 4 KiB tile through LDS so
//    every load and store instruction moves 1 KiB contiguously; N = 256 and
//    512 run persistent one-wave workgroups, four per CU (one wave per SIMD)
//    at large batches; N = 4096 loads 1 KiB-contiguous columns directly.
#include "../../include/fftoc_device.hpp"
#include "../../include/fftoc_mfma.hpp"
#include "fftoc_internal.hpp"

namespace fftoc {
namespace {

template <bool INV>
__device__ __forceinline__ c2 ld_in(const c2 *p) {
  c2 v = ld_stream(p);
  if constexpr (INV) v = cswap(v);
  return v;
}

// SC: the launch has a scale other than 1 (FFTOC_NORMALIZE).  A compile-time
// choice: a run-time test costs two selects and two multiplies per value.
template <bool INV, bool SC>
__device__ __forceinline__ void st_out(c2 *p, c2 v, double scale) {
  if constexpr (SC) v = cscale(v, scale);
  if constexpr (INV) v = cswap(v);
  st_stream(p, v);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int64_t wave_global() {
  return (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
}

constexpr int kWaves = 4;  // waves per workgroup for the one-unit-per-wave kernels

// ---------------------------------------------------------------------------
// Every kernel below takes G units per wave (the reference's FFTs per unit):
// all of a wave's HBM loads are issued before its first matrix product, so G
// multiplies the bytes each wave has in flight.

// N = 8: 32 signals per unit.  A = [x_a | x_b] (16 rows, K = 16: chunks 0-1
// hold signal i's 8 inputs, chunks 2-3 those of signal 16+i), B = diag(F8,F8).
template <int G, bool INV, bool SC>
__global__ __launch_bounds__(64 * kWaves) void tensor8_kernel(const c2 *in, c2 *out, int64_t batch,
                                                              const c2 *__restrict__ tw,
                                                              double scale) {
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t sw = wave_global() * (32 * G);
  if (sw >= batch) return;
  const Frag f = dft8x2_frag(tw, lane);
  c2 a[G][4];
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t s = sw + 32 * u + (c < 2 ? 0 : 16) + j;
      const int e = 4 * (c & 1) + g;
      a[u][c] = (s < batch) ? ld_in<INV>(in + s * 8 + e) : make_double2(0.0, 0.0);
    }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    v4d yr, yi;
    kgemm(a[u], f, yr, yi);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t s = sw + 32 * u + (j < 8 ? 0 : 16) + 4 * r + g;
      if (s < batch) st_out<INV, SC>(out + s * 8 + (j & 7), make_double2(yr[r], yi[r]), scale);
    }
  }
}

// ---------------------------------------------------------------------------
// N = 16: 16 signals per unit, one DFT16 per row.
template <int G, bool INV, bool SC>
__global__ __launch_bounds__(64 * kWaves) void tensor16_kernel(const c2 *in, c2 *out, int64_t batch,
                                                               const c2 *__restrict__ tw,
                                                               double scale) {
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t sw = wave_global() * (16 * G);
  if (sw >= batch) return;
  const Frag f = dft16_frag(tw, lane);
  c2 a[G][4];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const int64_t sa = sw + 16 * u + j;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      a[u][c] = (sa < batch) ? ld_in<INV>(in + sa * 16 + 4 * c + g) : make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int u = 0; u < G; ++u) {
    v4d yr, yi;
    kgemm(a[u], f, yr, yi);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t s = sw + 16 * u + 4 * r + g;
      if (s < batch) st_out<INV, SC>(out + s * 16 + j, make_double2(yr[r], yi[r]), scale);
    }
  }
}

// ---------------------------------------------------------------------------
// N = 8 and 16 staged through LDS: a wave loads its 4 KiB tile (32 signals at
// N = 8, 16 at N = 16) with four fully contiguous 1 KiB load instructions,
// moves it into the matrix-core layout through a wave-private LDS buffer,
// runs the same DFT (diag(F8, F8), or F16, on the matrix cores), and stores
// the results through the buffer again, 1 KiB contiguous per instruction.
// The direct kernels above read and write 64-256 B pieces per instruction.
// LDS element e lives at swz16(e).
template <int N>
__device__ __forceinline__ int staged_in(int c, int g, int j) {
  if constexpr (N == 8) return ((c < 2 ? 0 : 16) + j) * 8 + 4 * (c & 1) + g;  // signal, element
  else return j * 16 + 4 * c + g;
}
template <int N>
__device__ __forceinline__ int staged_out(int r, int g, int j) {
  if constexpr (N == 8) return ((j < 8 ? 0 : 16) + 4 * r + g) * 8 + (j & 7);
  else return (4 * r + g) * 16 + j;
}

template <int N, int WAVES, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor_staged_kernel(const c2 *in, c2 *out,
                                                                   int64_t batch,
                                                                   const c2 *__restrict__ tw,
                                                                   double scale) {
  static_assert(N == 8 || N == 16, "N = 8 or 16");
  __shared__ c2 lds[WAVES * 256];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t e0 = wave_global() * 256;  // first element of the wave's tile
  const int64_t total = batch * N;
  if (e0 >= total) return;
  c2 *buf = lds + (threadIdx.x >> 6) * 256;
  const Frag f = N == 8 ? dft8x2_frag(tw, lane) : dft16_frag(tw, lane);
  c2 v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    // past the end: read the last element (no branch) and use zero
    const int64_t e = e0 + 64 * c + lane;
    v[c] = ld_in<INV>(in + (e < total ? e : total - 1));
    if (e >= total) v[c] = make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) buf[swz16(64 * c + lane)] = v[c];
  wave_lds_sync();
  c2 a[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) a[c] = buf[swz16(staged_in<N>(c, g, j))];
  v4d yr, yi;
  kgemm(a, f, yr, yi);
  wave_lds_sync();  // every lane has read its operands before they are overwritten
#pragma unroll
  for (int r = 0; r < 4; ++r) buf[swz16(staged_out<N>(r, g, j))] = make_double2(yr[r], yi[r]);
  wave_lds_sync();
  c2 o[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) o[c] = buf[swz16(64 * c + lane)];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int64_t e = e0 + 64 * c + lane;
    if (e < total) st_out<INV, SC>(out + e, o[c], scale);
  }
}

// ---------------------------------------------------------------------------
// N = 16 * M with M in {2, 4, 8}: n = M n1 + n2.  Rows of the MFMA are
// (signal, n2) pairs, SPW = 16/M signals per wave, row i <-> (signal i % SPW,
// n2 = i / SPW).  Stage 1: DFT16 over n1 on MFMA; then twiddle W_N^{n2 k1}
// and a DFT_M over n2 on the VALU.  D register r of lane l is row 4r + g:
//   M = 2: signal (4r + g) % 8, n2 = r / 2  -> both n2 of two signals in-lane
//   M = 4: signal g, n2 = r                 -> all four n2 in-lane
//   M = 8: signal g % 2, n2 = 2r + g / 2    -> even n2 here, odd n2 in lane
//          l ^ 32; DFT4 in-lane, then one radix-2 exchange across the pair.
// Loads the A operand of the unit whose first signal is s0.
template <int M, bool INV>
__device__ __forceinline__ void load16xm(const c2 *in, int64_t s0, int64_t batch, int g, int j,
                                         c2 (&a)[4]) {
  constexpr int N = 16 * M;
  constexpr int SPW = 16 / M;
  const int64_t s = s0 + j % SPW;
  const int n2 = j / SPW;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    a[c] = (s < batch) ? ld_in<INV>(in + s * N + M * (4 * c + g) + n2) : make_double2(0.0, 0.0);
}

// W_N^k of a unit16xm from the 4096-entry table in HBM, or from a copy of
// W_N^k, k < N, in LDS (whose reads do not wait for HBM loads in flight).
// get(r, k): the twiddle W_N^k of result register r.
template <int N>
struct TwNGlobal {
  const c2 *__restrict__ tw;
  __device__ c2 get(int, int k) const { return tw[k * (kTwN / N)]; }
};
struct TwNLds {
  const c2 *t;
  __device__ c2 get(int, int k) const { return t[k]; }
};
// Loaded into registers ahead (the same for every unit of a lane).
struct TwNRegs {
  c2 w[4];
  template <int M>
  __device__ void fetch(const c2 *__restrict__ tw, int g, int k1) {
    constexpr int SPW = 16 / M;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n2 = (4 * r + g) / SPW;
      w[r] = n2 != 0 ? tw[(n2 * k1) * (kTwN / (16 * M))] : make_double2(1.0, 0.0);
    }
  }
  __device__ c2 get(int r, int) const { return w[r]; }
};

// Where unit16xm puts result k of the unit's signal t (t < 16 / M): the
// output in HBM (signal s0 + t, when it exists), or a wave's LDS tile.
template <int N, bool INV, bool SC>
struct HbmSink {
  c2 *out;
  int64_t s0, batch;
  double scale;
  __device__ bool has(int t) const { return s0 + t < batch; }
  __device__ void operator()(int t, int k, c2 v) const {
    st_out<INV, SC>(out + (s0 + t) * N + k, v, scale);
  }
};
template <int N>
struct LdsSink {
  c2 *buf;  // the unit's 16 / M * N = 256 values, element e at swz16(e)
  __device__ bool has(int) const { return true; }
  __device__ void operator()(int t, int k, c2 v) const { buf[swz16(t * N + k)] = v; }
};

// Transforms the unit whose first signal is s0 and hands its results to
// `put`: the DFT16 over n1 on the matrix cores (contraction D), whose lane
// column holds frequency k1 = d.col, then the twiddle W_N^{n2 k1} and the
// DFT_M over n2.
template <int M, class TW, class D, class SINK>
__device__ __forceinline__ void unit16xm_to(const c2 (&a)[4], const D &d, const TW &tw,
                                            const SINK &put, int g) {
  constexpr int SPW = 16 / M;
  const int j = d.col;  // the frequency k1 of this lane's results
  c2 y[4];
  d.run(a, y);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int n2 = (4 * r + g) / SPW;
    if (n2 != 0) y[r] = cmul(y[r], tw.get(r, n2 * j));
  }
  if constexpr (M == 2) {
    // registers (0, 2) -> signal g, (1, 3) -> signal 4 + g; DFT2 over n2
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int t = 4 * h + g;
      c2 u0 = y[h], u1 = y[h + 2];
      dft2(u0, u1);
      if (put.has(t)) {
        put(t, j, u0);
        put(t, j + 16, u1);
      }
    }
  } else if constexpr (M == 4) {
    dft4(y[0], y[1], y[2], y[3]);
    if (put.has(g)) {
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) put(g, j + 16 * k2, y[k2]);
    }
  } else {
    static_assert(M == 8, "M must be 2, 4 or 8");
    // DIT over n2 = 2m + h (h = g / 2, m = r): E/O = DFT4 over m of the
    // even / odd n2, X[k] = E[k] + W8^k O[k], X[k + 4] = E[k] - W8^k O[k].
    const int h = g >> 1;
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
    if (put.has(g & 1)) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const c2 v = (h == 0) ? cadd(y[k], p[k]) : csub(p[k], y[k]);
        put(g & 1, j + 16 * (4 * h + k), v);
      }
    }
  }
}

// Transforms and stores the unit whose first signal is s0 (signals past
// `batch` are not stored).
template <int M, bool INV, bool SC, class TW = TwNGlobal<16 * M>, class D = KaratsubaDft16>
__device__ __forceinline__ void unit16xm(const c2 (&a)[4], const D &d, const TW &tw, c2 *out,
                                         int64_t s0, int64_t batch, double scale, int g) {
  unit16xm_to<M>(a, d, tw, HbmSink<16 * M, INV, SC>{out, s0, batch, scale}, g);
}

// The 16 x M kernel staged through LDS (as tensor_staged_kernel): a unit's
// 16 / M signals are 256 consecutive values, so a wave's G units are one
// contiguous tile of 256 G values, loaded and stored 1 KiB per instruction
// and moved between that order and the matrix-core layout in its LDS buffer.
template <int M, int G, int WAVES, class D, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor16xm_staged_kernel(
    const c2 *in, c2 *out, int64_t batch, const c2 *__restrict__ tw, double scale) {
  constexpr int N = 16 * M, SPW = 16 / M, TILE = 256 * G;
  __shared__ c2 lds[WAVES * TILE];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t e0 = wave_global() * TILE;  // first value of the wave's tile
  const int64_t total = batch * N;
  if (e0 >= total) return;
  c2 *buf = lds + (threadIdx.x >> 6) * TILE;
  // fragment and twiddles before the data (the vector-memory counter
  // retires in order)
  D d;
  d.load(tw, lane);
  TwNRegs w;
  w.fetch<M>(tw, g, d.col);
  c2 v[G][4];
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // past the end: read the last value (no branch) and use zero
      const int64_t e = e0 + 256 * u + 64 * c + lane;
      v[u][c] = ld_in<INV>(in + (e < total ? e : total - 1));
      if (e >= total) v[u][c] = make_double2(0.0, 0.0);
    }
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) buf[swz16(256 * u + 64 * c + lane)] = v[u][c];
  wave_lds_sync();
  // A operand: row j <-> (signal j % SPW, n2 = j / SPW), K index n1 = 4c + g
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      v[u][c] = buf[swz16(256 * u + (j % SPW) * N + M * (4 * c + g) + j / SPW)];
  wave_lds_sync();  // every lane has read its operands before they are overwritten
#pragma unroll
  for (int u = 0; u < G; ++u) unit16xm_to<M>(v[u], d, w, LdsSink<N>{buf + 256 * u}, g);
  wave_lds_sync();
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) v[u][c] = buf[swz16(256 * u + 64 * c + lane)];
#pragma unroll
  for (int u = 0; u < G; ++u)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t e = e0 + 256 * u + 64 * c + lane;
      if (e < total) st_out<INV, SC>(out + e, v[u][c], scale);
    }
}

template <int M, int G, class D, bool INV, bool SC>
__global__ __launch_bounds__(64 * kWaves) void tensor16xm_kernel(const c2 *in, c2 *out,
                                                                 int64_t batch,
                                                                 const c2 *__restrict__ tw,
                                                                 double scale) {
  constexpr int SPW = 16 / M;  // signals per unit
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t sw = wave_global() * (SPW * G);
  if (sw >= batch) return;
  D d;
  d.load(tw, lane);
  if constexpr (G > 1) {
    // the twiddles (the same for every unit) before the data: loaded after
    // it, their wait (the vector-memory counter retires in order) would hold
    // the first unit until every unit's data had arrived
    TwNRegs w;
    w.fetch<M>(tw, g, d.col);
    c2 au[G][4];
    __builtin_amdgcn_s_setprio(3);  // loads at raised priority (see tensor4096h_kernel)
#pragma unroll
    for (int u = 0; u < G; ++u) load16xm<M, INV>(in, sw + SPW * u, batch, g, j, au[u]);
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int u = 0; u < G; ++u)
      unit16xm<M, INV, SC>(au[u], d, w, out, sw + SPW * u, batch, scale, g);
  } else {
    c2 a[4];
    load16xm<M, INV>(in, sw, batch, g, j, a);
    unit16xm<M, INV, SC>(a, d, TwNGlobal<16 * M>{tw}, out, sw, batch, scale, g);
  }
}

// A unit of a small-N kernel as a type, for the persistent form below:
// kSignals signals per unit, frag() the B operand, load() the A operand of
// the unit starting at signal s0, run() its transform and stores.  (Measured
// persistent forms of N = 8, 16, 128 and 256 were no faster than their
// one-pass kernels; N = 64 below 2^17 signals gains 6 %.)
template <int M, class D = KaratsubaDft16>
struct Unit16xM {
  static constexpr int kSignals = 16 / M;
  __device__ static D frag(const c2 *__restrict__ tw, int lane) {
    D d;
    d.load(tw, lane);
    return d;
  }
  template <bool INV>
  __device__ static void load(const c2 *in, int64_t s0, int64_t batch, int g, int j, c2 (&a)[4]) {
    load16xm<M, INV>(in, s0, batch, g, j, a);
  }
  // W_N^k, k < N, copied into LDS once per workgroup (fill_tw) and read
  // from there by run()
  static constexpr int kTwElems = 16 * M;
  __device__ static void fill_tw(c2 *t, const c2 *__restrict__ tw) {
    for (int k = threadIdx.x; k < kTwElems; k += blockDim.x) t[k] = tw[k * (kTwN / (16 * M))];
  }
  template <bool INV, bool SC>
  __device__ static void run(const c2 (&a)[4], const D &f, const c2 *twl, c2 *out, int64_t s0,
                             int64_t batch, double scale, int g, int j) {
    unit16xm<M, INV, SC>(a, f, TwNLds{twl}, out, s0, batch, scale, g);
  }
};

// Persistent form: wave w handles groups of G units w, w + (waves in the
// grid), ...; the next group's loads are issued before the current group is
// transformed, so they are in flight during its matrix products and stores.
// The twiddles come from a copy in LDS: read from HBM after those loads,
// each twiddle's wait (the vector-memory counter retires in order) would
// also wait for the whole prefetch.
template <class U, int G, int WAVES, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor_persistent_kernel(
    const c2 *in, c2 *out, int64_t batch, const c2 *__restrict__ tw, double scale) {
  __shared__ c2 twl[U::kTwElems];
  constexpr int SPG = U::kSignals * G;  // signals per group
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t groups = (batch + SPG - 1) / SPG;
  const int64_t stride = (int64_t)gridDim.x * WAVES;
  int64_t q = wave_global();
  const bool active = q < groups;  // uniform over the wave
  // The first group's loads go out before the twiddle table is copied into
  // LDS: copied first, the table's round trip (from HBM at the kernel's
  // start) and the barrier came before any data was requested.
  // Loads are issued at raised wave priority (see tensor4096h_kernel).
  c2 cur[G][4];
  if (active) {
    __builtin_amdgcn_s_setprio(3);
#pragma unroll
    for (int u = 0; u < G; ++u)
      U::template load<INV>(in, q * SPG + U::kSignals * u, batch, g, j, cur[u]);
    __builtin_amdgcn_s_setprio(0);
  }
  const auto f = U::frag(tw, lane);
  U::fill_tw(twl, tw);
  __syncthreads();  // before any wave leaves
  if (!active) return;
  for (;;) {
    const int64_t qn = q + stride;
    c2 nxt[G][4];
    if (qn < groups) {
      __builtin_amdgcn_s_setprio(3);
#pragma unroll
      for (int u = 0; u < G; ++u)
        U::template load<INV>(in, qn * SPG + U::kSignals * u, batch, g, j, nxt[u]);
      __builtin_amdgcn_s_setprio(0);
    }
#pragma unroll
    for (int u = 0; u < G; ++u)
      U::template run<INV, SC>(cur[u], f, twl, out, q * SPG + U::kSignals * u, batch, scale, g, j);
    if (qn >= groups) break;
#pragma unroll
    for (int u = 0; u < G; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) cur[u][c] = nxt[u][c];
    q = qn;
  }
}

template <class U, int G, int WAVES = kWaves>
int launch_tensor_persistent(const LaunchArgs &a) {
  if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
  if (a.occupancy) {
    *a.occupancy =
        occupancy<tensor_persistent_kernel<U, G, WAVES, false, false>, 64 * WAVES>().per_cu;
    return FFTOC_SUCCESS;
  }
  if (a.batch <= 0) return FFTOC_SUCCESS;
  constexpr int SPG = U::kSignals * G;
  const int64_t groups = (a.batch + SPG - 1) / SPG;
  const int64_t resident =
      resident_workgroups<tensor_persistent_kernel<U, G, WAVES, false, false>, 64 * WAVES>(a.per_cu);
  const int64_t need = (groups + WAVES - 1) / WAVES;
  const dim3 gd((unsigned)(need < resident ? need : resident)), bd(64 * WAVES);
  const c2 *in = static_cast<const c2 *>(a.in);
  c2 *out = static_cast<c2 *>(a.out);
  const c2 *tw = static_cast<const c2 *>(a.twiddles);
  const bool sc = a.scale != 1.0;
  if (a.inverse && sc)
    hipLaunchKernelGGL((tensor_persistent_kernel<U, G, WAVES, true, true>), gd, bd, 0, a.stream, in, out,
                       a.batch, tw, a.scale);
  else if (a.inverse)
    hipLaunchKernelGGL((tensor_persistent_kernel<U, G, WAVES, true, false>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  else if (sc)
    hipLaunchKernelGGL((tensor_persistent_kernel<U, G, WAVES, false, true>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  else
    hipLaunchKernelGGL((tensor_persistent_kernel<U, G, WAVES, false, false>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

// N = 256: one signal per unit, two MFMA stages, output re-ordered through
// 4 KiB of LDS per wave so every store instruction writes 1 KiB contiguously.
template <int G, class CORE, bool INV, bool SC>
__global__ __launch_bounds__(64 * kWaves) void tensor256_kernel(const c2 *in, c2 *out,
                                                                int64_t batch,
                                                                const c2 *__restrict__ tw,
                                                                double scale) {
  __shared__ c2 lds[kWaves * 256];
  const int lane = lane_id();
  const int64_t sw = wave_global() * G;
  if (sw >= batch) return;
  c2 *buf = lds + (threadIdx.x >> 6) * 256;
  // with several units the twiddles (the same for every unit) come before
  // the data: loaded after it, their wait (the vector-memory counter retires
  // in order) would hold the first unit until every unit's data had arrived
  // (G = 2: 115 -> 103 us, G = 4: 113 -> 108 us at 65536 signals; G = 1
  // keeps them after its data, 104 against 107 us)
  CORE core;
  if constexpr (G > 1) core.load(tw, lane);
  c2 a[G][4];
#pragma unroll
  for (int u = 0; u < G; ++u) {
    const int64_t s = sw + u < batch ? sw + u : sw;  // a short tail re-reads signal sw
#pragma unroll
    for (int c = 0; c < 4; ++c) a[u][c] = ld_in<INV>(in + s * 256 + 64 * c + lane);
  }
  if constexpr (G == 1) core.load(tw, lane);
#pragma unroll
  for (int u = 0; u < G; ++u) {
    c2 z[4];
    core.run(a[u], z);
    if (u > 0) wave_lds_sync();  // the previous unit's LDS reads are done
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[swz16(core.out_index(r))] = z[r];
    wave_lds_sync();
    if (sw + u < batch) {
      c2 *y = out + (sw + u) * 256;
      c2 o[4];  // every read before the first store (see tensor4096h_kernel)
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c] = buf[swz16(64 * c + lane)];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < 4; ++c) st_out<INV, SC>(y + 64 * c + lane, o[c], scale);
    }
  }
}

// Persistent N = 256: wave w handles signals w, w + (waves in the grid), ...;
// the next signal's loads are issued before the current one is transformed.
// Registered with one-wave workgroups, the fftoc_tensor_fft default from 2^14
// signals at four workgroups per CU (one wave per SIMD): 0.75-0.80 of HBM
// against 0.69-0.74 for the two-signal kernel above (three or five per CU:
// 0.62-0.74; the one-matrix-stage and two-product cores within 1 % at 2^16
// signals, 3 % lower at 2^18; at N = 128 the same shape pays through
// tensor_persistent_kernel with one 16x8 unit per iteration; at N = 32 it did
// not gain).
// The core's DFT fragment and twiddles live in registers, loaded once before
// the first signal's data, so no vector-memory load inside the loop waits
// behind the prefetch.  No workgroup barrier: a wave without a signal leaves.
template <class CORE, int WAVES, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor256p_kernel(const c2 *in, c2 *out,
                                                                int64_t batch,
                                                                const c2 *__restrict__ tw,
                                                                double scale) {
  __shared__ c2 lds[WAVES * 256];
  const int lane = lane_id();
  const int64_t stride = (int64_t)gridDim.x * WAVES;
  int64_t s = wave_global();
  if (s >= batch) return;  // uniform over the wave
  c2 *buf = lds + (threadIdx.x >> 6) * 256;
  CORE core;
  core.load(tw, lane);
  c2 cur[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) cur[c] = ld_in<INV>(in + s * 256 + 64 * c + lane);
  for (bool reuse = false;; reuse = true) {
    const int64_t sn = s + stride;
    c2 nxt[4];
    if (sn < batch) {
#pragma unroll
      for (int c = 0; c < 4; ++c) nxt[c] = ld_in<INV>(in + sn * 256 + 64 * c + lane);
    }
    c2 z[4];
    core.run(cur, z);
    if (reuse) wave_lds_sync();  // the previous signal's LDS reads are done
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[swz16(core.out_index(r))] = z[r];
    wave_lds_sync();
    c2 *y = out + s * 256;
    c2 o[4];  // every read before the first store (see tensor4096h_kernel)
#pragma unroll
    for (int c = 0; c < 4; ++c) o[c] = buf[swz16(64 * c + lane)];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) st_out<INV, SC>(y + 64 * c + lane, o[c], scale);
    if (sn >= batch) break;
#pragma unroll
    for (int c = 0; c < 4; ++c) cur[c] = nxt[c];
    s = sn;
  }
}

template <class CORE, int WAVES>
int launch_t256p(const LaunchArgs &a) {
  if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;
  constexpr auto K0 = tensor256p_kernel<CORE, WAVES, false, false>;
  if (a.occupancy) {
    *a.occupancy = occupancy<K0, 64 * WAVES>().per_cu;
    return FFTOC_SUCCESS;
  }
  if (a.batch <= 0) return FFTOC_SUCCESS;
  const int64_t resident = resident_workgroups<K0, 64 * WAVES>(a.per_cu);
  const int64_t need = (a.batch + WAVES - 1) / WAVES;
  const dim3 gd((unsigned)(need < resident ? need : resident)), bd(64 * WAVES);
  const c2 *in = static_cast<const c2 *>(a.in);
  c2 *out = static_cast<c2 *>(a.out);
  const c2 *tw = static_cast<const c2 *>(a.twiddles);
  const bool sc = a.scale != 1.0;
  if (a.inverse && sc)
    hipLaunchKernelGGL((tensor256p_kernel<CORE, WAVES, true, true>), gd, bd, 0, a.stream, in, out,
                       a.batch, tw, a.scale);
  else if (a.inverse)
    hipLaunchKernelGGL((tensor256p_kernel<CORE, WAVES, true, false>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  else if (sc)
    hipLaunchKernelGGL((tensor256p_kernel<CORE, WAVES, false, true>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  else
    hipLaunchKernelGGL((tensor256p_kernel<CORE, WAVES, false, false>), gd, bd, 0, a.stream, in,
                       out, a.batch, tw, a.scale);
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

// ---------------------------------------------------------------------------
// N = 512 = 16 x 32, one signal per unit.  n = 32 n1 + n2, n2 = 16 h + n2'.
// Stage 1 (two MFMA groups h = 0, 1): DFT16 over n1 -> Y_h[n2'][k1];
// twiddle W_512^{n2 k1}; radix-2 over h in registers with W_32^{n2' k2a};
// stage 2 (two groups k2a = 0, 1): DFT16 over n2' with the free transpose.
// X[k1 + 16 k2a + 32 k2b]: lane l register r of group k2a holds
// k1 = 4r + g, k2b = j.
template <bool INV>
__device__ __forceinline__ void load512(const c2 *in, int64_t s, int g, int j, c2 (&a)[2][4]) {
  // n1 = 4c + g, n2 = 16h + j
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < 4; ++c) a[h][c] = ld_in<INV>(in + s * 512 + 32 * (4 * c + g) + 16 * h + j);
}

// Transforms signal s from its operands and stores it (if s < batch) through
// the wave's 512-entry LDS buffer; `reuse`: the buffer still serves the
// previous unit's stores.
// Twiddles of unit512 from the 4096-entry table in HBM.
struct Tw512Global {
  const c2 *__restrict__ tw;
  __device__ c2 w512(int k) const { return tw[k << 3]; }   // W_512^k
  __device__ c2 w32(int m) const { return tw[m << 7]; }    // W_32^m
};
// The same values copied into LDS (fill_tw512_lds): their reads count on the
// LDS counter only, so waiting for one does not wait for HBM loads in flight.
struct Tw512Lds {
  const c2 *t;  // [512] W_512^k, then [32] W_32^m
  __device__ c2 w512(int k) const { return t[k]; }
  __device__ c2 w32(int m) const { return t[512 + m]; }
};
constexpr int kTw512LdsElems = 512 + 32;
__device__ __forceinline__ void fill_tw512_lds(c2 *t, const c2 *__restrict__ tw) {
  for (int k = threadIdx.x; k < kTw512LdsElems; k += blockDim.x)
    t[k] = k < 512 ? tw[k << 3] : tw[(k - 512) << 7];
}

template <bool INV, bool SC, class TW>
__device__ __forceinline__ void unit512(const c2 (&a)[2][4], const Frag &f, const TW &tw, c2 *buf,
                                        c2 *out, int64_t s, int64_t batch, double scale, int lane,
                                        bool reuse) {
  const int g = lane >> 4, j = lane & 15;
  c2 yh[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    v4d yr, yi;
    kgemm(a[h], f, yr, yi);
    // reg r: n2' = 4r + g (n2 = 16h + n2'), k1 = j
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n2 = 16 * h + 4 * r + g;
      yh[h][r] = cmul(make_double2(yr[r], yi[r]), tw.w512(n2 * j));
    }
  }
  // radix-2 over h, then W_32^{n2' k2a}
  c2 b[2][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int n2p = 4 * r + g;
    b[0][r] = cadd(yh[0][r], yh[1][r]);
    b[1][r] = cmul(csub(yh[0][r], yh[1][r]), tw.w32(n2p));
  }
  if (reuse) wave_lds_sync();  // the previous unit's LDS reads are done
#pragma unroll
  for (int k2a = 0; k2a < 2; ++k2a) {
    v4d zr, zi;
    kgemm(b[k2a], f, zr, zi);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = 4 * r + g + 16 * k2a + 32 * j;
      buf[e ^ ((e >> 5) & 15)] = make_double2(zr[r], zi[r]);
    }
  }
  wave_lds_sync();
  if (s < batch) {
    c2 *y = out + s * 512;
    c2 o[8];  // every read before the first store (see tensor4096h_kernel)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int e = 64 * c + lane;
      o[c] = buf[e ^ ((e >> 5) & 15)];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) st_out<INV, SC>(y + 64 * c + lane, o[c], scale);
  }
}

template <int G, int WAVES, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor512_kernel(const c2 *in, c2 *out,
                                                               int64_t batch,
                                                               const c2 *__restrict__ tw,
                                                               double scale) {
  __shared__ c2 lds[WAVES * 512];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t sw = wave_global() * G;
  if (sw >= batch) return;
  c2 *buf = lds + (threadIdx.x >> 6) * 512;
  const Frag f = dft16_frag(tw, lane);
  c2 au[G][2][4];
#pragma unroll
  for (int u = 0; u < G; ++u)
    load512<INV>(in, sw + u < batch ? sw + u : sw, g, j, au[u]);  // a short tail re-reads sw
#pragma unroll
  for (int u = 0; u < G; ++u)
    unit512<INV, SC>(au[u], f, Tw512Global{tw}, buf, out, sw + u, batch, scale, lane, u > 0);
}

// Persistent form: wave w handles signals w, w + (waves in the grid), ...;
// the next signal's loads are issued before the current one is transformed.
// The twiddles are read from a copy in LDS: read from HBM after those loads,
// each twiddle's wait (the vector-memory counter retires in order) would
// also wait for the whole prefetch.
template <int WAVES, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor512p_kernel(const c2 *in, c2 *out,
                                                                int64_t batch,
                                                                const c2 *__restrict__ tw,
                                                                double scale) {
  __shared__ c2 lds[WAVES * 512];
  __shared__ c2 twl[kTw512LdsElems];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t stride = (int64_t)gridDim.x * WAVES;
  int64_t s = wave_global();
  const bool active = s < batch;  // uniform over the wave
  // first signal's loads before the twiddle copy (see tensor_persistent_kernel)
  c2 cur[2][4];
  if (active) load512<INV>(in, s, g, j, cur);
  const Frag f = dft16_frag(tw, lane);
  fill_tw512_lds(twl, tw);
  __syncthreads();  // before any wave leaves
  if (!active) return;
  c2 *buf = lds + (threadIdx.x >> 6) * 512;
  const Tw512Lds twt{twl};
  for (bool reuse = false;; reuse = true) {
    const int64_t sn = s + stride;
    c2 nxt[2][4];
    if (sn < batch) load512<INV>(in, sn, g, j, nxt);
    unit512<INV, SC>(cur, f, twt, buf, out, s, batch, scale, lane, reuse);
    if (sn >= batch) break;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 4; ++c) cur[h][c] = nxt[h][c];
    s = sn;
  }
}
// ---------------------------------------------------------------------------
// N = 4096 = 16 x 256 (16 x 16 x 16), one signal per 512-thread workgroup.
// n = 256 n1 + m, k = k1 + 16 k2.
//  Stage A (each wave: 2 groups of 16 columns m): DFT16 over n1 on MFMA,
//   twiddle W_4096^{m k1}, write Z[k1][m] to LDS.
//  Stage B (each wave: 2 rows k1): 256-point core over m -> X[k1 + 16 k2],
//   results kept in VGPRs until every wave has read Z, then written to LDS
//   in natural order and streamed out with contiguous 1 KiB stores.
constexpr int k4096Waves = 8;

__device__ __forceinline__ int swz_z(int k1, int m) { return k1 * 256 + (m ^ k1); }
__device__ __forceinline__ int swz_x(int e) { return e ^ ((e >> 8) & 15); }

template <bool INV, bool SC>
__global__ __launch_bounds__(64 * k4096Waves) void tensor4096_kernel(const c2 *in, c2 *out,
                                                                     int64_t batch,
                                                                     const c2 *__restrict__ tw,
                                                                     double scale) {
  __shared__ c2 lds[4096];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int w = threadIdx.x >> 6;
  const int64_t s = blockIdx.x;
  if (s >= batch) return;  // uniform over the workgroup
  const c2 *x = in + s * 4096;
  const Frag f = dft16_frag(tw, lane);

  // ---- stage A
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int m0 = 16 * (2 * w + q);
    c2 a[4];
    // row i = j -> m = m0 + j; k = n1 = 4c + g
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = ld_in<INV>(x + 256 * (4 * c + g) + m0 + j);
    v4d yr, yi;
    kgemm(a, f, yr, yi);
    // reg r: m = m0 + 4r + g, k1 = j
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 4 * r + g;
      lds[swz_z(j, m)] = cmul(make_double2(yr[r], yi[r]), tw[m * j]);
    }
  }
  __syncthreads();

  // ---- stage B: rows k1 = 2w + q
  c2 z[2][4];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k1 = 2 * w + q;
    c2 a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = lds[swz_z(k1, 64 * c + lane)];
    core256(a, f, tw, lane, z[q]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k1 = 2 * w + q;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k2 = 4 * r + g + 16 * j;
      lds[swz_x(k1 + 16 * k2)] = z[q][r];
    }
  }
  __syncthreads();
  c2 *y = out + s * 4096;
  c2 o[8];  // every read before the first store (see tensor4096h_kernel)
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = lds[swz_x(512 * c + threadIdx.x)];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < 8; ++c) st_out<INV, SC>(y + 512 * c + threadIdx.x, o[c], scale);
}

// N = 4096 = 16 x 16 x 16 with all three DFT16 stages on the matrix cores and
// the two-product contraction: 384 MFMAs per signal (Karatsuba's three stages
// take 576).  Stage A (each wave: 2 groups of 16 columns m) has the data as
// the B operand (K = n1, column = m), so its combine stays in the lane: o[q]
// = Z'[k1 = hk2(g, q)][m = m0 + j], twiddled by W_4096^{m k1} into LDS.
// Stage B is the two-product 256-point core of each row k1.  The same LDS
// layouts and output order as tensor4096_kernel.
template <bool INV, bool SC>
__global__ __launch_bounds__(64 * k4096Waves) void tensor4096t_kernel(const c2 *in, c2 *out,
                                                                      int64_t batch,
                                                                      const c2 *__restrict__ tw,
                                                                      double scale) {
  __shared__ c2 lds[4096];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int w = threadIdx.x >> 6;
  const int64_t s = blockIdx.x;
  if (s >= batch) return;  // uniform over the workgroup
  const c2 *x = in + s * 4096;
  TwoProductCore core;
  core.load(tw, lane);
  // ---- stage A
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int m = 16 * (2 * w + q) + j;
    c2 wa[4], a[4], o[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) wa[r] = tw[m * hk2(g, r)];
    // K chunk c, slot g: n1 = 4c + g; column j: m
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = ld_in<INV>(x + 256 * (4 * c + g) + m);
    hgemm_b(core.f, a, g, o);
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[swz_z(hk2(g, r), m)] = cmul(o[r], wa[r]);
  }
  __syncthreads();
  // ---- stage B: rows k1 = 2w + q
  c2 z[2][4];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k1 = 2 * w + q;
    c2 b[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) b[c] = lds[swz_z(k1, 64 * c + lane)];
    core.run(b, z[q]);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[swz_x(2 * w + q + 16 * core.out_index(r))] = z[q][r];
  __syncthreads();
  c2 *y = out + s * 4096;
  c2 ov[8];  // every read before the first store (see tensor4096h_kernel)
#pragma unroll
  for (int c = 0; c < 8; ++c) ov[c] = lds[swz_x(512 * c + threadIdx.x)];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < 8; ++c) st_out<INV, SC>(y + 512 * c + threadIdx.x, ov[c], scale);
}

// N = 4096 with stage A on the vector ALUs (threads 0..255; thread t owns
// column m = t: a radix-16 butterfly over n1 in registers, twiddles
// W_4096^{m k1}) and stage B, the 256-point core of each row, on the matrix
// cores (wave w of WAVES: rows (16/WAVES) w + q).  Two thirds of the MFMA work of
// tensor4096_kernel; the VALU stage runs while other waves use the matrix
// cores.  Each wave writes its stage-B results X[k1 + 16 k2] back into the
// LDS row k1 it alone read (k2 permuted within the row against bank
// conflicts), so only two barriers are needed.
__device__ __forceinline__ int swz_row(int k1, int k2) {
  return k1 * 256 + (k2 ^ (((k2 >> 4) ^ k1) & 15));
}

template <int WAVES, class CORE, bool INV, bool SC>
__global__ __launch_bounds__(64 * WAVES) void tensor4096h_kernel(const c2 *in, c2 *out,
                                                                 int64_t batch,
                                                                 const c2 *__restrict__ tw,
                                                                 double scale) {
  constexpr int kRowsPerWave = 16 / WAVES;
  __shared__ c2 lds[4096];
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int w = t >> 6;
  const int64_t s = blockIdx.x;
  if (s >= batch) return;  // uniform over the workgroup
  const c2 *x = in + s * 4096;
  // The waves of a starting workgroup issue their loads at raised priority:
  // the other resident workgroup is usually in its matrix stage, and its
  // instructions would otherwise delay these loads, whose HBM round trip is
  // half of a workgroup's time (+2-3 % at N = 4096 x 16384).
  __builtin_amdgcn_s_setprio(3);
  // Every table value the thread needs is loaded before the signal: the DFT16
  // fragment and the core's twiddles (the same for every row of stage B) and
  // the stage-A twiddle bases of its column.  Loaded after the signal, each
  // wait for them would also wait for it (the vector-memory counter retires in
  // order); loaded after the barrier (as before), they held stage B's first
  // matrix product for a cache round trip in every wave.
  CORE core;
  core.load(tw, lane);
  TwBase<16, c2> ta;
  if (t < 256) ta.fetch(tw, t);
  if (t < 256) {
    // ---- stage A (VALU): column m = t
    c2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) v[n1] = ld_in<INV>(x + 256 * n1 + t);
    __builtin_amdgcn_s_setprio(0);
    dft<16>(v);
    apply_twiddle_base<16>(v, ta);  // v[k1] *= W_4096^{t k1}
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) lds[swz_z(k1, t)] = v[k1];
  }
  __builtin_amdgcn_s_setprio(0);  // waves without a column (WAVES > 4)
  __syncthreads();
  // ---- stage B (MFMA): rows k1 = kRowsPerWave w + q, results back into the same row
#pragma unroll
  for (int q = 0; q < kRowsPerWave; ++q) {
    const int k1 = kRowsPerWave * w + q;
    c2 a[4], z[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = lds[swz_z(k1, 64 * c + lane)];
    core.run(a, z);
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[swz_row(k1, core.out_index(r))] = z[r];
  }
  __syncthreads();
  // All of the thread's outputs are read from LDS before the first store: the
  // compiler otherwise pairs each read with its store and waits for that read
  // alone, and the sixteen LDS round trips run one after another (about a
  // microsecond per signal).
  constexpr int kOut = 4096 / (64 * WAVES);
  c2 o[kOut];
#pragma unroll
  for (int c = 0; c < kOut; ++c) {
    const int e = 64 * WAVES * c + t;
    o[c] = lds[swz_row(e & 15, e >> 4)];
  }
  __builtin_amdgcn_sched_barrier(0);
  c2 *y = out + s * 4096;
#pragma unroll
  for (int c = 0; c < kOut; ++c) st_out<INV, SC>(y + 64 * WAVES * c + t, o[c], scale);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Persistent launch of a one-signal-per-wave kernel: one wave per signal up
// to the resident count.

#define FFTOC_PERSISTENT_LAUNCHER(FN, KERNEL, WAVES)                                          \
  int FN(const LaunchArgs &a) {                                                               \
    if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;                       \
    if (a.occupancy) {                                                                        \
      *a.occupancy = occupancy<KERNEL<WAVES, false, false>, 64 * (WAVES)>().per_cu;           \
      return FFTOC_SUCCESS;                                                                   \
    }                                                                                         \
    if (a.batch <= 0) return FFTOC_SUCCESS;                                                   \
    const int64_t resident =                                                                  \
        resident_workgroups<KERNEL<WAVES, false, false>, 64 * (WAVES)>(a.per_cu);            \
    const int64_t need = ceil_div(a.batch, (WAVES));                                          \
    const dim3 gd((unsigned)(need < resident ? need : resident)), bd(64 * (WAVES));           \
    const c2 *in = static_cast<const c2 *>(a.in);                                             \
    c2 *out = static_cast<c2 *>(a.out);                                                       \
    const c2 *tw = static_cast<const c2 *>(a.twiddles);                                       \
    const bool sc = a.scale != 1.0;                                                           \
    if (a.inverse && sc)                                                                      \
      hipLaunchKernelGGL((KERNEL<WAVES, true, true>), gd, bd, 0, a.stream, in, out, a.batch, tw,     \
                         a.scale);                                                            \
    else if (a.inverse)                                                                       \
      hipLaunchKernelGGL((KERNEL<WAVES, true, false>), gd, bd, 0, a.stream, in, out, a.batch, tw,    \
                         a.scale);                                                            \
    else if (sc)                                                                              \
      hipLaunchKernelGGL((KERNEL<WAVES, false, true>), gd, bd, 0, a.stream, in, out, a.batch, tw,    \
                         a.scale);                                                            \
    else                                                                                      \
      hipLaunchKernelGGL((KERNEL<WAVES, false, false>), gd, bd, 0, a.stream, in, out, a.batch, tw,   \
                         a.scale);                                                            \
    return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;              \
  }
FFTOC_PERSISTENT_LAUNCHER(launch_t512p, tensor512p_kernel, kWaves)
// one wave (one 8 KiB signal) per workgroup
FFTOC_PERSISTENT_LAUNCHER(launch_t512p1, tensor512p_kernel, 1)
#undef FFTOC_PERSISTENT_LAUNCHER

#define FFTOC_TENSOR_LAUNCHER(FN, KERNEL, UNITS_PER_WG, BLOCK, ...)                              \
  int FN(const LaunchArgs &a) {                                                                \
    if (a.precision != FFTOC_PREC_FP64) return FFTOC_ERROR_UNSUPPORTED;                        \
    constexpr auto K0 = KERNEL<__VA_ARGS__ __VA_OPT__(, ) false, false>;                       \
    if (a.occupancy) {                                                                         \
      *a.occupancy = occupancy<K0, (BLOCK)>().per_cu;                                          \
      return FFTOC_SUCCESS;                                                                    \
    }                                                                                          \
    const c2 *in = static_cast<const c2 *>(a.in);                                              \
    c2 *out = static_cast<c2 *>(a.out);                                                        \
    const c2 *tw = static_cast<const c2 *>(a.twiddles);                                        \
    const int64_t grid = ceil_div(a.batch, (UNITS_PER_WG));                                    \
    if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;                                 \
    const dim3 gd((unsigned)grid), bd(BLOCK);                                                \
    const size_t lds = occupancy_lds<K0, (BLOCK)>(a.per_cu);                                   \
    const bool sc = a.scale != 1.0;                                                            \
    if (a.inverse && sc)                                                                       \
      hipLaunchKernelGGL((KERNEL<__VA_ARGS__ __VA_OPT__(, ) true, true>), gd, bd, lds,         \
                         a.stream, in, out, a.batch, tw, a.scale);                             \
    else if (a.inverse)                                                                        \
      hipLaunchKernelGGL((KERNEL<__VA_ARGS__ __VA_OPT__(, ) true, false>), gd, bd, lds,        \
                         a.stream, in, out, a.batch, tw, a.scale);                             \
    else if (sc)                                                                               \
      hipLaunchKernelGGL((KERNEL<__VA_ARGS__ __VA_OPT__(, ) false, true>), gd, bd, lds,        \
                         a.stream, in, out, a.batch, tw, a.scale);                             \
    else                                                                                       \
      hipLaunchKernelGGL((KERNEL<__VA_ARGS__ __VA_OPT__(, ) false, false>), gd, bd, lds,       \
                         a.stream, in, out, a.batch, tw, a.scale);                             \
    return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;               \
  }

FFTOC_TENSOR_LAUNCHER(launch_t8, tensor8_kernel, 32 * kWaves, 64 * kWaves, 1)
FFTOC_TENSOR_LAUNCHER(launch_t8g2, tensor8_kernel, 64 * kWaves, 64 * kWaves, 2)
FFTOC_TENSOR_LAUNCHER(launch_t8g4, tensor8_kernel, 128 * kWaves, 64 * kWaves, 4)
FFTOC_TENSOR_LAUNCHER(launch_t16, tensor16_kernel, 16 * kWaves, 64 * kWaves, 1)
FFTOC_TENSOR_LAUNCHER(launch_t16g2, tensor16_kernel, 32 * kWaves, 64 * kWaves, 2)
FFTOC_TENSOR_LAUNCHER(launch_t16g4, tensor16_kernel, 64 * kWaves, 64 * kWaves, 4)
FFTOC_TENSOR_LAUNCHER(launch_t32, tensor16xm_kernel, 8 * kWaves, 64 * kWaves, 2, 1, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t32g2, tensor16xm_kernel, 16 * kWaves, 64 * kWaves, 2, 2, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t32g4, tensor16xm_kernel, 32 * kWaves, 64 * kWaves, 2, 4, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t64, tensor16xm_kernel, 4 * kWaves, 64 * kWaves, 4, 1, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t64g2, tensor16xm_kernel, 8 * kWaves, 64 * kWaves, 4, 2, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t64g2h, tensor16xm_kernel, 8 * kWaves, 64 * kWaves, 4, 2, TwoProductDft16)
FFTOC_TENSOR_LAUNCHER(launch_t64g4, tensor16xm_kernel, 16 * kWaves, 64 * kWaves, 4, 4, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t128, tensor16xm_kernel, 2 * kWaves, 64 * kWaves, 8, 1, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t128g2, tensor16xm_kernel, 4 * kWaves, 64 * kWaves, 8, 2, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t128g2h, tensor16xm_kernel, 4 * kWaves, 64 * kWaves, 8, 2, TwoProductDft16)
FFTOC_TENSOR_LAUNCHER(launch_t128g4, tensor16xm_kernel, 8 * kWaves, 64 * kWaves, 8, 4, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t256, tensor256_kernel, kWaves, 64 * kWaves, 1, KaratsubaCore)
FFTOC_TENSOR_LAUNCHER(launch_t256g2, tensor256_kernel, 2 * kWaves, 64 * kWaves, 2, KaratsubaCore)
FFTOC_TENSOR_LAUNCHER(launch_t256g2h, tensor256_kernel, 2 * kWaves, 64 * kWaves, 2, TwoProductCore)
FFTOC_TENSOR_LAUNCHER(launch_t256g4, tensor256_kernel, 4 * kWaves, 64 * kWaves, 4, KaratsubaCore)
FFTOC_TENSOR_LAUNCHER(launch_t256g2q, tensor256_kernel, 2 * kWaves, 64 * kWaves, 2, KaratsubaHalfCore)
FFTOC_TENSOR_LAUNCHER(launch_t256g4q, tensor256_kernel, 4 * kWaves, 64 * kWaves, 4, KaratsubaHalfCore)
FFTOC_TENSOR_LAUNCHER(launch_t512, tensor512_kernel, kWaves, 64 * kWaves, 1, kWaves)
FFTOC_TENSOR_LAUNCHER(launch_t512g2, tensor512_kernel, 2 * kWaves, 64 * kWaves, 2, kWaves)
FFTOC_TENSOR_LAUNCHER(launch_t512w1, tensor512_kernel, 1, 64, 1, 1)
FFTOC_TENSOR_LAUNCHER(launch_t4096, tensor4096_kernel, 1, 64 * k4096Waves)
// LDS-staged N = 8 / 16 (tile: 256 values per wave), one or four waves per
// workgroup
FFTOC_TENSOR_LAUNCHER(launch_t8s1, tensor_staged_kernel, 32, 64, 8, 1)
FFTOC_TENSOR_LAUNCHER(launch_t8s4, tensor_staged_kernel, 128, 256, 8, 4)
FFTOC_TENSOR_LAUNCHER(launch_t16s1, tensor_staged_kernel, 16, 64, 16, 1)
FFTOC_TENSOR_LAUNCHER(launch_t16s4, tensor_staged_kernel, 64, 256, 16, 4)
// LDS-staged 16 x 2 (N = 32; tile: 256 values per wave, one-wave workgroups)
FFTOC_TENSOR_LAUNCHER(launch_t32s1, tensor16xm_staged_kernel, 8, 64, 2, 1, 1, KaratsubaDft16)
FFTOC_TENSOR_LAUNCHER(launch_t4096t, tensor4096t_kernel, 1, 64 * k4096Waves)
FFTOC_TENSOR_LAUNCHER(launch_t4096h, tensor4096h_kernel, 1, 256, 4, KaratsubaCore)
FFTOC_TENSOR_LAUNCHER(launch_t4096hh, tensor4096h_kernel, 1, 256, 4, TwoProductCore)
FFTOC_TENSOR_LAUNCHER(launch_t4096h1, tensor4096h_kernel, 1, 256, 4, KaratsubaHalfCore)
FFTOC_TENSOR_LAUNCHER(launch_t4096h1h, tensor4096h_kernel, 1, 256, 4, TwoProductHalfCore)
#undef FFTOC_TENSOR_LAUNCHER

constexpr unsigned TF = FAM_TENSOR | FAM_ANY;
// tile: signals per workgroup (4 waves x signals per wave)
// N = 8, 16, 32: the LDS-staged kernels first (the defaults): 1 KiB contiguous
// per load and store instruction instead of 64-256 B pieces, measured
// against the direct kernels in one process: N = 8 x 2^21 0.794 of HBM
// against 0.737, N = 16 x 2^20 0.792 against 0.721, N = 32 x 2^19 0.798
// against 0.780 (2^17: 0.704 against 0.659).  The same staging at N = 64
// and 128 did not beat their defaults.
const Variant kT8[] = {{"tensor:diag(F8,F8)/32 per wave/LDS-staged", TF, &launch_t8s4, 128},
                       {"tensor:diag(F8,F8)/32 per wave/LDS-staged/1 wave", TF, &launch_t8s1, 32},
                       {"tensor:diag(F8,F8)/32 per wave", TF, &launch_t8, 128},
                       {"tensor:diag(F8,F8)/64 per wave", TF, &launch_t8g2, 256},
                       {"tensor:diag(F8,F8)/128 per wave", TF, &launch_t8g4, 512}};
const Variant kT16[] = {{"tensor:16/16 per wave/LDS-staged/1 wave", TF, &launch_t16s1, 16},
                        {"tensor:16/16 per wave/LDS-staged", TF, &launch_t16s4, 64},
                        {"tensor:16/16 per wave", TF, &launch_t16, 64},
                        {"tensor:16/32 per wave", TF, &launch_t16g2, 128},
                        {"tensor:16/64 per wave", TF, &launch_t16g4, 256}};
const Variant kT32[] = {{"tensor:16x2/8 per wave/LDS-staged/1 wave", TF, &launch_t32s1, 8},
                        {"tensor:16x2/16 per wave", TF, &launch_t32g2, 64},
                        {"tensor:16x2/8 per wave", TF, &launch_t32, 32},
                        {"tensor:16x2/32 per wave", TF, &launch_t32g4, 128}};
const Variant kT64[] = {{"tensor:16x4/8 per wave", TF, &launch_t64g2, 32},
                        {"tensor:16x4/8 per wave/persistent", TF, &launch_tensor_persistent<Unit16xM<4>, 2>, 32},
                        {"tensor:16x4/8 per wave/persistent/2-product", TF, &launch_tensor_persistent<Unit16xM<4, TwoProductDft16>, 2>, 32},
                        {"tensor:16x4/8 per wave/2-product", TF, &launch_t64g2h, 32},
                        {"tensor:16x4/4 per wave", TF, &launch_t64, 16},
                        {"tensor:16x4/16 per wave", TF, &launch_t64g4, 64}};
const Variant kT128[] = {{"tensor:16x8/4 per wave", TF, &launch_t128g2, 16},
                         {"tensor:16x8/4 per wave/2-product", TF, &launch_t128g2h, 16},
                         {"tensor:16x8/2 per wave", TF, &launch_t128, 8},
                         {"tensor:16x8/8 per wave", TF, &launch_t128g4, 32},
                         {"tensor:16x8/2 per wave/persistent/1 wave", TF, &launch_tensor_persistent<Unit16xM<8>, 1, 1>, 2}};
const Variant kT256[] = {{"tensor:16x16/2 per wave", TF, &launch_t256g2, 8},
                         {"tensor:16x16/2 per wave/2-product", TF, &launch_t256g2h, 8},
                         {"tensor:16x16/4 per wave", TF, &launch_t256g4, 16},
                         {"tensor:16x16/1 per wave", TF, &launch_t256, 4},
                         {"tensor:16x16/2 per wave/VALU stage 2", TF, &launch_t256g2q, 8},
                         {"tensor:16x16/4 per wave/VALU stage 2", TF, &launch_t256g4q, 16},
                         {"tensor:16x16/1 per wave/persistent/1 wave", TF, &launch_t256p<KaratsubaCore, 1>, 1}};
const Variant kT512[] = {{"tensor:16x2x16/1 per wave/persistent", TF, &launch_t512p, 4},
                         {"tensor:16x2x16/1 per wave", TF, &launch_t512, 4},
                         {"tensor:16x2x16/2 per wave", TF, &launch_t512g2, 8},
                         {"tensor:16x2x16/1 per wave/persistent/1 wave", TF, &launch_t512p1, 1},
                         {"tensor:16x2x16/1 per wave/1 wave", TF, &launch_t512w1, 1}};
// The first variant is the default of fftoc_tensor_fft at N = 4096: one
// matrix-core DFT16 stage (Karatsuba) between two vector-ALU radix-16 stages,
// 0.73-0.75 of HBM on random data, against 0.68 for two matrix-core stages
// and 0.52 for three (FP64 matrix products on random data are power-bound:
// the two-stage form runs at 0.785 on zeros).
const Variant kT4096[] = {{"tensor:16x16x16/VALU stages A and C, 4 waves", TF, &launch_t4096h1, 1},
                          {"tensor:16x16x16/VALU stages A and C, 4 waves/2-product", TF, &launch_t4096h1h, 1},
                          {"tensor:16x16x16/VALU stage A, 4 waves", TF, &launch_t4096h, 1},
                          {"tensor:16x16x16/VALU stage A, 4 waves/2-product", TF, &launch_t4096hh, 1},
                          {"tensor:16x16x16/8 waves", TF, &launch_t4096, 1},
                          {"tensor:16x16x16/8 waves/2-product", TF, &launch_t4096t, 1}};

template <size_t K>
int set_out(const Variant (&v)[K], const Variant **out) {
  *out = v;
  return (int)K;
}

}  // namespace

int tensor_variants(int n, bool c64, const Variant **out) {
  if (c64) {
    *out = nullptr;
    return 0;
  }
  switch (n) {
    case 8: return set_out(kT8, out);
    case 16: return set_out(kT16, out);
    case 32: return set_out(kT32, out);
    case 64: return set_out(kT64, out);
    case 128: return set_out(kT128, out);
    case 256: return set_out(kT256, out);
    case 512: return set_out(kT512, out);
    case 4096: return set_out(kT4096, out);
    default: *out = nullptr; return 0;
  }
}

}  // namespace fftoc
