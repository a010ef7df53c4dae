// Split-precision MFMA ("tensor") FFT kernels for gfx950: the DFT stage
// GEMMs run in FP32 (v_mfma_f32_16x16x4_f32) or in 3-term split BF16
// (v_mfma_f32_16x16x16_bf16) with FP32 accumulation, complex products with
// the three-real-product (Karatsuba) form of includes/tensor_utils.cuh:25-36.
// Data stay complex128 in HBM; they are rounded to FP32 on load, the
// inter-stage twiddles are the FP64 table rounded to FP32, results are
// widened to FP64 on store and the 1/N normalisation is applied in FP64.
// At N = 4096 the first of the three 16-point stages is an FP64 butterfly on
// the vector ALUs (rounded to FP32 into LDS) and the two GEMM stages of each
// row's 256-point core run in the chosen precision.
//
// Operand layouts on gfx950 (wave64, lane l, g = l/16, j = l%16), measured:
//  f32 16x16x4:   A[i][k] = lane(i=j, k=g); B[k][c] = lane(k=g, c=j);
//                 D register r = row 4g + r, column j.
//  bf16 16x16x16: lane l holds A[j][4g + t] (t = 0..3) and B[4g + t][j];
//                 D register r = row 4g + r, column j.
// With the K order k = 4g + c, the four D registers of stage 1 are exactly
// the A operand of stage 2 in both forms (FP32: chunk c = register c; BF16:
// element t = register t), so a 256-point transform still needs no data
// movement between its stages and one B fragment serves both.
//
// 3-term split BF16: x = x0 + x1 + x2 with x0 = bf16(x), x1 = bf16(x - x0),
// x2 = bf16(x - x0 - x1) (24 significant bits); a product keeps the six terms
// a_p b_q with p + q <= 2, smallest first, accumulated in FP32.
#include <algorithm>

#include "../../include/fftoc_device.hpp"
#include "fftoc_internal.hpp"

namespace fftoc {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

constexpr int kFP32 = 1;
constexpr int kBF16X3 = 2;

__device__ __forceinline__ v4f fzero() { return v4f{0.f, 0.f, 0.f, 0.f}; }

// float -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ short bf16_bits(float x) {
  const unsigned u = __float_as_uint(x);
  return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_value(short b) {
  return __uint_as_float(((unsigned)(unsigned short)b) << 16);
}
// x ~= t0 + t1 + t2 (each a bf16)
__device__ __forceinline__ void split3(float x, short &t0, short &t1, short &t2) {
  t0 = bf16_bits(x);
  const float r1 = x - bf16_value(t0);
  t1 = bf16_bits(r1);
  const float r2 = r1 - bf16_value(t1);
  t2 = bf16_bits(r2);
}

template <int P>
struct Stage;

// ---- FP32 -------------------------------------------------------------------
template <>
struct Stage<kFP32> {
  struct Frag {
    float r[4], i[4], s[4];
  };
  // B[k][col] = W_16^{n col}, K-chunk c, slot g -> k = 4g + c; the input
  // sample is n = k, or n = 4c + g when SAMPLES_4CG (a single-stage transform
  // whose loads then cover 16 consecutive samples per signal)
  template <bool SAMPLES_4CG = false>
  __device__ static Frag frag(const c2 *__restrict__ tw, int lane) {
    const int g = lane >> 4, j = lane & 15;
    Frag f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int n = SAMPLES_4CG ? 4 * c + g : 4 * g + c;
      const c2 w = tw[((n * j) & 15) << 8];
      f.r[c] = (float)w.x;
      f.i[c] = (float)w.y;
      f.s[c] = (float)(w.x + w.y);
    }
    return f;
  }
  // Y = X * F; a[c] = A operand of K-chunk c
  __device__ static void gemm(const f2 (&a)[4], const Frag &f, v4f &yr, v4f &yi) {
    v4f p1 = fzero(), p2 = fzero(), p3 = fzero();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      p1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c].x, f.r[c], p1, 0, 0, 0);
      p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c].y, f.i[c], p2, 0, 0, 0);
      p3 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[c].x + a[c].y, f.s[c], p3, 0, 0, 0);
    }
    yr = p1 - p2;
    yi = p3 - p1 - p2;
  }
};

// ---- 3-term split BF16 --------------------------------------------------------
template <>
struct Stage<kBF16X3> {
  struct Frag {
    v4s r[3], i[3], s[3];  // [term]: B elements t = 0..3 <-> k = 4g + t
  };
  // input sample n = k = 4g + t, or n = 4t + g when SAMPLES_4CG
  template <bool SAMPLES_4CG = false>
  __device__ static Frag frag(const c2 *__restrict__ tw, int lane) {
    const int g = lane >> 4, j = lane & 15;
    Frag f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int n = SAMPLES_4CG ? 4 * t + g : 4 * g + t;
      const c2 w = tw[((n * j) & 15) << 8];
      short a0, a1, a2;
      split3((float)w.x, a0, a1, a2);
      f.r[0][t] = a0; f.r[1][t] = a1; f.r[2][t] = a2;
      split3((float)w.y, a0, a1, a2);
      f.i[0][t] = a0; f.i[1][t] = a1; f.i[2][t] = a2;
      split3((float)(w.x + w.y), a0, a1, a2);
      f.s[0][t] = a0; f.s[1][t] = a1; f.s[2][t] = a2;
    }
    return f;
  }
  // acc += A * B with both split into three BF16 terms (six products)
  __device__ static v4f prod3(const v4s (&a)[3], const v4s (&b)[3], v4f acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[2], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b[0], acc, 0, 0, 0);
    return acc;
  }
  __device__ static void gemm(const f2 (&a)[4], const Frag &f, v4f &yr, v4f &yi) {
    v4s ar[3], ai[3], as[3];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      short x0, x1, x2;
      split3(a[t].x, x0, x1, x2);
      ar[0][t] = x0; ar[1][t] = x1; ar[2][t] = x2;
      split3(a[t].y, x0, x1, x2);
      ai[0][t] = x0; ai[1][t] = x1; ai[2][t] = x2;
      split3(a[t].x + a[t].y, x0, x1, x2);
      as[0][t] = x0; as[1][t] = x1; as[2][t] = x2;
    }
    const v4f p1 = prod3(ar, f.r, fzero());
    const v4f p2 = prod3(ai, f.i, fzero());
    const v4f p3 = prod3(as, f.s, fzero());
    yr = p1 - p2;
    yi = p3 - p1 - p2;
  }
};

template <bool INV>
__device__ __forceinline__ f2 ld_in(const c2 *p) {
  c2 v = ld_stream(p);
  if constexpr (INV) v = cswap(v);
  return make_float2((float)v.x, (float)v.y);
}

template <bool INV>
__device__ __forceinline__ void st_out(c2 *p, f2 v, double scale) {
  c2 w = make_double2((double)v.x, (double)v.y);
  if (scale != 1.0) w = cscale(w, scale);
  if constexpr (INV) w = cswap(w);
  st_stream(p, w);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int64_t wave_global() {
  return (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
}

constexpr int kWaves = 4;
// Units per wave (all of a wave's HBM loads are issued before its first
// product, as in the FP64 tensor kernels).
constexpr int kG64 = 2;
constexpr int kG256 = 4;

// N = 64 = 16 x 4, four signals per unit; row i <-> (signal i / 4, n2 = i % 4),
// chunk / element c of lane group g holds n1 = 4c + g (16 consecutive samples
// of a signal per load instruction).  D register r of lane l: signal g,
// n2 = r, k1 = j; twiddle W_64^{n2 k1} and DFT4 over r in FP32, output
// X_g[j + 16 k2].
template <int P, bool INV>
__device__ __forceinline__ void load64p(const c2 *in, int64_t s0, int64_t batch, int g, int j,
                                        f2 (&a)[4]) {
  const int64_t s = s0 + (j >> 2);
  const int n2 = j & 3;
#pragma unroll
  for (int c = 0; c < 4; ++c)
    a[c] = (s < batch) ? ld_in<INV>(in + s * 64 + 4 * (4 * c + g) + n2) : make_float2(0.f, 0.f);
}

template <int P, bool INV>
__device__ __forceinline__ void unit64p(const f2 (&a)[4], const typename Stage<P>::Frag &f,
                                        const f2 *__restrict__ tw32, c2 *out, int64_t s0,
                                        int64_t batch, double scale, int g, int j) {
  v4f yr, yi;
  Stage<P>::gemm(a, f, yr, yi);
  f2 y[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    f2 v = make_float2(yr[r], yi[r]);
    if (r != 0) v = cmul(v, tw32[(r * j) << 6]);  // W_64^{r j}
    y[r] = v;
  }
  dft4(y[0], y[1], y[2], y[3]);
  const int64_t s = s0 + g;
  if (s < batch) {
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) st_out<INV>(out + s * 64 + j + 16 * k2, y[k2], scale);
  }
}

// Persistent: wave w handles groups of kG64 units w, w + (waves in the
// grid), ...; the next group's loads are issued before the current group is
// transformed (as the FP64 N = 64 kernel).
template <int P, bool INV>
__global__ __launch_bounds__(64 * kWaves) void tensor64p_kernel(const c2 *in, c2 *out,
                                                                int64_t batch,
                                                                const c2 *__restrict__ tw,
                                                                const f2 *__restrict__ tw32,
                                                                double scale) {
  constexpr int SPG = 4 * kG64;  // signals per group
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t groups = (batch + SPG - 1) / SPG;
  const int64_t stride = (int64_t)gridDim.x * kWaves;
  int64_t q = wave_global();
  if (q >= groups) return;
  const auto f = Stage<P>::template frag<true>(tw, lane);
  f2 cur[kG64][4];
#pragma unroll
  for (int u = 0; u < kG64; ++u) load64p<P, INV>(in, q * SPG + 4 * u, batch, g, j, cur[u]);
  for (;;) {
    const int64_t qn = q + stride;
    f2 nxt[kG64][4];
    if (qn < groups) {
#pragma unroll
      for (int u = 0; u < kG64; ++u) load64p<P, INV>(in, qn * SPG + 4 * u, batch, g, j, nxt[u]);
    }
#pragma unroll
    for (int u = 0; u < kG64; ++u)
      unit64p<P, INV>(cur[u], f, tw32, out, q * SPG + 4 * u, batch, scale, g, j);
    if (qn >= groups) break;
#pragma unroll
    for (int u = 0; u < kG64; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) cur[u][c] = nxt[u][c];
    q = qn;
  }
}

// 256-point core on one wave in precision P: a[c] = x[64g + 16c + j]
// (n = 16 n1 + n2, n1 = 4g + c, n2 = j); z[r] = X[4r + g + 16 j]... with the
// D rows 4g + r: z[r] = X[4g + r + 16 j].  w[r] = W_256^{(4g + r) j} rounded
// to FP32 (core256p_twiddles).
__device__ __forceinline__ void core256p_twiddles(const f2 *__restrict__ tw32, int lane,
                                                  f2 (&w)[4]) {
  const int g = lane >> 4, j = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = tw32[((4 * g + r) * j) << 4];
}
template <int P>
__device__ __forceinline__ void core256p(f2 (&a)[4], const typename Stage<P>::Frag &f,
                                         const f2 (&w)[4], f2 (&z)[4]) {
  v4f yr, yi;
  Stage<P>::gemm(a, f, yr, yi);
  // D register r = Y[n2 = 4g + r][k1 = j]; twiddle W_256^{n2 k1}
  f2 b[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) b[r] = cmul(make_float2(yr[r], yi[r]), w[r]);
  // stage 2: A = b (row k1 = j, K = n2 = 4g + r), same fragment
  Stage<P>::gemm(b, f, yr, yi);
#pragma unroll
  for (int r = 0; r < 4; ++r) z[r] = make_float2(yr[r], yi[r]);
}
template <int P>
__device__ __forceinline__ void core256p(f2 (&a)[4], const typename Stage<P>::Frag &f,
                                         const f2 *__restrict__ tw32, int lane, f2 (&z)[4]) {
  f2 w[4];
  core256p_twiddles(tw32, lane, w);
  core256p<P>(a, f, w, z);
}

__device__ __forceinline__ int swz16(int e) { return e ^ ((e >> 4) & 15); }

template <int P, bool INV>
__global__ __launch_bounds__(64 * kWaves) void tensor256p_kernel(const c2 *in, c2 *out,
                                                                 int64_t batch,
                                                                 const c2 *__restrict__ tw,
                                                                 const f2 *__restrict__ tw32,
                                                                 double scale) {
  __shared__ f2 lds[kWaves * 256];
  const int lane = lane_id(), g = lane >> 4, j = lane & 15;
  const int64_t sw = wave_global() * kG256;
  if (sw >= batch) return;
  f2 *buf = lds + (threadIdx.x >> 6) * 256;
  const auto f = Stage<P>::frag(tw, lane);
  f2 au[kG256][4];
#pragma unroll
  for (int u = 0; u < kG256; ++u) {
    const int64_t s = sw + u < batch ? sw + u : sw;  // a short tail re-reads signal sw
#pragma unroll
    for (int c = 0; c < 4; ++c) au[u][c] = ld_in<INV>(in + s * 256 + 64 * g + 16 * c + j);
  }
#pragma unroll
  for (int u = 0; u < kG256; ++u) {
    f2 z[4];
    core256p<P>(au[u], f, tw32, lane, z);
    if (u > 0) wave_lds_sync();  // the previous unit's LDS reads are done
#pragma unroll
    for (int r = 0; r < 4; ++r) buf[swz16(4 * g + r + 16 * j)] = z[r];
    wave_lds_sync();
    if (sw + u < batch) {
      c2 *y = out + (sw + u) * 256;
#pragma unroll
      for (int c = 0; c < 4; ++c) st_out<INV>(y + 64 * c + lane, buf[swz16(64 * c + lane)], scale);
    }
  }
}

// LDS element of row k1, column m for the N = 4096 transpose (XOR swizzle).
__device__ __forceinline__ int swz_z(int k1, int m) { return k1 * 256 + (m ^ k1); }

// N = 4096 with stage A as an FP64 radix-16 butterfly on the vector ALUs
// (thread t of 256 owns column m = t; twiddles W_4096^{m k1} from the FP64
// table; the result is rounded to FP32 into LDS) and stage B, the two DFT16
// GEMM stages of each row's 256-point core, in precision P on the matrix
// cores (wave w: rows 4w .. 4w + 3), as the FP64 default does.  Each wave
// writes its results X[k1 + 16 k2] back into the LDS row k1 it alone read (k2
// permuted within the row), so two barriers suffice.
__device__ __forceinline__ int swz_row(int k1, int k2) {
  return k1 * 256 + (k2 ^ (((k2 >> 4) ^ k1) & 15));
}

template <int P, bool INV>
__global__ __launch_bounds__(256) void tensor4096ph_kernel(const c2 *in, c2 *out, int64_t batch,
                                                           const c2 *__restrict__ tw,
                                                           const f2 *__restrict__ tw32,
                                                           double scale) {
  __shared__ f2 lds[4096];
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int w = t >> 6;
  const int64_t s = blockIdx.x;
  if (s >= batch) return;  // uniform over the workgroup
  const c2 *x = in + s * 4096;
  // The column's twiddle bases before the column, whose loads go out at
  // raised priority (as in the FP64 kernel: loaded after the data, each
  // table wait would also wait for it; the other resident workgroups are
  // usually in their matrix stage).
  TwBase<16, c2> ta;
  ta.fetch(tw, t);
  __builtin_amdgcn_s_setprio(3);
  {
    c2 v[16];
#pragma unroll
    for (int n1 = 0; n1 < 16; ++n1) {
      c2 e = ld_stream(x + 256 * n1 + t);
      if constexpr (INV) e = cswap(e);
      v[n1] = e;
    }
    __builtin_amdgcn_s_setprio(0);
    dft<16>(v);
    apply_twiddle_base<16>(v, ta);  // v[k1] *= W_4096^{t k1}
#pragma unroll
    for (int k1 = 0; k1 < 16; ++k1) lds[swz_z(k1, t)] = make_float2((float)v[k1].x, (float)v[k1].y);
  }
  // the core's table values (L2-resident) load while the workgroup meets at
  // the barrier
  const auto f = Stage<P>::frag(tw, lane);
  f2 cw[4];
  core256p_twiddles(tw32, lane, cw);
  const int g = lane >> 4, j = lane & 15;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k1 = 4 * w + q;
    f2 a[4], z[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = lds[swz_z(k1, 64 * g + 16 * c + j)];
    core256p<P>(a, f, cw, z);
    // z[r] = X[k1 + 16 k2], k2 = 4g + r + 16j
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[swz_row(k1, 4 * g + r + 16 * j)] = z[r];
  }
  __syncthreads();
  // every output read from LDS before the first store (the compiler would
  // otherwise pair each read with its store and wait for that read alone)
  f2 o[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const int e = 256 * c + t;
    o[c] = lds[swz_row(e & 15, e >> 4)];
  }
  __builtin_amdgcn_sched_barrier(0);
  c2 *y = out + s * 4096;
#pragma unroll
  for (int c = 0; c < 16; ++c) st_out<INV>(y + 256 * c + t, o[c], scale);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

template <int P>
int launch_p(int n, const LaunchArgs &a) {
  const c2 *in = static_cast<const c2 *>(a.in);
  c2 *out = static_cast<c2 *>(a.out);
  const c2 *tw = static_cast<const c2 *>(a.twiddles);
  const f2 *tw32 = static_cast<const f2 *>(a.twiddles32);
  int64_t grid;
  int block;
  switch (n) {
    case 64: {  // persistent: up to the resident workgroups
      const int64_t resident = resident_workgroups<tensor64p_kernel<P, false>, 64 * kWaves>();
      grid = std::min<int64_t>(ceil_div(a.batch, 4 * kG64 * kWaves), resident);
      block = 64 * kWaves;
      break;
    }
    case 256: grid = ceil_div(a.batch, kG256 * kWaves); block = 64 * kWaves; break;
    case 4096: grid = a.batch; block = 256; break;
    default: return FFTOC_ERROR_UNSUPPORTED;
  }
  if (grid > 0x7fffffffLL) return FFTOC_ERROR_INVALID_BATCH;
  // FFTOC_WGS_PER_CU on the one-tile kernels: an unused dynamic LDS request
  // that leaves room for at most that many workgroups per CU.  Default at
  // N = 4096 from 8192 signals: two per CU for FP32 stages, three for split
  // BF16 (measured at 16384 signals: 0.727 / 0.707 of HBM against 0.683 /
  // 0.660 at the occupancy, four per CU).
  int per_cu = a.per_cu;
  if (per_cu == 0 && n == 4096 && a.batch >= 8192) per_cu = P == kFP32 ? 2 : 3;
  size_t lds = 0;
  if (n == 256) lds = occupancy_lds<tensor256p_kernel<P, false>, 64 * kWaves>(per_cu);
  if (n == 4096) lds = occupancy_lds<tensor4096ph_kernel<P, false>, 256>(per_cu);
  const dim3 gd((unsigned)grid), bd(block);
#define FFTOC_P_LAUNCH(K)                                                                      \
  do {                                                                                        \
    if (a.inverse)                                                                            \
      hipLaunchKernelGGL((K<P, true>), gd, bd, lds, a.stream, in, out, a.batch, tw, tw32,      \
                         a.scale);                                                            \
    else                                                                                      \
      hipLaunchKernelGGL((K<P, false>), gd, bd, lds, a.stream, in, out, a.batch, tw, tw32,     \
                         a.scale);                                                            \
  } while (0)
  if (n == 64) FFTOC_P_LAUNCH(tensor64p_kernel);
  else if (n == 256) FFTOC_P_LAUNCH(tensor256p_kernel);
  else FFTOC_P_LAUNCH(tensor4096ph_kernel);
#undef FFTOC_P_LAUNCH
  return hipGetLastError() == hipSuccess ? FFTOC_SUCCESS : FFTOC_ERROR_LAUNCH;
}

}  // namespace

int launch_tensor_precision(int n, const LaunchArgs &a) {
  if (a.c64) return FFTOC_ERROR_UNSUPPORTED;
  switch (a.precision) {
    case FFTOC_PREC_FP32: return launch_p<kFP32>(n, a);
    case FFTOC_PREC_BF16X3: return launch_p<kBF16X3>(n, a);
    default: return FFTOC_ERROR_UNSUPPORTED;
  }
}

}  // namespace fftoc
