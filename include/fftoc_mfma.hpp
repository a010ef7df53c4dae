// FP64 matrix-core (MFMA) building blocks of the tensor FFT path for gfx950:
// DFT16 stages as dense complex 16 x 16 matrix products on
// v_mfma_f64_16x16x4_f64, in two contractions (Karatsuba's three real
// products, and the two-product form that packs the DFT matrix's distinct
// cosine and sine columns into one real matrix), and the 256-point two-stage
// core built from them.  Shared by the library kernels
// (fft_onchip_amd/csrc/mfma_kernels.hip) and the in-LDS functors of
// fftoc_lds.hpp.
//
// CDNA4 operand layout of v_mfma_f64_16x16x4_f64 (wave64, lane l, g = l/16,
// j = l%16), measured on MI355X:  A[i][k]: lane l holds A[j][g];
// B[k][col]: lane l holds B[g][j];  D[i][col]: lane l register r holds
// D[4r + g][j].
//
// Twiddles and DFT-matrix entries come from the library's FP64 table
// W_4096^q (fftoc_twiddle_table); only additions are done in the kernels.
#pragma once

#include "fftoc_device.hpp"

namespace fftoc {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4d vzero() { return v4d{0.0, 0.0, 0.0, 0.0}; }

// B fragment of the 16x16 DFT matrix M[k][col] = W_16^{k*col} in the
// Karatsuba form; K-chunk c, slot g -> k = 4c + g.
struct Frag {
  double r[4], i[4], s[4];
};

__device__ __forceinline__ Frag dft16_frag(const c2 *__restrict__ tw, int lane) {
  const int g = lane >> 4, j = lane & 15;
  Frag f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = 4 * c + g;
    const c2 w = tw[((k * j) & 15) << 8];  // W_16^{kj} = W_4096^{256 kj}
    f.r[c] = w.x;
    f.i[c] = w.y;
    f.s[c] = w.x + w.y;
  }
  return f;
}

// Block-diagonal diag(F8, F8) with K order k = 4c + g.
__device__ __forceinline__ Frag dft8x2_frag(const c2 *__restrict__ tw, int lane) {
  const int g = lane >> 4, j = lane & 15;
  Frag f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = 4 * c + g;
    if ((k < 8) == (j < 8)) {
      const c2 w = tw[(((k & 7) * (j & 7)) & 7) << 9];  // W_8^{..} = W_4096^{512 ..}
      f.r[c] = w.x;
      f.i[c] = w.y;
      f.s[c] = w.x + w.y;
    } else {
      f.r[c] = 0.0;
      f.i[c] = 0.0;
      f.s[c] = 0.0;
    }
  }
  return f;
}

// Complex 16x16x16 GEMM Y = X * M with 3 real MFMA chains (Karatsuba).
// a[c] = A operand of K-chunk c as (re, im).
__device__ __forceinline__ void kgemm(const c2 (&a)[4], const Frag &f, v4d &yr, v4d &yi) {
  v4d p1 = vzero(), p2 = vzero(), p3 = vzero();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    p1 = mfma16(a[c].x, f.r[c], p1);
    p2 = mfma16(a[c].y, f.i[c], p2);
    p3 = mfma16(a[c].x + a[c].y, f.s[c], p3);
  }
  yr = p1 - p2;
  yi = p3 - p1 - p2;
}

// ---------------------------------------------------------------------------
// Two-product DFT16 contraction.  The forward DFT16 matrix is C - iS with
// C[k][m] = cos(2 pi k m / 16) even and S[k][m] = sin(2 pi k m / 16) odd in m,
// so its 16 columns hold only 9 distinct cosine columns (m = 0..8) and 7
// distinct sine columns (m = 1..7).  Packed side by side they make one real
// 16 x 16 matrix P (columns 0..8 = C[.][0..8], columns 9..15 = S[.][1..7]),
// and a complex product X * F needs two real MFMA products, Xr P and Xi P,
// instead of Karatsuba's three:
//   X[m]      = (Cr + Si) + i (Ci - Sr)      m = 1..7
//   X[16 - m] = (Cr - Si) + i (Ci + Sr)      m = 1..7
//   X[0], X[8] = the cosine sums alone,
// with Cr, Ci (Sr, Si) the cosine (sine) sums of the real and imaginary parts.
// P is symmetric in role: lane l holds P[4c + g][j] in K-chunk c both as the
// B operand (data in A: packed frequency = column = lane j) and as the A
// operand (data in B, P^T: packed frequency = row 4r + g = register r).
struct HFrag {
  double p[4];
};

__device__ __forceinline__ HFrag hdft16_frag(const c2 *__restrict__ tw, int lane) {
  const int g = lane >> 4, j = lane & 15;
  const int m = j <= 8 ? j : j - 8;
  HFrag f;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const c2 w = tw[(((4 * c + g) * m) & 15) << 8];  // W_16^{km} = cos - i sin
    f.p[c] = j <= 8 ? w.x : -w.y;
  }
  return f;
}

// Frequency of packed column j after the combine: j for the cosine columns,
// 16 - (j - 8) for the sine columns.
__device__ __forceinline__ int hphi(int j) { return j <= 8 ? j : 24 - j; }

// Per-lane constants of the data-in-A combine: the lane's result is
//   re = a u + b u' + c v + d v',  im = a v + b v' - c u - d u'
// with (u, v) its own sums and (u', v') those of lane l ^ 8 (the other half of
// its frequency).  Every coefficient is 0 or +-1, so the sums are exact.
struct HCoef {
  double a, b, c, d;
};
__device__ __forceinline__ HCoef hcoef(int j) {
  HCoef k;
  const bool cosine = j <= 8, edge = (j & 7) == 0;  // j = 0, 8: no sine partner
  k.a = cosine ? 1.0 : 0.0;
  k.b = cosine ? 0.0 : 1.0;
  k.c = cosine ? 0.0 : -1.0;
  k.d = (cosine && !edge) ? 1.0 : 0.0;
  return k;
}

// DFT16 over K with the data as the A operand (rows = independent
// transforms, K chunk c = k = 4c + g): y[r] = X[row 4r + g][hphi(j)].
__device__ __forceinline__ void hgemm_a(const c2 (&a)[4], const HFrag &f, const HCoef &k,
                                        c2 (&y)[4]) {
  v4d u = vzero(), v = vzero();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    u = mfma16(a[c].x, f.p[c], u);
    v = mfma16(a[c].y, f.p[c], v);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double up = dpp<kDppRowRor8>(u[r]), vp = dpp<kDppRowRor8>(v[r]);
    y[r].x = fma(k.d, vp, fma(k.c, v[r], fma(k.b, up, k.a * u[r])));
    y[r].y = fma(-k.d, up, fma(-k.c, u[r], fma(k.b, vp, k.a * v[r])));
  }
}

// DFT16 over K with the data as the B operand (columns = independent
// transforms, K chunk c = k = 4c + g): o[q] = X[hk2(g, q)][column j].
__device__ __forceinline__ void hgemm_b(const HFrag &f, const c2 (&b)[4], int g, c2 (&o)[4]) {
  v4d u = vzero(), v = vzero();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    u = mfma16(f.p[c], b[c].x, u);
    v = mfma16(f.p[c], b[c].y, v);
  }
  // rows 4r + g: r = 0 cosine g, r = 1 cosine 4 + g, r = 2 sine g (cosine 8
  // when g = 0), r = 3 sine 4 + g
  const double e = g ? 1.0 : 0.0;
  o[0] = make_double2(fma(e, v[2], u[0]), fma(-e, u[2], v[0]));
  const c2 t = make_double2(u[0] - v[2], v[0] + u[2]);
  o[1] = g ? t : make_double2(u[2], v[2]);
  o[2] = make_double2(u[1] + v[3], v[1] - u[3]);
  o[3] = make_double2(u[1] - v[3], v[1] + u[3]);
}
// Frequency of hgemm_b's output o[q] in a lane of group g.
__device__ __forceinline__ int hk2(int g, int q) {
  return q == 0 ? g : q == 1 ? (g ? 16 - g : 8) : q == 2 ? 4 + g : 12 - g;
}

// A DFT16 contraction with the data as the A operand, as a type: load() its
// per-lane table values, run(a, y) -> y[r] = X[row 4r + g][frequency col].
struct KaratsubaDft16 {
  // frequency held by output column (lane j) or row i of a product
  __device__ static __forceinline__ int freq(int i) { return i; }
  Frag f;
  int col;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = dft16_frag(tw, lane);
    col = lane & 15;
  }
  __device__ __forceinline__ void run(const c2 (&a)[4], c2 (&y)[4]) const {
    v4d yr, yi;
    kgemm(a, f, yr, yi);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] = make_double2(yr[r], yi[r]);
  }
};
struct TwoProductDft16 {
  __device__ static __forceinline__ int freq(int i) { return hphi(i); }
  HFrag f;
  HCoef k;
  int col;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = hdft16_frag(tw, lane);
    k = hcoef(lane & 15);
    col = hphi(lane & 15);
  }
  __device__ __forceinline__ void run(const c2 (&a)[4], c2 (&y)[4]) const { hgemm_a(a, f, k, y); }
};

// ---------------------------------------------------------------------------
// 256-point core on one wave: input a[c] = x[64c + l] (n = 16 n1 + n2 with
// n1 = 4c + g, n2 = j); output z[r] = X[4r + g + 16 j].
// The inter-stage twiddles of core256 for this lane: w[r] = W_256^{n2 k1},
// n2 = 4r + g, k1 = j.
__device__ __forceinline__ void core256_twiddles(const c2 *__restrict__ tw, int lane, c2 (&w)[4]) {
  const int g = lane >> 4, j = lane & 15;
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = tw[((4 * r + g) * j) << 4];
}

__device__ __forceinline__ void core256(c2 (&a)[4], const Frag &f, const c2 (&w)[4], c2 (&z)[4]) {
  v4d yr, yi;
  kgemm(a, f, yr, yi);
  // D: reg r = Y[n2 = 4r + g][k1 = j]; twiddle W_256^{n2 k1}
  c2 b[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) b[r] = cmul(make_double2(yr[r], yi[r]), w[r]);
  // stage 2: A chunk r = b[r] (rows = k1 = j, K = n2 = 4r + g)
  kgemm(b, f, yr, yi);
#pragma unroll
  for (int r = 0; r < 4; ++r) z[r] = make_double2(yr[r], yi[r]);
}

__device__ __forceinline__ void core256(c2 (&a)[4], const Frag &f, const c2 *__restrict__ tw,
                                        int lane, c2 (&z)[4]) {
  c2 w[4];
  core256_twiddles(tw, lane, w);
  core256(a, f, w, z);
}

// 256-point core with the two-product contraction: input a[c] = x[64c + l]
// (n = 16 n1 + n2, n1 = 4c + g, n2 = j); stage 1 with the data in A gives
// Y[n2 = 4r + g][k1 = hphi(j)] in register r, which after the twiddle
// W_256^{n2 k1} is already the B operand of stage 2 (K = n2 = 4r + g, column
// j); output z[q] = X[hphi(j) + 16 hk2(g, q)].  16 MFMAs against 24.
__device__ __forceinline__ void core256h_twiddles(const c2 *__restrict__ tw, int lane,
                                                  c2 (&w)[4]) {
  const int g = lane >> 4, k1 = hphi(lane & 15);
#pragma unroll
  for (int r = 0; r < 4; ++r) w[r] = tw[((4 * r + g) * k1) << 4];
}

__device__ __forceinline__ void core256h(const c2 (&a)[4], const HFrag &f, const HCoef &k,
                                         const c2 (&w)[4], int g, c2 (&z)[4]) {
  c2 y[4];
  hgemm_a(a, f, k, y);
#pragma unroll
  for (int r = 0; r < 4; ++r) y[r] = cmul(y[r], w[r]);
  hgemm_b(f, y, g, z);
}

// A 256-point core as a type: its per-lane table values (load, once per
// kernel), the transform of one row (run: a[c] = x[64c + l] in, z[q] out) and
// the row position k1 + 16 k2 of z[q] (out_index).
struct KaratsubaCore {  // three real products per DFT16 (24 MFMAs per core)
  Frag f;
  c2 w[4];
  int g, j;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = dft16_frag(tw, lane);
    core256_twiddles(tw, lane, w);
    g = lane >> 4;
    j = lane & 15;
  }
  __device__ __forceinline__ void run(c2 (&a)[4], c2 (&z)[4]) const { core256(a, f, w, z); }
  __device__ __forceinline__ int out_index(int q) const { return 4 * q + g + 16 * j; }
  // out_index(q) of the lane with group g and column j
  __device__ static __forceinline__ int out_k2(int g, int q, int j) { return 4 * q + g + 16 * j; }
};
struct TwoProductCore {  // two real products per DFT16 (16 MFMAs per core)
  HFrag f;
  HCoef k;
  c2 w[4];
  int g, j;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = hdft16_frag(tw, lane);
    g = lane >> 4;
    j = lane & 15;
    k = hcoef(j);
    core256h_twiddles(tw, lane, w);
  }
  __device__ __forceinline__ void run(c2 (&a)[4], c2 (&z)[4]) const { core256h(a, f, k, w, g, z); }
  __device__ __forceinline__ int out_index(int q) const { return hphi(j) + 16 * hk2(g, q); }
  __device__ static __forceinline__ int out_k2(int g, int q, int j) { return hphi(j) + 16 * hk2(g, q); }
};

// Exchanges complex values between the lane pairs (l, l + D), D = 32 (wave
// halves) or 16 (rows 2m, 2m + 1), with one v_permlane{32,16}_swap per dword
// (p's upper lanes swapped with q's lower lanes): afterwards both lanes of a
// pair hold in p the lower lane's value and in q the upper lane's value, of
// register p in the lower lane and of register q in the upper lane.
template <int D>
__device__ __forceinline__ void swap_halves(double &p, double &q) {
  static_assert(D == 16 || D == 32, "lane distance 16 or 32");
  const unsigned long long pu = __builtin_bit_cast(unsigned long long, p);
  const unsigned long long qu = __builtin_bit_cast(unsigned long long, q);
  unsigned plo = (unsigned)pu, phi = (unsigned)(pu >> 32);
  unsigned qlo = (unsigned)qu, qhi = (unsigned)(qu >> 32);
  if constexpr (D == 32) {
    const auto rl = __builtin_amdgcn_permlane32_swap(plo, qlo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(phi, qhi, false, false);
    plo = rl[0], qlo = rl[1], phi = rh[0], qhi = rh[1];
  } else {
    const auto rl = __builtin_amdgcn_permlane16_swap(plo, qlo, false, false);
    const auto rh = __builtin_amdgcn_permlane16_swap(phi, qhi, false, false);
    plo = rl[0], qlo = rl[1], phi = rh[0], qhi = rh[1];
  }
  p = __builtin_bit_cast(double, ((unsigned long long)phi << 32) | plo);
  q = __builtin_bit_cast(double, ((unsigned long long)qhi << 32) | qlo);
}
template <int D>
__device__ __forceinline__ void swap_halves(c2 &p, c2 &q) {
  swap_halves<D>(p.x, q.x);
  swap_halves<D>(p.y, q.y);
}

// DFT16 over n = 4r + g held as register r of the lanes j, j + 16, j + 32,
// j + 48 (g = 0..3), on the vector ALUs as 4 x 4: a DFT4 over r in the lane,
// the twiddle W_16^{g ka} (t16[ka - 1]), and a DFT4 over g across the lanes
// (radix-2 over g1 = g / 2 between wave halves, then the twiddle
// W_4^{g0 kb0} and radix-2 over g0 = g % 2 between rows, each exchange one
// permlane swap per dword).  Register q of lane (g, j) then holds
// X[lanes_k(g, q)].
__device__ __forceinline__ int lanes_k(int g, int q) {
  return (g >> 1) + 2 * (q >> 1) + 4 * (g & 1) + 8 * (q & 1);
}
__device__ __forceinline__ void dft16_lanes(c2 (&y)[4], const c2 (&t16)[3], int g) {
  dft4(y[0], y[1], y[2], y[3]);  // over r: y[ka]
#pragma unroll
  for (int ka = 1; ka < 4; ++ka) y[ka] = cmul(y[ka], t16[ka - 1]);
  // radix-2 over g1: lanes l and l + 32 (registers paired (ka, ka + 1); the
  // lower lane keeps ka, the upper one ka + 1)
  swap_halves<32>(y[0], y[1]);
  swap_halves<32>(y[2], y[3]);
  dft2(y[0], y[1]);
  dft2(y[2], y[3]);
  // radix-2 over g0: rows 2m and 2m + 1; the odd row's kb0 = 1 value times -i
  const bool odd = g & 1;
  swap_halves<16>(y[0], y[1]);
  swap_halves<16>(y[2], y[3]);
#pragma unroll
  for (int p = 1; p < 4; p += 2) {
    // selects per double (a select of two c2 values compiled to a select
    // of their stack addresses)
    const double re = y[p].x, im = y[p].y;
    y[p].x = odd ? im : re;
    y[p].y = odd ? -re : im;
  }
  dft2(y[0], y[1]);
  dft2(y[2], y[3]);
}
// The lane's twiddles of dft16_lanes: t16[ka - 1] = W_16^{g ka}.
__device__ __forceinline__ void dft16_lanes_twiddles(const c2 *__restrict__ tw, int g,
                                                     c2 (&t16)[3]) {
#pragma unroll
  for (int ka = 1; ka < 4; ++ka) t16[ka - 1] = tw[((g * ka) & 15) << 8];
}

// 256-point core with one matrix-core stage: stage 1 (DFT16 over n1) on the
// matrix cores with Karatsuba's three real products, as in core256; stage 2
// (DFT16 over n2 = 4 r + g for each k1 = j) on the vector ALUs as 4 x 4: a
// DFT4 over r in the lane, the twiddle W_16^{g ka}, and a DFT4 over g across
// the lanes j, j + 16, j + 32, j + 48 (radix-2 over g1 = g / 2 between wave
// halves, then the twiddle W_4^{g0 kb0} and radix-2 over g0 = g % 2 between
// rows, each exchange one permlane swap per dword).  12 MFMAs per core
// instead of 24: FP64 matrix products on random data draw enough power to
// lower the clock, and the FFT factorization of stage 2 needs a quarter of
// the dense product's flops.  Output z[q] = X[j + 16 k2], k2 = g1 + 2 (q / 2)
// + 4 g0 + 8 (q % 2).
struct KaratsubaHalfCore {
  Frag f;
  c2 w[4];    // W_256^{(4r + g) j}
  c2 t16[3];  // W_16^{g ka}, ka = 1..3
  int g, j;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = dft16_frag(tw, lane);
    core256_twiddles(tw, lane, w);
    g = lane >> 4;
    j = lane & 15;
    dft16_lanes_twiddles(tw, g, t16);
  }
  __device__ __forceinline__ void run(c2 (&a)[4], c2 (&z)[4]) const {
    v4d yr, yi;
    kgemm(a, f, yr, yi);
#pragma unroll
    for (int r = 0; r < 4; ++r) z[r] = cmul(make_double2(yr[r], yi[r]), w[r]);
    dft16_lanes(z, t16, g);
  }
  __device__ static __forceinline__ int out_k2(int g, int q, int j) { return j + 16 * lanes_k(g, q); }
  __device__ __forceinline__ int out_index(int q) const { return out_k2(g, q, j); }
};

// The one-matrix-stage core with the two-product contraction on its matrix
// stage: stage 1 gives Y[n2 = 4r + g][k1 = hphi(j)] (hgemm_a), twiddled by
// W_256^{n2 k1}, then the same VALU stage 2.  8 MFMAs per core.
struct TwoProductHalfCore {
  HFrag f;
  HCoef k;
  c2 w[4];    // W_256^{(4r + g) hphi(j)}
  c2 t16[3];  // W_16^{g ka}
  int g, j;
  __device__ __forceinline__ void load(const c2 *__restrict__ tw, int lane) {
    f = hdft16_frag(tw, lane);
    g = lane >> 4;
    j = lane & 15;
    k = hcoef(j);
    core256h_twiddles(tw, lane, w);
    dft16_lanes_twiddles(tw, g, t16);
  }
  __device__ __forceinline__ void run(c2 (&a)[4], c2 (&z)[4]) const {
    hgemm_a(a, f, k, z);
#pragma unroll
    for (int r = 0; r < 4; ++r) z[r] = cmul(z[r], w[r]);
    dft16_lanes(z, t16, g);
  }
  __device__ static __forceinline__ int out_k2(int g, int q, int j) {
    return hphi(j) + 16 * lanes_k(g, q);
  }
  __device__ __forceinline__ int out_index(int q) const { return out_k2(g, q, j); }
};

// LDS element swizzle for the output transposes: XOR the low 4 bits of the
// element index with bits 4..7 (16 x 16-byte slots per 256 B bank window).
__device__ __forceinline__ int swz16(int e) { return e ^ ((e >> 4) & 15); }

}  // namespace fftoc
