// Device-side building blocks shared by the gfx950 FFT kernels: complex
// arithmetic, fixed-root multiplications and in-register DFT butterflies,
// generic over the complex type CT = double2 (complex128 data) or float2
// (complex64 data).
//
// Conventions: a complex value is {x = re, y = im}; every butterfly computes a
// forward DFT (exp(-2 pi i nk/R)) in place with natural-order output.
// Inverse transforms reuse the forward kernels by swapping re/im on load and
// on store (swap(fwd(swap(x))) = inverse DFT of x).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftoc {

using c2 = double2;
using f2 = float2;

template <class CT>
using real_t = decltype(CT{}.x);

// Twiddle table length: W_4096^j, j = 0..4095, covers every N <= 4096.
constexpr int kTwN = 4096;

// Correctly rounded constants of the 16th roots of unity in type R.
template <class R>
struct Roots {
  static constexpr R sqrt_half = (R)0.70710678118654752440L;  // cos(pi/4)
  static constexpr R cos8 = (R)0.92387953251128675613L;       // cos(pi/8)
  static constexpr R sin8 = (R)0.38268343236508977173L;       // sin(pi/8)
};

template <class CT>
__device__ __forceinline__ CT mk(real_t<CT> re, real_t<CT> im) {
  CT v;
  v.x = re;
  v.y = im;
  return v;
}

template <class CT>
__device__ __forceinline__ CT cadd(CT a, CT b) { return mk<CT>(a.x + b.x, a.y + b.y); }
template <class CT>
__device__ __forceinline__ CT csub(CT a, CT b) { return mk<CT>(a.x - b.x, a.y - b.y); }
template <class CT>
__device__ __forceinline__ CT cmul(CT a, CT w) {
  return mk<CT>(fma(a.x, w.x, -a.y * w.y), fma(a.x, w.y, a.y * w.x));
}
template <class CT>
__device__ __forceinline__ CT cswap(CT a) { return mk<CT>(a.y, a.x); }
template <class CT>
__device__ __forceinline__ CT cscale(CT a, real_t<CT> s) { return mk<CT>(a.x * s, a.y * s); }

// Multiplication by fixed roots W_16^j = exp(-2 pi i j / 16).
template <class CT>
__device__ __forceinline__ CT mul_mi(CT a) { return mk<CT>(a.y, -a.x); }  // * -i = W16^4
template <class CT>
__device__ __forceinline__ CT mul_w8_1(CT a) {  // * (1 - i)/sqrt2 = W16^2
  constexpr real_t<CT> r = Roots<real_t<CT>>::sqrt_half;
  return mk<CT>(r * (a.x + a.y), r * (a.y - a.x));
}
template <class CT>
__device__ __forceinline__ CT mul_w8_3(CT a) {  // * (-1 - i)/sqrt2 = W16^6
  constexpr real_t<CT> r = Roots<real_t<CT>>::sqrt_half;
  return mk<CT>(r * (a.y - a.x), -r * (a.x + a.y));
}
template <class CT>
__device__ __forceinline__ CT mul_w16_1(CT a) {  // * (c - i s)
  constexpr real_t<CT> c = Roots<real_t<CT>>::cos8, s = Roots<real_t<CT>>::sin8;
  return mk<CT>(fma(a.x, c, a.y * s), fma(a.y, c, -a.x * s));
}
template <class CT>
__device__ __forceinline__ CT mul_w16_3(CT a) {  // * (s - i c)
  constexpr real_t<CT> c = Roots<real_t<CT>>::cos8, s = Roots<real_t<CT>>::sin8;
  return mk<CT>(fma(a.x, s, a.y * c), fma(a.y, s, -a.x * c));
}
template <class CT>
__device__ __forceinline__ CT mul_w16_9(CT a) {  // * (-c + i s)
  constexpr real_t<CT> c = Roots<real_t<CT>>::cos8, s = Roots<real_t<CT>>::sin8;
  return mk<CT>(-fma(a.x, c, a.y * s), fma(a.x, s, -a.y * c));
}

// ---- in-register DFTs (natural order in, natural order out) ---------------
template <class CT>
__device__ __forceinline__ void dft2(CT &a0, CT &a1) {
  CT t = csub(a0, a1);
  a0 = cadd(a0, a1);
  a1 = t;
}

template <class CT>
__device__ __forceinline__ void dft4(CT &a0, CT &a1, CT &a2, CT &a3) {
  CT t0 = cadd(a0, a2), t1 = csub(a0, a2);
  CT t2 = cadd(a1, a3), t3 = mul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

// DFT8 = radix-2 DIT over two DFT4s: X[k] = E[k] + W8^k O[k].
template <class CT>
__device__ __forceinline__ void dft8(CT *a) {
  CT e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  CT o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  o1 = mul_w8_1(o1);
  o2 = mul_mi(o2);
  o3 = mul_w8_3(o3);
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

// DFT16 = 4 x 4: x[4 n1 + n2]; DFT4 over n1, twiddle W16^{n2 k1}, DFT4 over n2;
// X[k1 + 4 k2].
template <class CT>
__device__ __forceinline__ void dft16(CT *a) {
  CT y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    CT b0 = a[n2], b1 = a[4 + n2], b2 = a[8 + n2], b3 = a[12 + n2];
    dft4(b0, b1, b2, b3);
    y[4 * n2 + 0] = b0; y[4 * n2 + 1] = b1; y[4 * n2 + 2] = b2; y[4 * n2 + 3] = b3;
  }
  // y[4 n2 + k1] *= W16^{n2 k1}
  y[5] = mul_w16_1(y[5]);   // n2=1,k1=1 -> W^1
  y[6] = mul_w8_1(y[6]);    // W^2
  y[7] = mul_w16_3(y[7]);   // W^3
  y[9] = mul_w8_1(y[9]);    // n2=2,k1=1 -> W^2
  y[10] = mul_mi(y[10]);    // W^4
  y[11] = mul_w8_3(y[11]);  // W^6
  y[13] = mul_w16_3(y[13]); // n2=3,k1=1 -> W^3
  y[14] = mul_w8_3(y[14]);  // W^6
  y[15] = mul_w16_9(y[15]); // W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    CT b0 = y[k1], b1 = y[4 + k1], b2 = y[8 + k1], b3 = y[12 + k1];
    dft4(b0, b1, b2, b3);
    a[k1] = b0; a[k1 + 4] = b1; a[k1 + 8] = b2; a[k1 + 12] = b3;
  }
}

template <int R, class CT>
__device__ __forceinline__ void dft(CT *a) {
  if constexpr (R == 2) {
    dft2(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    dft8(a);
  } else {
    static_assert(R == 16, "radix must be 2, 4, 8 or 16");
    dft16(a);
  }
}

// The table values one radix-R butterfly's twiddles are built from: W^k,
// W^2k, W^4k, W^8k (as many as R needs), read at idx, 2 idx, 4 idx, 8 idx.
template <int R, class CT>
struct TwBase {
  static constexpr int K = R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4;
  CT w[K];
  __device__ __forceinline__ void fetch(const CT *__restrict__ tw, int idx) {
#pragma unroll
    for (int b = 0; b < K; ++b) w[b] = tw[idx << b];
  }
};

// Multiplies a[m] (m = 1..R-1) by W^{m k} given the base values: the powers
// that are not in the table are products of at most four of them (error a
// few ulp).
template <int R, class CT>
__device__ __forceinline__ void apply_twiddle_base(CT *a, const TwBase<R, CT> &b) {
  if constexpr (R == 2) {
    a[1] = cmul(a[1], b.w[0]);
  } else {
    const CT w1 = b.w[0], w2 = b.w[1];
    const CT w3 = cmul(w2, w1);
    a[1] = cmul(a[1], w1);
    a[2] = cmul(a[2], w2);
    a[3] = cmul(a[3], w3);
    if constexpr (R >= 8) {
      const CT w4 = b.w[2];
      a[4] = cmul(a[4], w4);
      a[5] = cmul(a[5], cmul(w4, w1));
      a[6] = cmul(a[6], cmul(w4, w2));
      const CT w7 = cmul(w4, w3);
      a[7] = cmul(a[7], w7);
      if constexpr (R == 16) {
        const CT w8 = b.w[3];
        a[8] = cmul(a[8], w8);
        a[9] = cmul(a[9], cmul(w8, w1));
        a[10] = cmul(a[10], cmul(w8, w2));
        a[11] = cmul(a[11], cmul(w8, w3));
        a[12] = cmul(a[12], cmul(w8, w4));
        a[13] = cmul(a[13], cmul(w8, cmul(w4, w1)));
        a[14] = cmul(a[14], cmul(w8, cmul(w4, w2)));
        a[15] = cmul(a[15], cmul(w8, w7));
      }
    }
  }
}

// Multiplies a[m] (m = 1..R-1) by W^{m k}, where tw[idx * m] = W^{m k}.
template <int R, class CT>
__device__ __forceinline__ void apply_twiddles(CT *a, const CT *__restrict__ tw, int idx) {
  TwBase<R, CT> b;
  b.fetch(tw, idx);
  apply_twiddle_base<R>(a, b);
}

// DPP control words (GFX9 encoding).  row_ror:n makes lane l of a 16-lane
// row read lane (l - n) mod 16 of the same row.
constexpr int kDppQuadXor1 = 0xB1;  // quad_perm [1,0,3,2]: lane l reads l ^ 1
constexpr int kDppQuadXor2 = 0x4E;  // quad_perm [2,3,0,1]: lane l reads l ^ 2
constexpr int kDppRowRor4 = 0x124;  // reads (l - 4) mod 16
constexpr int kDppRowRor8 = 0x128;  // reads (l + 8) mod 16 = l ^ 8
constexpr int kDppRowRor12 = 0x12C; // reads (l + 4) mod 16

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xF, 0xF, true);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int CTRL, class CT>
__device__ __forceinline__ CT dpp_c(CT v) {
  return mk<CT>(dpp<CTRL>(v.x), dpp<CTRL>(v.y));
}

// Streaming (non-temporal) global access for data touched exactly once:
// keeps the input/output stream from evicting the twiddle table and other
// reused lines from L2 / MALL.
template <class CT>
__device__ __forceinline__ CT ld_stream(const CT *p) {
  typedef real_t<CT> R;
  typedef R vec2 __attribute__((ext_vector_type(2)));
  const vec2 v = __builtin_nontemporal_load(reinterpret_cast<const vec2 *>(p));
  return mk<CT>(v.x, v.y);
}
template <class CT>
__device__ __forceinline__ void st_stream(CT *p, CT v) {
  typedef real_t<CT> R;
  typedef R vec2 __attribute__((ext_vector_type(2)));
  const vec2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<vec2 *>(p));
}

// Marks where a prefetched value is first used: the compiler may not move a
// use above this point (it would wait for the load there), and the value stays
// in its registers until here.
template <class CT>
__device__ __forceinline__ void pin_vgpr(CT &v) {
  asm volatile("" : "+v"(v.x), "+v"(v.y));
}

// A wave-uniform 64-bit value in scalar registers.  readfirstlane returns a
// signed 32-bit int: each half goes through uint32_t, or the low half would
// sign-extend into the high one when its bit 31 is set.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((unsigned)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Wave-scope ordering point for LDS exchanges among the lanes of one wave:
// LDS operations of a wave execute in issue order, so only compiler
// reordering has to be prevented.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace fftoc
