/*
 * fftoc_oracle.c -- CPU oracle for the batched small complex FFT.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against; it is never part of the product path.  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 *
 * Every transform here computes the forward DFT
 *     X[k] = sum_n x[n] * exp(-2*pi*i*n*k/N)        (or exp(+...) if inverse)
 * of contiguous interleaved (re, im) float64 signals, the data layout of the
 * reference (includes/config.hpp:6, thrust::complex<double>, signal-major).
 *
 * Functions (each is pinned by the tests, see tests/test_oracle.py):
 *   oracle_dft        O(N^2) brute force, long double      -- golden fixtures,
 *                                                             closed forms
 *   oracle_fft        radix-2 DIT, long double              -- vs oracle_dft
 *   oracle_fft_f64    radix-4/2 Stockham, float64, OpenMP   -- vs oracle_dft;
 *                                                             the CPU baseline
 *   oracle_legacy8    radix-8 Stockham, restating
 *                     includes/legacy8_fft.cuh:57-118       -- vs oracle_dft
 *   oracle_legacy16   radix-16 Stockham, restating
 *                     includes/legacy16_fft.cuh:99-172      -- vs oracle_dft
 *   oracle_karatsuba  DFT-as-matrix stages with 3-real-product complex GEMMs,
 *                     restating includes/c64_fft64.cuh:8-27 and
 *                     includes/tensor_utils.cuh:25-36 for any radix split
 *                                                           -- vs oracle_dft
 *
 * Long-double results are returned as a double-double pair (hi, lo) with
 * hi = (double)v and lo = (double)(v - hi), so callers keep ~64 bits.
 * Twiddles are reduced to the first octant before sinl/cosl so every table
 * entry is accurate to about one long-double ulp (1e-19).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef long double ld;

static const ld PI_L = 3.141592653589793238462643383279502884L;

/* exp(sign * 2*pi*i * q / n), q reduced mod n and then to the first octant. */
static void twiddle_ld(int64_t n, int64_t q, int sign, ld *re, ld *im) {
  q %= n;
  if (q < 0) q += n;
  /* angle = 2*pi*q/n; use 8q vs n comparisons to stay in exact integers */
  int64_t q8 = 8 * q;
  ld c, s;
  /* reduce to [0, pi/4] using symmetries of (cos, sin) */
  int64_t oct = q8 / n; /* 0..7 */
  int64_t rem = q8 - oct * n; /* 0..n-1, angle within octant = rem/(8n) * 2pi */
  ld a, ca, sa;
  /* within octant o, base angle o*pi/4 plus t = 2*pi*rem/(8n) */
  if (oct & 1) {
    /* odd octant: use reflected angle pi/4 - t measured from (o+1)*pi/4 */
    a = PI_L * (ld)(n - rem) / (ld)(4 * n);
  } else {
    a = PI_L * (ld)rem / (ld)(4 * n);
  }
  ca = cosl(a);
  sa = sinl(a);
  switch (oct) {
  case 0: c = ca; s = sa; break;             /* t */
  case 1: c = sa; s = ca; break;             /* pi/2 - a */
  case 2: c = -sa; s = ca; break;            /* pi/2 + a */
  case 3: c = -ca; s = sa; break;            /* pi - a */
  case 4: c = -ca; s = -sa; break;           /* pi + a */
  case 5: c = -sa; s = -ca; break;           /* 3pi/2 - a */
  case 6: c = sa; s = -ca; break;            /* 3pi/2 + a */
  default: c = ca; s = -sa; break;           /* 2pi - a */
  }
  /* forward transform uses exp(-i*angle) */
  *re = c;
  *im = (sign < 0) ? -s : s;
}

/* Exported for the tests: twiddle table exp(sign*2*pi*i*j/n), j = 0..n-1. */
int oracle_twiddles(int64_t n, int sign, double *out_hi, double *out_lo) {
  if (n <= 0) return 1;
  for (int64_t j = 0; j < n; ++j) {
    ld re, im;
    twiddle_ld(n, j, sign, &re, &im);
    double h0 = (double)re, h1 = (double)im;
    out_hi[2 * j] = h0;
    out_hi[2 * j + 1] = h1;
    if (out_lo) {
      out_lo[2 * j] = (double)(re - (ld)h0);
      out_lo[2 * j + 1] = (double)(im - (ld)h1);
    }
  }
  return 0;
}

static ld *twiddle_table_ld(int64_t n, int sign) {
  ld *t = (ld *)malloc(sizeof(ld) * 2 * (size_t)n);
  if (!t) return NULL;
  for (int64_t j = 0; j < n; ++j) twiddle_ld(n, j, sign, &t[2 * j], &t[2 * j + 1]);
  return t;
}

static void store_dd(const ld *v, int64_t count, double *hi, double *lo) {
  for (int64_t i = 0; i < count; ++i) {
    double h = (double)v[i];
    hi[i] = h;
    if (lo) lo[i] = (double)(v[i] - (ld)h);
  }
}

static int is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

/* ------------------------------------------------------------------------ */
/* O(N^2) brute-force DFT in long double.                                    */
int oracle_dft(int64_t n, int64_t batch, const double *in, double *out_hi,
               double *out_lo, int inverse) {
  if (n <= 0 || batch < 0) return 1;
  int sign = inverse ? +1 : -1;
  ld *w = twiddle_table_ld(n, sign);
  ld *acc = (ld *)malloc(sizeof(ld) * 2 * (size_t)n);
  if (!w || !acc) { free(w); free(acc); return 2; }
  for (int64_t b = 0; b < batch; ++b) {
    const double *x = in + 2 * n * b;
    for (int64_t k = 0; k < n; ++k) {
      ld sr = 0, si = 0;
      int64_t idx = 0;
      for (int64_t j = 0; j < n; ++j) {
        ld xr = x[2 * j], xi = x[2 * j + 1];
        ld wr = w[2 * idx], wi = w[2 * idx + 1];
        sr += xr * wr - xi * wi;
        si += xr * wi + xi * wr;
        idx += k;
        if (idx >= n) idx -= n;
      }
      acc[2 * k] = sr;
      acc[2 * k + 1] = si;
    }
    store_dd(acc, 2 * n, out_hi + 2 * n * b, out_lo ? out_lo + 2 * n * b : NULL);
  }
  free(w);
  free(acc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Radix-2 decimation-in-time FFT in long double (bit-reversed input).       */
static void fft_radix2_ld(int64_t n, ld *a, const ld *w /* n entries */) {
  /* bit reversal permutation */
  for (int64_t i = 1, j = 0; i < n; ++i) {
    int64_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      ld tr = a[2 * i], ti = a[2 * i + 1];
      a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
      a[2 * j] = tr; a[2 * j + 1] = ti;
    }
  }
  for (int64_t len = 2; len <= n; len <<= 1) {
    int64_t half = len >> 1, step = n / len;
    for (int64_t s = 0; s < n; s += len) {
      for (int64_t j = 0; j < half; ++j) {
        ld wr = w[2 * j * step], wi = w[2 * j * step + 1];
        ld *u = a + 2 * (s + j), *v = a + 2 * (s + j + half);
        ld vr = v[0] * wr - v[1] * wi, vi = v[0] * wi + v[1] * wr;
        v[0] = u[0] - vr; v[1] = u[1] - vi;
        u[0] += vr; u[1] += vi;
      }
    }
  }
}

int oracle_fft(int64_t n, int64_t batch, const double *in, double *out_hi,
               double *out_lo, int inverse) {
  if (!is_pow2(n) || batch < 0) return 1;
  ld *w = twiddle_table_ld(n, inverse ? +1 : -1);
  ld *a = (ld *)malloc(sizeof(ld) * 2 * (size_t)n);
  if (!w || !a) { free(w); free(a); return 2; }
  for (int64_t b = 0; b < batch; ++b) {
    for (int64_t i = 0; i < 2 * n; ++i) a[i] = in[2 * n * b + i];
    fft_radix2_ld(n, a, w);
    store_dd(a, 2 * n, out_hi + 2 * n * b, out_lo ? out_lo + 2 * n * b : NULL);
  }
  free(w);
  free(a);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Float64 Stockham autosort FFT (radix 4 passes, one radix-2 pass when      */
/* log2 N is odd).  Used as the timed CPU baseline.  Parallel over the batch */
/* with OpenMP; twiddles are long-double values rounded once to float64.     */
typedef struct { double re, im; } c64;

static void stockham_f64(int64_t n, const c64 *x, c64 *y, c64 *tmp,
                         const c64 *w /* n entries */) {
  const c64 *src = x;
  c64 *bufs[2] = {tmp, y};
  int64_t l = 1; /* product of radices done so far */
  int npass = 0;
  {
    int64_t m = n;
    while (m > 1) { m = (m % 4 == 0) ? m / 4 : m / 2; ++npass; }
  }
  int pass = 0;
  while (l < n) {
    int64_t r = ((n / l) % 4 == 0) ? 4 : 2;
    /* choose destination so that the final pass lands in y */
    c64 *dst = bufs[(npass - 1 - pass) % 2 == 0 ? 1 : 0];
    int64_t m = n / r; /* butterflies */
    int64_t tstep = n / (r * l);
    for (int64_t j = 0; j < m; ++j) {
      int64_t k = j % l;
      int64_t base = (j - k) * r + k;
      if (r == 4) {
        c64 a0 = src[j], a1 = src[j + m], a2 = src[j + 2 * m], a3 = src[j + 3 * m];
        if (k) {
          c64 w1 = w[k * tstep], w2 = w[2 * k * tstep], w3 = w[3 * k * tstep];
          c64 t;
          t.re = a1.re * w1.re - a1.im * w1.im; t.im = a1.re * w1.im + a1.im * w1.re; a1 = t;
          t.re = a2.re * w2.re - a2.im * w2.im; t.im = a2.re * w2.im + a2.im * w2.re; a2 = t;
          t.re = a3.re * w3.re - a3.im * w3.im; t.im = a3.re * w3.im + a3.im * w3.re; a3 = t;
        }
        /* forward radix-4: W4 = -i (sign folded into the table for inverse) */
        double s = w[n / 4].im; /* -1 for forward, +1 for inverse */
        c64 b0 = {a0.re + a2.re, a0.im + a2.im};
        c64 b1 = {a0.re - a2.re, a0.im - a2.im};
        c64 b2 = {a1.re + a3.re, a1.im + a3.im};
        c64 b3 = {a1.re - a3.re, a1.im - a3.im};
        /* b3 * (s*i) */
        c64 b3r = {-s * b3.im, s * b3.re};
        dst[base] = (c64){b0.re + b2.re, b0.im + b2.im};
        dst[base + l] = (c64){b1.re + b3r.re, b1.im + b3r.im};
        dst[base + 2 * l] = (c64){b0.re - b2.re, b0.im - b2.im};
        dst[base + 3 * l] = (c64){b1.re - b3r.re, b1.im - b3r.im};
      } else {
        c64 a0 = src[j], a1 = src[j + m];
        if (k) {
          c64 w1 = w[k * tstep], t;
          t.re = a1.re * w1.re - a1.im * w1.im; t.im = a1.re * w1.im + a1.im * w1.re; a1 = t;
        }
        dst[base] = (c64){a0.re + a1.re, a0.im + a1.im};
        dst[base + l] = (c64){a0.re - a1.re, a0.im - a1.im};
      }
    }
    src = dst;
    l *= r;
    ++pass;
  }
  if (n == 1) y[0] = x[0];
}

int oracle_fft_f64(int64_t n, int64_t batch, const double *in, double *out,
                   int inverse, int nthreads) {
  if (!is_pow2(n) || batch < 0) return 1;
  c64 *w = (c64 *)malloc(sizeof(c64) * (size_t)n);
  if (!w) return 2;
  for (int64_t j = 0; j < n; ++j) {
    ld re, im;
    twiddle_ld(n, j, inverse ? +1 : -1, &re, &im);
    w[j].re = (double)re;
    w[j].im = (double)im;
  }
  if (n < 4) { /* the radix-4 sign probe reads w[n/4]; keep it valid */
    c64 *w4 = (c64 *)realloc(w, sizeof(c64) * 4);
    if (!w4) { free(w); return 2; }
    w = w4;
    w[n / 4].im = inverse ? 1.0 : -1.0;
  }
  int err = 0;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
    c64 *tmp = (c64 *)malloc(sizeof(c64) * (size_t)n);
    if (!tmp) {
#pragma omp atomic write
      err = 2;
    } else {
#pragma omp for schedule(static)
      for (int64_t b = 0; b < batch; ++b) {
        const c64 *x = (const c64 *)(in + 2 * n * b);
        c64 *y = (c64 *)(out + 2 * n * b);
        if (n == 4) {
          /* the generic loop handles this, but keep n<4 table guard above */
        }
        stockham_f64(n, x, y, tmp, w);
      }
      free(tmp);
    }
  }
  free(w);
  return err;
}

/* ------------------------------------------------------------------------ */
/* complex long double helpers for the restatements below                    */
typedef struct { ld re, im; } cld;
static inline cld cl_mul(cld a, cld b) {
  return (cld){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
static inline cld cl_add(cld a, cld b) { return (cld){a.re + b.re, a.im + b.im}; }
static inline cld cl_sub(cld a, cld b) { return (cld){a.re - b.re, a.im - b.im}; }
static inline cld cl_tw(int64_t n, int64_t q) {
  cld r;
  twiddle_ld(n, q, -1, &r.re, &r.im);
  return r;
}

static const ld SQRT_1_2_L = 0.707106781186547524400844362104849039L;
static const ld COS_PI_8_L = 0.923879532511286756128183189396788933L;
static const ld SIN_PI_8_L = 0.382683432365089771728459984030398867L;

/* Multiplications by fixed roots used by the radix-8/16 butterflies of the
 * reference (includes/legacy8_fft.cuh:41-56, includes/legacy16_fft.cuh:47-90),
 * here with correctly rounded long-double constants instead of 12 digits. */
static inline cld b12(cld a) { return (cld){a.im, -a.re}; }              /* *W4^1 = -i */
static inline cld b14(cld a) { return (cld){SQRT_1_2_L * (a.re + a.im), SQRT_1_2_L * (-a.re + a.im)}; } /* *W8^1 */
static inline cld b34(cld a) { return (cld){SQRT_1_2_L * (-a.re + a.im), SQRT_1_2_L * (-a.re - a.im)}; } /* *W8^3 */
static inline cld b18(cld a) { return cl_mul((cld){COS_PI_8_L, -SIN_PI_8_L}, a); } /* *W16^1 */
static inline cld b38(cld a) { return cl_mul((cld){SIN_PI_8_L, -COS_PI_8_L}, a); } /* *W16^3 */
static inline cld b58(cld a) { return cl_mul((cld){-SIN_PI_8_L, -COS_PI_8_L}, a); } /* *W16^5 */
static inline cld b78(cld a) { return cl_mul((cld){-COS_PI_8_L, -SIN_PI_8_L}, a); } /* *W16^7 */

static int load_signal(int64_t n, const double *x, cld *a) {
  for (int64_t i = 0; i < n; ++i) { a[i].re = x[2 * i]; a[i].im = x[2 * i + 1]; }
  return 0;
}
static void store_signal(int64_t n, const cld *a, double *hi, double *lo) {
  for (int64_t i = 0; i < n; ++i) {
    double h;
    h = (double)a[i].re; hi[2 * i] = h; if (lo) lo[2 * i] = (double)(a[i].re - (ld)h);
    h = (double)a[i].im; hi[2 * i + 1] = h; if (lo) lo[2 * i + 1] = (double)(a[i].im - (ld)h);
  }
}

/* Radix-8 Stockham pass structure of includes/legacy8_fft.cuh:57-118:
 * p = 1, 8, 64, ...; thread t reads a[t + m*N/8], twiddles by W_{8p}^{k}
 * powers (k = t mod p) applied the way the reference does it (u1,u3,u5,u7
 * by W, u2,u3,u6,u7 by W^2, u4..u7 by W^4), butterflies, and writes
 * out[(t - k)*8 + k + q*p] in the reference's output order. */
int oracle_legacy8(int64_t n, int64_t batch, const double *in, double *out_hi,
                   double *out_lo) {
  int64_t m = n;
  while (m > 1 && m % 8 == 0) m /= 8;
  if (n < 8 || m != 1 || batch < 0) return 1;
  cld *a = (cld *)malloc(sizeof(cld) * (size_t)n);
  cld *o = (cld *)malloc(sizeof(cld) * (size_t)n);
  if (!a || !o) { free(a); free(o); return 2; }
  const int64_t bs = n / 8;
  for (int64_t b = 0; b < batch; ++b) {
    load_signal(n, in + 2 * n * b, a);
    for (int64_t p = 1; p < n; p *= 8) {
      for (int64_t t = 0; t < bs; ++t) {
        int64_t k = t & (p - 1);
        cld tw = cl_tw(8 * p, k);
        cld u[8];
        for (int j = 0; j < 8; ++j) u[j] = a[t + j * bs];
        u[1] = cl_mul(tw, u[1]); u[3] = cl_mul(tw, u[3]);
        u[5] = cl_mul(tw, u[5]); u[7] = cl_mul(tw, u[7]);
        cld tw2 = cl_mul(tw, tw);
        u[2] = cl_mul(tw2, u[2]); u[3] = cl_mul(tw2, u[3]);
        u[6] = cl_mul(tw2, u[6]); u[7] = cl_mul(tw2, u[7]);
        cld tw4 = cl_mul(tw2, tw2);
        u[4] = cl_mul(tw4, u[4]); u[5] = cl_mul(tw4, u[5]);
        u[6] = cl_mul(tw4, u[6]); u[7] = cl_mul(tw4, u[7]);
        cld v0 = cl_add(u[0], u[4]), v4 = cl_sub(u[0], u[4]);
        cld v1 = cl_add(u[1], u[5]), v5 = b14(cl_sub(u[1], u[5]));
        cld v2 = cl_add(u[2], u[6]), v6 = b12(cl_sub(u[2], u[6]));
        cld v3 = cl_add(u[3], u[7]), v7 = b34(cl_sub(u[3], u[7]));
        cld w0 = cl_add(v0, v2), w2 = cl_sub(v0, v2);
        cld w1 = cl_add(v1, v3), w3 = b12(cl_sub(v1, v3));
        cld w4 = cl_add(v4, v6), w6 = cl_sub(v4, v6);
        cld w5 = cl_add(v5, v7), w7 = b12(cl_sub(v5, v7));
        cld *dst = o + ((t - k) << 3) + k;
        dst[0] = cl_add(w0, w1);
        dst[p] = cl_add(w4, w5);
        dst[2 * p] = cl_add(w2, w3);
        dst[3 * p] = cl_add(w6, w7);
        dst[4 * p] = cl_sub(w0, w1);
        dst[5 * p] = cl_sub(w4, w5);
        dst[6 * p] = cl_sub(w2, w3);
        dst[7 * p] = cl_sub(w6, w7);
      }
      cld *t = a; a = o; o = t;
    }
    store_signal(n, a, out_hi + 2 * n * b, out_lo ? out_lo + 2 * n * b : NULL);
  }
  free(a);
  free(o);
  return 0;
}

/* Radix-16 Stockham pass structure of includes/legacy16_fft.cuh:99-172:
 * per pass, twiddle u[m] by W_{16p}^{m*k}, then four layers of radix-2
 * butterflies with fixed roots (distance 8, 4, 2, 1), and the reference's
 * bit-reversed store map out[(t-k)*16 + k + q*p]. */
static inline void dft2_tw(cld *a, cld *b, cld (*op)(cld)) {
  cld d = cl_sub(*a, *b);
  *a = cl_add(*a, *b);
  *b = op ? op(d) : d;
}
static cld b28(cld a) { return b14(a); } /* W16^2 = W8^1 */
static cld b48(cld a) { return b12(a); } /* W16^4 = -i */
static cld b68(cld a) { return b34(a); } /* W16^6 = W8^3 */

int oracle_legacy16(int64_t n, int64_t batch, const double *in,
                    double *out_hi, double *out_lo) {
  int64_t m = n;
  while (m > 1 && m % 16 == 0) m /= 16;
  if (n < 16 || m != 1 || batch < 0) return 1;
  cld *a = (cld *)malloc(sizeof(cld) * (size_t)n);
  cld *o = (cld *)malloc(sizeof(cld) * (size_t)n);
  if (!a || !o) { free(a); free(o); return 2; }
  const int64_t bs = n / 16;
  for (int64_t b = 0; b < batch; ++b) {
    load_signal(n, in + 2 * n * b, a);
    for (int64_t p = 1; p < n; p *= 16) {
      for (int64_t t = 0; t < bs; ++t) {
        int64_t k = t & (p - 1);
        cld u[16];
        for (int j = 0; j < 16; ++j) u[j] = a[t + j * bs];
        for (int j = 1; j < 16; ++j) u[j] = cl_mul(cl_tw(16 * p, j * k), u[j]);
        dft2_tw(&u[0], &u[8], NULL);  dft2_tw(&u[1], &u[9], b18);
        dft2_tw(&u[2], &u[10], b28);  dft2_tw(&u[3], &u[11], b38);
        dft2_tw(&u[4], &u[12], b48);  dft2_tw(&u[5], &u[13], b58);
        dft2_tw(&u[6], &u[14], b68);  dft2_tw(&u[7], &u[15], b78);
        dft2_tw(&u[0], &u[4], NULL);  dft2_tw(&u[1], &u[5], b14);
        dft2_tw(&u[2], &u[6], b12);   dft2_tw(&u[3], &u[7], b34);
        dft2_tw(&u[8], &u[12], NULL); dft2_tw(&u[9], &u[13], b14);
        dft2_tw(&u[10], &u[14], b12); dft2_tw(&u[11], &u[15], b34);
        dft2_tw(&u[0], &u[2], NULL);  dft2_tw(&u[1], &u[3], b12);
        dft2_tw(&u[4], &u[6], NULL);  dft2_tw(&u[5], &u[7], b12);
        dft2_tw(&u[8], &u[10], NULL); dft2_tw(&u[9], &u[11], b12);
        dft2_tw(&u[12], &u[14], NULL); dft2_tw(&u[13], &u[15], b12);
        cld *dst = o + ((t - k) << 4) + k;
        /* output q <- register pair, bit-reversed as in the reference */
        static const int src_of_q[8] = {0, 8, 4, 12, 2, 10, 6, 14};
        for (int q = 0; q < 8; ++q) {
          int r = src_of_q[q];
          dst[q * p] = cl_add(u[r], u[r + 1]);
          dst[(q + 8) * p] = cl_sub(u[r], u[r + 1]);
        }
      }
      cld *t = a; a = o; o = t;
    }
    store_signal(n, a, out_hi + 2 * n * b, out_lo ? out_lo + 2 * n * b : NULL);
  }
  free(a);
  free(o);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* DFT-as-matrix restatement (includes/c64_fft64.cuh:8-27 generalised).
 * N = r[0] * r[1] * ... ; stage s computes, for every column of the current
 * sub-problem, a DFT of length r[s] as a complex matrix product whose three
 * real products follow includes/tensor_utils.cuh:25-36:
 *   s1 = Ar*Br, s2 = Ai*Bi, s3 = (Ar+Ai)*(Br+Bi);  Re = s1 - s2,
 *   Im = s3 - s1 - s2
 * then multiplies by W_L^{m*k1} (m = column, L = current length) and recurses
 * on each output row.  Output index k = k1 + r0 * k2 (natural order), the
 * ordering of c64_fft64's second GEMM.                                      */
static void karatsuba_dft_rows(int64_t r, int64_t cols, const cld *x /* r x cols */,
                               cld *y /* r x cols */) {
  /* y[k][c] = sum_j F[k][j] x[j][c] with F = exp(-2 pi i jk / r) */
  for (int64_t k = 0; k < r; ++k) {
    for (int64_t c = 0; c < cols; ++c) {
      ld s1 = 0, s2 = 0, s3 = 0;
      for (int64_t j = 0; j < r; ++j) {
        cld f = cl_tw(r, j * k);
        cld v = x[j * cols + c];
        s1 += f.re * v.re;
        s2 += f.im * v.im;
        s3 += (f.re + f.im) * (v.re + v.im);
      }
      y[k * cols + c] = (cld){s1 - s2, s3 - s1 - s2};
    }
  }
}

static void karatsuba_fft_rec(int64_t n, const int *radix, int nr, const cld *x,
                              cld *out, cld *scratch) {
  if (n == 1) { out[0] = x[0]; return; }
  int64_t r0 = radix[0];
  int64_t m = n / r0;
  /* view x as r0 rows x m columns: x[n1*m + n2], n = m*n1 + n2 */
  cld *y = scratch;              /* r0 x m */
  cld *row = scratch + n;        /* m */
  cld *rowout = scratch + n + m; /* m */
  cld *deeper = scratch + n + 2 * m;
  karatsuba_dft_rows(r0, m, x, y);
  for (int64_t k1 = 0; k1 < r0; ++k1) {
    for (int64_t n2 = 0; n2 < m; ++n2) row[n2] = cl_mul(y[k1 * m + n2], cl_tw(n, n2 * k1));
    if (nr > 1) {
      karatsuba_fft_rec(m, radix + 1, nr - 1, row, rowout, deeper);
    } else {
      for (int64_t i = 0; i < m; ++i) rowout[i] = row[i];
    }
    for (int64_t k2 = 0; k2 < m; ++k2) out[k1 + r0 * k2] = rowout[k2];
  }
}

int oracle_karatsuba(int64_t n, const int *radices, int nradix, int64_t batch,
                     const double *in, double *out_hi, double *out_lo) {
  int64_t prod = 1;
  for (int i = 0; i < nradix; ++i) {
    if (radices[i] < 2) return 1;
    prod *= radices[i];
  }
  if (prod != n || batch < 0) return 1;
  /* scratch: sum over levels of (n_l + 2 m_l) <= 4n */
  cld *x = (cld *)malloc(sizeof(cld) * (size_t)n);
  cld *o = (cld *)malloc(sizeof(cld) * (size_t)n);
  cld *s = (cld *)malloc(sizeof(cld) * (size_t)(6 * n + 16));
  if (!x || !o || !s) { free(x); free(o); free(s); return 2; }
  for (int64_t b = 0; b < batch; ++b) {
    load_signal(n, in + 2 * n * b, x);
    karatsuba_fft_rec(n, radices, nradix, x, o, s);
    store_signal(n, o, out_hi + 2 * n * b, out_lo ? out_lo + 2 * n * b : NULL);
  }
  free(x);
  free(o);
  free(s);
  return 0;
}
