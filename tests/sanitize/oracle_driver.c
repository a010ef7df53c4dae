/* Runs every function of the C oracle (oracle/fftoc_oracle.c, test
 * infrastructure) on small transforms under AddressSanitizer and
 * UndefinedBehaviorSanitizer, and checks each against the O(N^2) DFT, so a
 * buffer overrun or undefined operation in the checker aborts the run. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

int oracle_twiddles(int64_t n, int sign, double *hi, double *lo);
int oracle_dft(int64_t n, int64_t batch, const double *in, double *hi, double *lo, int inv);
int oracle_fft(int64_t n, int64_t batch, const double *in, double *hi, double *lo, int inv);
int oracle_fft_f64(int64_t n, int64_t batch, const double *in, double *out, int inv, int threads);
int oracle_legacy8(int64_t n, int64_t batch, const double *in, double *hi, double *lo);
int oracle_legacy16(int64_t n, int64_t batch, const double *in, double *hi, double *lo);
int oracle_karatsuba(int64_t n, const int *radices, int nradix, int64_t batch, const double *in,
                     double *hi, double *lo);

static double rel_err(int64_t count, const double *a, const double *b) {
  double num = 0, den = 0;
  for (int64_t i = 0; i < count; ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return sqrt(num / (den > 0 ? den : 1));
}

static void fail(const char *what, int64_t n, double err) {
  fprintf(stderr, "oracle_asan: %s n=%lld err=%.3e\n", what, (long long)n, err);
  exit(1);
}

int main(void) {
  const int64_t batch = 3;
  int checks = 0;
  srand(7);
  for (int64_t n = 2; n <= 1024; n *= 2) {
    const int64_t m = 2 * n * batch;
    double *x = malloc(m * sizeof *x), *ref = malloc(m * sizeof *ref), *lo = malloc(m * sizeof *lo);
    double *hi = malloc(m * sizeof *hi), *f = malloc(m * sizeof *f);
    double *th = malloc(2 * n * sizeof *th), *tl = malloc(2 * n * sizeof *tl);
    for (int64_t i = 0; i < m; ++i) x[i] = 2.0 * rand() / RAND_MAX - 1.0;
    for (int inv = 0; inv < 2; ++inv) {
      if (oracle_dft(n, batch, x, ref, lo, inv)) fail("dft", n, 0);
      if (oracle_fft(n, batch, x, hi, lo, inv)) fail("fft", n, 0);
      if (rel_err(m, hi, ref) > 1e-15) fail("fft", n, rel_err(m, hi, ref));
      if (oracle_fft_f64(n, batch, x, f, inv, 2)) fail("fft_f64", n, 0);
      if (rel_err(m, f, ref) > 1e-14) fail("fft_f64", n, rel_err(m, f, ref));
      checks += 3;
    }
    if (oracle_dft(n, batch, x, ref, lo, 0)) fail("dft", n, 0);
    if (oracle_twiddles(n, -1, th, tl)) fail("twiddles", n, 0);
    if (fabs(th[0] - 1.0) > 0 || fabs(th[1]) > 0) fail("twiddles", n, th[0]);
    if (n == 8 || n == 64 || n == 512) {
      if (oracle_legacy8(n, batch, x, hi, lo)) fail("legacy8", n, 0);
      if (rel_err(m, hi, ref) > 1e-15) fail("legacy8", n, rel_err(m, hi, ref));
      ++checks;
    } else if (oracle_legacy8(n, batch, x, hi, lo) == 0) {
      fail("legacy8 accepted a length that is not a power of 8", n, 0);
    }
    if (n == 16 || n == 256) {
      if (oracle_legacy16(n, batch, x, hi, lo)) fail("legacy16", n, 0);
      if (rel_err(m, hi, ref) > 1e-15) fail("legacy16", n, rel_err(m, hi, ref));
      ++checks;
    }
    /* two-factor radix split n = r * (n / r) with r = 2^(floor(log2 n / 2)) */
    if (n < 4) {
      free(x); free(ref); free(lo); free(hi); free(f); free(th); free(tl);
      continue;
    }
    int r = 1;
    while ((int64_t)r * r * 4 <= n) r *= 2;
    const int radices[2] = {r, (int)(n / r)};
    if (oracle_karatsuba(n, radices, 2, batch, x, hi, lo)) fail("karatsuba", n, 0);
    if (rel_err(m, hi, ref) > 1e-15) fail("karatsuba", n, rel_err(m, hi, ref));
    const int bad[2] = {3, (int)n};
    if (oracle_karatsuba(n, bad, 2, batch, x, hi, lo) == 0) fail("karatsuba bad split", n, 0);
    checks += 2;
    free(x); free(ref); free(lo); free(hi); free(f); free(th); free(tl);
  }
  printf("oracle_asan: %d checks passed\n", checks);
  return 0;
}
