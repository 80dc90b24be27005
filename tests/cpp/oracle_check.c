/* Host-side check of the CPU oracle (oracle/oracle.c) built with
 * AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitizers.py:
 * every strategy of the FP32 and FP64 ports and the FP64 extension on a set
 * of lengths, against the naive FP64 DFT.  Non-zero exit on a sanitizer
 * report or an error above the bounds. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static double rel(const double* a, const double* b, size_t n) {
  double num = 0, den = 0;
  for (size_t i = 0; i < 2 * n; ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return sqrt(num / den);
}

int main(void) {
  const size_t lengths[] = {1, 2, 3, 4, 5, 7, 8, 12, 16, 60, 105, 128, 243, 625, 720, 1024,
                            11, 101, 1009, 22, 121, 1031, 2401, 4096, 6720};
  int bad = 0;
  for (size_t li = 0; li < sizeof(lengths) / sizeof(lengths[0]); ++li) {
    const size_t n = lengths[li];
    double* x = malloc(2 * n * sizeof(double));
    double* ref = malloc(2 * n * sizeof(double));
    double* y = malloc(2 * n * sizeof(double));
    float* xf = malloc(2 * n * sizeof(float));
    float* yf = malloc(2 * n * sizeof(float));
    double* yd = malloc(2 * n * sizeof(double));
    unsigned s = 12345u + (unsigned)n;
    for (size_t i = 0; i < 2 * n; ++i) {
      s = s * 1103515245u + 12345u;
      x[i] = ((s >> 8) & 0xffff) / 32768.0 - 1.0;
      xf[i] = (float)x[i];
      x[i] = (double)xf[i];
    }
    for (int inv = 0; inv < 2; ++inv) {
      fyo_dft_naive_f64(n, x, ref, inv);
      double e64 = -1, e32 = -1, ex = -1;
      if (fyo_fft_f64(n, x, y, inv) == 0) e64 = rel(y, ref, n);
      if (fyo_fft_f32(n, xf, yf, inv) == 0) {
        for (size_t i = 0; i < 2 * n; ++i) yd[i] = yf[i];
        e32 = rel(yd, ref, n);
      }
      if (fyo_fft_ext_f64(n, x, y, inv) == 0) ex = rel(y, ref, n);
      const double lg = n > 1 ? log2((double)n) : 1.0;
      if (e64 > 1e-15 * lg * 10 || ex > 1e-15 * lg * 10 || e32 > 2e-7 * lg || ex < 0) {
        fprintf(stderr, "N=%zu inv=%d: f64 %.2e f32 %.2e ext %.2e\n", n, inv, e64, e32, ex);
        ++bad;
      }
    }
    free(x); free(ref); free(y); free(xf); free(yf); free(yd);
  }
  printf("oracle check: %zu lengths, %d failures\n", sizeof(lengths) / sizeof(lengths[0]), bad);
  return bad ? 1 : 0;
}
