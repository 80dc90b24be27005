/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Planner integers restated from the reference; each function cites the
 * reference file:line it follows.  The transform restatement lives in
 * oracle_impl.inc, included once for double and once for float.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* The 38 directly supported lengths, ascending
 * (src/foye_FFT_radix_wrap.hpp:93-98). */
static const size_t kSmooth[38] = {
    2, 3, 4, 5, 7, 8, 9, 16, 25, 27, 32, 49, 64, 81,
    125, 128, 243, 256, 343, 512, 625, 729, 1024,
    2048, 2187, 2401, 3125, 4096, 6561, 8192, 15625,
    16384, 16807, 19683, 32768, 59049, 65536, 78125};

size_t fyo_smooth_length(int i) { return (i >= 0 && i < 38) ? kSmooth[i] : 0; }

/* binary search in the registry (src/foye_FFT_radix_wrap.hpp:41-44) */
int fyo_is_directly_supported(size_t n) {
  int lo = 0, hi = 37;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (kSmooth[mid] == n) return 1;
    if (kSmooth[mid] < n) lo = mid + 1; else hi = mid - 1;
  }
  return 0;
}

/* lower_bound, 0 past the end (src/foye_FFT_radix_wrap.hpp:75-79) */
size_t fyo_next_smooth(size_t n) {
  for (int i = 0; i < 38; ++i)
    if (kSmooth[i] >= n) return kSmooth[i];
  return 0;
}

/* radices with a batch codelet (src/foye_FFT_batch_kernel.hpp:768-781) */
static int has_batch_codelet(size_t r) {
  return r == 2 || r == 3 || r == 4 || r == 5 || r == 7;
}

/* Split choice of the six-step planner, restated literally including its
 * priority/cost bookkeeping (src/foye_FFT_mix_radix.hpp:253-352). */
size_t fyo_compute_n1(size_t n) {
  size_t best_n1 = 0;
  double best_cost = INFINITY;
  for (int i = 0; i < 38; ++i) {
    size_t f = kSmooth[i];
    if (f >= n) continue;
    if (n % f != 0) continue;
    size_t n1 = f, n2 = n / f;
    int n1_sup = fyo_is_directly_supported(n1);
    int n2_sup = fyo_is_directly_supported(n2);
    int n1_b = has_batch_codelet(n1), n2_b = has_batch_codelet(n2);
    double imb = fabs(log2((double)n1) - log2((double)n2));
    int prio = ((n1_sup && n1_b) || (n2_sup && n2_b)) ? 2 : ((n1_sup || n2_sup) ? 1 : 0);
    double cost = (double)n * (1.0 + imb * 0.2);
    if (n1_b) cost *= 0.85;
    if (n2_b) cost *= 0.85;
    if (prio > 0 || best_n1 == 0) {
      if (prio > 0) {
        int bar = (best_n1 == 0) ? -1 : 0;
        if (prio > bar) { best_n1 = n1; best_cost = cost; }
        else if (prio == bar && cost < best_cost) { best_n1 = n1; best_cost = cost; }
      } else if (cost < best_cost) {
        best_n1 = n1; best_cost = cost;
      }
    }
  }
  if (best_n1 == 0) {
    if (n % 8 == 0) return n / 8;
    if (n % 4 == 0) return n / 4;
    if (n % 2 == 0) return n / 2;
    return 0;
  }
  return best_n1;
}

/* greedy division by the largest listed divisor
 * (src/foye_FFT_mix_radix.hpp:618-643) */
int fyo_mix_is_processable(size_t n) {
  while (n > 1) {
    int found = 0;
    for (int i = 37; i >= 0; --i) {
      size_t f = kSmooth[i];
      if (f <= n && n % f == 0) { n /= f; found = 1; break; }
    }
    if (!found) return 0;
  }
  return 1;
}

/* (src/foye_FFT_mix_radix.hpp:577-598) */
int fyo_decomposition(size_t n, size_t out[10]) {
  memset(out, 0, 10 * sizeof(size_t));
  size_t cur = n;
  int idx = 0;
  while (!fyo_is_directly_supported(cur)) {
    size_t n1 = fyo_compute_n1(cur);
    if (n1 == 0 || idx >= 10) return -1;
    out[idx++] = n1;
    cur /= n1;
  }
  if (idx < 10) out[idx++] = cur;
  return idx;
}

/* trial division (src/foye_FFT_utility.hpp:392-418) */
int fyo_is_prime(size_t num) {
  if (num <= 1) return 0;
  if (num == 2 || num == 3) return 1;
  if (num % 2 == 0) return 0;
  for (size_t i = 3; i * i <= num; i += 2)
    if (num % i == 0) return 0;
  return 1;
}

/* square-and-multiply; 128-bit products so that moduli above 2^32 are safe
 * (src/foye_FFT_utility.hpp:420-435) */
size_t fyo_power_mod(size_t base, size_t e, size_t mod) {
  unsigned __int128 res = 1, b = base % mod;
  while (e > 0) {
    if (e & 1) res = (res * b) % mod;
    b = (b * b) % mod;
    e >>= 1;
  }
  return (size_t)res;
}

/* smallest generator: g^(phi/q) != 1 for every prime q | phi
 * (src/foye_FFT_utility.hpp:437-485) */
size_t fyo_primitive_root(size_t n) {
  if (n == 2) return 1;
  size_t phi = n - 1, m = phi;
  size_t factors[64];
  int nf = 0;
  for (size_t i = 2; i * i <= m; ++i) {
    if (m % i == 0) {
      factors[nf++] = i;
      while (m % i == 0) m /= i;
    }
  }
  if (m > 1) factors[nf++] = m;
  for (size_t g = 2; g <= n; ++g) {
    int ok = 1;
    for (int i = 0; i < nf; ++i)
      if (fyo_power_mod(g, phi / factors[i], n) == 1) { ok = 0; break; }
    if (ok) return g;
  }
  return 0;
}

/* prime, not directly supported, and 2N-3 padded length in the list
 * (src/foye_FFT_rader.hpp:403-423) */
int fyo_rader_is_processable(size_t n) {
  if (n < 2 || n > kSmooth[37]) return 0;
  return !fyo_is_directly_supported(n) && fyo_is_prime(n) && fyo_next_smooth(2 * n - 3) != 0;
}

/* M = next_smooth(2(N-1)-1) (src/foye_FFT_rader.hpp:166) */
size_t fyo_rader_M(size_t n) { return n >= 2 ? fyo_next_smooth(2 * (n - 1) - 1) : 0; }

/* N <= 39062 (src/foye_FFT_bluestein.hpp:189,212-215) */
int fyo_bluestein_is_processable(size_t n) { return n <= 39062; }

/* first listed length >= 2N-1 (src/foye_FFT_bluestein.hpp:93,173-183) */
size_t fyo_bluestein_M(size_t n) { return fyo_next_smooth(2 * n - 1); }

/* the ladder (foye_FFT.hpp:162-184) */
int fyo_strategy(size_t n) {
  if (fyo_is_directly_supported(n)) return FYO_KERNEL_DIRECT;
  if (fyo_mix_is_processable(n)) return FYO_MIX_RADIX;
  if (fyo_rader_is_processable(n)) return FYO_RADER;
  if (fyo_bluestein_is_processable(n)) return FYO_BLUESTEIN;
  return FYO_DFT;
}

/* exp(sign * 2*pi*i * m / n) for 0 <= m < n, in long double from an exact
 * integer argument, so each root is correct to well below double rounding. */
static void root_ld(size_t m, size_t n, int sign, long double* re, long double* im) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  m %= n;
  long double a = two_pi * (long double)m / (long double)n;
  *re = cosl(a);
  *im = (sign < 0 ? -1.0L : 1.0L) * sinl(a);
}

int fyo_dft_naive_f64(size_t n, const double* in, double* out, int inverse) {
  if (n == 0) return 1;
  long double* wr = (long double*)malloc(n * sizeof(long double));
  long double* wi = (long double*)malloc(n * sizeof(long double));
  long double* acc = (long double*)malloc(2 * n * sizeof(long double));
  if (!wr || !wi || !acc) { free(wr); free(wi); free(acc); return 2; }
  for (size_t m = 0; m < n; ++m) root_ld(m, n, inverse ? +1 : -1, &wr[m], &wi[m]);
  for (size_t k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    size_t idx = 0; /* (j*k) mod n, kept reduced by integer addition */
    for (size_t j = 0; j < n; ++j) {
      long double xr = in[2 * j], xi = in[2 * j + 1];
      sr += xr * wr[idx] - xi * wi[idx];
      si += xr * wi[idx] + xi * wr[idx];
      idx += k;
      if (idx >= n) idx -= n;
    }
    acc[2 * k] = sr;
    acc[2 * k + 1] = si;
  }
  long double s = inverse ? 1.0L / (long double)n : 1.0L;
  for (size_t k = 0; k < 2 * n; ++k) out[k] = (double)(acc[k] * s);
  free(wr); free(wi); free(acc);
  return 0;
}

/* ---- transform restatement, instantiated for double and float ---- */
#define R double
#define CX double complex
#define SUF(name) name##_f64
#include "oracle_impl.inc"
#undef R
#undef CX
#undef SUF

#define R float
#define CX float complex
#define SUF(name) name##_f32
#include "oracle_impl.inc"
#undef R
#undef CX
#undef SUF

int fyo_fft_f64(size_t n, const double* in, double* out, int inverse) {
  return fft_top_f64(n, in, out, inverse, 0);
}
int fyo_fft_f32(size_t n, const float* in, float* out, int inverse) {
  return fft_top_f32(n, in, out, inverse, 0);
}
int fyo_fft_ext_f64(size_t n, const double* in, double* out, int inverse) {
  return fft_top_f64(n, in, out, inverse, 1);
}

/* ---- batched FP32 port for the CPU baseline ---- */
typedef struct {
  size_t n, batch, first, step;
  const float* in;
  float* out;
  int inverse;
  int status;
} batch_job;

static void* batch_worker(void* arg) {
  batch_job* j = (batch_job*)arg;
  j->status = 0;
  for (size_t b = j->first; b < j->batch; b += j->step) {
    int s = fyo_fft_f32(j->n, j->in + 2 * j->n * b, j->out + 2 * j->n * b, j->inverse);
    if (s) { j->status = s; break; }
  }
  return NULL;
}

double fyo_fft_f32_batch_timed(size_t n, size_t batch, const float* in, float* out,
                               int inverse, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t tid[256];
  batch_job jobs[256];
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < threads; ++t) {
    jobs[t] = (batch_job){n, batch, (size_t)t, (size_t)threads, in, out, inverse, 0};
    if (threads == 1) batch_worker(&jobs[0]);
    else pthread_create(&tid[t], NULL, batch_worker, &jobs[t]);
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  for (int t = 0; t < threads; ++t)
    if (jobs[t].status) return -1.0;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
