/*
 * CPU oracle for the 1-D complex FFT path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker (or, for cpu_baseline, as the
 * timed CPU port of the reference algorithm).  The product path
 * (fyfft_amd, libfyfft.so) never links or calls it.
 *
 * Contents:
 *   - planner integer functions restated from the reference (strategy
 *     ladder, smooth-length registry, mixed-radix split, Rader/Bluestein
 *     pad lengths, number theory);
 *   - fyo_dft_naive_f64: O(N^2) DFT, exact integer reduction of n*k mod N,
 *     long-double accumulation -- the FP64 bar every other result is
 *     measured against;
 *   - fyo_fft_{f32,f64}: a scalar C restatement of the reference's
 *     strategy ladder (direct radix-r DIT, six-step mixed radix, Rader,
 *     Bluestein, naive DFT) in float or double;
 *   - fyo_fft_ext_f64: FP64 FFT for any N >= 1 (the ladder where the
 *     reference has a feasible strategy, Bluestein with a power-of-two pad
 *     elsewhere), used as the FP64 oracle beyond the reference's domain.
 *
 * Data is interleaved (re, im); forward uses exp(-2*pi*i*n*k/N); inverse
 * uses exp(+2*pi*i*n*k/N) and scales by 1/N (foye_FFT.hpp:128-151).
 */
#ifndef FYFFT_ORACLE_H
#define FYFFT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { FYO_KERNEL_DIRECT = 0, FYO_MIX_RADIX = 1, FYO_RADER = 2,
       FYO_BLUESTEIN = 3, FYO_DFT = 4 };

/* planner integers */
int    fyo_is_directly_supported(size_t n);
size_t fyo_next_smooth(size_t n);
size_t fyo_smooth_length(int i);           /* i-th entry of the 38-list, 0 past the end */
size_t fyo_compute_n1(size_t n);
int    fyo_mix_is_processable(size_t n);
int    fyo_decomposition(size_t n, size_t out[10]);
int    fyo_is_prime(size_t n);
size_t fyo_power_mod(size_t base, size_t exp, size_t mod);
size_t fyo_primitive_root(size_t n);
int    fyo_rader_is_processable(size_t n);
size_t fyo_rader_M(size_t n);
int    fyo_bluestein_is_processable(size_t n);
size_t fyo_bluestein_M(size_t n);
int    fyo_strategy(size_t n);

/* transforms; return 0 on success, nonzero if the length is unsupported */
int fyo_dft_naive_f64(size_t n, const double* in, double* out, int inverse);
int fyo_fft_f64(size_t n, const double* in, double* out, int inverse);
int fyo_fft_f32(size_t n, const float* in, float* out, int inverse);
int fyo_fft_ext_f64(size_t n, const double* in, double* out, int inverse);

/* Batched FP32 port on `threads` POSIX threads (one plan per thread, batch
 * split round-robin); used only for the CPU baseline timing.  Returns
 * elapsed seconds of the transform loop, or a negative value on error. */
double fyo_fft_f32_batch_timed(size_t n, size_t batch, const float* in,
                               float* out, int inverse, int threads);

#ifdef __cplusplus
}
#endif
#endif
