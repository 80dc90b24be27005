/*
 * fyfft.h -- C ABI of the MI355X-native 1-D complex FFT (libfyfft.so).
 *
 * Plain pointers and sizes only; no HIP or torch types in the signatures
 * (a stream is passed as `void*` holding a hipStream_t, NULL = the null
 * stream).
 *
 * The reference (kjear/fyFFT) exports a header-only C++ class and no C ABI:
 *   fy::fft::FFT1D_C2C_invoker(std::size_t N)          foye_FFT.hpp:101-126
 *   template<bool invert> dispatch(const float*, float*) foye_FFT.hpp:128-151
 *   acquired_execution_strategy(std::size_t)           foye_FFT.hpp:162-184
 *   min_input_length / max_input_length                foye_FFT.hpp:86-87
 * Each entry point below names the piece of that interface it replaces.
 * include/foye_FFT.hpp re-creates the class itself on top of this ABI.
 *
 * Conventions (same as the reference):
 *   - data is interleaved (re, im), std::complex<T>-compatible;
 *   - forward:  X[k] = sum_n x[n] exp(-2*pi*i*n*k/N)   (unnormalised);
 *   - inverse:  x[n] = (1/N) sum_k X[k] exp(+2*pi*i*n*k/N);
 *   - adjoint (this ABI only): the inverse without 1/N, i.e. the exact
 *     adjoint of the forward transform (= N * inverse);
 *   - in == out (in place) is allowed.
 * Errors are reported as status codes where the reference has undefined
 * behaviour (foye_FFT.hpp:39-43,103-104).
 */
#ifndef FYFFT_H
#define FYFFT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FYFFT_VERSION_MAJOR 0
#define FYFFT_VERSION_MINOR 1

/* Longest transform a plan accepts.  The reference stops at 80000
 * (foye_FFT.hpp:87); this ABI goes beyond it (Bluestein N = 1000003 and
 * smooth N = 2^20 are required workloads). */
#define FYFFT_MAX_LENGTH ((size_t)1 << 24)

typedef enum fyfft_status {
  FYFFT_SUCCESS = 0,
  FYFFT_INVALID_ARGUMENT = 1, /* null pointer, unknown enum value, batch 0 */
  FYFFT_INVALID_LENGTH = 2,   /* N < 1 or N > FYFFT_MAX_LENGTH            */
  FYFFT_UNSUPPORTED = 3,      /* valid request this build does not serve  */
  FYFFT_ALLOC_FAILED = 4,     /* device or host allocation failed         */
  FYFFT_DEVICE_ERROR = 5,     /* a HIP runtime call failed                */
  FYFFT_INVALID_PLAN = 6,     /* null or destroyed plan                   */
  FYFFT_NO_DEVICE = 7         /* no GPU visible to the process            */
} fyfft_status;

typedef enum fyfft_precision {
  FYFFT_F32 = 0,  /* complex<float>  (the reference's only precision) */
  FYFFT_F64 = 1,  /* complex<double>                                   */
  FYFFT_BF16 = 2, /* bf16 storage: complex as two bf16 (4 bytes)       */
  FYFFT_F16 = 3   /* fp16 storage: complex as two fp16 (4 bytes)       */
} fyfft_precision;

typedef enum fyfft_direction {
  FYFFT_FORWARD = 0, /* dispatch<false>                     */
  FYFFT_INVERSE = 1, /* dispatch<true> (normalised by 1/N)  */
  FYFFT_ADJOINT = 2  /* conjugate twiddles, no 1/N          */
} fyfft_direction;

/* Same order and meaning as FFT1D_C2C_invoker::FFT_execution_strategy
 * (foye_FFT.hpp:153-160).  The values report THIS library's choice:
 *   KERNEL_DIRECT - N is one of the reference's 38 directly supported
 *                   lengths (radix_wrap.hpp:93-98);
 *   MIX_RADIX     - any other 2/3/5/7-smooth N;
 *   RADER         - prime N whose N-1 is 2/3/5/7-smooth;
 *   BLUESTEIN     - every other N;
 *   DFT           - never chosen by this library (kept for ABI parity). */
typedef enum fyfft_strategy {
  FYFFT_KERNEL_DIRECT = 0,
  FYFFT_MIX_RADIX = 1,
  FYFFT_RADER = 2,
  FYFFT_BLUESTEIN = 3,
  FYFFT_DFT = 4
} fyfft_strategy;

/* Kernel form of a plan (fyfft_plan_create_ex):
 *   AUTO  - a compile-time specialised kernel where one exists for the
 *           length, else the generic kernels (what fyfft_plan_create does);
 *   PLAIN - always the generic ("plain GPU form") kernels, the baseline the
 *           other forms are measured against;
 *   TUNED - time candidate configurations on the device at plan time and
 *           keep the fastest: radix order, LDS padding, points per thread
 *           and specialised vs generic kernels for lengths held in one
 *           workgroup's LDS; the split and tiles of four-step lengths and
 *           of four-step Rader convolutions; Bluestein's convolution
 *           length.  The choice is remembered as wisdom per (GPU model, N,
 *           precision);
 *   MFMA  - BF16/F16 storage only: radix-16 stages on MFMA with split hi/lo
 *           fp16 operands (for BF16/F16 the other forms run FP32 arithmetic
 *           in the compile-time kernels, rounding once at the store). */
typedef enum fyfft_form {
  FYFFT_FORM_AUTO = 0,
  FYFFT_FORM_PLAIN = 1,
  FYFFT_FORM_TUNED = 2,
  FYFFT_FORM_MFMA = 3 /* BF16/F16 only: radix-16 stages on MFMA */
} fyfft_form;

typedef struct fyfft_options {
  int form;         /* fyfft_form */
  int reserved[7];  /* must be zero */
} fyfft_options;

typedef struct fyfft_plan fyfft_plan;

/* BF16 / FP16 plans store data as 16-bit complex pairs (4 bytes per
 * element).  By default they run the FP32 compile-time kernels with
 * widening loads and one rounding at the final store; FYFFT_FORM_MFMA runs
 * the radix-16 stages on MFMA with split hi/lo operands and FP32
 * accumulation instead.  Either way the internal error is at FP32 level and
 * the stored result is rounded once.  Available for N = 256 .. 16384 powers
 * of two; other lengths return FYFFT_UNSUPPORTED.  The reference has only
 * complex<float>. */

/* Replaces FFT1D_C2C_invoker::FFT1D_C2C_invoker(std::size_t)
 * (foye_FFT.hpp:101-126), extended with a batch of `batch` contiguous
 * transforms (transform b starts at element b*N) and a precision.
 * Plans the factorisation on the host, allocates device tables (generated
 * in FP64 on the GPU) and any workspace, on the current HIP device.
 * Synchronous. */
fyfft_status fyfft_plan_create(fyfft_plan** plan, size_t n, size_t batch,
                               fyfft_precision precision);

/* Replaces the reference's per-strategy invokers' constructors
 * (mix_radix::FFT1D_mixed_radix_invoker, rader::FFT1D_C2C_rader_invoker,
 * bluestein::FFT1D_bluestein_invoker, dft::FFT1D_C2C_dft_invoker:
 * foye_FFT.hpp:81-84; mix_radix.hpp:645, rader.hpp:339, bluestein.hpp:217,
 * dft.hpp:106): a plan of the given fyfft_strategy.
 *   FYFFT_BLUESTEIN: a Bluestein convolution for any N > 1;
 *   FYFFT_RADER: a Rader convolution when N is prime and N-1 is
 *     2/3/5/7-smooth; other primes run the library's choice (Bluestein), as
 *     this library has no zero-padded Rader (rader.hpp:166);
 *   FYFFT_MIX_RADIX / FYFFT_KERNEL_DIRECT: N must be 2/3/5/7-smooth / one of
 *     the reference's 38 direct lengths (FYFFT_UNSUPPORTED otherwise), then
 *     as fyfft_plan_create;
 *   FYFFT_DFT: the library's choice (the transform is the same DFT; no
 *     O(N^2) tables are built).
 * fyfft_plan_strategy reports what the plan runs.  The results equal those
 * of fyfft_plan_create up to rounding. */
fyfft_status fyfft_plan_create_strategy(fyfft_plan** plan, size_t n, size_t batch,
                                        fyfft_precision precision, int strategy);

/* The reference's eligibility predicate of each strategy invoker
 * (is_processable: radix_wrap.hpp:93-98 for KERNEL_DIRECT,
 * mix_radix.hpp:618-643, rader.hpp:341-345 + 403-423,
 * bluestein.hpp:212-215; DFT accepts every N >= 1), so caller code that
 * branches on it behaves as with the reference.  No GPU needed. */
fyfft_status fyfft_strategy_processable(int strategy, size_t n, int* processable);

/* Strided batches (no counterpart in the reference, which transforms one
 * contiguous array per call): `howmany` transforms of length n; element j
 * of transform b is read from in[b*idist + j*istride] and written to
 * out[b*odist + j*ostride] (element units).  Contiguous rows at one
 * distance run the batched kernels directly; interleaved transforms
 * (idist = odist = 1) run a column kernel that moves rows of adjacent
 * elements through LDS; every other layout goes through a contiguous
 * workspace (gather, transform, scatter).  In place requires identical
 * input and output layouts.  fyfft_plan_elements reports the elements the
 * input layout spans; fyfft_execute_host copies the spans of both layouts
 * (output elements outside the layout keep the values of `out`). */
typedef struct fyfft_layout {
  size_t istride, idist, ostride, odist;
} fyfft_layout;
fyfft_status fyfft_plan_create_many(fyfft_plan** plan, size_t n, size_t howmany,
                                    const fyfft_layout* layout, fyfft_precision precision);

/* Multi-dimensional C2C (no counterpart in the reference): `batch`
 * contiguous row-major arrays of rank 1..3 with dims[0] the slowest axis.
 * Executed as one batched 1-D pass per axis, innermost first: the innermost
 * axis as contiguous rows, the others by the column kernel when the axis
 * length fits one workgroup's LDS (gather/transform/scatter otherwise).
 * FYFFT_INVERSE scales by 1/prod(dims).  In place allowed. */
fyfft_status fyfft_plan_create_nd(fyfft_plan** plan, int rank, const size_t* dims, size_t batch,
                                  fyfft_precision precision);

/* Real-input / real-output plans (no counterpart in the reference, which
 * has complex<float> transforms only; callers there pass real data as a
 * complex row with zero imaginary parts).  `batch` rows of real length n:
 *   fyfft_execute(plan, x, X, FYFFT_FORWARD, stream): R2C, x is batch x n
 *     reals, X is batch x (n/2 + 1) complex, X[k] = sum_j x[j] e^{-2 pi i jk/n};
 *   fyfft_execute(plan, X, x, FYFFT_INVERSE, stream): C2R, the inverse of a
 *     Hermitian spectrum given by its bins 0..n/2, scaled by 1/n (the
 *     imaginary parts of bin 0, and of bin n/2 for even n, are ignored);
 *   FYFFT_ADJOINT returns FYFFT_UNSUPPORTED.
 * Element type: float/double for the real side, fyfft's complex layout for
 * the spectrum.  In and out must not overlap.  Even n runs through one
 * complex transform of length n/2 (half the HBM bytes of a C2C transform),
 * odd n through one of length n; lengths whose transform fits one
 * workgroup's LDS run as one fused kernel, others as the planned complex
 * transform plus elementwise passes.  fyfft_plan_elements reports batch*n;
 * fyfft_execute_host copies batch*n reals one way and batch*(n/2+1)
 * complex values the other. */
fyfft_status fyfft_plan_create_real(fyfft_plan** plan, size_t n, size_t batch,
                                    fyfft_precision precision);

/* Grouped ragged plans (no counterpart in the reference, which plans one
 * length at a time): `count` transforms of lengths lengths[0..count-1],
 * packed back to back (transform i starts at element sum_{j<i} lengths[j]),
 * executed by fyfft_execute as persistent grouped launches that read a
 * device-side task table: per phase (LDS-resident transforms and first
 * four-step passes; row passes; closing passes of four-step Rader and
 * Bluestein), one launch per task family present, the families of a phase
 * running as parallel branches.  FYFFT_BF16 / FYFFT_F16 store the data as
 * 16-bit pairs and compute in FP32 (one rounding at the final store).
 *   FYFFT_GROUPED_GRAPH: capture the phases into a hipGraph the first time
 *     a (in, out, direction, stream) combination runs on a non-null stream,
 *     and replay the graph afterwards;
 *   FYFFT_GROUPED_PER_TRANSFORM: the baseline form -- one launch of the
 *     plain (non-specialised) kernels per transform (16-bit plans widen
 *     each transform into an FP32 copy around it). */
#define FYFFT_GROUPED_GRAPH 1u
#define FYFFT_GROUPED_PER_TRANSFORM 2u
/*   FYFFT_GROUPED_TUNED: autotune the sub-plans per length (radix order,
 *     LDS padding and points per thread; four-step split and tiles;
 *     Bluestein convolution length), lengths carrying the most elements
 *     first, within about 20 s of plan time; choices are recorded as wisdom
 *     (fyfft_wisdom_export), so a later plan of the same lengths tunes
 *     instantly. */
#define FYFFT_GROUPED_TUNED 4u
fyfft_status fyfft_plan_create_grouped(fyfft_plan** plan, const size_t* lengths, size_t count,
                                       fyfft_precision precision, unsigned flags);

/* Elements (complex values) one execution reads and writes: N * batch, or
 * the sum of the lengths of a grouped plan. */
fyfft_status fyfft_plan_elements(const fyfft_plan* plan, size_t* elements);

/* fyfft_plan_create with options (NULL = defaults, same as
 * fyfft_plan_create).  No counterpart in the reference. */
fyfft_status fyfft_plan_create_ex(fyfft_plan** plan, size_t n, size_t batch,
                                  fyfft_precision precision, const fyfft_options* options);

/* Autotuner wisdom as text: the line "fyfft-wisdom 2", then per GPU model a
 * line "device <arch> <compute units>" (e.g. "device gfx950 256") followed
 * by its choices, one line each:
 *   lds N prec fixed ppt pad tpb radices   (LDS-resident lengths; "ldsg":
 *                                            runtime-pass kernels only, the
 *                                            sub-plans of grouped plans)
 *   fs N prec kind n1 ccol crow             (four-step split and tiles; kind
 *                                            0 smooth N, 1 Rader convolution)
 *   bluestein N prec M                      (Bluestein convolution length).
 * export writes it to buf (NUL-terminated; *needed receives the size
 * including the NUL, buf may be NULL to query), import merges it into the
 * process's wisdom -- FYFFT_INVALID_ARGUMENT on malformed text, or when a
 * device line does not match the current GPU (wisdom is only applied on the
 * model it was tuned on) -- and forget clears it.  Plans look up the entries
 * of their own device.  The reference has no plan serialisation. */
fyfft_status fyfft_wisdom_export(char* buf, size_t len, size_t* needed);
fyfft_status fyfft_wisdom_import(const char* text);
fyfft_status fyfft_wisdom_forget(void);

/* Releases every device allocation of the plan; NULL is accepted. */
fyfft_status fyfft_plan_destroy(fyfft_plan* plan);

/* Replaces dispatch<invert>(const float* in, float* out)
 * (foye_FFT.hpp:128-151) on DEVICE pointers: enqueues the batch on
 * `stream` and returns without waiting.  `direction` is a fyfft_direction.
 * A plan may be executed repeatedly; executions that overlap in time on
 * different streams must use different plans (the reference invoker is
 * not reentrant either: mix_radix.hpp:538, rader.hpp:363,
 * bluestein.hpp:240). */
fyfft_status fyfft_execute(const fyfft_plan* plan, const void* in, void* out,
                           int direction, void* stream);

/* fyfft_execute without any allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Single-
 * kernel plans (LDS-resident smooth, Rader and Bluestein lengths, 16-bit
 * storage) run exactly as fyfft_execute.  Four-step plans (smooth lengths
 * beyond one workgroup's LDS, four-step Rader and Bluestein) run their
 * passes through a static device scratch array of FYFFT_NOALLOC_SCRATCH_BYTES
 * (64 MiB per GPU, held from library load until the process ends) instead
 * of the plan's workspace, in chunks of as many transforms as it holds;
 * FYFFT_INVALID_LENGTH if one transform's workspace exceeds it.  Must be
 * called with the plan's device current.  All calls on one GPU share the
 * array: calls that overlap in time on one GPU (other streams, other
 * threads) race.  Grouped, real, strided and multi-dimensional plans return
 * FYFFT_UNSUPPORTED. */
#define FYFFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
fyfft_status fyfft_execute_noalloc(const fyfft_plan* plan, const void* in, void* out,
                                   int direction, void* stream);

/* dispatch<invert> with the reference's HOST-pointer semantics: copies the
 * batch to the device, runs it, copies back and synchronises before
 * returning.  in == out is allowed. */
fyfft_status fyfft_execute_host(const fyfft_plan* plan, const void* in, void* out,
                                int direction);

/* Replaces acquired_execution_strategy(std::size_t) (foye_FFT.hpp:162-184)
 * for a planned transform, and for a bare length (no GPU needed). */
fyfft_status fyfft_plan_strategy(const fyfft_plan* plan, int* strategy);
fyfft_status fyfft_length_strategy(size_t n, int* strategy);

/* Radix passes the planner chooses for a 2/3/5/7-smooth length held in one
 * workgroup's LDS, or for each factor of a four-step split (no GPU needed):
 * each in {2,3,4,5,7,8,9,16}, threes paired as 9.
 * Writes up to `max` radices, sets *count; FYFFT_UNSUPPORTED if N is not
 * smooth.  The counterpart of the reference's split planner
 * (mix_radix.hpp:253-352, 577-598). */
fyfft_status fyfft_length_radices(size_t n, int* radices, int max, int* count);

/* Spectral-convolution transforms (no counterpart in the reference, which
 * has neither real-input transforms nor differentiation; they serve the
 * fused spectral convolution and its hand-written adjoint).  `plan` is an
 * F32 or F64 plan of length L whose batch is the number of rows.  Even L
 * with k <= L/2 + 1 runs one complex transform of length L/2 per row (the
 * real row read as L/2 complex pairs); L/2 must fit one workgroup's LDS.
 * Otherwise L itself must fit (FYFFT_UNSUPPORTED if neither holds).  The
 * first call builds the plan's half-length tables (synchronous, allocates).
 * c_0 = scale0 and c_j = scale for j >= 1; 1 <= k <= L.  Asynchronous on
 * `stream`.
 *
 * fyfft_rfft_modes:  x is rows x L real (float/double), modes is rows x k
 *   complex:  modes[r][j] = c_j * sum_n x[r][n] exp(-2*pi*i*n*j/L).
 * fyfft_irfft_modes: modes is rows x k complex, y is rows x L real:
 *   y[r][n] = Re( sum_{j<k} c_j * modes[r][j] * exp(+2*pi*i*n*j/L) ).
 *   scale0 = 1/L, scale = 2/L gives the real inverse of a Hermitian
 *   spectrum holding the k kept modes (k <= L/2); scale0 = scale = 1 gives
 *   the real part of the adjoint transform. */
fyfft_status fyfft_rfft_modes(const fyfft_plan* plan, const void* x, void* modes, size_t k,
                              double scale0, double scale, void* stream);
fyfft_status fyfft_irfft_modes(const fyfft_plan* plan, const void* modes, void* y, size_t k,
                               double scale0, double scale, void* stream);

/* The same transforms with the modes laid out mode-major when layout is
 * FYFFT_MODES_MODE_MAJOR: mode j of row r at modes[j*rows + r] (rows = the
 * plan's batch), the layout of one GEMM per mode.  FYFFT_MODES_ROW_MAJOR
 * is the layout of fyfft_rfft_modes / fyfft_irfft_modes. */
#define FYFFT_MODES_ROW_MAJOR 0
#define FYFFT_MODES_MODE_MAJOR 1
fyfft_status fyfft_rfft_modes_ex(const fyfft_plan* plan, const void* x, void* modes, size_t k,
                                 double scale0, double scale, int layout, void* stream);
fyfft_status fyfft_irfft_modes_ex(const fyfft_plan* plan, const void* modes, void* y, size_t k,
                                  double scale0, double scale, int layout, void* stream);

/* Mode mixing of the spectral convolution (no counterpart in the reference):
 * one complex GEMM per kept mode on the matrix cores, with the mode-major
 * layouts of fyfft_rfft_modes_ex: modes [k][rows][channels], weights
 * [k][cin][cout] (complex, interleaved re/im).  FYFFT_F32 (complex float,
 * FP32 MFMA) or FYFFT_F64 (complex double, FP64 MFMA); k <= 65535.
 * Asynchronous on `stream`.
 *   FYFFT_MIX_FORWARD:   out[j][r][o] = sum_i p[j][r][i] * q[j][i][o]
 *                        (p: input modes, q: weights, out: output modes)
 *   FYFFT_MIX_ADJOINT_X: out[j][r][i] = sum_o p[j][r][o] * conj(q[j][i][o])
 *                        (p: output-mode gradient, q: weights)
 *   FYFFT_MIX_GRAD_W:    out[j][i][o] = sum_r conj(p[j][r][i]) * q[j][r][o]
 *                        (p: input modes, q: output-mode gradient) */
#define FYFFT_MIX_FORWARD 0
#define FYFFT_MIX_ADJOINT_X 1
#define FYFFT_MIX_GRAD_W 2
fyfft_status fyfft_mix_modes(int kind, fyfft_precision precision, const void* p, const void* q, void* out,
                             size_t k, size_t rows, size_t cin, size_t cout, void* stream);

/* Device bytes the plan holds (tables + workspace). */
fyfft_status fyfft_plan_workspace_bytes(const fyfft_plan* plan, size_t* bytes);

/* Human-readable one-line description of the plan's kernels (for logs). */
fyfft_status fyfft_plan_describe(const fyfft_plan* plan, char* buf, size_t len);

/* Static text for a status code; never NULL. */
const char* fyfft_status_string(fyfft_status status);

/* Text of the last HIP error seen on this thread ("" if none). */
const char* fyfft_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* FYFFT_H */
