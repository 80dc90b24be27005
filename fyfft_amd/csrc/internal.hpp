// Declarations shared between the C ABI (api.hip) and the kernel
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "planner.hpp"

namespace fyfft {
namespace dev {
// Parameters of an LDS-resident single-kernel transform (smooth, Rader or
// Bluestein) of one length.
template <typename V> struct LdsBody {
  int n = 0;          // user length
  int m = 0;          // LDS FFT length: n (smooth), n-1 (Rader), M (Bluestein)
  int threads = 0;    // threads per transform
  int ps = 30;        // LDS pad shift
  int stride = 0;     // LDS elements per transform slot
  PassDesc pd{};
  const V* tw = nullptr;
  const int* perm_in = nullptr;
  const int* perm_out = nullptr;
  const int* pos_out = nullptr;  // Rader: pos_out[s] = i with g^i = s + 1 (mod n)
  const V* chirp = nullptr;
  const V* kern = nullptr;
  S<V> scale = 1;
};
}  // namespace dev

enum Prec { kF32 = 0, kF64 = 1, kBF16 = 2, kF16 = 3 };
enum Dir { kForward = 0, kInverse = 1, kAdjoint = 2 };

inline size_t elem_bytes(int prec) { return prec == kF64 ? 16 : (prec >= kBF16 ? 4 : 8); }

// Everything one LDS-resident launch needs.
struct LdsLaunch {
  LdsGeom geom;
  dev::PassDesc pd;
  const void* tw = nullptr;  // Stockham twiddle table (N-1 entries)
  int fixed = -1;            // index of a compile-time specialised kernel, -1 = generic
};

// True if a runtime-pass kernel instantiated for at most `ppt` points per
// thread (without its radix-16 case when no16) covers every pass of pd for a
// transform of n points on `threads` threads: each pass of radix R then has
// at most cdiv(ppt, R) butterflies per thread.  The launchers check it so a
// geometry that disagrees with the chosen instantiation fails loudly instead
// of leaving outputs unwritten.
inline bool generic_fits(const dev::PassDesc& pd, int n, int threads, int ppt, bool no16) {
  if (threads <= 0 || (long long)ppt * threads < (long long)n) return false;
  long long prod = 1;
  for (int p = 0; p < pd.npass; ++p) {
    if (no16 && pd.radix[p] == 16) return false;
    prod *= pd.radix[p];
  }
  return pd.npass == 0 || prod == n;
}

// --- kernels_lds.hip
// Batched LDS-resident transform: `batch` transforms of length N, transform
// b at in + b*dist (elements).  scale multiplies the output.
hipError_t launch_lds(int prec, int dir, const LdsLaunch& L, const void* in, void* out,
                      size_t batch, size_t dist, double scale, hipStream_t s);
// Index of a compile-time specialised kernel for (prec, radices), or -1; on
// a hit the geometry is replaced by the kernel's own.
int find_fixed_kernel(int prec, const dev::PassDesc& pd, LdsGeom* g);
// Workgroups per CU of a compile-time kernel of the split-exchange form, 0
// for the other forms.
int fixed_kernel_split_w(int idx);

// --- kernels_fourstep.hip
enum FsMode {
  kFsPlainA = 0,  // column pass of a plain transform (x -> tmp, twiddled)
  kFsRaderA = 1,  // column pass reading the Rader gather x[g^-j]
  kFsBlueA = 2,   // column pass reading x * chirp, zero padded
  kFsRaderC = 3,  // inverse column pass, Rader scatter + x0
  kFsBlueC = 4,   // inverse column pass, post-chirp, first n_orig outputs
  kFsPlainB = 5,  // row pass of a plain transform (tmp -> out, transposed)
  kFsConvB = 6,   // row pass of a convolution (fwd, * kernel, inv, conj twiddle)
  kFsRaderC2 = 8  // inverse column pass writing bucket runs (zw) + x0
};
// Column passes that apply the four-step twiddle on their store.  Their
// workgroups keep C rows of roots in LDS after the tile and the Stockham
// table: entry j of column c is w_N^(n2 * s * j), j below the points per
// thread, and one entry of padding against bank conflicts.
constexpr bool fs_col_twiddled(int mode) { return mode == kFsPlainA || mode == kFsRaderA || mode == kFsBlueA; }
constexpr int fs_col_tw_row(int ppt) { return ppt + 1; }
// The convolution row pass applies the conjugate four-step twiddle
// w_N^(-k1*k2) at its last store as hi[k2 >> 4] * lo[k2 & 15]: per row a
// table of fs_row_tw_hi(n2) + 16 roots in LDS after the tile and the
// Stockham table.
constexpr int fs_row_tw_hi(int n2) { return (n2 - 1) / 16 + 1; }
constexpr int fs_row_tw_len(int n2) { return fs_row_tw_hi(n2) + 16; }

template <typename V> struct FsArgs {
  const V* in = nullptr;
  V* out = nullptr;
  V* tmp = nullptr;
  long long in_dist = 0, out_dist = 0, tmp_dist = 0;  // per transform, elements
  int n1 = 0, n2 = 0;          // column length, row length
  int log2c = 0;               // columns (rows) per workgroup = 1 << log2c
  int nsub = 0;                // grouped plans: columns (rows) per workgroup when nonzero, any count
  int tmp_row = 0;             // twiddled column pass: workspace row stride when nonzero (else n2)
  // column pass: transform t at in + (t / group) * in_gdist + (t % group) *
  // in_dist when group > 0 (likewise tmp), else in + t * in_dist
  int group = 0;
  long long in_gdist = 0, tmp_gdist = 0;
  int win_bits = 0;            // Rader scatter (kFsRaderC): destinations in windows of 2^win_bits, 0: one sweep
  int threads = 0;             // threads per sub-transform
  int ps = 4, stride = 0;      // LDS pad shift and per-sub-transform stride
  dev::PassDesc pd{};          // passes of the sub-length this kernel runs
  const V* tw = nullptr;       // Stockham table of that sub-length
  const V* tw_hi = nullptr;    // w_N^(i << lo_bits)
  const V* tw_lo = nullptr;    // w_N^i, i < 2^lo_bits
  int lo_bits = 0;
  float inv_n = 0;             // 1 / (n1 * n2): FP32 twiddles are computed (dev::root_f32) for n1 * n2 <= 2^24
  int n_orig = 0;              // Rader/Bluestein: the user length
  const int* perm = nullptr;   // Rader gather (A) or scatter (C) permutation
  const V* chirp = nullptr;    // Bluestein pre/post chirp
  const V* kern = nullptr;     // kernel spectrum in [k1][k2] order, 1/L folded in
  V* x0buf = nullptr;          // Rader: x[0] per transform
  dev::S<V> scale = 1;
  // bucketed Rader output (kFsRaderC2): element i of column group b is
  // read from LDS element lidx[b*N1*C + i] and written to zw[b*N1*C + i]
  V* zw = nullptr;
  long long zw_dist = 0;
  const int* lidx = nullptr;
  // row passes: stage the workgroup's rows through LDS with coalesced loads
  // (rows that are not a whole number of 128-byte lines)
  int stage_rows = 0;
  // column pass: input rows that are not a whole number of 128-byte lines
  // are read through the caches (a workgroup's row segments then share
  // their first and last lines with the neighbouring column groups);
  // row pass: the same for the transposed output when N1 is not
  int col_cached = 0, out_cached = 0;
};

// One four-step pass; `args` points at FsArgs<float2> or FsArgs<double2>.
// g gives threads per sub-transform, sub-transforms per workgroup and LDS
// bytes.  dir (-1/+1) only matters for the plain modes.
// fixed_n: a length from find_fs_fixed selects the compile-time
// sub-transform, 0 the generic one.
hipError_t launch_fs_pass(int prec, bool col, int mode, int dir, const void* args, size_t batch,
                          const LdsGeom& g, int fixed_n, hipStream_t s);
// Length of a compile-time sub-transform matching (prec, passes, n, threads), or 0.
// all_modes: the caller runs modes other than the plain column/row passes
// (convolution plans, the layout kernels), which the odd-radix compile-time
// forms do not serve.
int find_fs_fixed(int prec, const dev::PassDesc& pd, int n, int threads, bool all_modes);
// Plain column (col) or row pass with a compile-time odd-radix sub-transform
// of length n (a plain-only length of find_fs_fixed).
hipError_t launch_fs_odd(int prec, int n, bool col, int dir, const void* args, size_t batch, size_t lds,
                         int threads, hipStream_t s);

// --- kernels_conv.hip
template <typename V> struct ConvArgs {
  const V* in = nullptr;
  V* out = nullptr;
  long long batch = 0, dist = 0;
  int n_orig = 0;              // user length N
  int m = 0;                   // convolution length (Rader: N-1, Bluestein: M)
  int threads = 0, per_block = 1, ps = 4, stride = 0;
  dev::PassDesc pd{};
  const V* tw = nullptr;       // Stockham table of m
  const int* perm_in = nullptr;   // Rader: g^-j mod N
  const int* perm_out = nullptr;  // Rader: g^i mod N
  const int* pos_out = nullptr;   // Rader: i with g^i = s + 1 mod N, s < N - 1
  const V* chirp = nullptr;    // Bluestein chirp for this direction
  const V* kern = nullptr;     // kernel spectrum / m, natural order
  dev::S<V> scale = 1;
};
// Fused single-kernel Rader (rader=true) or Bluestein, whole convolution in LDS.
// --- kernels_mix.hip
// Mode mixing of the spectral convolution (complex64 or complex128 by prec,
// mode-major operands: modes [K][rows][C], weights [K][cin][cout]).
enum MixKind { kMixForward = 0, kMixAdjointX = 1, kMixGradW = 2 };
hipError_t launch_mix_modes(int prec, int kind, const void* p, const void* q, void* out, long long K,
                            long long rows, long long cin, long long cout, hipStream_t s);

// fixed: index from find_conv_fixed, or -1 for the runtime passes
hipError_t launch_conv_lds(int prec, bool rader, const void* args, const LdsGeom& g, int fixed, hipStream_t s);
// Index of a compile-time convolution kernel of length g->n with the pass
// list pd (g takes its geometry), or -1.
int find_conv_fixed(int prec, const dev::PassDesc& pd, LdsGeom* g);

// Table generation (FP64 results in double2 buffers unless noted).
// c[n] = exp(sign * i*pi*(n^2 mod 2N)/N), n < count
hipError_t launch_gen_chirp(double2* c, long long n_orig, long long count, int sign, hipStream_t s);
// Bluestein kernel sequence b of length M from c (optionally conjugated)
hipError_t launch_gen_bluestein_b(double2* b, const double2* c, long long n_orig, long long m,
                                  bool conj, hipStream_t s);
// Rader kernel sequence b[j] = exp(sign*2*pi*i*perm_out[j]/N), j < L
hipError_t launch_gen_rader_b(double2* b, const int* perm_out, long long n_orig, long long l,
                              int sign, hipStream_t s);
// out[i] = in[src(i)] * scale stored as prec; src(i) = i, or for a [k1][k2]
// permutation of an N1 x N2 spectrum src(k1*N2 + k2) = k1 + N1*k2.
hipError_t launch_convert(int prec, void* out, const double2* in, long long count, double scale,
                          long long n1, long long n2, bool permute, hipStream_t s);
// out[i] = exp(-2*pi*i * i*step / D) stored as prec, i < count
hipError_t launch_gen_roots(int prec, void* out, long long count, long long step, long long d,
                            hipStream_t s);

// --- kernels_spectral.hip
template <typename V> struct SpecArgs {
  const dev::S<V>* real_in = nullptr;  // rfft_modes input, rows x n
  V* cplx_out = nullptr;               // rfft_modes output, rows x k
  const V* cplx_in = nullptr;          // irfft_modes input, rows x k
  dev::S<V>* real_out = nullptr;       // irfft_modes output, rows x n
  long long rows = 0;
  long long rdist = 0, kdist = 1;      // mode j of row r at r*rdist + j*kdist
  int n = 0, k = 0;
  int threads = 0, per_block = 1, ps = 4, stride = 0;
  dev::PassDesc pd{};
  const V* tw = nullptr;
  dev::S<V> a0 = 1, a = 1;             // scale of mode 0 and of modes 1..k-1
  const V* rtw = nullptr;              // half-length forms: exp(-2 pi i k / n), k <= n/2
  unsigned mode_bytes = 0;             // inverse: bytes the modes array spans (under 2 GiB, a != 0), else 0
};
hipError_t launch_spectral(int prec, bool inverse, const void* args, const LdsGeom& g, hipStream_t s);
// Half-length forms (even n, k <= n/2 + 1): pd/tw/g describe the length n/2;
// fixed_h = n/2 selects compile-time passes matching the FP32/FP64 fixed
// kernels' geometry, 0 the runtime passes.
hipError_t launch_spectral_half(int prec, bool inverse, const void* args, const LdsGeom& g, int fixed_h,
                                hipStream_t s);

// --- kernels_rader.hip
// The Rader output permutation of four-step plans through LDS: chunk k of y
// (Q consecutive elements after y[0]) gathers its elements from zw (runs
// written by the closing column pass) at pos[i].x, places them at offset
// pos[i].y and is stored contiguously.  One workgroup per (transform, chunk).
hipError_t launch_rader_bucket(int prec, const void* zw, void* y, long long rows, long long L, int Q,
                               const int2* pos, long long zw_dist, long long y_dist, hipStream_t s);

// --- kernels_layout.hip
// Strided transforms.  Element j of transform (plane p, column c) sits at
// p*pdist + c + j*stride (columns adjacent in memory, as the non-contiguous
// axes of a row-major multi-dimensional array): a workgroup transforms
// 1 << log2c adjacent columns of one plane in LDS, reading and writing rows
// of adjacent elements.  In place when in == out with equal layouts.
template <typename V> struct ColArgs {
  const V* in = nullptr;
  V* out = nullptr;
  long long istride = 0, ostride = 0, ipdist = 0, opdist = 0;
  int n = 0;           // transform length
  int ncols = 0;       // columns per plane
  int log2c = 0;       // columns per workgroup = 1 << log2c
  int threads = 0;     // threads per column transform
  int ps = 4, stride = 0;
  dev::PassDesc pd{};
  const V* tw = nullptr;
  dev::S<V> scale = 1;
  // inner planes (the two passes of a long column axis split as A x B):
  // launch plane q is outer plane q / inner, inner plane b = q % inner at
  // extra offsets b * ipdist2 / b * opdist2
  int inner = 1;
  long long ipdist2 = 0, opdist2 = 0;
  // if set, output element k of inner plane b is multiplied by w_N^(b*k)
  // (w_N^m = tw_hi[m >> lo_bits] * tw_lo[m & (2^lo_bits - 1)], conjugated
  // for the inverse direction)
  const V* tw_hi = nullptr;
  const V* tw_lo = nullptr;
  int lo_bits = 0;
};
hipError_t launch_columns(int prec, int dir, const void* args, size_t planes, const LdsGeom& g, int fixed_n,
                          hipStream_t s);
// Gather transforms t0 .. t0+count-1 (t = p*ncols + c) of a general strided
// layout into contiguous rows of ws (scatter: the reverse).  Element j of
// transform (p, c) at p*pdist + c*cdist + j*stride.
hipError_t launch_gather(int prec, bool scatter, const void* strided, void* ws, long long n,
                         long long stride, long long cdist, long long ncols, long long pdist,
                         long long t0, long long count, hipStream_t s);

// --- kernels_real.hip
// Fused LDS-resident real transforms (one row per workgroup slot):
//   kRealR2cHalf / kRealC2rHalf: even n through one complex FFT of n/2,
//   kRealR2cFull / kRealC2rFull: any n through one complex FFT of n.
// R2C rows: n reals in, n/2+1 complex out; C2R the reverse.
enum RealMode { kRealR2cHalf = 0, kRealC2rHalf = 1, kRealR2cFull = 2, kRealC2rFull = 3 };
template <typename V> struct RealArgs {
  const void* in = nullptr;
  void* out = nullptr;
  long long rows = 0;
  int n = 0;  // real length
  int threads = 0, per_block = 1, ps = 4, stride = 0;
  dev::PassDesc pd{};
  const V* tw = nullptr;   // Stockham table of the complex FFT (n/2 or n)
  const V* rtw = nullptr;  // exp(-2*pi*i*k/n), k <= n/2 (half forms)
  dev::S<V> scale = 1;
};
// fixed_len: the complex length when g is the geometry of a compile-time
// kernel of that length (compile-time passes), 0 for the runtime passes.
hipError_t launch_real_lds(int prec, int mode, const void* args, const LdsGeom& g, int fixed_len, hipStream_t s);
// Elementwise passes of the multi-pass real forms over `rows` rows of real
// length n: Post (half spectrum Z -> X[0..n/2], times scale), Pre (X -> Z),
// Expand (real -> complex), Truncate (full spectrum -> bins 0..n/2, times
// scale), HermExpand (bins 0..n/2 -> full spectrum), RealPart (times scale).
enum RealOp { kRealOpPost = 0, kRealOpPre = 1, kRealOpExpand = 2, kRealOpTruncate = 3,
              kRealOpHermExpand = 4, kRealOpRealPart = 5 };
hipError_t launch_real_pass(int prec, int op, const void* src, void* dst, long long rows, int n,
                            const void* w, double scale, hipStream_t s);

// --- kernels_grouped.hip
enum GroupedKind { kGkLds = 0, kGkRader = 1, kGkBlue = 2, kGkFsPlain = 3, kGkFsRader = 4, kGkFsBlue = 5 };

// One workgroup's work: phase 0 LDS tasks run `count` transforms of one
// length starting at list position `first`; four-step tasks run the column
// (or row) group starting at `count` of the transform at list position
// `first`.
struct GTask {
  int len;    // index into the GLen table
  int phase;  // 0, 1, 2 (see kernels_grouped.hip)
  int first;
  int count;
};

// Device descriptor of one distinct length of a grouped plan; forward tables
// are stored in lds / col / row, the inverse chirp and kernel spectrum apart.
template <typename V> struct GLen {
  int kind = kGkLds;
  int n = 0;
  int lds_ppt = 16, col_ppt = 16, row_ppt = 16;
  dev::LdsBody<V> lds;
  FsArgs<V> col, row;
  const V* chirp_inv = nullptr;
  const V* kern_inv = nullptr;
};

template <typename V> struct GroupedArgs {
  const GTask* tasks = nullptr;
  int xbase[dev::kXcds + 1] = {0};  // tasks of XCD x are [xbase[x], xbase[x+1])
  const GLen<V>* lens = nullptr;
  const long long* off = nullptr;   // element offset of each transform (list order)
  const long long* woff = nullptr;  // workspace offset of each transform (list order)
  const void* in = nullptr;         // storage elements: V, or 16-bit pairs for BF16 / FP16 plans
  void* out = nullptr;
  V* ws = nullptr;
  V* x0buf = nullptr;
  int tables = 0;     // 0 forward, 1 inverse/adjoint tables
  int normalize = 0;  // scale by 1/n
};
// Task families of the grouped kernels (grouped.hpp): the three LDS-resident
// kinds and the four-step passes.  Phase 0: kGfLds .. kGfColBlueA, phase 1:
// the row passes, phase 2: the closing column passes.
enum GroupFamily {
  kGfLds = 0, kGfRader = 1, kGfBlue = 2, kGfColPlain = 3, kGfColRaderA = 4, kGfColBlueA = 5,
  kGfRowPlain = 6, kGfRowConv = 7, kGfColRaderC = 8, kGfColBlueC = 9, kGroupFamilies = 10
};
inline int family_phase(int fam) { return fam <= kGfColBlueA ? 0 : (fam <= kGfRowConv ? 1 : 2); }
// Grouped launches come in workgroup-size classes of 256, 512 and 1024
// threads with LDS budgets of 1/4, 1/2 and all of a CU's LDS.
constexpr int kGroupClasses = 3;
constexpr int kGroupClassThreads[kGroupClasses] = {256, 512, 1024};
// One family kernel of a grouped plan: storage kF32 / kF64 (data in the
// compute type) or kBF16 / kF16 (FP32 arithmetic, 16-bit data); hi: for
// the LDS families the doubled points-per-thread bound (1024 threads only),
// for the four-step families the form that reads the sub-length's Stockham
// table from global memory (the others copy it into LDS after the tile);
// `args` points at GroupedArgs<float2> or GroupedArgs<double2>.  With
// prepare, only sets the kernel's dynamic-LDS limit (plan time, so the
// launch path holds nothing but launches and can be captured in a graph).
hipError_t group_family(int storage, bool prepare, int fam, int dir, bool hi, int block, const void* args,
                        unsigned grid, size_t lds, hipStream_t s);
hipError_t group_family_lowp(bool bf16, bool prepare, int fam, int dir, bool hi, int block, const void* args,
                             unsigned grid, size_t lds, hipStream_t s);
// n complex values between 16-bit storage and complex float (to_f32: from
// 16-bit); the per-transform baseline of 16-bit grouped plans.
hipError_t launch_convert16(bool bf16, bool to_f32, const void* src, void* dst, long long n, hipStream_t s);

// --- kernels_lowp.hip
// BF16/FP16 storage with FP32 arithmetic for the lengths of the FP32
// compile-time kernels (N = 256 .. 16384 powers of two); pd receives the
// pass list the FP32 twiddle table must be built for.
bool lowp_fixed_supported(size_t n, dev::PassDesc* pd);
hipError_t launch_lowp_fixed(bool bf16, int dir, size_t n, const void* in, void* out, size_t batch,
                             const void* tw, double scale, hipStream_t s);

// --- kernels_mfma.hip
// BF16/FP16-storage transforms with radix-16 stages on MFMA: N = R0 * 16^P
// for the supported powers of two (256 .. 16384); *r0 receives R0.
bool mfma16_supported(size_t n, int* r0);
hipError_t launch_mfma16(bool bf16, int dir, size_t n, const void* in, void* out, size_t batch,
                         const void* tw, double scale, hipStream_t s);

// --- kernels_tables.hip
// Stockham twiddles for pass list pd (N-1 entries) in FP64, stored as prec.
hipError_t launch_gen_stockham_twiddles(int prec, void* tw, int n, const dev::PassDesc& pd,
                                        hipStream_t s);

}  // namespace fyfft
