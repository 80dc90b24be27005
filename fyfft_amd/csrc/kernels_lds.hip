// LDS-resident batched transforms.
//
// k_lds_generic: any 2/3/5/7-smooth N that fits one workgroup's LDS.  The
//   pass list is a kernel argument; each pass switches on its radix.  HBM is
//   touched once per element in each direction (coalesced copy in, passes in
//   LDS, coalesced copy out with the 1/N scale folded in).
// k_lds_fixed: the same passes with N, radices and geometry fixed at compile
//   time, for the lengths that carry the headline workloads.  Pass 0 reads
//   HBM straight into registers and the last pass writes HBM straight from
//   registers, so a P-pass transform costs P-1 LDS round trips.
#include <cstdio>

#include "bodies.hpp"
#include "split.hpp"

namespace fyfft {
using namespace dev;

template <typename V, int PPT, int DIR>
__global__ __launch_bounds__(1024) void k_lds_generic(const V* in, V* out, long long batch,
                                                      long long dist, int N, int T, int TPB,
                                                      int ps, int lds_stride, PassDesc pd,
                                                      const V* __restrict__ tw, S<V> scale,
                                                      int staged) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int slot = threadIdx.x / T;
  const int tid = threadIdx.x - slot * T;
  const long long t = (long long)blockIdx.x * TPB + slot;
  const bool active = t < batch;
  if (staged) {
    // Short transforms (a wavefront spans many of them; dist == N): HBM is
    // read and written as the workgroup's one contiguous block, through LDS,
    // instead of each thread touching its transform's few elements (lanes N
    // elements apart: a cache line per lane per instruction).  Element e of
    // the block is transform e / N, found as floor((e + 0.5) / N) in FP32
    // (exact for blocks below 2^16 elements, which the host guarantees).
    V* lds = reinterpret_cast<V*>(smem_raw);
    const long long t0 = (long long)blockIdx.x * TPB;
    const int total = (int)min((long long)TPB, batch - t0) * N;
    const float inv = 1.0f / (float)N;
    const V* src = in + t0 * N;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
      const int s = (int)(((float)e + 0.5f) * inv), i = e - s * N;
      lds[s * lds_stride + i + (i >> ps)] = gload(src + e);
    }
    __syncthreads();
    lds_fft<V, PPT, DIR>(lds + slot * lds_stride, N, T, tid, pd, tw, active, ps);
    V* dst = out + t0 * N;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
      const int s = (int)(((float)e + 0.5f) * inv), i = e - s * N;
      gstore(dst + e, cscale(lds[s * lds_stride + i + (i >> ps)], scale));
    }
    return;
  }
  LdsBody<V> b;
  b.n = N;
  b.m = N;
  b.threads = T;
  b.ps = ps;
  b.stride = lds_stride;
  b.pd = pd;
  b.tw = tw;
  b.scale = scale;
  const long long off = (active ? t : 0) * dist;
  body_lds_smooth<V, PPT, DIR>(b, in + off, out + off,
                               reinterpret_cast<V*>(smem_raw) + slot * lds_stride, tid, active);
}

// ------------------------------------------------------------ fixed kernels
template <typename V, int DIR, int N, int T, int PS, int Ls, int R, int... Rest>
__device__ __forceinline__ void run_fixed(const V* gin, V* lds, V* gout, int tid,
                                          const V* __restrict__ tw, bool active, S<V> scale) {
  constexpr bool first = (Ls == 1);
  constexpr bool last = (sizeof...(Rest) == 0);
  constexpr int MAXB = cdiv(N / R, T);
  const V* src = first ? gin : lds;
  V* dst = last ? gout : lds;
  stockham_pass<V, R, MAXB, DIR, first, last>(src, dst, N, Ls, T, tid, tw, active, scale, PS);
  if constexpr (!last) run_fixed<V, DIR, N, T, PS, Ls * R, Rest...>(gin, lds, gout, tid, tw, active, scale);
}

template <typename V, int DIR, int N, int T, int TPB, int PS, int... Rs>
__global__ __launch_bounds__(T * TPB) void k_lds_fixed(const V* in, V* out, long long batch,
                                                       long long dist, const V* __restrict__ tw,
                                                       S<V> scale) {
  constexpr int stride = lds_len(N, PS);
  __shared__ V lds_all[TPB * stride];
  const int slot = threadIdx.x / T;
  const int tid = threadIdx.x - slot * T;
  const long long t = (long long)blockIdx.x * TPB + slot;
  // one transform per workgroup: the grid is exactly the batch, and a
  // compile-time true lets every pass drop its predicate (about 3 % faster)
  const bool active = TPB == 1 || t < batch;
  const V* gin = in + (active ? t : 0) * dist;
  V* gout = out + (active ? t : 0) * dist;
  run_fixed<V, DIR, N, T, PS, 1, Rs...>(gin, lds_all + slot * stride, gout, tid, tw, active, scale);
}

using FixedLauncher = hipError_t (*)(int dir, const void* in, void* out, size_t batch,
                                     size_t dist, const void* tw, double scale, hipStream_t s);

template <typename V, int N, int T, int PS, int W, int... Rs>
static hipError_t launch_fixed_split(int dir, const void* in, void* out, size_t batch, size_t dist,
                                     const void* tw, double scale, hipStream_t s) {
  if (batch > 0xffffffffull) return hipErrorInvalidValue;  // one workgroup per transform
  const dim3 grid((unsigned)batch), block(T);
  if (dir == kForward)
    hipLaunchKernelGGL((k_lds_fixed_split<V, V, -1, N, T, PS, W, Rs...>), grid, block, 0, s, (const V*)in, (V*)out,
                       (long long)dist, (const V*)tw, (S<V>)scale);
  else
    hipLaunchKernelGGL((k_lds_fixed_split<V, V, +1, N, T, PS, W, Rs...>), grid, block, 0, s, (const V*)in, (V*)out,
                       (long long)dist, (const V*)tw, (S<V>)scale);
  return hipGetLastError();
}

template <typename V, int N, int T, int TPB, int PS, int... Rs>
static hipError_t launch_fixed(int dir, const void* in, void* out, size_t batch, size_t dist,
                               const void* tw, double scale, hipStream_t s) {
  const dim3 grid((unsigned)((batch + TPB - 1) / TPB)), block(T * TPB);
  if (dir == kForward)
    hipLaunchKernelGGL((k_lds_fixed<V, -1, N, T, TPB, PS, Rs...>), grid, block, 0, s,
                       (const V*)in, (V*)out, (long long)batch, (long long)dist, (const V*)tw,
                       (S<V>)scale);
  else
    hipLaunchKernelGGL((k_lds_fixed<V, +1, N, T, TPB, PS, Rs...>), grid, block, 0, s,
                       (const V*)in, (V*)out, (long long)batch, (long long)dist, (const V*)tw,
                       (S<V>)scale);
  return hipGetLastError();
}

// Short lengths with compile-time passes: a wavefront spans several
// transforms, so the workgroup's TPB transforms (stored back to back, dist ==
// N) are copied through LDS as one contiguous block, every pass runs LDS to
// LDS, and the block is copied out; the per-transform LDS stride is odd so
// that lanes reading one element of successive transforms hit distinct banks.
template <typename V, int DIR, int N, int T, int TPB, int... Rs>
__global__ __launch_bounds__(T * TPB) void k_lds_fixed_staged(const V* in, V* out, long long batch,
                                                              const V* __restrict__ tw, S<V> scale) {
  constexpr int kS = N | 1, kIter = cdiv(N, T);
  __shared__ V lds[TPB * kS];
  const int slot = threadIdx.x / T;
  const int tid = threadIdx.x - slot * T;
  const long long t0 = (long long)blockIdx.x * TPB;
  const int nt = (int)min((long long)TPB, batch - t0);
  const int total = nt * N;
  const V* src = in + t0 * N;
  V v[kIter];
#pragma unroll
  for (int k = 0; k < kIter; ++k) {
    const int e = threadIdx.x + k * T * TPB;
    if (e < total) v[k] = gload(src + e);
  }
#pragma unroll
  for (int k = 0; k < kIter; ++k) {
    const int e = threadIdx.x + k * T * TPB;
    if (e < total) lds[(e / N) * kS + e % N] = v[k];
  }
  __syncthreads();
  fixed_lds_passes<V, DIR, N, T, 1, Rs...>(lds + slot * kS, tid, tw, slot < nt, 30);
  V* dst = out + t0 * N;
#pragma unroll
  for (int k = 0; k < kIter; ++k) {
    const int e = threadIdx.x + k * T * TPB;
    if (e < total) gstore(dst + e, cscale(lds[(e / N) * kS + e % N], scale));
  }
}

template <typename V, int N, int T, int TPB, int... Rs>
static hipError_t launch_fixed_staged(int dir, const void* in, void* out, size_t batch, size_t,
                                      const void* tw, double scale, hipStream_t s) {
  const dim3 grid((unsigned)((batch + TPB - 1) / TPB)), block(T * TPB);
  if (dir == kForward)
    hipLaunchKernelGGL((k_lds_fixed_staged<V, -1, N, T, TPB, Rs...>), grid, block, 0, s, (const V*)in, (V*)out,
                       (long long)batch, (const V*)tw, (S<V>)scale);
  else
    hipLaunchKernelGGL((k_lds_fixed_staged<V, +1, N, T, TPB, Rs...>), grid, block, 0, s, (const V*)in, (V*)out,
                       (long long)batch, (const V*)tw, (S<V>)scale);
  return hipGetLastError();
}

struct FixedEntry {
  int prec, n, threads, per_block, pad_shift, npass;
  int radix[6];
  FixedLauncher launch;
  bool staged;  // k_lds_fixed_staged: transforms stored back to back (dist == n) only
  int split_w;  // k_lds_fixed_split with split_w workgroups per CU, 0 otherwise
};

// PS = 30 stores the transform unpadded: for these power-of-two radix-8/16
// pass orders the access patterns stay within two LDS cycles per wave, and
// the smaller footprint lets five 4096-point workgroups share a CU.  For
// N = 6720 (FP32) the unpadded 52.5 KB lets three workgroups share a CU
// (padded, only two fit), and 512 threads (8 full waves, 13-14 points per
// thread) beat 420 (6.6 waves): measured 0.35 against 0.46 ms for 16384
// transforms.  Thread counts need not divide N: every pass guards its
// butterfly index.
#define FIXED(P, V, N, T, TPB, PS, NP, ...) \
  {P, N, T, TPB, PS, NP, {__VA_ARGS__}, &launch_fixed<V, N, T, TPB, PS, __VA_ARGS__>, false, 0}
// Split-exchange entries: the geometry (T, TPB, PS) the plan records (which
// the real and spectral kernels of the same length reuse), then the split
// kernel's own pad shift KPS and workgroups per CU W.
#define XFIXED(P, V, N, T, TPB, PS, KPS, W, NP, ...) \
  {P, N, T, TPB, PS, NP, {__VA_ARGS__}, &launch_fixed_split<V, N, T, KPS, W, __VA_ARGS__>, false, W}
#define SFIXED(P, V, N, T, TPB, NP, ...) \
  {P, N, T, TPB, 30, NP, {__VA_ARGS__}, &launch_fixed_staged<V, N, T, TPB, __VA_ARGS__>, true, 0}

// Radix lists here must equal radices_for(N) (planner.cpp) so that the plan's
// twiddle table matches; find_fixed_kernel checks this at plan time.  The
// real-transform and spectral kernels reuse the geometry of 256...8192 (FP32)
// and 1024 / 4096 (FP64) for their compile-time forms, so those entries keep
// the pad shift those forms are built with (padding them measured 1-5 %
// faster for the complex transform alone).
static const FixedEntry kFixed[] = {
    FIXED(kF32, float2, 256, 16, 16, 30, 2, 16, 16),
    FIXED(kF32, float2, 512, 32, 8, 30, 3, 8, 8, 8),
    FIXED(kF32, float2, 1024, 64, 4, 30, 3, 16, 8, 8),
    FIXED(kF32, float2, 2048, 128, 2, 30, 3, 16, 16, 8),
    FIXED(kF32, float2, 4096, 256, 1, 30, 3, 16, 16, 16),
    // 8192 and 16384 through the split-exchange kernel (registers between
    // passes, an exchange buffer of N floats, W workgroups per CU): 8192 x
    // 32768 0.755 -> 0.736 ms (71 -> 73 % of 8 TB/s), 16384 x 16384 0.937 ->
    // 0.743 (57 -> 72 %); the same results to the bit.  2048 and 4096 keep
    // the whole transform in LDS: split (12 / 8 workgroups per CU) they ran
    // faster on one host (0.712 -> 0.698, 0.723 -> 0.707 ms) and slower on
    // the others (0.719 -> 0.755, 0.730 -> 0.741 ms); 8192 at 3 workgroups
    // per CU measured the same as at 4
    XFIXED(kF32, float2, 8192, 512, 1, 30, 4, 4, 4, 16, 8, 8, 8),
    XFIXED(kF32, float2, 16384, 1024, 1, 4, 4, 2, 4, 16, 16, 8, 8),
    FIXED(kF32, float2, 6720, 512, 1, 30, 5, 8, 8, 7, 5, 3),
    // odd-radix lengths of the reference's table, one butterfly of the
    // largest radix per thread and pass, unpadded (odd strides spread over
    // the banks): 56-67 % of 8 TB/s against 34-51 % for the runtime-pass
    // kernel; half the threads measured slower.  Up to 729 the staged block
    // copy is faster again (odd-length rows straddle cache lines when each
    // pass reads its own: N = 343 0.211 -> 0.183 ms, FP64 243 0.225 -> 0.183);
    // from 2187 on (one transform per workgroup) it measured slower
    SFIXED(kF32, float2, 125, 25, 10, 3, 5, 5, 5),
    SFIXED(kF32, float2, 243, 27, 9, 3, 9, 9, 3),
    SFIXED(kF32, float2, 343, 49, 5, 3, 7, 7, 7),
    SFIXED(kF32, float2, 625, 125, 2, 4, 5, 5, 5, 5),
    SFIXED(kF32, float2, 729, 81, 3, 3, 9, 9, 9),
    FIXED(kF32, float2, 2187, 243, 1, 30, 4, 9, 9, 9, 3),
    FIXED(kF32, float2, 2401, 343, 1, 30, 4, 7, 7, 7, 7),
    FIXED(kF32, float2, 3125, 625, 1, 30, 5, 5, 5, 5, 5, 5),
    FIXED(kF32, float2, 6561, 729, 1, 30, 4, 9, 9, 9, 9),
    // the long odd-radix lengths of the reference's table: compile-time
    // passes avoid the runtime-pass kernel's call frames (2.9-3.0x faster for
    // 16807 and 19683, 1.3x for 15625 at 1024 threads, measured against 625-896)
    FIXED(kF32, float2, 15625, 1024, 1, 4, 6, 5, 5, 5, 5, 5, 5),
    FIXED(kF32, float2, 16807, 1024, 1, 4, 5, 7, 7, 7, 7, 7),
    FIXED(kF32, float2, 19683, 1024, 1, 30, 5, 9, 9, 9, 9, 3),
    // short lengths of the reference's table, staged through LDS: 70-80 % of
    // 8 TB/s against 37-71 % for the runtime-pass kernel's staged path
    SFIXED(kF32, float2, 2, 1, 256, 1, 2),
    SFIXED(kF32, float2, 3, 1, 256, 1, 3),
    SFIXED(kF32, float2, 4, 1, 256, 1, 4),
    SFIXED(kF32, float2, 5, 1, 256, 1, 5),
    SFIXED(kF32, float2, 7, 1, 256, 1, 7),
    SFIXED(kF32, float2, 8, 1, 256, 1, 8),
    SFIXED(kF32, float2, 9, 1, 256, 1, 9),
    SFIXED(kF32, float2, 128, 8, 32, 2, 16, 8),
    SFIXED(kF64, double2, 2, 1, 128, 1, 2),
    SFIXED(kF64, double2, 3, 1, 128, 1, 3),
    SFIXED(kF64, double2, 4, 1, 128, 1, 4),
    SFIXED(kF64, double2, 5, 1, 128, 1, 5),
    SFIXED(kF64, double2, 7, 1, 128, 1, 7),
    SFIXED(kF64, double2, 8, 1, 128, 1, 8),
    SFIXED(kF64, double2, 9, 1, 128, 1, 9),
    SFIXED(kF64, double2, 128, 16, 16, 2, 16, 8),
    SFIXED(kF32, float2, 16, 1, 256, 1, 16),
    SFIXED(kF32, float2, 25, 5, 51, 2, 5, 5),
    SFIXED(kF32, float2, 27, 3, 85, 2, 9, 3),
    SFIXED(kF32, float2, 32, 4, 64, 2, 8, 4),
    SFIXED(kF32, float2, 49, 7, 36, 2, 7, 7),
    SFIXED(kF32, float2, 64, 8, 32, 2, 8, 8),
    SFIXED(kF32, float2, 81, 9, 28, 2, 9, 9),
    SFIXED(kF64, double2, 16, 1, 128, 1, 16),
    SFIXED(kF64, double2, 25, 5, 51, 2, 5, 5),
    SFIXED(kF64, double2, 27, 3, 85, 2, 9, 3),
    SFIXED(kF64, double2, 32, 4, 64, 2, 8, 4),
    SFIXED(kF64, double2, 49, 7, 36, 2, 7, 7),
    SFIXED(kF64, double2, 64, 8, 32, 2, 8, 8),
    SFIXED(kF64, double2, 81, 9, 28, 2, 9, 9),
    FIXED(kF64, double2, 1024, 64, 4, 30, 3, 16, 8, 8),
    FIXED(kF64, double2, 4096, 256, 1, 30, 3, 16, 16, 16),  // split exchange at 4 per CU: no faster (74 %)
    // FP64 lengths whose transform fills most of a CU's LDS: split exchange
    // (k_lds_fixed_split), two workgroups per CU: 6720 x 4096 0.216 ->
    // 0.186 ms (51 -> 59 % of 8 TB/s), 6561 x 8192 0.416 -> 0.349, 8192 x
    // 4096 0.211 -> 0.181 (the same results to the bit).  FP32 16384 above
    // likewise; the FP32 odd-radix lengths of that size (15625, 16807,
    // 19683: 20-27 points per thread at 1024 threads) spill at the register
    // budget of two workgroups per CU and ran 28-64 % slower split.
    XFIXED(kF64, double2, 6720, 480, 1, 4, 4, 2, 5, 8, 8, 7, 5, 3),
    // FP64 forms of the same lengths and 256 / 512 / 2048 / 8192 (the
    // runtime-pass FP64 kernel keeps call frames in scratch): 1.4-2.9x
    // faster, 52-74 % of 8 TB/s
    SFIXED(kF64, double2, 125, 25, 10, 3, 5, 5, 5),
    SFIXED(kF64, double2, 243, 27, 9, 3, 9, 9, 3),
    FIXED(kF64, double2, 256, 32, 8, 30, 2, 16, 16),
    FIXED(kF64, double2, 512, 64, 4, 30, 3, 8, 8, 8),
    SFIXED(kF64, double2, 343, 49, 5, 3, 7, 7, 7),
    SFIXED(kF64, double2, 625, 125, 2, 4, 5, 5, 5, 5),
    SFIXED(kF64, double2, 729, 81, 3, 3, 9, 9, 9),
    FIXED(kF64, double2, 2048, 256, 1, 30, 3, 16, 16, 8),
    FIXED(kF64, double2, 2187, 243, 1, 30, 4, 9, 9, 9, 3),
    FIXED(kF64, double2, 2401, 343, 1, 30, 4, 7, 7, 7, 7),
    FIXED(kF64, double2, 3125, 625, 1, 30, 5, 5, 5, 5, 5, 5),
    XFIXED(kF64, double2, 6561, 729, 1, 30, 30, 2, 4, 9, 9, 9, 9),
    XFIXED(kF64, double2, 8192, 512, 1, 4, 4, 2, 4, 16, 8, 8, 8),
};
#undef FIXED
#undef XFIXED
#undef SFIXED

int fixed_kernel_split_w(int idx) {
  return idx >= 0 && idx < (int)(sizeof(kFixed) / sizeof(kFixed[0])) ? kFixed[idx].split_w : 0;
}

int find_fixed_kernel(int prec, const PassDesc& pd, LdsGeom* g) {
  for (int i = 0; i < (int)(sizeof(kFixed) / sizeof(kFixed[0])); ++i) {
    const FixedEntry& e = kFixed[i];
    if (e.prec != prec || e.n != g->n || e.npass != pd.npass) continue;
    bool same = true;
    for (int p = 0; p < pd.npass; ++p) same = same && (e.radix[p] == pd.radix[p]);
    if (!same) continue;
    g->threads = e.threads;
    g->per_block = e.per_block;
    g->ppt = cdiv(e.n, e.threads);  // the generic kernels reuse this geometry
    // the kernel's own geometry replaces the planner's: its points per thread
    // are not those of the lean small-radix instantiation
    g->small_radix = false;
    g->pad_shift = e.pad_shift;
    g->lds_bytes = (size_t)lds_len(e.n, e.pad_shift) * e.per_block * elem_bytes(prec);
    return i;
  }
  return -1;
}

// ------------------------------------------------------------ launch
template <typename V, int PPT>
static hipError_t launch_generic_t(int dir, const LdsLaunch& L, const void* in, void* out,
                                   size_t batch, size_t dist, double scale, hipStream_t s) {
  const LdsGeom& g = L.geom;
  if (!generic_fits(L.pd, g.n, g.threads, PPT, kNo16<V, PPT>)) return hipErrorInvalidValue;
  // staged copies for transforms of 3 or more points and fewer than 8
  // threads stored back to back (measured: 2.5-6x faster for N = 3..32;
  // slower from 8 threads on, where each thread's runs already fill cache
  // lines, and for N = 2); the per-transform LDS stride is made odd so that
  // lanes reading one element of successive transforms hit distinct banks
  const int staged = g.n >= 3 && g.threads < 8 && dist == (size_t)g.n && (long long)g.n * g.per_block < 65536;
  const int stride = lds_len(g.n, g.pad_shift) | staged;
  const dim3 grid((unsigned)((batch + g.per_block - 1) / g.per_block));
  const dim3 block(g.threads * g.per_block);
  const size_t lds = (size_t)stride * g.per_block * sizeof(V);
  if (dir == kForward) {
    auto k = k_lds_generic<V, PPT, -1>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, grid, block, lds, s, (const V*)in, (V*)out, (long long)batch,
                       (long long)dist, g.n, g.threads, g.per_block, g.pad_shift, stride, L.pd,
                       (const V*)L.tw, (S<V>)scale, staged);
  } else {
    auto k = k_lds_generic<V, PPT, +1>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, grid, block, lds, s, (const V*)in, (V*)out, (long long)batch,
                       (long long)dist, g.n, g.threads, g.per_block, g.pad_shift, stride, L.pd,
                       (const V*)L.tw, (S<V>)scale, staged);
  }
  return hipGetLastError();
}

hipError_t launch_lds(int prec, int dir, const LdsLaunch& L, const void* in, void* out,
                      size_t batch, size_t dist, double scale, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  // a staged compile-time kernel serves back-to-back transforms only; other
  // distances run the runtime-pass kernel on the same geometry and table
  if (L.fixed >= 0 && (!kFixed[L.fixed].staged || dist == (size_t)L.geom.n))
    return kFixed[L.fixed].launch(dir, in, out, batch, dist, L.tw, scale, s);
  if (prec == kF64 && L.geom.small_radix) return launch_generic_t<double2, 4>(dir, L, in, out, batch, dist, scale, s);
  if (prec == kF64)
    return L.geom.ppt <= 8 ? launch_generic_t<double2, 8>(dir, L, in, out, batch, dist, scale, s)
                           : launch_generic_t<double2, 16>(dir, L, in, out, batch, dist, scale, s);
  if (L.geom.small_radix) return launch_generic_t<float2, 8>(dir, L, in, out, batch, dist, scale, s);
  return L.geom.ppt <= 16 ? launch_generic_t<float2, 16>(dir, L, in, out, batch, dist, scale, s)
                          : launch_generic_t<float2, 32>(dir, L, in, out, batch, dist, scale, s);
}

}  // namespace fyfft
