// Four-step passes for lengths beyond one workgroup's LDS, and the
// convolution pipelines (Rader, Bluestein) built on them.
//
// Length N = N1 * N2, input index n = n1*N2 + n2, output index
// k = k1 + N1*k2:
//   X[k1 + N1*k2] = sum_n2 w_N2^(n2*k2) * [ w_N^(n2*k1) * sum_n1 x[n1*N2 + n2] w_N1^(n1*k1) ]
// Column pass: C consecutive columns n2 per workgroup (coalesced rows of C
//   elements), length-N1 FFTs in LDS, twiddle w_N^(n2*k1) on the way out,
//   tmp[k1*N2 + n2].
// Row pass: C consecutive rows k1 (one contiguous span), length-N2 FFTs in
//   LDS, then either the transposed store X[k1 + N1*k2] (plain transform) or,
//   for a convolution, pointwise multiply by the kernel spectrum (stored in
//   the same [k1][k2] order), the inverse length-N2 FFT and the conjugate
//   twiddle w_N^(-n2*k1), back into tmp -- the first half of the inverse
//   four-step, so the spectrum never needs reordering.
// The inverse column pass then finishes the convolution and applies the
// Rader scatter or the Bluestein post-chirp.
//
// Twiddles w_N^m (m < N) come from two FP64-generated tables,
// w_N^m = hi[m >> lo_bits] * lo[m & (2^lo_bits - 1)], each about sqrt(N)
// entries, so they stay cache resident for any N.
#include "internal.hpp"

namespace fyfft {
using namespace dev;

template <int DIR, typename V>
__device__ __forceinline__ V big_twiddle(const FsArgs<V>& a, long long m) {
  const V w = cmul(a.tw_hi[m >> a.lo_bits], a.tw_lo[m & ((1LL << a.lo_bits) - 1)]);
  return DIR < 0 ? w : V{w.x, -w.y};
}

// ------------------------------------------------------------ column pass
template <typename V, int PPT, int DIR, int MODE>
__global__ __launch_bounds__(1024) void k_fs_col(FsArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int N1 = a.n1, N2 = a.n2, C = 1 << a.log2c;
  const int groups = (N2 + C - 1) >> a.log2c;
  const long long t = blockIdx.x / groups;
  const int c0 = (int)(blockIdx.x - t * groups) << a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int total = N1 << a.log2c;

  const V* x = a.in + t * a.in_dist;
  V* tmp = a.tmp + t * a.tmp_dist;
  V* out = a.out + t * a.out_dist;

  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int n1 = i >> a.log2c, c = i & (C - 1), n2 = c0 + c;
    if (n2 >= N2) continue;
    const int n = n1 * N2 + n2;
    V v;
    if constexpr (MODE == kFsPlainA) {
      v = x[n];
    } else if constexpr (MODE == kFsRaderA) {
      v = x[a.perm[n]];
    } else if constexpr (MODE == kFsBlueA) {
      v = n < a.n_orig ? cmul(x[n], a.chirp[n]) : V{0, 0};
    } else {  // inverse column pass of a convolution: input is tmp
      v = tmp[n];
    }
    lds_all[c * stride + n1 + (n1 >> ps)] = v;
  }
  if constexpr (MODE == kFsRaderA) {
    if (c0 == 0 && threadIdx.x == 0) a.x0buf[t] = x[0];
  }
  __syncthreads();
  lds_fft<V, PPT, DIR>(lds_all + slot * stride, N1, a.threads, tid, a.pd, a.tw,
                       c0 + slot < N2, ps);

  V x0 = V{0, 0};
  if constexpr (MODE == kFsRaderC) x0 = a.x0buf[t];
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int k1 = i >> a.log2c, c = i & (C - 1), n2 = c0 + c;
    if (n2 >= N2) continue;
    V v = lds_all[c * stride + k1 + (k1 >> ps)];
    if constexpr (MODE == kFsPlainA || MODE == kFsRaderA || MODE == kFsBlueA) {
      tmp[(long long)k1 * N2 + n2] = cmul(v, big_twiddle<DIR>(a, (long long)n2 * k1));
    } else if constexpr (MODE == kFsRaderC) {
      // here k1 plays the role of n1 of the convolution output
      const int n = k1 * N2 + n2;
      out[a.perm[n]] = cscale(cadd(v, x0), a.scale);
    } else {  // kFsBlueC
      const int n = k1 * N2 + n2;
      if (n < a.n_orig) out[n] = cscale(cmul(v, a.chirp[n]), a.scale);
    }
  }
}

// ------------------------------------------------------------ row pass
template <typename V, int PPT, int DIR, int MODE>
__global__ __launch_bounds__(1024) void k_fs_row(FsArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int N1 = a.n1, N2 = a.n2, C = 1 << a.log2c;
  const int groups = (N1 + C - 1) >> a.log2c;
  const long long t = blockIdx.x / groups;
  const int r0 = (int)(blockIdx.x - t * groups) << a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int rows = min(C, N1 - r0);
  const int total = rows * N2;

  V* tmp = a.tmp + t * a.tmp_dist;
  V* out = a.out + t * a.out_dist;
  const V* src = tmp + (long long)r0 * N2;

  // rows r0 .. r0+rows-1 are one contiguous span of tmp
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int c = i / N2, n2 = i - c * N2;
    lds_all[c * stride + n2 + (n2 >> ps)] = src[i];
  }
  __syncthreads();
  const bool active = slot < rows;
  V* lds = lds_all + slot * stride;
  const int k1 = r0 + slot;

  if constexpr (MODE == kFsPlainB) {
    lds_fft<V, PPT, DIR>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    // transposed store X[k1 + N1*k2]: C consecutive k1 per k2
    for (int i = threadIdx.x; i < (N2 << a.log2c); i += blockDim.x) {
      const int k2 = i >> a.log2c, c = i & (C - 1);
      if (c >= rows) continue;
      out[(long long)(r0 + c) + (long long)N1 * k2] = cscale(lds_all[c * stride + k2 + (k2 >> ps)], a.scale);
    }
  } else {  // kFsConvB: forward, multiply by kernel spectrum, inverse, conj twiddle
    lds_fft<V, PPT, -1>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    if (a.x0buf != nullptr && r0 == 0 && threadIdx.x == 0) {
      // Rader: X[0] = x0 + sum_{n>=1} x[n] = x0 + A[0] (row k1 = 0, k2 = 0)
      const V x0 = a.x0buf[t];
      a.out[t * a.out_dist] = cscale(cadd(x0, lds_all[0]), a.scale);
    }
    __syncthreads();
    if (active) {
      const V* kr = a.kern + (long long)k1 * N2;
      for (int k2 = tid; k2 < N2; k2 += a.threads) {
        const int li = k2 + (k2 >> ps);
        lds[li] = cmul(lds[li], kr[k2]);
      }
    }
    __syncthreads();
    lds_fft<V, PPT, +1>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    V* dst = tmp + (long long)r0 * N2;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
      const int c = i / N2, n2 = i - c * N2;
      const V v = lds_all[c * stride + n2 + (n2 >> ps)];
      dst[i] = cmul(v, big_twiddle<+1>(a, (long long)n2 * (r0 + c)));
    }
  }
}

// ------------------------------------------------------------ launch
template <typename V, int PPT, int DIR, int MODE>
static hipError_t go_col(const FsArgs<V>& a, size_t batch, size_t lds, int threads, hipStream_t s) {
  auto k = k_fs_col<V, PPT, DIR, MODE>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const size_t groups = ((size_t)a.n2 + (1 << a.log2c) - 1) >> a.log2c;
  hipLaunchKernelGGL(k, dim3((unsigned)(groups * batch)), dim3(threads), lds, s, a);
  return hipGetLastError();
}
template <typename V, int PPT, int DIR, int MODE>
static hipError_t go_row(const FsArgs<V>& a, size_t batch, size_t lds, int threads, hipStream_t s) {
  auto k = k_fs_row<V, PPT, DIR, MODE>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const size_t groups = ((size_t)a.n1 + (1 << a.log2c) - 1) >> a.log2c;
  hipLaunchKernelGGL(k, dim3((unsigned)(groups * batch)), dim3(threads), lds, s, a);
  return hipGetLastError();
}

template <typename V, int PPT>
static hipError_t dispatch_fs(bool col, int mode, int dir, const FsArgs<V>& a, size_t batch,
                              size_t lds, int threads, hipStream_t s) {
  if (col) {
    switch (mode) {
      case kFsPlainA:
        return dir < 0 ? go_col<V, PPT, -1, kFsPlainA>(a, batch, lds, threads, s)
                       : go_col<V, PPT, +1, kFsPlainA>(a, batch, lds, threads, s);
      case kFsRaderA: return go_col<V, PPT, -1, kFsRaderA>(a, batch, lds, threads, s);
      case kFsBlueA: return go_col<V, PPT, -1, kFsBlueA>(a, batch, lds, threads, s);
      case kFsRaderC: return go_col<V, PPT, +1, kFsRaderC>(a, batch, lds, threads, s);
      case kFsBlueC: return go_col<V, PPT, +1, kFsBlueC>(a, batch, lds, threads, s);
    }
  } else {
    switch (mode) {
      case kFsPlainB:
        return dir < 0 ? go_row<V, PPT, -1, kFsPlainB>(a, batch, lds, threads, s)
                       : go_row<V, PPT, +1, kFsPlainB>(a, batch, lds, threads, s);
      case kFsConvB: return go_row<V, PPT, -1, kFsConvB>(a, batch, lds, threads, s);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fs_pass(int prec, bool col, int mode, int dir, const void* args, size_t batch,
                          const LdsGeom& g, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const int threads = g.threads * g.per_block;
  if (prec == kF64) {
    const auto& a = *static_cast<const FsArgs<double2>*>(args);
    return g.ppt <= 8 ? dispatch_fs<double2, 8>(col, mode, dir, a, batch, g.lds_bytes, threads, s)
                      : dispatch_fs<double2, 16>(col, mode, dir, a, batch, g.lds_bytes, threads, s);
  }
  const auto& a = *static_cast<const FsArgs<float2>*>(args);
  return g.ppt <= 16 ? dispatch_fs<float2, 16>(col, mode, dir, a, batch, g.lds_bytes, threads, s)
                     : dispatch_fs<float2, 32>(col, mode, dir, a, batch, g.lds_bytes, threads, s);
}

}  // namespace fyfft
