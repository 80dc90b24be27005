// Workgroup-level bodies of every transform kernel, shared by the batched
// kernels (one launch per plan) and the grouped ragged kernel (one task per
// workgroup from a device task table).  Each body takes pointers already
// positioned at its transform, so the callers only differ in how they find
// the transform: batch index * distance, or a per-transform offset table.
//
// All bodies call __syncthreads(); every thread of the workgroup must call
// them, with `active` false for threads that have no transform.
#pragma once

#include "internal.hpp"

namespace fyfft {
namespace dev {

// Global <-> LDS transfers issue up to 16 loads per thread before the first
// dependent store: element i = start + j * step for j < P, guarded by
// i < total.  P must bound the per-thread element count.
template <int P, typename LoadF, typename StoreF>
__device__ __forceinline__ void staged(int start, int step, int total, LoadF ld, StoreF st) {
  using T = decltype(ld(0));
  constexpr int K = P < 16 ? P : 16;
#pragma unroll
  for (int j0 = 0; j0 < P; j0 += K) {
    T buf[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = start + (j0 + j) * step;
      if (i < total) buf[j] = ld(i);
    }
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int i = start + (j0 + j) * step;
      if (i < total) st(i, buf[j]);
    }
  }
}

// FFT policies for the bodies: how the LDS-resident sub-transform is run.
// GenericFft: runtime pass list (any smooth length), passes out of line.
template <typename V, int PPT> struct GenericFft {
  static constexpr int kPpt = PPT;
  template <int D>
  static __device__ __forceinline__ void run(V* lds, int n, int t, int tid, const PassDesc& pd,
                                             const V* tw, bool active, int ps) {
    lds_fft<V, PPT, D>(lds, n, t, tid, pd, tw, active, ps);
  }
};

template <typename V, int D, int N, int T, int Ls, int R, int... Rest>
__device__ __forceinline__ void fixed_lds_passes(V* lds, int tid, const V* tw, bool active, int ps) {
  constexpr int MAXB = cdiv(N / R, T);
  stockham_pass<V, R, MAXB, D, false, false>(lds, lds, N, Ls, T, tid, tw, active, 1, ps);
  if constexpr (sizeof...(Rest) > 0) fixed_lds_passes<V, D, N, T, Ls * R, Rest...>(lds, tid, tw, active, ps);
}

// FixedFft: length, threads per transform and radices known at compile time,
// passes inlined.  The plan must use exactly this geometry and pass list.
template <typename V, int N, int T, int... Rs> struct FixedFft {
  static constexpr int kPpt = N / T;
  template <int D>
  static __device__ __forceinline__ void run(V* lds, int, int, int tid, const PassDesc&,
                                             const V* tw, bool active, int ps) {
    fixed_lds_passes<V, D, N, T, 1, Rs...>(as_lds(lds), tid, tw, active, ps);
  }
};

// ------------------------------------------------------------ smooth
template <typename V, int PPT, int DIR>
__device__ __forceinline__ void body_lds_smooth(const LdsBody<V>& a, const V* x, V* y, V* lds_in,
                                                int tid, bool active) {
  V* lds = as_lds(lds_in);
  const int N = a.m, T = a.threads, ps = a.ps;
  if (active)
    staged<PPT>(tid, T, N, [&](int i) { return gload(x + i); },
                   [&](int i, V v) { lds[i + (i >> ps)] = v; });
  __syncthreads();
  lds_fft<V, PPT, DIR>(lds, N, T, tid, a.pd, a.tw, active, ps);
  if (active)
    for (int i = tid; i < N; i += T) gstore(y + i, cscale(lds[i + (i >> ps)], a.scale));
}

// ------------------------------------------------------------ Rader
// X[0] = x0 + sum_{n>=1} x[n];  X[g^i] = x0 + (a (*) b)[i],  a[j] = x[g^-j]
template <typename V, int PPT>
__device__ __forceinline__ void body_rader(const LdsBody<V>& a, const V* x, V* y, V* lds_in, int tid,
                                           bool active) {
  V* lds = as_lds(lds_in);
  const int T = a.threads, L = a.m, ps = a.ps;
  V x0 = V{0, 0};
  if (active) {
    x0 = x[0];
    staged<PPT>(tid, T, L, [&](int j) { return x[a.perm_in[j]]; },
                   [&](int j, V v) { lds[j + (j >> ps)] = v; });
  }
  __syncthreads();
  lds_fft<V, PPT, -1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  const V sum_rest = lds[0];
  __syncthreads();
  if (active)
    staged<PPT>(tid, T, L, [&](int j) { return a.kern[j]; },
                   [&](int j, V k) {
                     const int li = j + (j >> ps);
                     lds[li] = cmul(lds[li], k);
                   });
  __syncthreads();
  lds_fft<V, PPT, +1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  if (active) {
    staged<PPT>(tid, T, L, [&](int i) { return a.perm_out[i]; },
                   [&](int i, int d) { y[d] = cscale(cadd(lds[i + (i >> ps)], x0), a.scale); });
    if (tid == 0) y[0] = cscale(cadd(x0, sum_rest), a.scale);
  }
}

// ------------------------------------------------------------ Bluestein
// X[k] = chirp[k] * IFFT(FFT(x * chirp, zero-padded to M) * kernel)[k]
template <typename V, int PPT>
__device__ __forceinline__ void body_blue(const LdsBody<V>& a, const V* x, V* y, V* lds_in, int tid,
                                          bool active) {
  V* lds = as_lds(lds_in);
  const int T = a.threads, M = a.m, N = a.n, ps = a.ps;
  if (active)
    staged<PPT>(tid, T, M, [&](int n) { return n < N ? cmul(x[n], a.chirp[n]) : V{0, 0}; },
                   [&](int n, V v) { lds[n + (n >> ps)] = v; });
  __syncthreads();
  lds_fft<V, PPT, -1>(lds, M, T, tid, a.pd, a.tw, active, ps);
  if (active)
    staged<PPT>(tid, T, M, [&](int j) { return a.kern[j]; },
                   [&](int j, V k) {
                     const int li = j + (j >> ps);
                     lds[li] = cmul(lds[li], k);
                   });
  __syncthreads();
  lds_fft<V, PPT, +1>(lds, M, T, tid, a.pd, a.tw, active, ps);
  if (active)
    staged<PPT>(tid, T, N, [&](int k) { return a.chirp[k]; },
                   [&](int k, V c) { y[k] = cscale(cmul(lds[k + (k >> ps)], c), a.scale); });
}

// ------------------------------------------------------------ four-step
template <int DIR, typename V>
__device__ __forceinline__ V big_twiddle(const FsArgs<V>& a, long long m) {
  const V w = cmul(a.tw_hi[m >> a.lo_bits], a.tw_lo[m & ((1LL << a.lo_bits) - 1)]);
  return DIR < 0 ? w : V{w.x, -w.y};
}

// Column pass over columns [c0, c0 + C) of one transform (C = 1 << log2c).
// x: the transform's input, tmp: its workspace, out: its output, x0: its
// Rader x[0] slot.  Element i of the workgroup's N1 x C tile is row
// n1 = i >> log2c, column c = i & (C - 1), so a wavefront reads C-element
// row segments.
template <typename V, typename F, int DIR, int MODE>
__device__ __forceinline__ void body_fs_col(const FsArgs<V>& a, const V* x, V* tmp, V* out, V* x0,
                                            int c0, V* lds_in) {
  constexpr int P = F::kPpt;
  V* lds_all = as_lds(lds_in);
  const int N1 = a.n1, N2 = a.n2, C = 1 << a.log2c, lc = a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int total = N1 << lc;
  const int ncols = min(C, N2 - c0);
  staged<P>(threadIdx.x, blockDim.x, total,
               [&](int i) {
                 const int n1 = i >> lc, c = i & (C - 1);
                 const int n = n1 * N2 + c0 + (c < ncols ? c : 0);
                 V v;
                 if constexpr (MODE == kFsPlainA) v = gload(x + n);
                 else if constexpr (MODE == kFsRaderA) v = x[a.perm[n]];
                 else if constexpr (MODE == kFsBlueA) v = n < a.n_orig ? cmul(x[n], a.chirp[n]) : V{0, 0};
                 else v = tmp[n];
                 return v;
               },
               [&](int i, V v) {
                 const int n1 = i >> lc, c = i & (C - 1);
                 if (c < ncols) lds_all[c * stride + n1 + (n1 >> ps)] = v;
               });
  if constexpr (MODE == kFsRaderA) {
    if (c0 == 0 && threadIdx.x == 0) *x0 = x[0];
  }
  __syncthreads();
  const bool active = slot < C && slot < ncols;
  F::template run<DIR>(lds_all + (slot < C ? slot : 0) * stride, N1, a.threads, tid, a.pd, a.tw, active, ps);
  V xz = V{0, 0};
  if constexpr (MODE == kFsRaderC) xz = *x0;
  if constexpr (MODE == kFsPlainA || MODE == kFsRaderA || MODE == kFsBlueA) {
    staged<P>(threadIdx.x, blockDim.x, total,
                 [&](int i) {
                   const int k1 = i >> lc, c = i & (C - 1), n2 = c0 + c;
                   return big_twiddle<DIR>(a, (long long)(c < ncols ? n2 : c0) * k1);
                 },
                 [&](int i, V w) {
                   const int k1 = i >> lc, c = i & (C - 1);
                   if (c < ncols) tmp[(long long)k1 * N2 + c0 + c] = cmul(lds_all[c * stride + k1 + (k1 >> ps)], w);
                 });
  } else if constexpr (MODE == kFsRaderC) {
    // k1 runs over n1 of the convolution output here
    staged<P>(threadIdx.x, blockDim.x, total,
                 [&](int i) {
                   const int k1 = i >> lc, c = i & (C - 1);
                   return a.perm[k1 * N2 + c0 + (c < ncols ? c : 0)];
                 },
                 [&](int i, int d) {
                   const int k1 = i >> lc, c = i & (C - 1);
                   if (c < ncols) out[d] = cscale(cadd(lds_all[c * stride + k1 + (k1 >> ps)], xz), a.scale);
                 });
  } else {  // kFsBlueC
    staged<P>(threadIdx.x, blockDim.x, total,
                 [&](int i) {
                   const int k1 = i >> lc, c = i & (C - 1);
                   const int n = k1 * N2 + c0 + (c < ncols ? c : 0);
                   return n < a.n_orig ? a.chirp[n] : V{0, 0};
                 },
                 [&](int i, V ch) {
                   const int k1 = i >> lc, c = i & (C - 1);
                   const int n = k1 * N2 + c0 + c;
                   if (c < ncols && n < a.n_orig) out[n] = cscale(cmul(lds_all[c * stride + k1 + (k1 >> ps)], ch), a.scale);
                 });
  }
}

// Row pass over rows [r0, r0 + C) of one transform's workspace.
template <typename V, typename F, int DIR, int MODE>
__device__ __forceinline__ void body_fs_row(const FsArgs<V>& a, V* tmp, V* out, const V* x0,
                                            int r0, V* lds_in) {
  constexpr int P = F::kPpt;
  V* lds_all = as_lds(lds_in);
  const int N1 = a.n1, N2 = a.n2, C = 1 << a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int rows = min(C, N1 - r0);
  const int total = rows * N2;
  const V* src = tmp + (long long)r0 * N2;
  staged<P>(threadIdx.x, blockDim.x, total, [&](int i) { return src[i]; },
               [&](int i, V v) {
                 const int c = i / N2, n2 = i - c * N2;
                 lds_all[c * stride + n2 + (n2 >> ps)] = v;
               });
  __syncthreads();
  const bool active = slot < rows;
  V* lds = lds_all + (slot < C ? slot : 0) * stride;
  const int k1 = r0 + slot;
  if constexpr (MODE == kFsPlainB) {
    F::template run<DIR>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    for (int i = threadIdx.x; i < (N2 << a.log2c); i += blockDim.x) {
      const int k2 = i >> a.log2c, c = i & (C - 1);
      if (c >= rows) continue;
      gstore(out + (long long)(r0 + c) + (long long)N1 * k2,
             cscale(lds_all[c * stride + k2 + (k2 >> ps)], a.scale));
    }
  } else {  // kFsConvB
    F::template run<-1>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    if (x0 != nullptr && r0 == 0 && threadIdx.x == 0) {
      // Rader: X[0] = x0 + sum_{n>=1} x[n] = x0 + A[0] (row k1 = 0, k2 = 0)
      out[0] = cscale(cadd(*x0, lds_all[0]), a.scale);
    }
    __syncthreads();
    if (active) {
      const V* kr = a.kern + (long long)k1 * N2;
      staged<P>(tid, a.threads, N2, [&](int k2) { return kr[k2]; },
                   [&](int k2, V k) {
                     const int li = k2 + (k2 >> ps);
                     lds[li] = cmul(lds[li], k);
                   });
    }
    __syncthreads();
    F::template run<+1>(lds, N2, a.threads, tid, a.pd, a.tw, active, ps);
    V* dst = tmp + (long long)r0 * N2;
    staged<P>(threadIdx.x, blockDim.x, total,
                 [&](int i) {
                   const int c = i / N2, n2 = i - c * N2;
                   return big_twiddle<+1>(a, (long long)n2 * (r0 + c));
                 },
                 [&](int i, V w) {
                   const int c = i / N2, n2 = i - c * N2;
                   dst[i] = cmul(lds_all[c * stride + n2 + (n2 >> ps)], w);
                 });
  }
}

}  // namespace dev
}  // namespace fyfft
