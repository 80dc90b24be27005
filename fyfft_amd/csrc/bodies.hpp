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
template <int P, typename LoadF, typename StoreF, int KMAX = 16>
__device__ __forceinline__ void staged(int start, int step, int total, LoadF ld, StoreF st) {
  using T = decltype(ld(0));
  constexpr int K = P < KMAX ? P : KMAX;
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

// The same value, opaque to the optimiser: index arithmetic recomputed from
// it is not merged with identical arithmetic before a long stretch of code
// (the compiler otherwise keeps every element's LDS offset of a load loop
// alive across the FFT for the store loop, and spills).
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// FFT policies for the bodies: how the LDS-resident sub-transform is run.
// GenericFft: runtime pass list (any smooth length), passes out of line.
// TL: the four-step bodies stage the sub-length's Stockham table in LDS
// after their tile (n - 1 entries) and the passes read it there: the
// runtime passes' twiddle loads from global memory sat on the critical path
// of every pass (N = 96000 = 300 x 320 column pass 643 us with them, 444 us
// with none at all).
// CMP: the pass list may hold the composite radices (FP32 four-step
// sub-transforms at 16 points per thread, see radices_composite).
template <typename V, int PPT, bool TL = false, bool CMP = false> struct GenericFft {
  static constexpr int kPpt = PPT;
  static constexpr bool kTwLds = TL;
  template <int D>
  static __device__ __forceinline__ void run(V* lds, int n, int t, int tid, const PassDesc& pd,
                                             const V* tw, bool active, int ps) {
    lds_fft<V, PPT, D, TL, CMP>(lds, n, t, tid, pd, tw, active, ps);
  }
  // the same transform with the first pass reading src and the last writing dst
  template <int D, typename Src, typename Dst>
  static __device__ __forceinline__ void run_io(Src src, Dst dst, V* lds, int n, int t, int tid,
                                                const PassDesc& pd, const V* tw, bool active, int ps) {
    lds_fft_io<V, PPT, D, TL, CMP>(src, dst, lds, n, t, tid, pd, tw, active, ps);
  }
};
// Four-step runtime-pass sub-transforms of plain transforms: composite
// radices for FP32 at 16 points per thread.
template <typename V, int PPT> constexpr bool kFsComposite = sizeof(V) == 8 && PPT == 16;

// Copy of the n-entry (n - 1 used) Stockham table into LDS at `dst` when the
// policy reads it there; the caller's next barrier publishes it (the first
// pass of a transform uses no twiddles).  Returns the table the passes read.
template <typename F, typename V>
__device__ __forceinline__ const V* stage_table(const V* tw, int n, V* dst) {
  if constexpr (F::kTwLds) {
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) dst[i] = tw[i];
    return dst;
  } else {
    return tw;
  }
}

template <typename V, int D, int N, int T, int Ls, int R, int... Rest>
__device__ __forceinline__ void fixed_lds_passes(V* lds, int tid, const V* tw, bool active, int ps) {
  constexpr int MAXB = cdiv(N / R, T);
  stockham_pass<V, R, MAXB, D, false, false>(lds, lds, N, Ls, T, tid, tw, active, 1, ps);
  if constexpr (sizeof...(Rest) > 0) fixed_lds_passes<V, D, N, T, Ls * R, Rest...>(lds, tid, tw, active, ps);
}

// FixedFft: length, threads per transform and radices known at compile time,
// passes inlined.  The plan must use exactly this geometry and pass list.
template <typename V, int N, int T, int... Rs> struct FixedFft {
  static constexpr int kPpt = (N + T - 1) / T;
  static constexpr bool kTwLds = false;
  template <int D>
  static __device__ __forceinline__ void run(V* lds, int, int, int tid, const PassDesc&,
                                             const V* tw, bool active, int ps) {
    fixed_lds_passes<V, D, N, T, 1, Rs...>(as_lds(lds), tid, tw, active, ps);
  }
  template <int D, typename Src, typename Dst>
  static __device__ __forceinline__ void run_io(Src src, Dst dst, V* lds, int, int, int tid, const PassDesc&,
                                                const V* tw, bool active, int ps) {
    fixed_io_passes<V, D, N, T, 1, Src, Dst, Rs...>(src, dst, as_lds(lds), ps, tid, tw, active);
  }
};

// I/O policies for whole LDS-resident transforms with source/sink functors:
// runtime pass list, or compile-time passes (N, threads T, pad shift PS and
// radices fixed; the launch geometry must match).
template <typename V, int PPT> struct GenericIo {
  static constexpr int kPpt = PPT;  // bound on points per thread
  static constexpr int kN = 0, kThreads = 0, kPs = -1;  // geometry known at run time only
  template <int D, typename Src, typename Dst>
  static __device__ __forceinline__ void run(Src src, Dst dst, V* lds, int n, int T, int tid,
                                             const PassDesc& pd, const V* tw, bool active, int ps) {
    lds_fft_io<V, PPT, D>(src, dst, lds, n, T, tid, pd, tw, active, ps);
  }
};
template <typename V, int N, int T, int PS, int... Rs> struct FixedIo {
  static constexpr int kPpt = (N + T - 1) / T;
  static constexpr int kN = N, kThreads = T, kPs = PS;
  template <int D, typename Src, typename Dst>
  static __device__ __forceinline__ void run(Src src, Dst dst, V* lds, int, int, int tid, const PassDesc&,
                                             const V* tw, bool active, int) {
    fixed_io_passes<V, D, N, T, 1, Src, Dst, Rs...>(src, dst, as_lds(lds), PS, tid, tw, active);
  }
};

// w^k for k <= h from the table, or (TW > 0, a power of two) as the product
// w^(k mod TW) * w^(k - k mod TW): in a first pass whose point indices are
// thread + r*TW the first factor is the same for all of a thread's points
// and the second is uniform across the workgroup, so the table traffic per
// point drops from one vector load to a scalar one (one more rounding).
template <int TW, typename V> __device__ __forceinline__ V root_at(const V* w, int k) {
  if constexpr (TW > 0) {
    const int lo = k & (TW - 1);
    return lo == k ? w[k] : cmul(w[lo], w[k - lo]);
  } else {
    return w[k];
  }
}

// ------------------------------------------------------------ storage
// Element storage of a transform's input and output: E == V (complex float
// or double), or a pair of 16-bit floats (cplx16<H>) widened to / rounded
// from V = float2, so the 16-bit forms run the same bodies in FP32.  ld/st
// go through the caches (gathers, scatters); ldnt/stnt are the streaming
// accesses of contiguous reads and writes.
template <typename V, typename E> struct Elem {
  static_assert(sizeof(E) == sizeof(V), "E == V expected");
  static __device__ __forceinline__ V ld(const E* p) { return *p; }
  static __device__ __forceinline__ V ldnt(const E* p) { return gload(p); }
  static __device__ __forceinline__ void st(E* p, V v) { *p = v; }
  static __device__ __forceinline__ void stnt(E* p, V v) { gstore(p, v); }
};
template <typename H> struct Elem<float2, cplx16<H>> {
  static __device__ __forceinline__ float2 ld(const cplx16<H>* p) {
    const cplx16<H> c = *p;
    return float2{(float)c.re, (float)c.im};
  }
  static __device__ __forceinline__ float2 ldnt(const cplx16<H>* p) { return load16(p); }
  static __device__ __forceinline__ void st(cplx16<H>* p, float2 v) { *p = cplx16<H>{(H)v.x, (H)v.y}; }
  static __device__ __forceinline__ void stnt(cplx16<H>* p, float2 v) { store16(p, v); }
};

// ------------------------------------------------------------ functors
template <typename V> using gbl_ptr = __attribute__((address_space(1))) V*;

template <typename V, typename E = V> struct GlobalSrc {  // x[i]
  static constexpr bool kFromLds = false;
  gbl_cptr<E> x;
  __device__ V operator()(int i) const { return Elem<V, E>::ldnt((const E*)x + i); }
};
template <typename V, typename E = V> struct GlobalSink {  // y[i] = v * scale
  static constexpr bool kToLds = false;
  gbl_ptr<E> y;
  S<V> scale;
  __device__ void operator()(int i, V v) const { Elem<V, E>::stnt((E*)y + i, cscale(v, scale)); }
};
// Spectrum times the kernel spectrum, back into LDS.  With y0 set, bin 0
// (the sum of the Rader sequence) also gives y[0] = (x0 + bin 0) * scale.
template <typename V, typename E = V> struct KernSink {
  static constexpr bool kToLds = true;
  lds_ptr<V> p;
  int ps;
  gbl_cptr<V> kern;
  gbl_ptr<E> y0;
  V x0;
  S<V> scale;
  __device__ void operator()(int i, V v) const {
    if (y0 != nullptr && i == 0) Elem<V, E>::st((E*)y0, cscale(cadd(x0, v), scale));
    ((V*)p)[i + (i >> ps)] = cmul(v, ((const V*)kern)[i]);
  }
};
template <typename V, typename E = V> struct BlueSrc {  // x[n] * chirp[n], zero past N
  static constexpr bool kFromLds = false;
  gbl_cptr<E> x;
  gbl_cptr<V> chirp;
  int n;
  __device__ V operator()(int i) const {
    return i < n ? cmul(Elem<V, E>::ld((const E*)x + i), ((const V*)chirp)[i]) : V{0, 0};
  }
};
template <typename V, typename E = V> struct BlueSink {  // y[k] = v * chirp[k] * scale, k < N
  static constexpr bool kToLds = false;
  gbl_ptr<E> y;
  gbl_cptr<V> chirp;
  int n;
  S<V> scale;
  __device__ void operator()(int k, V v) const {
    if (k < n) Elem<V, E>::st((E*)y + k, cscale(cmul(v, ((const V*)chirp)[k]), scale));
  }
};

// Conjugate four-step twiddle at the store: y[k2] = v * conj(w_N^(k2*k1)),
// w_N^m = hi[m >> lo_bits] * lo[m & (2^lo_bits - 1)].
// y[k2] = v * w_N^(k1*k2) (the inverse direction), the root as the product
// of two entries of the row's table in LDS: hi[k2 >> 4] = w^(16*k1*(k2 >> 4)),
// lo[k2 & 15] = w^(k1*(k2 & 15)) (lo at tab + h).
template <typename V> struct TwSink {
  static constexpr bool kToLds = false;
  gbl_ptr<V> y;
  lds_cptr<V> tab;
  int h;
  __device__ void operator()(int k2, V v) const {
    const V* t = (const V*)tab;
    ((V*)y)[k2] = cmul(v, cmul(t[k2 >> 4], t[h + (k2 & 15)]));
  }
};

// ------------------------------------------------------------ smooth
template <typename V, int PPT, int DIR, typename E>
__device__ __forceinline__ void body_lds_smooth(const LdsBody<V>& a, const E* x, E* y, V* lds,
                                                int tid, bool active) {
  lds_fft_io<V, PPT, DIR>(GlobalSrc<V, E>{(gbl_cptr<E>)x}, GlobalSink<V, E>{(gbl_ptr<E>)y, a.scale}, lds,
                          a.m, a.threads, tid, a.pd, a.tw, active, a.ps);
}

// Convolutions whose compile-time transform runs 640 or more threads stage
// their gathers (Rader permutation, Bluestein chirp) through LDS eight points
// per thread at a time: as first-pass sources and last-pass sinks the
// gathers kept every point's index and both factors live at once and spilled
// (M = 20480: 168 B per lane; staged, none), 1.23 against 1.52 ms for 8192
// transforms of N = 10007.  Smaller and runtime-pass forms read and write
// straight from the passes (the extra LDS round trip measured 15-30 %
// slower there).
template <typename F> constexpr bool kStageConv = F::kThreads >= 640;

// ------------------------------------------------------------ Rader
// X[0] = x0 + sum_{n>=1} x[n];  X[g^i] = x0 + (a (*) b)[i],  a[j] = x[g^-j]
// F: the I/O policy of the LDS transforms (GenericIo or FixedIo).
// Both permutations run inside LDS, so every HBM access is coalesced:
// x[1..L] is copied into LDS slots 0..L-1 as it lies, the first FFT pass
// reads a[j] from slot g^-j - 1 (an in-place pass: all its reads precede its
// writes), and after the inverse FFT y[1..L] is written in order, each y[s+1]
// read from the cyclic position pos_out[s].  Gathering x[g^-j] and
// scattering to y[g^i] straight from and to HBM cost an L2 request per
// element (and partial-line writes) instead.
template <typename V> struct LdsPermSrc {  // a[j] = LDS slot perm[j] - 1
  static constexpr bool kFromLds = true;
  lds_ptr<V> p;
  int ps;
  gbl_cptr<int> perm;
  __device__ V operator()(int j) const {
    const int q = ((const int*)perm)[j] - 1;
    return ((const V*)p)[q + (q >> ps)];
  }
};

template <typename V, typename F, typename E>
__device__ __forceinline__ void body_rader(const LdsBody<V>& a, const E* x, E* y, V* lds_in, int tid,
                                           bool active) {
  using El = Elem<V, E>;
  V* lds = as_lds(lds_in);
  const int L = a.m, ps = a.ps;
  const V x0 = active ? El::ld(x) : V{0, 0};
  // x[1..L] into LDS, eight loads per thread in flight
  if (active) {
    for (int s0 = tid; s0 < L; s0 += 8 * a.threads) {
      V v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int s = s0 + q * a.threads;
        if (s < L) v[q] = El::ldnt(x + 1 + s);
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int s = s0 + q * a.threads;
        if (s < L) lds[s + (s >> ps)] = v[q];
      }
    }
  }
  // every read of x (x[0] included) precedes this barrier, so y may alias x
  __syncthreads();
  F::template run<-1>(LdsPermSrc<V>{(lds_ptr<V>)lds, ps, (gbl_cptr<int>)a.perm_in},
                      KernSink<V, E>{(lds_ptr<V>)lds, ps, (gbl_cptr<V>)a.kern, (gbl_ptr<E>)y, x0, a.scale},
                      lds, L, a.threads, tid, a.pd, a.tw, active, ps);
  F::template run<+1>(LdsSrc<V>{(lds_ptr<V>)lds, ps}, LdsSink<V>{(lds_ptr<V>)lds, ps}, lds, L, a.threads,
                      tid, a.pd, a.tw, active, ps);
  // y[s + 1] = (c[pos_out[s]] + x0) * scale, in order
  if (active) {
    const int* pos = a.pos_out;
    for (int s0 = tid; s0 < L; s0 += 8 * a.threads) {
      int d[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int s = s0 + q * a.threads;
        d[q] = s < L ? pos[s] : 0;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int s = s0 + q * a.threads;
        if (s < L) El::stnt(y + 1 + s, cscale(cadd(lds[d[q] + (d[q] >> ps)], x0), a.scale));
      }
    }
  }
}

// ------------------------------------------------------------ Bluestein
// X[k] = chirp[k] * IFFT(FFT(x * chirp, zero-padded to M) * kernel)[k]
template <typename V, typename F, typename E>
__device__ __forceinline__ void body_blue(const LdsBody<V>& a, const E* x, E* y, V* lds_in, int tid,
                                          bool active) {
  using El = Elem<V, E>;
  V* lds = as_lds(lds_in);
  if constexpr (!kStageConv<F>) {
    F::template run<-1>(BlueSrc<V, E>{(gbl_cptr<E>)x, (gbl_cptr<V>)a.chirp, a.n},
                        KernSink<V>{(lds_ptr<V>)lds, a.ps, (gbl_cptr<V>)a.kern, nullptr, V{0, 0}, 1},
                        lds, a.m, a.threads, tid, a.pd, a.tw, active, a.ps);
    F::template run<+1>(LdsSrc<V>{(lds_ptr<V>)lds, a.ps},
                        BlueSink<V, E>{(gbl_ptr<E>)y, (gbl_cptr<V>)a.chirp, a.n, a.scale},
                        lds, a.m, a.threads, tid, a.pd, a.tw, active, a.ps);
    return;
  }
  // x * chirp staged into LDS (zero past N)
  if (active) {
    const V* chirp = a.chirp;
    for (int i0 = tid; i0 < a.m; i0 += 8 * a.threads) {
      V v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int i = i0 + j * a.threads;
        v[j] = i < a.n ? cmul(El::ld(x + i), chirp[i]) : V{0, 0};
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int i = i0 + j * a.threads;
        if (i < a.m) lds[i + (i >> a.ps)] = v[j];
      }
    }
  }
  __syncthreads();
  F::template run<-1>(LdsSrc<V>{(lds_ptr<V>)lds, a.ps},
                      KernSink<V>{(lds_ptr<V>)lds, a.ps, (gbl_cptr<V>)a.kern, nullptr, V{0, 0}, 1},
                      lds, a.m, a.threads, tid, a.pd, a.tw, active, a.ps);
  F::template run<+1>(LdsSrc<V>{(lds_ptr<V>)lds, a.ps}, LdsSink<V>{(lds_ptr<V>)lds, a.ps}, lds, a.m, a.threads,
                      tid, a.pd, a.tw, active, a.ps);
  // y = result * chirp * scale for k < N, staged like the input
  if (active) {
    const V* chirp = a.chirp;
    for (int k0 = tid; k0 < a.n; k0 += 8 * a.threads) {
      V v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + j * a.threads;
        if (k < a.n) v[j] = cmul(lds[k + (k >> a.ps)], chirp[k]);
      }
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = k0 + j * a.threads;
        if (k < a.n) El::st(y + k, cscale(v[j], a.scale));
      }
    }
  }
}

// ------------------------------------------------------------ four-step
// The four-step twiddle w_N^m (N = n1 * n2): computed in FP32 (one
// element's worth of VALU work instead of two scattered cache lookups; as
// accurate as the product of two rounded table entries), from the hi/lo
// tables in FP64 and beyond 2^24 points.
template <int DIR, typename V>
__device__ __forceinline__ V big_twiddle(const FsArgs<V>& a, long long m) {
  if constexpr (sizeof(V) == 8) {
    if (a.inv_n != 0.0f) return root_f32<DIR>((int)m, a.n1 * a.n2, a.inv_n);
  }
  const V w = cmul(a.tw_hi[m >> a.lo_bits], a.tw_lo[m & ((1LL << a.lo_bits) - 1)]);
  return DIR < 0 ? w : V{w.x, -w.y};
}

// Column pass over columns [c0, c0 + C) of one transform (C = 1 << log2c,
// or nsub).  x: the transform's input, tmp: its workspace, out: its output,
// x0: its Rader x[0] slot.  Thread t < s*C moves column c = t mod C of the
// N1 x C tile, rows n1 = t / C + j*s with s = blockDim / C, so a wavefront
// reads and writes C-element row segments and a thread's offsets advance by
// constant steps.  The twiddled modes
// take w_N^(n2*k1) as w^(n2*k0) * w^(n2*s*j): the first factor once per
// thread, the second from the workgroup's table of C x P roots in LDS
// (filled before the FFT, a few roots per thread) -- one complex product
// per element instead of one computed root (the pass was bound by VALU
// issue, half of it the per-element roots and index arithmetic).
template <typename V, typename F, int DIR, int MODE, typename E>
__device__ __forceinline__ void body_fs_col(const FsArgs<V>& a, const E* x, V* tmp, E* out, V* x0,
                                            int c0, V* lds_in, V* zw = nullptr) {
  using El = Elem<V, E>;
  constexpr int P = F::kPpt;
  constexpr bool kTw = fs_col_twiddled(MODE);
  constexpr int TR = fs_col_tw_row(P);  // root table row length
  V* lds_all = as_lds(lds_in);
  const int N1 = a.n1, N2 = a.n2, C = a.nsub ? a.nsub : 1 << a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int ncols = min(C, N2 - c0);
  const int s = blockDim.x / C, kw = threadIdx.x / C, c = threadIdx.x - kw * C;
  const int kr0 = min(kw, s - 1);  // threads past s*C walk nothing
  const bool cok = kw < s && c < ncols;
  const int n2 = c0 + (cok ? c : 0);  // the thread's column of the transform (clamped)
  const int gs = s * N2;              // offset step from row to row
  V* lcol = lds_all + c * stride;
  {
    const int k0 = kr0, g0 = k0 * N2 + n2;
    V buf[P];
    if constexpr (MODE == kFsRaderA) {
      // every source index first, then every gather, so all of a thread's
      // gathers are in flight together (N = 65537 x 2048: 2.57 -> 2.52 ms)
      int src[P];
#pragma unroll
      for (int j = 0; j < P; ++j)
        if (k0 + j * s < N1 && kw < s) src[j] = a.perm[g0 + j * gs];
#pragma unroll
      for (int j = 0; j < P; ++j)
        if (k0 + j * s < N1 && kw < s) buf[j] = El::ld(x + src[j]);
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int n = g0 + j * gs;
      if (k0 + j * s < N1 && kw < s) {
        if constexpr (MODE == kFsPlainA) buf[j] = a.col_cached ? El::ld(x + n) : El::ldnt(x + n);
        else if constexpr (MODE == kFsRaderA) {}
        else if constexpr (MODE == kFsBlueA) buf[j] = n < a.n_orig ? cmul(El::ld(x + n), a.chirp[n]) : V{0, 0};
        else buf[j] = tmp[n];
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int n1 = k0 + j * s;
      if (n1 < N1 && cok) lcol[n1 + (n1 >> ps)] = buf[j];
    }
  }
  const int g0 = opaque(kr0 * N2 + n2);
  if constexpr (MODE == kFsRaderA) {
    if (c0 == 0 && threadIdx.x == 0) *x0 = El::ld(x);
  }
  const V* tw = stage_table<F>(a.tw, N1, lds_all + C * stride);
  V* ptab = lds_all + C * stride + (F::kTwLds ? max(N1, 2) - 1 : 0);
  if constexpr (kTw) {
    for (int e = threadIdx.x; e < C * P; e += blockDim.x) {
      const int ec = e / P, j = e - ec * P, k = j * s;
      ptab[ec * TR + j] = k < N1 ? big_twiddle<DIR>(a, (long long)(c0 + (ec < ncols ? ec : 0)) * k)
                                             : V{1, 0};
    }
  }
  __syncthreads();
  const bool active = slot < C && slot < ncols;
  F::template run<DIR>(lds_all + (slot < C ? slot : 0) * stride, N1, a.threads, tid, a.pd, tw, active, ps);
  V xz = V{0, 0};
  if constexpr (MODE == kFsRaderC || MODE == kFsRaderC2) xz = *x0;
  const int k0 = opaque(kr0);
  if constexpr (MODE == kFsRaderC2) {
    // runs ordered by destination chunk, written contiguously
    const long long bstart = (long long)c0 * N1;  // this column group's first element
    staged<P>(threadIdx.x, blockDim.x, N1 * ncols, [&](int i) { return a.lidx[bstart + i]; },
              [&](int i, int l) { zw[bstart + i] = cscale(cadd(lds_all[l], xz), a.scale); });
  } else {
    // every element's LDS value (and table entry) read before the first
    // store, so the reads overlap; rows past N1 read row N1 - 1 and store
    // nothing.  The twiddled modes store row k1 of the workspace at k1 *
    // tmp_row (a three-pass plan's middle pass writes its rows transposed).
    V val[P];
    [[maybe_unused]] V aux[P];
    [[maybe_unused]] int dst[P];
    [[maybe_unused]] V wb = V{1, 0};
    if constexpr (kTw) wb = big_twiddle<DIR>(a, (long long)n2 * min(k0, N1 - 1));
    const V* prow = ptab + c * TR;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int k1 = min(k0 + j * s, N1 - 1), n = g0 + j * gs;
      val[j] = lcol[k1 + (k1 >> ps)];
      if constexpr (kTw) aux[j] = prow[j];
      else if (k0 + j * s < N1 && cok) {
        // closing modes (k1 runs over n1 of the convolution output)
        if constexpr (MODE == kFsRaderC) dst[j] = a.perm[n];
        else aux[j] = n < a.n_orig ? a.chirp[n] : V{0, 0};
      }
    }
    if constexpr (MODE == kFsRaderC) {
      if (a.win_bits > 0) {
        // destinations window by window: the workgroups of one transform
        // (adjacent on one XCD) sweep the windows together, so a window's
        // lines are completed in L2 before they are evicted instead of being
        // written back piecemeal (8-byte scatters into a transform of up to
        // a megabyte)
        const int nw = ((N1 * N2) >> a.win_bits) + 1;
        for (int w = 0; w < nw; ++w) {
#pragma unroll
          for (int j = 0; j < P; ++j)
            if (k0 + j * s < N1 && cok && (dst[j] >> a.win_bits) == w)
              El::st(out + dst[j], cscale(cadd(val[j], xz), a.scale));
        }
        return;
      }
    }
    const int trow = a.tmp_row ? a.tmp_row : N2;
    const int t0 = opaque(kr0 * trow + n2), ts = s * trow;
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int n = g0 + j * gs;
      if (k0 + j * s < N1 && cok) {
        if constexpr (kTw) tmp[t0 + j * ts] = cmul(val[j], cmul(wb, aux[j]));
        else if constexpr (MODE == kFsRaderC) El::st(out + dst[j], cscale(cadd(val[j], xz), a.scale));
        else if (n < a.n_orig) El::st(out + n, cscale(cmul(val[j], aux[j]), a.scale));
      }
    }
  }
}

// The workgroup's rows [r0, r0 + rows) of the workspace (one contiguous
// span) into LDS row slots with coalesced loads, up to P per thread.
template <typename V, int P>
__device__ __forceinline__ void stage_rows(const V* span, int rows, int N2, V* lds_all, int stride, int ps) {
  const int total = rows * N2, nt = blockDim.x;
  V buf[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int i = threadIdx.x + j * nt;
    if (i < total) buf[j] = gload(span + i);
  }
  int r = threadIdx.x / N2, c = threadIdx.x - (threadIdx.x / N2) * N2;
  const int dr = nt / N2, dc = nt - dr * N2;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int i = threadIdx.x + j * nt;
    if (i < total) lds_all[r * stride + c + (c >> ps)] = buf[j];
    r += dr;
    c += dc;
    if (c >= N2) {
      c -= N2;
      ++r;
    }
  }
  __syncthreads();
}

// Row pass over rows [r0, r0 + C) of one transform's workspace (C = 1 <<
// log2c, or nsub for convolution rows).  Rows whose
// length is a whole number of 128-byte lines are read by each slot's first
// pass straight from HBM (5-11 % faster than staging them); other rows start
// mid-line, so a slot's loads touch two lines per segment and the pass read
// 1.7x its bytes (N = 40500 = 162 x 250): those are staged (a.stage_rows)
// as the workgroup's one contiguous span with coalesced loads.
template <typename V, typename F, int DIR, int MODE, typename E>
__device__ __forceinline__ void body_fs_row(const FsArgs<V>& a, V* tmp, E* out, const V* x0,
                                            int r0, V* lds_in) {
  V* lds_all = as_lds(lds_in);
  const int N1 = a.n1, N2 = a.n2, C = a.nsub ? a.nsub : 1 << a.log2c;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int rows = min(C, N1 - r0);
  const bool active = slot < rows;
  V* lds = lds_all + (slot < C ? slot : 0) * stride;
  // the table copy is published by the barrier of stage_rows or of the
  // first pass, which reads no twiddles
  const V* tw = stage_table<F>(a.tw, N2, lds_all + C * stride);
  // convolution rows: each row's table of conjugate four-step roots (the
  // inverse FFT's last store reads it, many barriers later)
  V* rtab = lds_all + C * stride + (F::kTwLds ? max(N2, 2) - 1 : 0);
  const int th = fs_row_tw_hi(N2), tl = fs_row_tw_len(N2);
  if constexpr (MODE == kFsConvB) {
    for (int e = threadIdx.x; e < C * tl; e += blockDim.x) {
      const int r = e / tl, i = e - r * tl;
      const long long k1 = min(r0 + r, N1 - 1);
      const int b = i - th;
      rtab[e] = i < th ? big_twiddle<+1>(a, k1 * 16 * i) : (b < N2 ? big_twiddle<+1>(a, k1 * b) : V{1, 0});
    }
  }
  if (a.stage_rows) stage_rows<V, F::kPpt>(tmp + (long long)r0 * N2, rows, N2, lds_all, stride, ps);
  if constexpr (MODE == kFsPlainB) {
    // each slot's first pass reads its (contiguous) row straight from HBM
    // (or LDS, staged); the last pass leaves the row in LDS for the
    // transposed store
    const V* src = tmp + (long long)(r0 + (active ? slot : 0)) * N2;
    if (a.stage_rows)
      F::template run<DIR>(lds, N2, a.threads, tid, a.pd, tw, active, ps);
    else
      F::template run_io<DIR>(GlobalSrc<V>{(gbl_cptr<V>)src}, LdsSink<V>{(lds_ptr<V>)lds, ps}, lds, N2,
                              a.threads, tid, a.pd, tw, active, ps);
    // thread t < s*C stores row c = t mod C at k2 = t / C + j*s (s =
    // blockDim / C), so a wavefront writes C-element segments of the output
    // (all of a thread's LDS reads first, so they overlap; positions past
    // N2 read position N2 - 1 and store nothing)
    constexpr int P = F::kPpt;
    const int s = blockDim.x / C, kw = threadIdx.x / C, c = threadIdx.x - kw * C, k0 = min(kw, s - 1);
    const bool wok = kw < s && c < rows;
    const V* lrow = lds_all + c * stride;
    const int o0 = r0 + c + N1 * k0, os = N1 * s;
    V val[P];
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int k2 = min(k0 + j * s, N2 - 1);
      val[j] = cscale(lrow[k2 + (k2 >> ps)], a.scale);
    }
    if (a.out_cached) {
#pragma unroll
      for (int j = 0; j < P; ++j)
        if (k0 + j * s < N2 && wok) Elem<V, E>::st(out + (o0 + j * os), val[j]);
    } else {
#pragma unroll
      for (int j = 0; j < P; ++j)
        if (k0 + j * s < N2 && wok) Elem<V, E>::stnt(out + (o0 + j * os), val[j]);
    }
  } else {  // kFsConvB: one row per slot, read and written in place, contiguous
    // forward FFT straight from HBM, kernel multiply as the last pass's sink,
    // inverse FFT with the conjugate four-step twiddle applied at the store.
    // Rader: X[0] = x0 + sum_{n>=1} x[n] = x0 + A[0] (row 0, bin 0).
    const long long k1 = r0 + (active ? slot : 0);
    V* row = tmp + k1 * N2;
    const bool bin0 = x0 != nullptr && r0 == 0 && slot == 0;
    const V xz = bin0 ? *x0 : V{0, 0};
    const KernSink<V, E> ks{(lds_ptr<V>)lds, ps, (gbl_cptr<V>)(a.kern + k1 * N2), bin0 ? (gbl_ptr<E>)out : nullptr,
                            xz, a.scale};
    if (a.stage_rows)
      F::template run_io<-1>(LdsSrc<V>{(lds_ptr<V>)lds, ps}, ks, lds, N2, a.threads, tid, a.pd, tw, active, ps);
    else
      F::template run_io<-1>(GlobalSrc<V>{(gbl_cptr<V>)row}, ks, lds, N2, a.threads, tid, a.pd, tw, active, ps);
    F::template run_io<+1>(LdsSrc<V>{(lds_ptr<V>)lds, ps},
                           TwSink<V>{(gbl_ptr<V>)row, (lds_cptr<V>)(rtab + (slot < C ? slot : 0) * tl), th},
                           lds, N2, a.threads, tid, a.pd, tw, active, ps);
  }
}

}  // namespace dev
}  // namespace fyfft
