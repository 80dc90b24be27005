// Fused transforms for spectral convolution (rows of real signals whose
// spectrum is kept only in its first K modes).
//
// Mode k of row r sits at r*rdist + k*kdist: row-major (rdist K, kdist 1)
// or mode-major (rdist 1, kdist rows, the layout of per-mode GEMMs).
// Workgroups are remapped so that consecutive rows run on one XCD, which
// lets its L2 combine the mode-major writes of neighbouring rows.
//
// k_rfft_modes:  X[k] = c_k * sum_n x[n] exp(-2*pi*i*n*k/L),  k < K
//   reads a real row (the C2C transform of real input, computed in LDS),
//   stores only the K kept modes; c_0 = a0, c_k = a (k >= 1).
// k_irfft_modes: y[n] = Re( sum_{k<K} c_k * Y[k] exp(+2*pi*i*n*k/L) )
//   reads K modes, zero-fills the rest in LDS, stores only the real part.
// With a0 = 1/L, a = 2/L the second is the real inverse of a Hermitian
// spectrum whose upper half mirrors the kept modes (irfft of the first K
// bins); with a0 = a = 1 it is the real part of the adjoint transform.
//
// Even L with K <= L/2 + 1 runs the half-length forms: the real row is read
// as h = L/2 complex values, one FFT of length h (compile-time passes where
// the plan has them) and the two-for-one combination
//   X[k] = (Z[k] + conj Z[h-k])/2 + w^k (Z[k] - conj Z[h-k])/(2i)
// give the kept modes.  The inverse writes
//   Re(sum_{k<K} c_k Y[k] e^{+2 pi i nk/L}) = sum_k H[k] e^{+2 pi i nk/L}
// with the Hermitian H[0] = Re(c_0 Y[0]), H[k] = H*[L-k] = c_k Y[k]/2
// (0 < k < K; H[h] = Re(c_h Y[h]) if kept), runs the same combination
// backwards into one inverse FFT of length h and stores the h complex
// results as the 2h reals of the row.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_rfft_modes(SpecArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int T = a.threads, L = a.n, K = a.k, ps = a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * a.per_block + slot;
  const bool active = row < a.rows;
  V* lds = lds_all + slot * a.stride;
  if (active) {
    const S<V>* x = a.real_in + row * L;
    staged<PPT>(tid, T, L, [&](int i) { return __builtin_nontemporal_load(x + i); },
                [&](int i, S<V> v) { lds[i + (i >> ps)] = V{v, 0}; });
  }
  __syncthreads();
  lds_fft<V, PPT, -1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  if (active) {
    V* out = a.cplx_out + row * a.rdist;
    for (int k = tid; k < K; k += T) out[k * a.kdist] = cscale(lds[k + (k >> ps)], k == 0 ? a.a0 : a.a);
  }
}

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_irfft_modes(SpecArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int T = a.threads, L = a.n, K = a.k, ps = a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * a.per_block + slot;
  const bool active = row < a.rows;
  V* lds = lds_all + slot * a.stride;
  if (active) {
    const V* in = a.cplx_in + row * a.rdist;
    staged<PPT>(tid, T, L, [&](int i) { return i < K ? cscale(in[i * a.kdist], i == 0 ? a.a0 : a.a) : V{0, 0}; },
                [&](int i, V v) { lds[i + (i >> ps)] = v; });
  }
  __syncthreads();
  lds_fft<V, PPT, +1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  if (active) {
    S<V>* y = a.real_out + row * L;
    for (int n = tid; n < L; n += T) __builtin_nontemporal_store(lds[n + (n >> ps)].x, y + n);
  }
}

// ------------------------------------------------------------ half-length forms
template <typename V> __device__ __forceinline__ V conj_s(V a) { return V{a.x, -a.y}; }

// Source of the inverse: Z[k] = E[k] + i O[k] from the Hermitian H built
// on the fly from the kept modes.
template <typename V, int TW> struct ModesHermSrc {
  static constexpr bool kFromLds = false;
  gbl_cptr<V> m;  // kept modes of the row, mode j at m[j * kdist]
  gbl_cptr<V> w;  // exp(-2 pi i k / L), k <= h
  long long kdist;
  int K, h;
  S<V> a0, a;
  __device__ V herm(int j) const {
    const V* M = (const V*)m;
    if (j >= K) return V{0, 0};
    if (j == 0) return V{a0 * M[0].x, 0};
    if (j == h) return V{a * M[j * kdist].x, 0};
    return cscale(M[j * kdist], a * (S<V>)0.5);
  }
  __device__ V operator()(int k) const {
    const V hk = herm(k), hm = conj_s(herm(h - k));
    const V e = cscale(cadd(hk, hm), (S<V>)0.5);
    const V o = cmulc(cscale(csub(hk, hm), (S<V>)0.5), root_at<TW>((const V*)w, k));
    return V{e.x - o.y, e.y + o.x};
  }
};

// Loads through a buffer resource (raw, stride 0): an offset at or past
// the resource's size returns zero and makes no memory access.  The LLVM
// intrinsics are declared directly: this compiler's
// __builtin_amdgcn_raw_buffer_load_b64 is emitted as a 32-bit load.
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef double f64x2_t __attribute__((ext_vector_type(2)));
__device__ f32x2_t raw_buffer_load_f32x2(i32x4_t rsrc, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.buffer.load.v2f32");
__device__ f64x2_t raw_buffer_load_f64x2(i32x4_t rsrc, int voff, int soff, int aux)
    __asm("llvm.amdgcn.raw.buffer.load.v2f64");
constexpr unsigned kOob = 0x80000000u;  // past every resource built here (< 2 GiB)
// Resource over `bytes` bytes from p (uniform); word 3 is the gfx9 default
// for untyped dword accesses.
__device__ __forceinline__ i32x4_t buffer_rsrc(const void* p, unsigned bytes) {
  const unsigned long long a = (unsigned long long)p;
  return i32x4_t{(int)(unsigned)a, (int)((unsigned)(a >> 32) & 0xffffu), (int)bytes, 0x00020000};
}
template <typename V> __device__ __forceinline__ V bload(i32x4_t r, unsigned off) {
  if constexpr (sizeof(V) == 8) {
    const f32x2_t u = raw_buffer_load_f32x2(r, (int)off, 0, 0);
    return V{u.x, u.y};
  } else {
    const f64x2_t u = raw_buffer_load_f64x2(r, (int)off, 0, 0);
    return V{u.x, u.y};
  }
}

// A per-thread value the compiler may not re-associate with the uniform
// offsets added to it (it would otherwise fold base + (h - j)*s - u*s into
// one multiply per point); identical uses still merge.
__device__ __forceinline__ unsigned keep(unsigned v) { asm("" : "+v"(v)); return v; }

// The same source for the first pass when the modes array spans less than
// 2 GiB, in the two phases of a split source (j per thread, u uniform):
// load() reads the kept modes k = j + u and h - k through a buffer resource
// over the array (a lane whose mode is at or past K gets the out-of-range
// offset and reads zero), each load only where some lane of the wave needs
// it and with no use behind its branch, so a thread's loads are all in
// flight before finish() combines them; w^(j+u) = w^j w^u with w^u a scalar
// load.  The values are those of ModesHermSrc divided by q = a/4 (the sink
// scales by 2q): c0 = 2 a0 / a for mode 0 and 2 for mode h, whose
// imaginary parts are dropped.
template <typename V> struct ModesHermBuf {
  static constexpr bool kFromLds = false, kSplit = true;
  typedef S<V> e2 __attribute__((ext_vector_type(2)));
  struct Raw { V mk, mm, t; };
  i32x4_t res;                                    // the modes array
  gbl_cptr<V> w;                                  // exp(-2 pi i k / L), k <= h
  const __attribute__((address_space(4))) e2* wu; // the same table, scalar loads
  unsigned base, kstep;                           // bytes: mode 0 of the row, between modes
  int K, h;
  S<V> c0;
  __device__ Raw load(int j, int u) const {
    const unsigned kb = keep(base + (unsigned)j * kstep), mb = keep(base + (unsigned)(h - j) * kstep);
    const unsigned us = (unsigned)u * kstep;
    const bool ink = j < K - u, inm = j > h - K - u;
    Raw r{V{0, 0}, V{0, 0}, ((const V*)w)[j]};  // w^j: the same load for every u
    if (__builtin_amdgcn_ballot_w64(ink)) r.mk = bload<V>(res, ink ? kb + us : kOob);
    if (__builtin_amdgcn_ballot_w64(inm)) r.mm = bload<V>(res, inm ? mb - us : kOob);
    return r;
  }
  __device__ V finish(Raw r, int j, int u) const {
    if (!__builtin_amdgcn_ballot_w64(j < K - u || j > h - K - u)) return V{0, 0};
    V mk = r.mk, mm = r.mm;
    if (u == 0) {
      const bool z = j == 0;
      mk = z ? V{mk.x * c0, 0} : mk;
      mm = z ? V{mm.x * 2, 0} : mm;
    }
    V t = r.t;
    if (u != 0) {
      const e2 s = wu[u];
      t = cmul(t, V{s.x, s.y});
    }
    const V e{mk.x + mm.x, mk.y - mm.y}, d{mk.x - mm.x, mk.y + mm.y};
    const V o = cmulc(d, t);
    return V{e.x - o.y, e.y + o.x};
  }
  __device__ V operator()(int i) const { return finish(load(i, 0), i, 0); }
};

// One workgroup slot per row.  TPB > 0: F is a compile-time policy (FixedIo)
// with TPB rows per workgroup: geometry and LDS (static) are compile-time and
// for TPB == 1 the row guard is too; w^k is then formed as a product of a
// per-thread and a uniform factor (root_at) instead of a table load per bin.
template <typename V, typename F, int TPB>
__global__ __launch_bounds__(TPB > 0 ? F::kThreads * TPB : 1024) void k_rfft_modes_half(SpecArgs<V> a) {
  constexpr bool kFixed = TPB > 0;
  constexpr int kStride = kFixed ? lds_len(F::kN, F::kPs) : 1;
  __shared__ __align__(16) V lds_static[kFixed ? kStride * TPB : 1];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int T = kFixed ? F::kThreads : a.threads, h = kFixed ? F::kN : a.n / 2, K = a.k;
  const int ps = kFixed ? F::kPs : a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * (kFixed ? TPB : a.per_block) + slot;
  const bool active = TPB == 1 || row < a.rows;
  V* lds = as_lds((kFixed ? lds_static : reinterpret_cast<V*>(smem_raw)) + slot * (kFixed ? kStride : a.stride));
  const V* z = reinterpret_cast<const V*>(a.real_in) + (active ? row : 0) * h;
  F::template run<-1>(GlobalSrc<V>{(gbl_cptr<V>)z}, LdsSink<V>{(lds_ptr<V>)lds, ps}, lds, h, T, tid, a.pd,
                      a.tw, active, ps);
  if (active) {
    V* out = a.cplx_out + row * a.rdist;
    auto put = [&](int k, V w) {
      const int k0 = k == h ? 0 : k, km = k == 0 ? 0 : h - k;
      const V zk = lds[k0 + (k0 >> ps)], zm = conj_s(lds[km + (km >> ps)]);
      const V e = cscale(cadd(zk, zm), (S<V>)0.5);
      const V d = cscale(csub(zk, zm), (S<V>)0.5);
      const V X = cadd(e, cmul(w, V{d.y, -d.x}));
      out[k * a.kdist] = cscale(X, k == 0 ? a.a0 : a.a);
    };
    constexpr int kTw = (kFixed && (F::kThreads & (F::kThreads - 1)) == 0) ? F::kThreads : 0;
    if constexpr (kTw > 0) {
      // bins k = tid + j*T (K <= h + 1): w^k = w^tid * w^(jT), the second a
      // scalar load through the constant address space (read-only table)
      typedef S<V> e2 __attribute__((ext_vector_type(2)));
      const __attribute__((address_space(4))) e2* Wc = (const __attribute__((address_space(4))) e2*)a.rtw;
      auto W = [&](int k) { const e2 v = Wc[k]; return V{v.x, v.y}; };
      const V wt = W(tid);
#pragma unroll
      for (int j = 0; j <= F::kN / kTw; ++j) {
        const int k = tid + j * kTw;
        if (k < K) put(k, j == 0 ? wt : cmul(wt, W(j * kTw)));
      }
    } else {
      // K <= h + 1 modes: at most kPpt + 1 per thread, twiddle loads issued together
      staged<F::kPpt + 1>(tid, T, K, [&](int k) { return a.rtw[k]; }, put);
    }
  }
}

// BUF: the modes come through ModesHermBuf (a.mode_bytes > 0: the array
// spans less than 2 GiB and the scale a is not zero), else through
// ModesHermSrc.
template <typename V, typename F, int TPB, bool BUF>
__global__ __launch_bounds__(TPB > 0 ? F::kThreads * TPB : 1024) void k_irfft_modes_half(SpecArgs<V> a) {
  constexpr bool kFixed = TPB > 0;
  constexpr int kStride = kFixed ? lds_len(F::kN, F::kPs) : 1;
  __shared__ __align__(16) V lds_static[kFixed ? kStride * TPB : 1];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int T = kFixed ? F::kThreads : a.threads, h = kFixed ? F::kN : a.n / 2;
  const int ps = kFixed ? F::kPs : a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)xcd_remap(blockIdx.x, gridDim.x) * (kFixed ? TPB : a.per_block) + slot;
  const bool active = TPB == 1 || row < a.rows;
  V* lds = (kFixed ? lds_static : reinterpret_cast<V*>(smem_raw)) + slot * (kFixed ? kStride : a.stride);
  const long long r = active ? row : 0;
  V* y = reinterpret_cast<V*>(a.real_out) + r * h;
  constexpr int kTw = (kFixed && (F::kThreads & (F::kThreads - 1)) == 0 && F::kThreads + 16 <= F::kN)
                          ? F::kThreads : 0;
  if constexpr (BUF) {
    using Src = ModesHermBuf<V>;
    const Src src{buffer_rsrc(a.cplx_in, a.mode_bytes), (gbl_cptr<V>)a.rtw, (decltype(Src::wu))a.rtw,
                  (unsigned)(r * a.rdist * (long long)sizeof(V)), (unsigned)(a.kdist * (long long)sizeof(V)),
                  a.k, h, 2 * a.a0 / a.a};
    F::template run<+1>(src, GlobalSink<V>{(gbl_ptr<V>)y, a.a * (S<V>)0.5}, lds, h, T, tid, a.pd, a.tw, active, ps);
  } else {
    F::template run<+1>(ModesHermSrc<V, kTw>{(gbl_cptr<V>)(a.cplx_in + r * a.rdist), (gbl_cptr<V>)a.rtw, a.kdist,
                                             a.k, h, a.a0, a.a},
                        GlobalSink<V>{(gbl_ptr<V>)y, (S<V>)2}, lds, h, T, tid, a.pd, a.tw, active, ps);
  }
}

template <typename V, typename F, int TPB>
static hipError_t go_half(bool inverse, const SpecArgs<V>& a, const LdsGeom& g, hipStream_t s) {
  auto k = !inverse ? k_rfft_modes_half<V, F, TPB>
           : a.mode_bytes ? k_irfft_modes_half<V, F, TPB, true> : k_irfft_modes_half<V, F, TPB, false>;
  size_t lds = (size_t)a.stride * g.per_block * sizeof(V);
  if constexpr (TPB > 0) {
    // compile-time geometry: the plan's must be the same, LDS is static
    if (g.threads != F::kThreads || g.per_block != TPB || g.pad_shift != F::kPs || a.n != 2 * F::kN)
      return hipErrorInvalidValue;
    lds = 0;
  }
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const unsigned grid = (unsigned)((a.rows + g.per_block - 1) / g.per_block);
  hipLaunchKernelGGL(k, dim3(grid), dim3(g.threads * g.per_block), lds, s, a);
  return hipGetLastError();
}

// fixed_h: the half length when the plan's half-length table has a
// compile-time kernel (its geometry is g), 0 for the runtime passes.
hipError_t launch_spectral_half(int prec, bool inverse, const void* args, const LdsGeom& g, int fixed_h,
                                hipStream_t s) {
  if (prec == kF64) {
    const auto& a = *static_cast<const SpecArgs<double2>*>(args);
    if (a.rows == 0) return hipSuccess;
    switch (fixed_h) {
      case 1024: return go_half<double2, FixedIo<double2, 1024, 64, 30, 16, 8, 8>, 4>(inverse, a, g, s);
      case 4096: return go_half<double2, FixedIo<double2, 4096, 256, 30, 16, 16, 16>, 1>(inverse, a, g, s);
    }
    return g.ppt <= 8 ? go_half<double2, GenericIo<double2, 8>, 0>(inverse, a, g, s)
                      : go_half<double2, GenericIo<double2, 16>, 0>(inverse, a, g, s);
  }
  const auto& a = *static_cast<const SpecArgs<float2>*>(args);
  if (a.rows == 0) return hipSuccess;
  switch (fixed_h) {
    case 256: return go_half<float2, FixedIo<float2, 256, 16, 30, 16, 16>, 16>(inverse, a, g, s);
    case 512: return go_half<float2, FixedIo<float2, 512, 32, 30, 8, 8, 8>, 8>(inverse, a, g, s);
    case 1024: return go_half<float2, FixedIo<float2, 1024, 64, 30, 16, 8, 8>, 4>(inverse, a, g, s);
    case 2048: return go_half<float2, FixedIo<float2, 2048, 128, 30, 16, 16, 8>, 2>(inverse, a, g, s);
    case 4096: return go_half<float2, FixedIo<float2, 4096, 256, 30, 16, 16, 16>, 1>(inverse, a, g, s);
    case 8192: return go_half<float2, FixedIo<float2, 8192, 512, 30, 16, 8, 8, 8>, 1>(inverse, a, g, s);
  }
  return g.ppt <= 16 ? go_half<float2, GenericIo<float2, 16>, 0>(inverse, a, g, s)
                     : go_half<float2, GenericIo<float2, 32>, 0>(inverse, a, g, s);
}

template <typename V, int PPT>
static hipError_t go_spec(bool inverse, const SpecArgs<V>& a, const LdsGeom& g, hipStream_t s) {
  auto k = inverse ? k_irfft_modes<V, PPT> : k_rfft_modes<V, PPT>;
  const size_t lds = (size_t)a.stride * g.per_block * sizeof(V);
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const unsigned grid = (unsigned)((a.rows + g.per_block - 1) / g.per_block);
  hipLaunchKernelGGL(k, dim3(grid), dim3(g.threads * g.per_block), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_spectral(int prec, bool inverse, const void* args, const LdsGeom& g, hipStream_t s) {
  if (prec == kF64) {
    const auto& a = *static_cast<const SpecArgs<double2>*>(args);
    if (a.rows == 0) return hipSuccess;
    return g.ppt <= 8 ? go_spec<double2, 8>(inverse, a, g, s) : go_spec<double2, 16>(inverse, a, g, s);
  }
  const auto& a = *static_cast<const SpecArgs<float2>*>(args);
  if (a.rows == 0) return hipSuccess;
  return g.ppt <= 16 ? go_spec<float2, 16>(inverse, a, g, s) : go_spec<float2, 32>(inverse, a, g, s);
}

}  // namespace fyfft
