// Real-input (R2C) and real-output (C2R) transforms.
//
// Even n = 2h.  The real row is read as h complex values z[j] = x[2j] +
// i*x[2j+1] (the same bytes), one complex FFT of length h gives Z, and
//   E[k] = (Z[k] + conj(Z[h-k])) / 2          (spectrum of x[2j])
//   O[k] = (Z[k] - conj(Z[h-k])) / (2i)       (spectrum of x[2j+1])
//   X[k] = E[k] + w^k O[k],  w = exp(-2*pi*i/n),  k = 0..h  (Z[h] = Z[0]).
// The inverse runs the same algebra backwards from a Hermitian half
// spectrum X[0..h]:
//   E[k] = (X[k] + conj(X[h-k])) / 2,  O[k] = (X[k] - conj(X[h-k])) conj(w^k) / 2,
//   z = IFFT_h(E + i*O) / h,  x[2j] = Re z[j],  x[2j+1] = Im z[j],
// which is x = IDFT_n(X) / n with the upper half of X implied by symmetry.
// HBM traffic is 4n bytes in and 8(h+1) out for FP32 (half the C2C bytes).
//
// Odd n: one complex FFT of length n over the real row (imaginary parts
// zero), keeping bins 0..(n-1)/2; the inverse rebuilds the full spectrum
// from the half (X[n-k] = conj(X[k])) and keeps the real part.
//
// LDS-resident kernels fuse everything into one pass over HBM; the
// elementwise kernels below serve the multi-pass forms for lengths whose
// transform does not fit one workgroup.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

namespace {

template <typename V> __device__ __forceinline__ V conj_(V a) { return V{a.x, -a.y}; }
template <typename V> __device__ __forceinline__ V half_(V a) { return V{a.x * (S<V>)0.5, a.y * (S<V>)0.5}; }

// X[k] from the half-length spectrum (zk = Z[k mod h], zm = Z[(h-k) mod h])
template <typename V> __device__ __forceinline__ V r2c_combine(V zk, V zm, V w) {
  const V e = half_(cadd(zk, conj_(zm)));
  const V d = half_(csub(zk, conj_(zm)));
  const V o = V{d.y, -d.x};  // d / i
  return cadd(e, cmul(w, o));
}

// Z[k] = E[k] + i*O[k] from the Hermitian half spectrum (xk = X[k], xm = X[h-k])
template <typename V> __device__ __forceinline__ V c2r_combine(V xk, V xm, V w) {
  const V e = half_(cadd(xk, conj_(xm)));
  const V o = cmulc(half_(csub(xk, conj_(xm))), w);  // * conj(w^k)
  return V{e.x - o.y, e.y + o.x};
}

// ---- source and sink functors of the fused kernels
template <typename V, int TW> struct HermHalfSrc {  // Z[k] of an even-length C2R
  static constexpr bool kFromLds = false;
  gbl_cptr<V> x;  // X[0..h]
  gbl_cptr<V> w;  // w^k, k <= h
  int h;
  __device__ V operator()(int k) const {
    const V* X = (const V*)x;
    V xk = X[k], xm = X[h - k];
    if (k == 0) xk.y = xm.y = 0;  // bins 0 and n/2 of a real signal are real
    return c2r_combine(xk, xm, root_at<TW>((const V*)w, k));
  }
};
// Z[k] of an even-length C2R from X[0..h-1] staged in the transform's LDS
// buffer (X[h] passed in): the first pass reads both X[k] and X[h-k] from
// LDS instead of two global loads per point from rows of h+1 bins.
template <typename V, int TW> struct HermLdsSrc {
  static constexpr bool kFromLds = true;
  lds_ptr<V> p;
  int ps;
  gbl_cptr<V> w;  // w^k, k <= h
  V xh;           // X[h]
  int h;
  __device__ V operator()(int k) const {
    const V* L = (const V*)p;
    const int m = k == 0 ? 0 : h - k;  // X[h] (k = 0) is not in LDS: read bin 0, then select
    V xk = L[k + (k >> ps)];
    const V lm = L[m + (m >> ps)];
    V xm = V{k == 0 ? xh.x : lm.x, k == 0 ? xh.y : lm.y};
    if (k == 0) xk.y = xm.y = 0;  // bins 0 and n/2 of a real signal are real
    return c2r_combine(xk, xm, root_at<TW>((const V*)w, k));
  }
};
template <typename V> struct RealSrc {  // (x[i], 0)
  static constexpr bool kFromLds = false;
  const __attribute__((address_space(1))) S<V>* x;
  __device__ V operator()(int i) const { return V{((const S<V>*)x)[i], 0}; }
};
template <typename V> struct HalfSink {  // keep bins 0..kmax
  static constexpr bool kToLds = false;
  gbl_ptr<V> y;
  int kmax;
  __device__ void operator()(int k, V v) const {
    if (k <= kmax) ((V*)y)[k] = v;
  }
};
template <typename V> struct HermFullSrc {  // full spectrum from bins 0..n/2
  static constexpr bool kFromLds = false;
  gbl_cptr<V> x;
  int n;
  __device__ V operator()(int k) const {
    const V* X = (const V*)x;
    return 2 * k <= n ? X[k] : conj_(X[n - k]);
  }
};
template <typename V> struct RealSink {  // y[i] = Re(v) * scale
  static constexpr bool kToLds = false;
  __attribute__((address_space(1))) S<V>* y;
  S<V> scale;
  __device__ void operator()(int i, V v) const { ((S<V>*)y)[i] = v.x * scale; }
};

}  // namespace

// One workgroup slot per row; MODE selects the form, F the pass policy.
// TPB > 0: F is a compile-time policy (FixedIo) and the launch has TPB
// transforms per workgroup, so geometry, LDS (static) and, for TPB == 1,
// the row guard are compile-time as well.
template <typename V, typename F, int MODE, int TPB>
__global__ __launch_bounds__(TPB > 0 ? F::kThreads * TPB : 1024) void k_real_lds(RealArgs<V> a) {
  constexpr bool kFixed = TPB > 0;
  constexpr int kStride = kFixed ? lds_len(F::kN, F::kPs) : 1;
  __shared__ __align__(16) V lds_static[kFixed ? kStride * TPB : 1];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int T = kFixed ? F::kThreads : a.threads;
  const int ps = kFixed ? F::kPs : a.ps;
  const int per_block = kFixed ? TPB : a.per_block;
  const int stride = kFixed ? kStride : a.stride;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long t = (long long)xcd_remap(blockIdx.x, gridDim.x) * per_block + slot;
  const bool active = TPB == 1 || t < a.rows;
  const long long row = active ? t : 0;
  V* lds = (kFixed ? lds_static : reinterpret_cast<V*>(smem_raw)) + slot * stride;
  const lds_ptr<V> ldsp = (lds_ptr<V>)as_lds(lds);
  constexpr bool kHalf = MODE == kRealR2cHalf || MODE == kRealC2rHalf;
  const int n = kFixed ? (kHalf ? 2 * F::kN : F::kN) : a.n, h = n / 2;
  // product twiddles need compile-time power-of-two threads with T + 16 <= h
  constexpr int kTw = (kFixed && (F::kThreads & (F::kThreads - 1)) == 0 && F::kThreads + 16 <= F::kN)
                          ? F::kThreads : 0;
  if constexpr (MODE == kRealR2cHalf) {
    const V* z = (const V*)a.in + row * h;
    F::template run<-1>(GlobalSrc<V>{(gbl_cptr<V>)z}, LdsSink<V>{ldsp, ps}, lds, h, T, tid, a.pd, a.tw, active,
                        ps);
    if (active) {
      // Bins 0..h of the row in 128-byte aligned spans: thread i of the pass
      // writes bin k = i + ka, wrapping past h to 0..ka-1, where bin ka is the
      // row's first one on a 128-byte boundary (rows of h+1 bins are not
      // aligned, and unaligned spans cost partial-line writes).
      V* X = (V*)a.out + row * (h + 1);
      const V* L = (const V*)ldsp;
      constexpr int kLine = 128 / (int)sizeof(V);
      const int ka0 = (int)((kLine - (row * (h + 1)) % kLine) % kLine);
      const int ka = ka0 <= h ? ka0 : 0;  // rows shorter than a line: no rotation
      auto put = [&](int k, V w) {
        const int k0 = k == h ? 0 : k, km = k == 0 ? 0 : h - k;
        gstore(X + k, cscale(r2c_combine(L[k0 + (k0 >> ps)], L[km + (km >> ps)], w), a.scale));
      };
      if constexpr (kTw > 0) {
        // the table is read-only here: through the constant address space
        // the uniform factors become scalar loads issued ahead of the stores
        typedef S<V> e2 __attribute__((ext_vector_type(2)));
        const __attribute__((address_space(4))) e2* Wc = (const __attribute__((address_space(4))) e2*)a.rtw;
        auto W = [&](int k) { const e2 v = Wc[k]; return V{v.x, v.y}; };
        const V wt = W(tid + ka);
#pragma unroll
        for (int j = 0; j <= F::kN / kTw; ++j) {
          const int i = tid + j * kTw;
          if (i <= h) {
            int k = i + ka;
            V w;
            if (k > h) {
              k -= h + 1;
              w = W(k);
            } else {
              w = j == 0 ? wt : cmul(wt, W(j * kTw));
            }
            put(k, w);
          }
        }
      } else {
        staged<F::kPpt + 1>(tid, T, h + 1,
                            [&](int i) { const int k = i + ka; return a.rtw[k > h ? k - (h + 1) : k]; },
                            [&](int i, V w) { const int k = i + ka; put(k > h ? k - (h + 1) : k, w); });
      }
    }
  } else if constexpr (MODE == kRealC2rHalf) {
    // X[0..h-1] is staged into LDS with loads in 128-byte aligned spans (bin
    // k = i + ka as in the R2C store), then the first pass combines X[k] and
    // X[h-k] from LDS: one aligned load per bin instead of two unaligned ones
    const V* X = (const V*)a.in + row * (h + 1);
    V* z = (V*)a.out + row * h;
    if (active) {
      constexpr int kLine = 128 / (int)sizeof(V);
      const int ka0 = (int)((kLine - (row * (h + 1)) % kLine) % kLine);
      const int ka = ka0 <= h ? ka0 : 0;
      // loads in flight per thread: up to 16 FP32 / 8 FP64 bins (registers)
      constexpr int P = F::kPpt + 1, kChunk = sizeof(V) == 16 ? 8 : 16;
#pragma unroll
      for (int j0 = 0; j0 < P; j0 += kChunk) {
        V buf[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int i = tid + (j0 + j) * T;
          if (j0 + j < P && i <= h) {
            const int k = i + ka;
            buf[j] = gload(X + (k > h ? k - (h + 1) : k));
          }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int i = tid + (j0 + j) * T;
          if (j0 + j < P && i <= h) {
            const int k = i + ka > h ? i + ka - (h + 1) : i + ka;
            if (k < h) lds[k + (k >> ps)] = buf[j];
          }
        }
      }
    }
    const V xh = X[h];
    __syncthreads();
    F::template run<+1>(HermLdsSrc<V, kTw>{ldsp, ps, (gbl_cptr<V>)a.rtw, xh, h},
                        GlobalSink<V>{(gbl_ptr<V>)z, a.scale}, lds, h, T, tid, a.pd, a.tw, active, ps);
  } else if constexpr (MODE == kRealR2cFull) {
    const S<V>* x = (const S<V>*)a.in + row * n;
    V* X = (V*)a.out + row * (n / 2 + 1);
    F::template run<-1>(RealSrc<V>{(const __attribute__((address_space(1))) S<V>*)x},
                        HalfSink<V>{(gbl_ptr<V>)X, n / 2}, lds, n, T, tid, a.pd, a.tw, active, ps);
  } else {  // kRealC2rFull
    const V* X = (const V*)a.in + row * (n / 2 + 1);
    S<V>* y = (S<V>*)a.out + row * n;
    F::template run<+1>(HermFullSrc<V>{(gbl_cptr<V>)X, n},
                        RealSink<V>{(__attribute__((address_space(1))) S<V>*)y, a.scale}, lds, n, T, tid, a.pd,
                        a.tw, active, ps);
  }
}

// ---- elementwise helpers of the multi-pass forms (one thread per output)
template <typename V>
__global__ void k_real_post(const V* z, V* X, long long rows, int h, const V* w, S<V> scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * (h + 1)) return;
  const long long r = i / (h + 1);
  const int k = (int)(i - r * (h + 1));
  const V* Z = z + r * h;
  const int k0 = k == h ? 0 : k, km = k == 0 ? 0 : h - k;
  X[i] = cscale(r2c_combine(Z[k0], Z[km], w[k]), scale);
}
template <typename V>
__global__ void k_real_pre(const V* X, V* z, long long rows, int h, const V* w) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * h) return;
  const long long r = i / h;
  const int k = (int)(i - r * h);
  const V* Xr = X + r * (h + 1);
  V xk = Xr[k], xm = Xr[h - k];
  if (k == 0) xk.y = xm.y = 0;  // bins 0 and n/2 of a real signal are real
  z[i] = c2r_combine(xk, xm, w[k]);
}
template <typename V>
__global__ void k_real_expand(const S<V>* x, V* W, long long count) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) W[i] = V{x[i], 0};
}
template <typename V>
__global__ void k_real_truncate(const V* W, V* X, long long rows, int n, S<V> scale) {
  const int hk = n / 2 + 1;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * hk) return;
  const long long r = i / hk;
  const int k = (int)(i - r * hk);
  X[i] = cscale(W[r * n + k], scale);
}
template <typename V>
__global__ void k_herm_expand(const V* X, V* W, long long rows, int n) {
  const int hk = n / 2 + 1;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * n) return;
  const long long r = i / n;
  const int k = (int)(i - r * n);
  const V* Xr = X + r * hk;
  W[i] = 2 * k <= n ? Xr[k] : conj_(Xr[n - k]);
}
template <typename V>
__global__ void k_real_part(const V* W, S<V>* y, long long count, S<V> scale) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) y[i] = W[i].x * scale;
}

template <typename V, typename F, int TPB>
static hipError_t go_real(int mode, const RealArgs<V>& a, const LdsGeom& g, hipStream_t s) {
  void (*k)(RealArgs<V>) = nullptr;
  switch (mode) {
    case kRealR2cHalf: k = k_real_lds<V, F, kRealR2cHalf, TPB>; break;
    case kRealC2rHalf: k = k_real_lds<V, F, kRealC2rHalf, TPB>; break;
    case kRealR2cFull: k = k_real_lds<V, F, kRealR2cFull, TPB>; break;
    case kRealC2rFull: k = k_real_lds<V, F, kRealC2rFull, TPB>; break;
    default: return hipErrorInvalidValue;
  }
  size_t dyn = g.lds_bytes;
  if constexpr (TPB > 0) {
    // compile-time geometry: the plan's must be the same, LDS is static
    if (g.threads != F::kThreads || g.per_block != TPB || g.pad_shift != F::kPs) return hipErrorInvalidValue;
    dyn = 0;
  }
  if (dyn > 65536)
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  const unsigned grid = (unsigned)((a.rows + g.per_block - 1) / g.per_block);
  hipLaunchKernelGGL(k, dim3(grid), dim3(g.threads * g.per_block), dyn, s, a);
  return hipGetLastError();
}

// The compile-time kernels take the geometry of the batched fixed kernel of
// the same length (threads, transforms per block, pad shift).
hipError_t launch_real_lds(int prec, int mode, const void* args, const LdsGeom& g, int fixed_len, hipStream_t s) {
  if (prec == kF64) {
    const auto& a = *static_cast<const RealArgs<double2>*>(args);
    if (a.rows == 0) return hipSuccess;
    switch (fixed_len) {
      case 1024: return go_real<double2, FixedIo<double2, 1024, 64, 30, 16, 8, 8>, 4>(mode, a, g, s);
      case 4096: return go_real<double2, FixedIo<double2, 4096, 256, 30, 16, 16, 16>, 1>(mode, a, g, s);
    }
    return g.ppt <= 8 ? go_real<double2, GenericIo<double2, 8>, 0>(mode, a, g, s)
                      : go_real<double2, GenericIo<double2, 16>, 0>(mode, a, g, s);
  }
  const auto& a = *static_cast<const RealArgs<float2>*>(args);
  if (a.rows == 0) return hipSuccess;
  switch (fixed_len) {
    case 256: return go_real<float2, FixedIo<float2, 256, 16, 30, 16, 16>, 16>(mode, a, g, s);
    case 512: return go_real<float2, FixedIo<float2, 512, 32, 30, 8, 8, 8>, 8>(mode, a, g, s);
    case 1024: return go_real<float2, FixedIo<float2, 1024, 64, 30, 16, 8, 8>, 4>(mode, a, g, s);
    case 2048: return go_real<float2, FixedIo<float2, 2048, 128, 30, 16, 16, 8>, 2>(mode, a, g, s);
    case 4096: return go_real<float2, FixedIo<float2, 4096, 256, 30, 16, 16, 16>, 1>(mode, a, g, s);
    case 8192: return go_real<float2, FixedIo<float2, 8192, 512, 30, 16, 8, 8, 8>, 1>(mode, a, g, s);
  }
  return g.ppt <= 16 ? go_real<float2, GenericIo<float2, 16>, 0>(mode, a, g, s)
                     : go_real<float2, GenericIo<float2, 32>, 0>(mode, a, g, s);
}

static inline dim3 grid_of(long long count) { return dim3((unsigned)((count + 255) / 256)); }

template <typename V>
static hipError_t go_real_pass(int op, const void* src, void* dst, long long rows, int n, const void* w,
                               double scale, hipStream_t s) {
  const int h = n / 2;
  switch (op) {
    case kRealOpPost:
      hipLaunchKernelGGL(k_real_post<V>, grid_of(rows * (h + 1)), dim3(256), 0, s, (const V*)src, (V*)dst,
                         rows, h, (const V*)w, (S<V>)scale);
      break;
    case kRealOpPre:
      hipLaunchKernelGGL(k_real_pre<V>, grid_of(rows * h), dim3(256), 0, s, (const V*)src, (V*)dst, rows, h,
                         (const V*)w);
      break;
    case kRealOpExpand:
      hipLaunchKernelGGL(k_real_expand<V>, grid_of(rows * n), dim3(256), 0, s, (const S<V>*)src, (V*)dst,
                         rows * n);
      break;
    case kRealOpTruncate:
      hipLaunchKernelGGL(k_real_truncate<V>, grid_of(rows * (n / 2 + 1)), dim3(256), 0, s, (const V*)src,
                         (V*)dst, rows, n, (S<V>)scale);
      break;
    case kRealOpHermExpand:
      hipLaunchKernelGGL(k_herm_expand<V>, grid_of(rows * n), dim3(256), 0, s, (const V*)src, (V*)dst, rows, n);
      break;
    case kRealOpRealPart:
      hipLaunchKernelGGL(k_real_part<V>, grid_of(rows * n), dim3(256), 0, s, (const V*)src, (S<V>*)dst,
                         rows * n, (S<V>)scale);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_real_pass(int prec, int op, const void* src, void* dst, long long rows, int n,
                            const void* w, double scale, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  if (prec == kF64) return go_real_pass<double2>(op, src, dst, rows, n, w, scale, s);
  return go_real_pass<float2>(op, src, dst, rows, n, w, scale, s);
}

}  // namespace fyfft
