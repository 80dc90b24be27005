// Fused transforms for spectral convolution (rows of real signals whose
// spectrum is kept only in its first K modes).
//
// k_rfft_modes:  X[k] = c_k * sum_n x[n] exp(-2*pi*i*n*k/L),  k < K
//   reads a real row (the C2C transform of real input, computed in LDS),
//   stores only the K kept modes; c_0 = a0, c_k = a (k >= 1).
// k_irfft_modes: y[n] = Re( sum_{k<K} c_k * Y[k] exp(+2*pi*i*n*k/L) )
//   reads K modes, zero-fills the rest in LDS, stores only the real part.
// With a0 = 1/L, a = 2/L the second is the real inverse of a Hermitian
// spectrum whose upper half mirrors the kept modes (irfft of the first K
// bins); with a0 = a = 1 it is the real part of the adjoint transform.
#include "internal.hpp"

namespace fyfft {
using namespace dev;

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_rfft_modes(SpecArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int T = a.threads, L = a.n, K = a.k, ps = a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)blockIdx.x * a.per_block + slot;
  const bool active = row < a.rows;
  V* lds = lds_all + slot * a.stride;
  if (active) {
    const S<V>* x = a.real_in + row * L;
    for (int i = tid; i < L; i += T) lds[i + (i >> ps)] = V{__builtin_nontemporal_load(x + i), 0};
  }
  __syncthreads();
  lds_fft<V, PPT, -1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  if (active) {
    V* out = a.cplx_out + row * K;
    for (int k = tid; k < K; k += T) gstore(out + k, cscale(lds[k + (k >> ps)], k == 0 ? a.a0 : a.a));
  }
}

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_irfft_modes(SpecArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds_all = reinterpret_cast<V*>(smem_raw);
  const int T = a.threads, L = a.n, K = a.k, ps = a.ps;
  const int slot = threadIdx.x / T, tid = threadIdx.x - slot * T;
  const long long row = (long long)blockIdx.x * a.per_block + slot;
  const bool active = row < a.rows;
  V* lds = lds_all + slot * a.stride;
  if (active) {
    const V* in = a.cplx_in + row * K;
    for (int i = tid; i < L; i += T)
      lds[i + (i >> ps)] = i < K ? cscale(gload(in + i), i == 0 ? a.a0 : a.a) : V{0, 0};
  }
  __syncthreads();
  lds_fft<V, PPT, +1>(lds, L, T, tid, a.pd, a.tw, active, ps);
  if (active) {
    S<V>* y = a.real_out + row * L;
    for (int n = tid; n < L; n += T) __builtin_nontemporal_store(lds[n + (n >> ps)].x, y + n);
  }
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
