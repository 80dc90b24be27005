// Fused Rader and Bluestein transforms whose convolution length fits one
// workgroup's LDS, plus the FP64 table generators both pipelines use.
//
// Rader, prime N with L = N-1 smooth (g a primitive root mod N):
//   X[0]     = x[0] + sum_{n>=1} x[n]
//   X[g^i]   = x[0] + sum_j x[g^-j] * w_N^(g^(i-j))      (cyclic, length L)
// One kernel: gather a[j] = x[g^-j] into LDS, FFT_L, read X[0] - x[0] from
// bin 0, multiply by the kernel spectrum (1/L folded in), inverse FFT_L,
// add x[0] and scatter to g^i.  No padding, unlike the reference's
// zero-padded 2N-3 convolution (rader.hpp:166,357-401).
//
// Bluestein, any N with convolution length M >= 2N-1:
//   X[k] = c'[k] * sum_n (x[n] c'[n]) * b[k-n],  c'[n] = exp(-i*pi*n^2/N)
// One kernel: pre-chirp and zero-pad into LDS, FFT_M, multiply, inverse
// FFT_M, post-chirp the first N outputs (bluestein.hpp:235-354 does the same
// steps as separate passes over host memory).
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

template <typename V> __device__ __forceinline__ LdsBody<V> conv_body(const ConvArgs<V>& a) {
  LdsBody<V> b;
  b.n = a.n_orig;
  b.m = a.m;
  b.threads = a.threads;
  b.ps = a.ps;
  b.stride = a.stride;
  b.pd = a.pd;
  b.tw = a.tw;
  b.perm_in = a.perm_in;
  b.perm_out = a.perm_out;
  b.chirp = a.chirp;
  b.kern = a.kern;
  b.scale = a.scale;
  return b;
}

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_rader_lds(ConvArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int slot = threadIdx.x / a.threads, tid = threadIdx.x - slot * a.threads;
  const long long t = (long long)blockIdx.x * a.per_block + slot;
  const bool active = t < a.batch;
  const long long off = (active ? t : 0) * a.dist;
  body_rader<V, PPT>(conv_body(a), a.in + off, a.out + off,
                     reinterpret_cast<V*>(smem_raw) + slot * a.stride, tid, active);
}

template <typename V, int PPT>
__global__ __launch_bounds__(1024) void k_blue_lds(ConvArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int slot = threadIdx.x / a.threads, tid = threadIdx.x - slot * a.threads;
  const long long t = (long long)blockIdx.x * a.per_block + slot;
  const bool active = t < a.batch;
  const long long off = (active ? t : 0) * a.dist;
  body_blue<V, PPT>(conv_body(a), a.in + off, a.out + off,
                    reinterpret_cast<V*>(smem_raw) + slot * a.stride, tid, active);
}

template <typename V, int PPT>
static hipError_t go_conv(bool rader, const ConvArgs<V>& a, const LdsGeom& g, hipStream_t s) {
  auto k = rader ? k_rader_lds<V, PPT> : k_blue_lds<V, PPT>;
  const size_t lds = g.lds_bytes;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const unsigned grid = (unsigned)((a.batch + g.per_block - 1) / g.per_block);
  hipLaunchKernelGGL(k, dim3(grid), dim3(g.threads * g.per_block), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_conv_lds(int prec, bool rader, const void* args, const LdsGeom& g, hipStream_t s) {
  if (prec == kF64) {
    const auto& a = *static_cast<const ConvArgs<double2>*>(args);
    if (a.batch == 0) return hipSuccess;
    return g.ppt <= 8 ? go_conv<double2, 8>(rader, a, g, s) : go_conv<double2, 16>(rader, a, g, s);
  }
  const auto& a = *static_cast<const ConvArgs<float2>*>(args);
  if (a.batch == 0) return hipSuccess;
  return g.ppt <= 16 ? go_conv<float2, 16>(rader, a, g, s) : go_conv<float2, 32>(rader, a, g, s);
}

// ------------------------------------------------------------ tables
__global__ void k_gen_chirp(double2* c, long long n_orig, long long count, int sign) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= count) return;
  const unsigned long long two_n = 2ull * (unsigned long long)n_orig;
  const unsigned long long m = (unsigned long long)(((unsigned __int128)n * n) % two_n);
  c[n] = root_fp64((long long)m, (long long)two_n, sign);
}

__global__ void k_gen_bluestein_b(double2* b, const double2* c, long long n_orig, long long m,
                                  bool conj) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double2 v = double2{0.0, 0.0};
  if (i < n_orig) v = c[i];
  else if (i > m - n_orig) v = c[m - i];
  if (conj) v.y = -v.y;
  b[i] = v;
}

__global__ void k_gen_rader_b(double2* b, const int* perm_out, long long n_orig, long long l, int sign) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= l) return;
  b[j] = root_fp64(perm_out[j], n_orig, sign);
}

template <typename V>
__global__ void k_convert(V* out, const double2* in, long long count, double scale, long long n1,
                          long long n2, bool permute) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  long long src = i;
  if (permute) {
    const long long k1 = i / n2, k2 = i - k1 * n2;
    src = k1 + n1 * k2;
  }
  const double2 v = in[src];
  out[i] = V{(S<V>)(v.x * scale), (S<V>)(v.y * scale)};
}

template <typename V>
__global__ void k_gen_roots(V* out, long long count, long long step, long long d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double2 w = root_fp64((long long)(((unsigned __int128)i * step) % d), d, -1);
  out[i] = V{(S<V>)w.x, (S<V>)w.y};
}

static inline dim3 grid_for(long long count) { return dim3((unsigned)((count + 255) / 256)); }

hipError_t launch_gen_chirp(double2* c, long long n_orig, long long count, int sign, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_chirp, grid_for(count), dim3(256), 0, s, c, n_orig, count, sign);
  return hipGetLastError();
}
hipError_t launch_gen_bluestein_b(double2* b, const double2* c, long long n_orig, long long m,
                                  bool conj, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_bluestein_b, grid_for(m), dim3(256), 0, s, b, c, n_orig, m, conj);
  return hipGetLastError();
}
hipError_t launch_gen_rader_b(double2* b, const int* perm_out, long long n_orig, long long l,
                              int sign, hipStream_t s) {
  hipLaunchKernelGGL(k_gen_rader_b, grid_for(l), dim3(256), 0, s, b, perm_out, n_orig, l, sign);
  return hipGetLastError();
}
hipError_t launch_convert(int prec, void* out, const double2* in, long long count, double scale,
                          long long n1, long long n2, bool permute, hipStream_t s) {
  if (prec == kF64)
    hipLaunchKernelGGL(k_convert<double2>, grid_for(count), dim3(256), 0, s, (double2*)out, in,
                       count, scale, n1, n2, permute);
  else
    hipLaunchKernelGGL(k_convert<float2>, grid_for(count), dim3(256), 0, s, (float2*)out, in,
                       count, scale, n1, n2, permute);
  return hipGetLastError();
}
hipError_t launch_gen_roots(int prec, void* out, long long count, long long step, long long d,
                            hipStream_t s) {
  if (prec == kF64)
    hipLaunchKernelGGL(k_gen_roots<double2>, grid_for(count), dim3(256), 0, s, (double2*)out, count, step, d);
  else
    hipLaunchKernelGGL(k_gen_roots<float2>, grid_for(count), dim3(256), 0, s, (float2*)out, count, step, d);
  return hipGetLastError();
}

}  // namespace fyfft
