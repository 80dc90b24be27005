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
  b.pos_out = a.pos_out;
  b.chirp = a.chirp;
  b.kern = a.kern;
  b.scale = a.scale;
  return b;
}

// One transform per workgroup slot.  TPB > 0: F is a compile-time policy
// (FixedIo) whose geometry the launch has (TPB transforms per workgroup,
// static LDS, and for TPB == 1 no row guard); TPB == 0: runtime passes.
template <typename V, typename F, int TPB, bool RADER>
__global__ __launch_bounds__(TPB > 0 ? F::kThreads * TPB : 1024) void k_conv_lds(ConvArgs<V> a) {
  constexpr bool kFixed = TPB > 0;
  constexpr int kStride = kFixed ? lds_len(F::kN, F::kPs) : 1;
  __shared__ __align__(16) V lds_static[kFixed ? kStride * TPB : 1];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  LdsBody<V> b = conv_body(a);
  if constexpr (kFixed) {
    b.m = F::kN;
    b.threads = F::kThreads;
    b.ps = F::kPs;
    b.stride = kStride;
  }
  const int slot = threadIdx.x / b.threads, tid = threadIdx.x - slot * b.threads;
  const long long t = (long long)blockIdx.x * (kFixed ? TPB : a.per_block) + slot;
  const bool active = TPB == 1 || t < a.batch;
  const long long off = (active ? t : 0) * a.dist;
  V* lds = (kFixed ? lds_static : reinterpret_cast<V*>(smem_raw)) + slot * b.stride;
  if constexpr (RADER) body_rader<V, F>(b, a.in + off, a.out + off, lds, tid, active);
  else body_blue<V, F>(b, a.in + off, a.out + off, lds, tid, active);
}

template <typename V, typename F, int TPB>
static hipError_t go_conv(bool rader, const ConvArgs<V>& a, const LdsGeom& g, hipStream_t s) {
  auto k = rader ? k_conv_lds<V, F, TPB, true> : k_conv_lds<V, F, TPB, false>;
  size_t lds = g.lds_bytes;
  if constexpr (TPB > 0) {
    if (g.threads != F::kThreads || g.per_block != TPB || g.pad_shift != F::kPs || a.m != F::kN)
      return hipErrorInvalidValue;
    lds = 0;
  }
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const unsigned grid = (unsigned)((a.batch + g.per_block - 1) / g.per_block);
  hipLaunchKernelGGL(k, dim3(grid), dim3(g.threads * g.per_block), lds, s, a);
  return hipGetLastError();
}

// Compile-time convolution transforms for the power-of-two lengths and the
// 5 * 2^k lengths (20480 = 2^12 * 5 serves Bluestein N = 10007); radices as
// radices_for (planner.cpp), geometry per entry (the plan takes it).
using ConvLauncher = hipError_t (*)(bool rader, const void* args, const LdsGeom& g, hipStream_t s);
template <typename V, int N, int T, int TPB, int PS, int... Rs>
static hipError_t launch_conv_fixed(bool rader, const void* args, const LdsGeom& g, hipStream_t s) {
  const auto& a = *static_cast<const ConvArgs<V>*>(args);
  return go_conv<V, FixedIo<V, N, T, PS, Rs...>, TPB>(rader, a, g, s);
}
struct ConvFixed {
  int prec, n, threads, per_block, pad_shift, npass;
  int radix[4];
  ConvLauncher launch;
};
#define CFIXED(P, V, N, T, TPB, PS, NP, ...) \
  {P, N, T, TPB, PS, NP, {__VA_ARGS__}, &launch_conv_fixed<V, N, T, TPB, PS, __VA_ARGS__>}
static const ConvFixed kConvFixed[] = {
    CFIXED(kF32, float2, 256, 16, 16, 4, 2, 16, 16),
    CFIXED(kF32, float2, 512, 32, 8, 4, 3, 8, 8, 8),
    CFIXED(kF32, float2, 1024, 64, 4, 4, 3, 16, 8, 8),
    CFIXED(kF32, float2, 1280, 80, 2, 4, 3, 16, 16, 5),
    CFIXED(kF32, float2, 2048, 128, 2, 4, 3, 16, 16, 8),
    CFIXED(kF32, float2, 2560, 160, 1, 4, 4, 8, 8, 8, 5),
    CFIXED(kF32, float2, 4096, 256, 1, 4, 3, 16, 16, 16),
    CFIXED(kF32, float2, 5120, 320, 1, 4, 4, 16, 8, 8, 5),
    CFIXED(kF32, float2, 8192, 512, 1, 4, 4, 16, 8, 8, 8),
    CFIXED(kF32, float2, 10240, 640, 1, 4, 4, 16, 16, 8, 5),
    CFIXED(kF32, float2, 16384, 1024, 1, 4, 4, 16, 16, 8, 8),
    CFIXED(kF32, float2, 20480, 1024, 1, 30, 4, 16, 16, 16, 5),
    // FP64: the runtime-pass convolution kernel keeps call frames in scratch
    CFIXED(kF64, double2, 256, 32, 8, 4, 2, 16, 16),
    CFIXED(kF64, double2, 512, 64, 4, 4, 3, 8, 8, 8),
    CFIXED(kF64, double2, 1024, 128, 2, 4, 3, 16, 8, 8),
    CFIXED(kF64, double2, 1280, 160, 1, 4, 3, 16, 16, 5),
    CFIXED(kF64, double2, 2048, 256, 1, 4, 3, 16, 16, 8),
    CFIXED(kF64, double2, 2560, 320, 1, 4, 4, 8, 8, 8, 5),
    CFIXED(kF64, double2, 4096, 512, 1, 4, 3, 16, 16, 16),
    CFIXED(kF64, double2, 5120, 640, 1, 4, 4, 16, 8, 8, 5),
    CFIXED(kF64, double2, 8192, 1024, 1, 4, 4, 16, 8, 8, 8),
    CFIXED(kF64, double2, 10240, 1024, 1, 30, 4, 16, 16, 8, 5),
};
#undef CFIXED

int find_conv_fixed(int prec, const PassDesc& pd, LdsGeom* g) {
  for (int i = 0; i < (int)(sizeof(kConvFixed) / sizeof(kConvFixed[0])); ++i) {
    const ConvFixed& e = kConvFixed[i];
    if (e.prec != prec || e.n != g->n || e.npass != pd.npass) continue;
    bool same = true;
    for (int p = 0; p < pd.npass; ++p) same = same && e.radix[p] == pd.radix[p];
    if (!same) continue;
    g->threads = e.threads;
    g->per_block = e.per_block;
    g->pad_shift = e.pad_shift;
    g->ppt = cdiv(e.n, e.threads);
    g->lds_bytes = (size_t)lds_len(e.n, e.pad_shift) * e.per_block * elem_bytes(prec);
    return i;
  }
  return -1;
}

hipError_t launch_conv_lds(int prec, bool rader, const void* args, const LdsGeom& g, int fixed, hipStream_t s) {
  if (fixed >= 0) {
    if (fixed >= (int)(sizeof(kConvFixed) / sizeof(kConvFixed[0])) || kConvFixed[fixed].prec != prec)
      return hipErrorInvalidValue;
    const bool empty = prec == kF64 ? static_cast<const ConvArgs<double2>*>(args)->batch == 0
                                    : static_cast<const ConvArgs<float2>*>(args)->batch == 0;
    return empty ? hipSuccess : kConvFixed[fixed].launch(rader, args, g, s);
  }
  if (prec == kF64) {
    const auto& a = *static_cast<const ConvArgs<double2>*>(args);
    if (a.batch == 0) return hipSuccess;
    return g.ppt <= 8 ? go_conv<double2, GenericIo<double2, 8>, 0>(rader, a, g, s)
                      : go_conv<double2, GenericIo<double2, 16>, 0>(rader, a, g, s);
  }
  const auto& a = *static_cast<const ConvArgs<float2>*>(args);
  if (a.batch == 0) return hipSuccess;
  return g.ppt <= 16 ? go_conv<float2, GenericIo<float2, 16>, 0>(rader, a, g, s)
                     : go_conv<float2, GenericIo<float2, 32>, 0>(rader, a, g, s);
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
