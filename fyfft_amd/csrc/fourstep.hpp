// Four-step pass kernels and their launchers, shared by the translation
// units that instantiate them (power-of-two and odd sub-transform lengths).
#pragma once
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

template <typename V, typename F, int DIR, int MODE>
__global__ __launch_bounds__(1024) void k_fs_col(FsArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int groups = (a.n2 + (1 << a.log2c) - 1) >> a.log2c;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long t = lb / groups;
  const int c0 = (int)(lb - t * groups) << a.log2c;
  const long long tg = a.group ? t / a.group : 0, ti = a.group ? t - tg * a.group : t;
  body_fs_col<V, F, DIR, MODE>(a, a.in + tg * a.in_gdist + ti * a.in_dist, a.tmp + tg * a.tmp_gdist + ti * a.tmp_dist,
                               a.out + t * a.out_dist, a.x0buf ? a.x0buf + t : nullptr, c0,
                               reinterpret_cast<V*>(smem_raw), a.zw ? a.zw + t * a.zw_dist : nullptr);
}

template <typename V, typename F, int DIR, int MODE>
__global__ __launch_bounds__(1024) void k_fs_row(FsArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int groups = (a.n1 + (1 << a.log2c) - 1) >> a.log2c;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long t = lb / groups;
  const int r0 = (int)(lb - t * groups) << a.log2c;
  body_fs_row<V, F, DIR, MODE>(a, a.tmp + t * a.tmp_dist, a.out + t * a.out_dist,
                               a.x0buf ? a.x0buf + t : nullptr, r0,
                               reinterpret_cast<V*>(smem_raw));
}

// ------------------------------------------------------------ launch
template <typename V, typename F, int DIR, int MODE>
static hipError_t go_col(const FsArgs<V>& a, size_t batch, size_t lds, int threads, hipStream_t s) {
  auto k = k_fs_col<V, F, DIR, MODE>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const size_t groups = ((size_t)a.n2 + (1 << a.log2c) - 1) >> a.log2c;
  hipLaunchKernelGGL(k, dim3((unsigned)(groups * batch)), dim3(threads), lds, s, a);
  return hipGetLastError();
}
template <typename V, typename F, int DIR, int MODE>
static hipError_t go_row(const FsArgs<V>& a, size_t batch, size_t lds, int threads, hipStream_t s) {
  auto k = k_fs_row<V, F, DIR, MODE>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const size_t groups = ((size_t)a.n1 + (1 << a.log2c) - 1) >> a.log2c;
  hipLaunchKernelGGL(k, dim3((unsigned)(groups * batch)), dim3(threads), lds, s, a);
  return hipGetLastError();
}

// The plain (non-convolution) modes only.
template <typename V, typename F>
static hipError_t dispatch_fs_plain(bool col, int dir, const FsArgs<V>& a, size_t batch, size_t lds, int threads,
                                    hipStream_t s) {
  if (col)
    return dir < 0 ? go_col<V, F, -1, kFsPlainA>(a, batch, lds, threads, s)
                   : go_col<V, F, +1, kFsPlainA>(a, batch, lds, threads, s);
  return dir < 0 ? go_row<V, F, -1, kFsPlainB>(a, batch, lds, threads, s)
                 : go_row<V, F, +1, kFsPlainB>(a, batch, lds, threads, s);
}

template <typename V, typename F>
static hipError_t dispatch_fs(bool col, int mode, int dir, const FsArgs<V>& a, size_t batch,
                              size_t lds, int threads, hipStream_t s) {
  if (col) {
    switch (mode) {
      case kFsPlainA:
        return dir < 0 ? go_col<V, F, -1, kFsPlainA>(a, batch, lds, threads, s)
                       : go_col<V, F, +1, kFsPlainA>(a, batch, lds, threads, s);
      case kFsRaderA: return go_col<V, F, -1, kFsRaderA>(a, batch, lds, threads, s);
      case kFsBlueA: return go_col<V, F, -1, kFsBlueA>(a, batch, lds, threads, s);
      case kFsRaderC: return go_col<V, F, +1, kFsRaderC>(a, batch, lds, threads, s);
      case kFsRaderC2: return go_col<V, F, +1, kFsRaderC2>(a, batch, lds, threads, s);
      case kFsBlueC: return go_col<V, F, +1, kFsBlueC>(a, batch, lds, threads, s);
    }
  } else {
    switch (mode) {
      case kFsPlainB:
        return dir < 0 ? go_row<V, F, -1, kFsPlainB>(a, batch, lds, threads, s)
                       : go_row<V, F, +1, kFsPlainB>(a, batch, lds, threads, s);
      case kFsConvB: return go_row<V, F, -1, kFsConvB>(a, batch, lds, threads, s);
    }
  }
  return hipErrorInvalidValue;
}

}  // namespace fyfft
