// Strided layouts: the column kernel of multi-dimensional and interleaved
// batched transforms, and the gather/scatter pair that serves every other
// layout through a contiguous workspace.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

// Columns [c0, c0 + C) of one plane: C*n elements read as n rows of C
// adjacent values, C column transforms in LDS, written back the same way.
// Thread t moves column c = t & (C - 1), rows t >> log2c + j*s (s =
// blockDim >> log2c), with offsets advancing by constant steps; its LDS
// values are all read before its first store.  The twiddle w_N^(b*k) of
// row k in inner plane b (uniform across the workgroup) is w^(b*k0) *
// T[j], T[j] = w^(b*s*j) a table of P roots in LDS after the tile.
template <typename V, typename F, int DIR>
__device__ __forceinline__ void body_columns(const ColArgs<V>& a, const V* x, V* y, int c0, V* lds_in, int b) {
  constexpr int P = F::kPpt;
  V* lds_all = as_lds(lds_in);
  const int N = a.n, lc = a.log2c, C = 1 << lc;
  const int slot = threadIdx.x / a.threads;
  const int tid = threadIdx.x - slot * a.threads;
  const int ps = a.ps, stride = a.stride;
  const int ncols = min(C, a.ncols - c0);
  const int c = threadIdx.x & (C - 1), kr0 = threadIdx.x >> lc, s = blockDim.x >> lc;
  const bool cok = c < ncols;
  V* lcol = lds_all + c * stride;
  {
    const int k0 = kr0;
    const V* src = x + (long long)k0 * a.istride + c0 + (cok ? c : 0);
    const long long step = (long long)s * a.istride;
    V buf[P];
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (k0 + j * s < N) buf[j] = gload(src + j * step);
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const int k = k0 + j * s;
      if (k < N && cok) lcol[k + (k >> ps)] = buf[j];
    }
  }
  const bool twiddle = a.tw_hi != nullptr;
  const int lo_mask = (1 << a.lo_bits) - 1;
  auto root = [&](long long m) { return cmul(a.tw_hi[m >> a.lo_bits], a.tw_lo[m & lo_mask]); };
  V* ttab = lds_all + C * stride;
  if (twiddle)
    for (int j = threadIdx.x; j < P; j += blockDim.x) ttab[j] = j * s < N ? root((long long)b * s * j) : V{1, 0};
  __syncthreads();
  const bool active = slot < ncols;
  F::template run<DIR>(lds_all + (slot < C ? slot : 0) * stride, N, a.threads, tid, a.pd, a.tw, active, ps);
  const int k0 = opaque(kr0);
  V val[P];
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const int k = min(k0 + j * s, N - 1);
    val[j] = lcol[k + (k >> ps)];
  }
  const V wb = twiddle ? root((long long)b * min(k0, N - 1)) : V{1, 0};
  V* dst = y + (long long)k0 * a.ostride + c0 + c;
  const long long step = (long long)s * a.ostride;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    if (k0 + j * s < N && cok) {
      V v = val[j];
      if (twiddle) v = twmul<DIR>(v, cmul(wb, ttab[j]));
      gstore(dst + j * step, cscale(v, a.scale));
    }
  }
}

template <typename V, typename F, int DIR>
__global__ __launch_bounds__(1024) void k_columns(ColArgs<V> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int groups = (a.ncols + (1 << a.log2c) - 1) >> a.log2c;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const long long q = lb / groups;
  const int c0 = (int)(lb - q * groups) << a.log2c;
  const long long p = q / a.inner;
  const int b = (int)(q - p * a.inner);
  body_columns<V, F, DIR>(a, a.in + p * a.ipdist + b * a.ipdist2, a.out + p * a.opdist + b * a.opdist2, c0,
                          reinterpret_cast<V*>(smem_raw), b);
}

template <typename V, typename F>
static hipError_t go_columns(int dir, const ColArgs<V>& a, size_t planes, size_t lds, int threads,
                             hipStream_t s) {
  auto k = dir < 0 ? k_columns<V, F, -1> : k_columns<V, F, +1>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  const size_t groups = ((size_t)a.ncols + (1 << a.log2c) - 1) >> a.log2c;
  hipLaunchKernelGGL(k, dim3((unsigned)(groups * planes)), dim3(threads), lds, s, a);
  return hipGetLastError();
}

template <typename V, int PPT>
static hipError_t columns_len(int fixed_n, int dir, const ColArgs<V>& a, size_t planes, size_t lds, int threads,
                              hipStream_t s) {
  constexpr int P = sizeof(V) == 8 ? 16 : 8;  // points per thread of the fixed forms
  switch (fixed_n) {
    case 64: return go_columns<V, FixedFft<V, 64, 64 / P, 8, 8>>(dir, a, planes, lds, threads, s);
    case 128: return go_columns<V, FixedFft<V, 128, 128 / P, 16, 8>>(dir, a, planes, lds, threads, s);
    case 256: return go_columns<V, FixedFft<V, 256, 256 / P, 16, 16>>(dir, a, planes, lds, threads, s);
    case 512: return go_columns<V, FixedFft<V, 512, 512 / P, 8, 8, 8>>(dir, a, planes, lds, threads, s);
    case 1024: return go_columns<V, FixedFft<V, 1024, 1024 / P, 16, 8, 8>>(dir, a, planes, lds, threads, s);
    case 2048: return go_columns<V, FixedFft<V, 2048, 2048 / P, 16, 16, 8>>(dir, a, planes, lds, threads, s);
    case 4096:
      if constexpr (sizeof(V) == 8)
        return go_columns<V, FixedFft<V, 4096, 4096 / P, 16, 16, 16>>(dir, a, planes, lds, threads, s);
      break;
  }
  return go_columns<V, GenericFft<V, PPT>>(dir, a, planes, lds, threads, s);
}

hipError_t launch_columns(int prec, int dir, const void* args, size_t planes, const LdsGeom& g, int fixed_n,
                          hipStream_t s) {
  if (planes == 0) return hipSuccess;
  const int threads = g.threads * g.per_block;
  // the tile, then the table of g.ppt twiddle roots (body_columns)
  const size_t lds = g.lds_bytes + (size_t)g.ppt * (prec == kF64 ? 16 : 8);
  if (prec == kF64) {
    const auto& a = *static_cast<const ColArgs<double2>*>(args);
    return g.ppt <= 8 ? columns_len<double2, 8>(fixed_n, dir, a, planes, lds, threads, s)
                      : columns_len<double2, 16>(fixed_n, dir, a, planes, lds, threads, s);
  }
  const auto& a = *static_cast<const ColArgs<float2>*>(args);
  return g.ppt <= 16 ? columns_len<float2, 16>(fixed_n, dir, a, planes, lds, threads, s)
                     : columns_len<float2, 32>(fixed_n, dir, a, planes, lds, threads, s);
}

// ------------------------------------------------------------ gather / scatter
template <typename V, bool SCATTER>
__global__ void k_gather(const V* src, V* dst, long long n, long long stride, long long cdist, long long ncols,
                         long long pdist, long long t0, long long count) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= count * n) return;
  const long long r = e / n, j = e - r * n;  // row r of the workspace, element j
  const long long t = t0 + r;
  const long long p = t / ncols, c = t - p * ncols;
  const long long g = p * pdist + c * cdist + j * stride;
  if (SCATTER) dst[g] = src[e];
  else dst[e] = src[g];
}

// Adjacent transforms (cdist 1): 32 transforms x 32 elements per tile
// through LDS, so both the strided side (32 adjacent columns of one row)
// and the contiguous side (32 consecutive elements of one transform) move
// in 256-byte (FP32) segments.
template <typename V, bool SCATTER>
__global__ __launch_bounds__(256) void k_gather_tiled(const V* src, V* dst, long long n, long long stride,
                                                      long long ncols, long long pdist, long long t0,
                                                      long long count, long long jtiles) {
  __shared__ V tile[32][33];
  const long long jb = (blockIdx.x % jtiles) * 32, rb = (blockIdx.x / jtiles) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (!SCATTER) {
    for (int jj = ty; jj < 32; jj += 8) {
      const long long j = jb + jj, r = rb + tx;
      if (j < n && r < count) {
        const long long t = t0 + r, p = t / ncols, c = t - p * ncols;
        tile[jj][tx] = src[p * pdist + c + j * stride];
      }
    }
    __syncthreads();
    for (int rr = ty; rr < 32; rr += 8) {
      const long long r = rb + rr, j = jb + tx;
      if (r < count && j < n) dst[r * n + j] = tile[tx][rr];
    }
  } else {
    for (int rr = ty; rr < 32; rr += 8) {
      const long long r = rb + rr, j = jb + tx;
      if (r < count && j < n) tile[tx][rr] = src[r * n + j];
    }
    __syncthreads();
    for (int jj = ty; jj < 32; jj += 8) {
      const long long j = jb + jj, r = rb + tx;
      if (j < n && r < count) {
        const long long t = t0 + r, p = t / ncols, c = t - p * ncols;
        dst[p * pdist + c + j * stride] = tile[jj][tx];
      }
    }
  }
}

template <typename V>
static hipError_t go_gather(bool scatter, const void* strided, void* ws, long long n, long long stride,
                            long long cdist, long long ncols, long long pdist, long long t0, long long count,
                            hipStream_t s) {
  if (cdist == 1) {
    const long long jt = (n + 31) / 32, rt = (count + 31) / 32;
    const dim3 grid((unsigned)(jt * rt)), block(256);
    if (scatter)
      hipLaunchKernelGGL((k_gather_tiled<V, true>), grid, block, 0, s, (const V*)ws, (V*)strided, n, stride,
                         ncols, pdist, t0, count, jt);
    else
      hipLaunchKernelGGL((k_gather_tiled<V, false>), grid, block, 0, s, (const V*)strided, (V*)ws, n, stride,
                         ncols, pdist, t0, count, jt);
    return hipGetLastError();
  }
  const dim3 grid((unsigned)((count * n + 255) / 256)), block(256);
  if (scatter)
    hipLaunchKernelGGL((k_gather<V, true>), grid, block, 0, s, (const V*)ws, (V*)strided, n, stride, cdist,
                       ncols, pdist, t0, count);
  else
    hipLaunchKernelGGL((k_gather<V, false>), grid, block, 0, s, (const V*)strided, (V*)ws, n, stride, cdist,
                       ncols, pdist, t0, count);
  return hipGetLastError();
}

hipError_t launch_gather(int prec, bool scatter, const void* strided, void* ws, long long n, long long stride,
                         long long cdist, long long ncols, long long pdist, long long t0, long long count,
                         hipStream_t s) {
  if (count == 0) return hipSuccess;
  if (prec == kF64) return go_gather<double2>(scatter, strided, ws, n, stride, cdist, ncols, pdist, t0, count, s);
  return go_gather<float2>(scatter, strided, ws, n, stride, cdist, ncols, pdist, t0, count, s);
}

}  // namespace fyfft
