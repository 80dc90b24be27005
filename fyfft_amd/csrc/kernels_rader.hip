// The Rader output permutation of four-step plans through LDS.
//
// y[g^i mod N] = x[0] + c[i] scatters the convolution result with no
// locality: writing element by element leaves partial cache lines that are
// evicted from L2 several times over.  Instead the closing column pass
// writes its results as runs ordered by destination chunk (zw, contiguous
// writes), and this kernel gathers the runs that land in one chunk of Q
// consecutive outputs, places them in LDS and stores the chunk
// contiguously: every HBM access is a run, the permutation happens in LDS.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

// One workgroup per (transform, chunk); each thread moves up to 16
// elements with all table and data loads issued before the LDS stores.
template <typename V>
__global__ __launch_bounds__(512) void k_rader_out(const V* __restrict__ zw, V* __restrict__ y, long long L,
                                                   int Q, int K, const int2* __restrict__ pos, long long zw_dist,
                                                   long long y_dist) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  const long long t = blockIdx.x / K;
  const int k = (int)(blockIdx.x - t * K);
  const long long base = (long long)k * Q;
  const int size = (int)min((long long)Q, L - base);
  const V* w = zw + t * zw_dist;
  staged<16>(threadIdx.x, blockDim.x, size, [&](int i) { return pos[base + i]; },
             [&](int, int2 e) { lds[e.y] = w[e.x]; });
  __syncthreads();
  V* out = y + t * y_dist + 1 + base;
  for (int j = threadIdx.x; j < size; j += blockDim.x) out[j] = lds[j];
}

hipError_t launch_rader_bucket(int prec, const void* zw, void* y, long long rows, long long L, int Q,
                               const int2* pos, long long zw_dist, long long y_dist, hipStream_t s) {
  if (rows == 0) return hipSuccess;
  const int K = (int)((L + Q - 1) / Q);
  const dim3 grid((unsigned)(rows * K)), block(512);
  if (prec == kF64)
    hipLaunchKernelGGL(k_rader_out<double2>, grid, block, (size_t)Q * sizeof(double2), s, (const double2*)zw,
                       (double2*)y, L, Q, K, pos, zw_dist, y_dist);
  else
    hipLaunchKernelGGL(k_rader_out<float2>, grid, block, (size_t)Q * sizeof(float2), s, (const float2*)zw,
                       (float2*)y, L, Q, K, pos, zw_dist, y_dist);
  return hipGetLastError();
}

}  // namespace fyfft
