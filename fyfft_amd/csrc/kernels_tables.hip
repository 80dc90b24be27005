// Table generation on the GPU in FP64: every root is exp(-2*pi*i*m/D) with
// m reduced to a quadrant in integers before sincospi (root_fp64), then
// rounded once to the storage precision.
#include "internal.hpp"

namespace fyfft {
using namespace dev;

template <typename V>
__global__ void k_gen_stockham(V* tw, int n, PassDesc pd) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n - 1) return;
  int Ls = 1;
  for (int p = 0; p < pd.npass; ++p) {
    const int R = pd.radix[p];
    if (e < Ls * R - 1) {
      const int off = e - (Ls - 1);
      const int k = off / (R - 1);
      const int r = off % (R - 1) + 1;
      const double2 w = root_fp64((long long)k * r, (long long)Ls * R, -1);
      tw[e] = V{(S<V>)w.x, (S<V>)w.y};
      return;
    }
    Ls *= R;
  }
}

hipError_t launch_gen_stockham_twiddles(int prec, void* tw, int n, const PassDesc& pd,
                                        hipStream_t s) {
  if (n < 2) return hipSuccess;
  const int threads = 256, blocks = (n - 1 + threads - 1) / threads;
  if (prec == kF64)
    hipLaunchKernelGGL(k_gen_stockham<double2>, dim3(blocks), dim3(threads), 0, s, (double2*)tw, n, pd);
  else
    hipLaunchKernelGGL(k_gen_stockham<float2>, dim3(blocks), dim3(threads), 0, s, (float2*)tw, n, pd);
  return hipGetLastError();
}

}  // namespace fyfft
