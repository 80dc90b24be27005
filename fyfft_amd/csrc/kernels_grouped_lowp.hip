// Grouped ragged transforms with BF16 / FP16 storage: the family kernels of
// grouped.hpp with FP32 arithmetic and 16-bit input and output, and the
// 16-bit <-> complex float conversion the per-transform baseline of those
// plans runs around each FP32 transform.
#include <algorithm>

#include "grouped.hpp"

namespace fyfft {

hipError_t group_family_lowp(bool bf16, bool prepare, int fam, int dir, bool hi, int block, const void* args,
                             unsigned grid, size_t lds, hipStream_t s) {
  return bf16 ? group_family_dispatch<float2, cplx16<__bf16>>(prepare, fam, dir, hi, block, args, grid, lds, s)
              : group_family_dispatch<float2, cplx16<_Float16>>(prepare, fam, dir, hi, block, args, grid, lds, s);
}

template <typename H, bool TO_F32>
__global__ __launch_bounds__(256) void k_convert16(const void* src, void* dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if constexpr (TO_F32)
      gstore(static_cast<float2*>(dst) + i, load16(static_cast<const cplx16<H>*>(src) + i));
    else
      store16(static_cast<cplx16<H>*>(dst) + i, gload(static_cast<const float2*>(src) + i));
  }
}

hipError_t launch_convert16(bool bf16, bool to_f32, const void* src, void* dst, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  if (bf16) {
    if (to_f32) hipLaunchKernelGGL((k_convert16<__bf16, true>), dim3(grid), dim3(256), 0, s, src, dst, n);
    else hipLaunchKernelGGL((k_convert16<__bf16, false>), dim3(grid), dim3(256), 0, s, src, dst, n);
  } else {
    if (to_f32) hipLaunchKernelGGL((k_convert16<_Float16, true>), dim3(grid), dim3(256), 0, s, src, dst, n);
    else hipLaunchKernelGGL((k_convert16<_Float16, false>), dim3(grid), dim3(256), 0, s, src, dst, n);
  }
  return hipGetLastError();
}

}  // namespace fyfft
