// BF16 / FP16 storage with FP32 arithmetic: the compile-time Stockham
// passes of the FP32 kernels, with the first pass widening 16-bit values
// from HBM straight into registers and the last pass rounding FP32 results
// once to the storage type.  HBM traffic is 8 bytes per complex element
// (4 in, 4 out); arithmetic, twiddles and LDS stay FP32, so the only
// rounding beyond FP32 level is the final store.
#include "bodies.hpp"
#include "split.hpp"

namespace fyfft {
using namespace dev;

namespace {

template <typename H> struct LowSrc {
  static constexpr bool kFromLds = false;
  const cplx16<H>* x;
  __device__ float2 operator()(int i) const { return load16(x + i); }
};
template <typename H> struct LowSink {
  static constexpr bool kToLds = false;
  cplx16<H>* y;
  float scale;
  __device__ void operator()(int i, float2 v) const { store16(y + i, cscale(v, scale)); }
};

}  // namespace

template <typename H, int DIR, int N, int T, int TPB, int PS, int... Rs>
__global__ __launch_bounds__(T * TPB) void k_lowp_fixed(const cplx16<H>* in, cplx16<H>* out, long long batch,
                                                        const float2* __restrict__ tw, float scale) {
  constexpr int stride = lds_len(N, PS);
  __shared__ float2 lds_all[TPB * stride];
  const int slot = threadIdx.x / T;
  const int tid = threadIdx.x - slot * T;
  const long long t = (long long)blockIdx.x * TPB + slot;
  const bool active = t < batch;
  const long long off = (active ? t : 0) * (long long)N;
  fixed_io_passes<float2, DIR, N, T, 1, LowSrc<H>, LowSink<H>, Rs...>(
      LowSrc<H>{in + off}, LowSink<H>{out + off, scale}, as_lds(lds_all + slot * stride), PS, tid, tw, active);
}

using LowpLauncher = hipError_t (*)(bool bf16, int dir, const void* in, void* out, size_t batch, const void* tw,
                                    double scale, hipStream_t s);

template <typename H, int N, int T, int TPB, int PS, int... Rs>
static hipError_t launch_lowp_h(int dir, const void* in, void* out, size_t batch, const void* tw, double scale,
                                hipStream_t s) {
  const dim3 grid((unsigned)((batch + TPB - 1) / TPB)), block(T * TPB);
  if (dir < 0)
    hipLaunchKernelGGL((k_lowp_fixed<H, -1, N, T, TPB, PS, Rs...>), grid, block, 0, s, (const cplx16<H>*)in,
                       (cplx16<H>*)out, (long long)batch, (const float2*)tw, (float)scale);
  else
    hipLaunchKernelGGL((k_lowp_fixed<H, +1, N, T, TPB, PS, Rs...>), grid, block, 0, s, (const cplx16<H>*)in,
                       (cplx16<H>*)out, (long long)batch, (const float2*)tw, (float)scale);
  return hipGetLastError();
}

template <int N, int T, int TPB, int PS, int... Rs>
static hipError_t launch_lowp_n(bool bf16, int dir, const void* in, void* out, size_t batch, const void* tw,
                                double scale, hipStream_t s) {
  return bf16 ? launch_lowp_h<__bf16, N, T, TPB, PS, Rs...>(dir, in, out, batch, tw, scale, s)
              : launch_lowp_h<_Float16, N, T, TPB, PS, Rs...>(dir, in, out, batch, tw, scale, s);
}

// The split-exchange form (split.hpp): registers between passes, LDS an
// exchange buffer of N floats, W workgroups per CU.
template <typename H, int N, int T, int PS, int W, int... Rs>
static hipError_t launch_lowp_split_h(int dir, const void* in, void* out, size_t batch, const void* tw,
                                      double scale, hipStream_t s) {
  if (batch > 0xffffffffull) return hipErrorInvalidValue;  // one workgroup per transform
  const dim3 grid((unsigned)batch), block(T);
  if (dir < 0)
    hipLaunchKernelGGL((k_lds_fixed_split<float2, cplx16<H>, -1, N, T, PS, W, Rs...>), grid, block, 0, s,
                       (const cplx16<H>*)in, (cplx16<H>*)out, (long long)N, (const float2*)tw, (float)scale);
  else
    hipLaunchKernelGGL((k_lds_fixed_split<float2, cplx16<H>, +1, N, T, PS, W, Rs...>), grid, block, 0, s,
                       (const cplx16<H>*)in, (cplx16<H>*)out, (long long)N, (const float2*)tw, (float)scale);
  return hipGetLastError();
}

template <int N, int T, int PS, int W, int... Rs>
static hipError_t launch_lowp_split_n(bool bf16, int dir, const void* in, void* out, size_t batch, const void* tw,
                                      double scale, hipStream_t s) {
  return bf16 ? launch_lowp_split_h<__bf16, N, T, PS, W, Rs...>(dir, in, out, batch, tw, scale, s)
              : launch_lowp_split_h<_Float16, N, T, PS, W, Rs...>(dir, in, out, batch, tw, scale, s);
}

struct LowpEntry {
  int n, npass;
  int radix[6];
  LowpLauncher launch;
};

// Geometry and pass lists of the FP32 compile-time kernels; the radix lists
// equal radices_for(N) so the plan's FP32 Stockham table matches.  4096,
// 8192 and 16384 take the split-exchange form (registers between passes, LDS
// an exchange buffer of N floats) at 8 / 4 / 2 workgroups per CU: BF16
// 4096 x 65536 0.484 -> 0.451 ms, 8192 x 32768 0.620 -> 0.505, 16384 x 16384
// 0.741 -> 0.581 (the same results to the bit; 6 workgroups per CU for 4096
// measured the same as 8).
#define LOWP(N, T, TPB, PS, NP, ...) {N, NP, {__VA_ARGS__}, &launch_lowp_n<N, T, TPB, PS, __VA_ARGS__>}
#define LOWPX(N, T, PS, W, NP, ...) {N, NP, {__VA_ARGS__}, &launch_lowp_split_n<N, T, PS, W, __VA_ARGS__>}
static const LowpEntry kLowp[] = {
    LOWP(256, 16, 16, 30, 2, 16, 16),      LOWP(512, 32, 8, 30, 3, 8, 8, 8),
    LOWP(1024, 64, 4, 30, 3, 16, 8, 8),    LOWP(2048, 128, 2, 30, 3, 16, 16, 8),
    LOWPX(4096, 256, 4, 8, 3, 16, 16, 16), LOWPX(8192, 512, 4, 4, 4, 16, 8, 8, 8),
    LOWPX(16384, 1024, 4, 2, 4, 16, 16, 8, 8),
};
#undef LOWP
#undef LOWPX

bool lowp_fixed_supported(size_t n, dev::PassDesc* pd) {
  for (const LowpEntry& e : kLowp)
    if ((size_t)e.n == n) {
      if (pd) {
        pd->npass = e.npass;
        for (int p = 0; p < e.npass; ++p) pd->radix[p] = e.radix[p];
      }
      return true;
    }
  return false;
}

hipError_t launch_lowp_fixed(bool bf16, int dir, size_t n, const void* in, void* out, size_t batch,
                             const void* tw, double scale, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  for (const LowpEntry& e : kLowp)
    if ((size_t)e.n == n) return e.launch(bf16, dir, in, out, batch, tw, scale, s);
  return hipErrorInvalidValue;
}

}  // namespace fyfft
