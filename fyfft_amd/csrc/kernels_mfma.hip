// Low-precision storage form: BF16 or FP16 complex in HBM (4 bytes per
// element), radix-16 stages on MFMA.
//
// One radix-16 stage for 16 butterflies is a real matrix product
//   [Y_0r Y_0i ... Y_15r Y_15i]^T = A (32 x 32) * X (32 x 16 butterflies)
// computed as two v_mfma_f32_16x16x32_f16 (rows Yr with A_r, rows Yi with
// A_i; K = 32 runs over the interleaved (re, im) inputs).  Lane l of a wave
// holds, for butterfly n = l % 16, inputs r = 4*(l/16) .. +3 (B operand,
// k = 8*(l/16) + i) and receives outputs m = 4*(l/16) .. +3 (C operand) --
// the operand layout of the 16x16x32 MFMA, checked on the device.
//
// Accuracy: fp16 operands carry 11 significant bits.  Each operand is split
// v = hi + lo (hi = fp16(v), lo = fp16(v - hi)), giving ~22 bits, and the
// product is A_hi*X_hi + A_hi*X_lo + A_lo*X_hi accumulated in FP32 (the
// omitted A_lo*X_lo is ~2^-22 relative).  Each butterfly is scaled by a
// power of two so its largest input component is in [0.5, 1) before the
// split (exact, undone after the product), so fp16's range never limits
// the data.  Twiddles are applied in FP32 from the FP64-generated table.
// The stored result is rounded once to BF16/FP16.
#include "internal.hpp"

namespace fyfft {
using namespace dev;

typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// cos / sin (2*pi*j/16), j = 0..15
__device__ __forceinline__ float cos16(int j) {
  const double c[4] = {1.0, kCos8, kSqrtHalf, kSin8};
  j &= 15;
  const int q = j >> 2, r = j & 3;
  // cos(pi/2*q + t) with t = 2*pi*r/16
  const double cr = c[r], sr = (r == 0) ? 0.0 : c[4 - r];
  switch (q) {
    case 0: return (float)cr;
    case 1: return (float)-sr;
    case 2: return (float)-cr;
    default: return (float)sr;
  }
}
__device__ __forceinline__ float sin16(int j) { return cos16(j - 4); }

struct Mfma16A {
  f16x8_t rh, rl, ih, il;
};

// This lane's part of the 16x32 A operands (row m = l % 16, k = 8*(l/16)+i).
template <int DIR>
__device__ __forceinline__ Mfma16A mfma16_consts() {
  const int l = threadIdx.x & 63, m = l & 15, g = l >> 4;
  Mfma16A a;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = 4 * g + (i >> 1);
    const int idx = (m * r) & 15;
    const float fr = cos16(idx), fi = (DIR < 0 ? -1.f : 1.f) * sin16(idx);
    const float vr = (i & 1) ? -fi : fr;  // Yr = sum Fr*Xr - Fi*Xi
    const float vi = (i & 1) ? fr : fi;   // Yi = sum Fi*Xr + Fr*Xi
    const _Float16 rh = (_Float16)vr, ih = (_Float16)vi;
    a.rh[i] = rh;
    a.rl[i] = (_Float16)(vr - (float)rh);
    a.ih[i] = ih;
    a.il[i] = (_Float16)(vi - (float)ih);
  }
  return a;
}

// 16 butterflies across the wave: x[0..3] are this lane's inputs r = 4g+i,
// y[0..3] its outputs m = 4g+i.  Every lane of the wave must call this.
__device__ __forceinline__ void mfma_dft16(const Mfma16A& A, const float2 x[4], float2 y[4]) {
  float mx = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i) mx = fmaxf(mx, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  // mx = f * 2^e with f in [0.5, 1), read from the exponent field without
  // branches (zero and subnormal maxima give e = -126, which is harmless);
  // s = 2^-e and inv = 2^e are built directly as IEEE powers of two.
  const int e = min((int)((__float_as_uint(mx) >> 23) & 0xff) - 126, 126);
  const float s = __uint_as_float((unsigned)(127 - e) << 23);
  const float inv = __uint_as_float((unsigned)(127 + e) << 23);
  f16x8_t bh, bl;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float v = ((i & 1) ? x[i >> 1].y : x[i >> 1].x) * s;
    const _Float16 h = (_Float16)v;
    bh[i] = h;
    bl[i] = (_Float16)(v - (float)h);
  }
  f32x4_t yr = {0.f, 0.f, 0.f, 0.f}, yi = {0.f, 0.f, 0.f, 0.f};
  yr = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.rh, bh, yr, 0, 0, 0);
  yr = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.rh, bl, yr, 0, 0, 0);
  yr = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.rl, bh, yr, 0, 0, 0);
  yi = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.ih, bh, yi, 0, 0, 0);
  yi = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.ih, bl, yi, 0, 0, 0);
  yi = __builtin_amdgcn_mfma_f32_16x16x32_f16(A.il, bh, yi, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = float2{yr[i] * inv, yi[i] * inv};
}

// One radix-16 Stockham pass on MFMA.  W waves x G groups x 16 butterflies
// cover the N/16 butterflies; src/dst are LDS (FP32 complex, index i at
// i + (i >> ps)) or, for the first/last pass, the 16-bit HBM buffers.
template <typename H, int N, int W, int G, int DIR, bool SRC_G, bool DST_G>
__device__ __forceinline__ void mfma_pass16(const cplx16<H>* gsrc, cplx16<H>* gdst, float2* lds, int Ls,
                                            const float2* __restrict__ tw, const Mfma16A& A,
                                            float scale, int ps) {
  constexpr int nb = N / 16;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, n = l & 15, g = l >> 4;
  float2 y[G][4];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = (w * G + q) * 16 + n;
    float2 x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = j + (4 * g + i) * nb;
      x[i] = SRC_G ? load16(gsrc + idx) : lds[idx + (idx >> ps)];
    }
    if (Ls > 1) {
      const int k = j % Ls;
      const float2* t = tw + (Ls - 1) + k * 15;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * g + i;  // r = 0 only on lanes g = 0: select, no branch
        const float2 w = t[r > 0 ? r - 1 : 0];
        const float2 xt = twmul<DIR>(x[i], w);
        x[i] = r > 0 ? xt : x[i];
      }
    }
    mfma_dft16(A, x, y[q]);
  }
  if (!SRC_G && !DST_G) __syncthreads();
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int j = (w * G + q) * 16 + n;
    const int k = (Ls == 1) ? 0 : j % Ls;
    const int base = (j - k) * 16 + k;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int idx = base + (4 * g + i) * Ls;
      if (DST_G) store16(gdst + idx, cscale(y[q][i], scale));
      else lds[idx + (idx >> ps)] = y[q][i];
    }
  }
  if (!DST_G) __syncthreads();
}

// A VALU pass of radix R (2, 4 or 8) reading the 16-bit input, as the first
// pass when N is not a power of 16.
template <typename H, int N, int T, int R, int DIR>
__device__ __forceinline__ void first_pass_valu(const cplx16<H>* gsrc, float2* lds, int ps) {
  constexpr int nb = N / R, MAXB = cdiv(nb, T);
  const int tid = threadIdx.x;
  float2 v[MAXB][R];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = load16(gsrc + j + r * nb);
      Dft<R, DIR, float2>::run(v[b]);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int idx = j * R + r;
        lds[idx + (idx >> ps)] = v[b][r];
      }
    }
  }
  __syncthreads();
}

template <typename H, int DIR, int N, int R0, int P, int W, int G>
__global__ __launch_bounds__(64 * W) void k_mfma16(const cplx16<H>* in, cplx16<H>* out, long long batch,
                                                   const float2* __restrict__ tw, float scale) {
  constexpr int PS = 4;
  __shared__ float2 lds_raw[lds_len(N, PS)];
  float2* lds = as_lds(lds_raw);
  const long long t = blockIdx.x;
  if (t >= batch) return;  // one transform per workgroup; the grid is exactly batch
  const cplx16<H>* x = in + t * N;
  cplx16<H>* y = out + t * N;
  const Mfma16A A = mfma16_consts<DIR>();
  int Ls = 1;
  if constexpr (R0 > 1) {
    first_pass_valu<H, N, 64 * W, R0, DIR>(x, lds, PS);
    Ls = R0;
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const bool first = (R0 == 1 && p == 0), last = (p == P - 1);
    if (first && last) mfma_pass16<H, N, W, G, DIR, true, true>(x, y, lds, Ls, tw, A, scale, PS);
    else if (first) mfma_pass16<H, N, W, G, DIR, true, false>(x, y, lds, Ls, tw, A, scale, PS);
    else if (last) mfma_pass16<H, N, W, G, DIR, false, true>(x, y, lds, Ls, tw, A, scale, PS);
    else mfma_pass16<H, N, W, G, DIR, false, false>(x, y, lds, Ls, tw, A, scale, PS);
    Ls *= 16;
  }
}

struct MfmaEntry {
  int n, r0, passes, waves, groups;
  hipError_t (*launch)(bool bf16, int dir, const void*, void*, size_t, const void*, double, hipStream_t);
};

template <int N, int R0, int P, int W, int G>
static hipError_t launch_m(bool bf16, int dir, const void* in, void* out, size_t batch, const void* tw,
                           double scale, hipStream_t s) {
  const dim3 grid((unsigned)batch), block(64 * W);
  const float sc = (float)scale;
  if (bf16) {
    using H = __bf16;
    if (dir < 0)
      hipLaunchKernelGGL((k_mfma16<H, -1, N, R0, P, W, G>), grid, block, 0, s, (const cplx16<H>*)in,
                         (cplx16<H>*)out, (long long)batch, (const float2*)tw, sc);
    else
      hipLaunchKernelGGL((k_mfma16<H, +1, N, R0, P, W, G>), grid, block, 0, s, (const cplx16<H>*)in,
                         (cplx16<H>*)out, (long long)batch, (const float2*)tw, sc);
  } else {
    using H = _Float16;
    if (dir < 0)
      hipLaunchKernelGGL((k_mfma16<H, -1, N, R0, P, W, G>), grid, block, 0, s, (const cplx16<H>*)in,
                         (cplx16<H>*)out, (long long)batch, (const float2*)tw, sc);
    else
      hipLaunchKernelGGL((k_mfma16<H, +1, N, R0, P, W, G>), grid, block, 0, s, (const cplx16<H>*)in,
                         (cplx16<H>*)out, (long long)batch, (const float2*)tw, sc);
  }
  return hipGetLastError();
}

// N = R0 * 16^P; W waves x G groups x 16 = N/16 butterflies per MFMA pass.
static const MfmaEntry kMfma[] = {
    {256, 1, 2, 1, 1, &launch_m<256, 1, 2, 1, 1>},
    {512, 2, 2, 2, 1, &launch_m<512, 2, 2, 2, 1>},
    {1024, 4, 2, 4, 1, &launch_m<1024, 4, 2, 4, 1>},
    {2048, 8, 2, 4, 2, &launch_m<2048, 8, 2, 4, 2>},
    {4096, 1, 3, 4, 4, &launch_m<4096, 1, 3, 4, 4>},
    {8192, 2, 3, 8, 4, &launch_m<8192, 2, 3, 8, 4>},
    {16384, 4, 3, 16, 4, &launch_m<16384, 4, 3, 16, 4>},
};

bool mfma16_supported(size_t n, int* r0) {
  for (const MfmaEntry& e : kMfma)
    if ((size_t)e.n == n) {
      if (r0) *r0 = e.r0;
      return true;
    }
  return false;
}

hipError_t launch_mfma16(bool bf16, int dir, size_t n, const void* in, void* out, size_t batch,
                         const void* tw, double scale, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  for (const MfmaEntry& e : kMfma)
    if ((size_t)e.n == n) return e.launch(bf16, dir, in, out, batch, tw, scale, s);
  return hipErrorInvalidValue;
}

}  // namespace fyfft
