// Split-exchange LDS transforms (compile-time passes), shared by the
// C2C kernels and the 16-bit storage kernels.
#pragma once
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

// Lengths whose transform fills most of a CU's LDS (one workgroup per CU,
// so nothing overlaps a transform's loads with another's passes) keep the
// transform in registers between passes and use LDS only as an exchange
// buffer of N scalars: each exchange writes the real parts of a pass's
// outputs, the next pass reads the real parts of its inputs, then the same
// for the imaginary parts.  Four barriers per exchange instead of two, half
// the LDS, so twice the workgroups share a CU.

// Butterfly j = tid + b*T of a radix-R pass reads points i = j + r*(N/R);
// f(b, r, i, l) also gets the point's padded LDS index l (pad shift PS),
// formed as one base plus a compile-time step where the strides allow it
// (then the LDS instructions take the steps as immediate offsets and a
// thread keeps one address register per butterfly, not one per point).
template <int N, int T, int R, int PS, typename F>
__device__ __forceinline__ void split_inputs(int tid, F f) {
  constexpr int nb = N / R, MAXB = cdiv(nb, T);
  constexpr bool lin = PS >= 30 || nb % (1 << PS) == 0;
  constexpr int step = PS >= 30 ? nb : nb + (nb >> PS);
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (nb % T == 0 || j < nb) {
      const int pj = PS >= 30 ? j : j + (j >> PS);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = j + r * nb;
        f(b, r, i, lin ? pj + r * step : i + (i >> PS));
      }
    }
  }
}
// ... and writes point r to i = q*Ls*R + k + r*Ls (k = j mod Ls, q = j / Ls).
template <int N, int T, int Ls, int R, int PS, typename F>
__device__ __forceinline__ void split_outputs(int tid, F f) {
  constexpr int nb = N / R, MAXB = cdiv(nb, T);
  constexpr bool wide = PS < 30 && Ls % (1 << PS) == 0;        // steps of Ls + Ls / 2^PS
  constexpr bool inside = PS < 30 && (1 << PS) % (Ls * R) == 0;  // the butterfly within one pad block
  constexpr bool lin = PS >= 30 || wide || inside;
  constexpr int step = wide ? Ls + (Ls >> PS) : Ls;
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (nb % T == 0 || j < nb) {
      const int q = j / Ls, k = j - q * Ls, base = q * Ls * R + k;
      const int pb = PS >= 30 ? base : base + (base >> PS);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = base + r * Ls;
        f(b, r, i, lin ? pb + r * step : i + (i >> PS));
      }
    }
  }
}
// Twiddles and butterflies of a radix-R pass at stride Ls on registers.
template <typename V, int DIR, int N, int T, int Ls, int R>
__device__ __forceinline__ void split_compute(V (&v)[cdiv(N / R, T)][R], int tid, const V* __restrict__ tw) {
  constexpr int nb = N / R, MAXB = cdiv(nb, T);
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int j = tid + b * T;
    if (nb % T == 0 || j < nb) {
      if constexpr (Ls > 1) apply_twiddles<R, DIR>(v[b], tw + (Ls - 1) + (j % Ls) * (R - 1));
      Dft<R, DIR, V>::run(v[b]);
    }
  }
}

// Exchange the outputs `prev` of the pass (RP, Ls / RP) into the inputs of
// the pass (R, Ls), run it, and go on; the last pass hands its outputs to
// the sink dst(i, v).
template <typename V, typename Dst, int DIR, int N, int T, int PS, int Ls, int RP, int R, int... Rest>
__device__ __forceinline__ void run_split(V (&prev)[cdiv(N / RP, T)][RP], S<V>* lds, const Dst& dst, int tid,
                                          const V* __restrict__ tw) {
  constexpr int LP = Ls / RP;
  V cur[cdiv(N / R, T)][R];
  __syncthreads();  // the previous exchange's reads are complete
  split_outputs<N, T, LP, RP, PS>(tid, [&](int b, int r, int, int l) { lds[l] = prev[b][r].x; });
  __syncthreads();
  split_inputs<N, T, R, PS>(tid, [&](int b, int r, int, int l) { cur[b][r].x = lds[l]; });
  __syncthreads();
  split_outputs<N, T, LP, RP, PS>(tid, [&](int b, int r, int, int l) { lds[l] = prev[b][r].y; });
  __syncthreads();
  split_inputs<N, T, R, PS>(tid, [&](int b, int r, int, int l) { cur[b][r].y = lds[l]; });
  split_compute<V, DIR, N, T, Ls, R>(cur, tid, tw);
  if constexpr (sizeof...(Rest) == 0)
    split_outputs<N, T, Ls, R, PS>(tid, [&](int b, int r, int i, int) { dst(i, cur[b][r]); });
  else
    run_split<V, Dst, DIR, N, T, PS, Ls * R, R, Rest...>(cur, lds, dst, tid, tw);
}

// A whole transform: the first pass reads its points from the source (a
// split source, kSplit, issues every load of the butterfly before the first
// use), the last pass writes through the sink.  Every thread of the
// workgroup calls it (barriers).
template <typename V, int DIR, int N, int T, int PS, typename Src, typename Dst, int R0, int... Rs>
__device__ __forceinline__ void split_fft_io(const Src& src, const Dst& dst, S<V>* lds, int tid,
                                             const V* __restrict__ tw) {
  constexpr int nb = N / R0;
  V v[cdiv(nb, T)][R0];
  if constexpr (SplitSrc<Src>::value) {
    typename Src::Raw raw[cdiv(nb, T)][R0];
    split_inputs<N, T, R0, PS>(tid, [&](int b, int r, int i, int) { raw[b][r] = src.load(i - r * nb, r * nb); });
    split_inputs<N, T, R0, PS>(tid, [&](int b, int r, int i, int) {
      v[b][r] = src.finish(raw[b][r], i - r * nb, r * nb);
    });
  } else {
    split_inputs<N, T, R0, PS>(tid, [&](int b, int r, int i, int) { v[b][r] = src(i); });
  }
  split_compute<V, DIR, N, T, 1, R0>(v, tid, tw);
  if constexpr (sizeof...(Rs) == 0)
    split_outputs<N, T, 1, R0, PS>(tid, [&](int b, int r, int i, int) { dst(i, v[b][r]); });
  else
    run_split<V, Dst, DIR, N, T, PS, R0, R0, Rs...>(v, lds, dst, tid, tw);
}

// One transform per workgroup (the grid is the batch), W workgroups per CU
// (the second launch bound is the waves per SIMD that W workgroups make).
// E: the element storage in HBM (V, or 16-bit pairs widened to V = float2 on
// load and rounded once on the store).
template <typename V, typename E, int DIR, int N, int T, int PS, int W, int R0, int... Rs>
__global__ __launch_bounds__(T, cdiv(W * cdiv(T, 64), 4)) void k_lds_fixed_split(const E* in, E* out, long long dist,
                                                          const V* __restrict__ tw, S<V> scale) {
  __shared__ S<V> lds[lds_len(N, PS)];
  const long long off = (long long)blockIdx.x * dist;
  split_fft_io<V, DIR, N, T, PS, GlobalSrc<V, E>, GlobalSink<V, E>, R0, Rs...>(
      GlobalSrc<V, E>{(gbl_cptr<E>)(in + off)}, GlobalSink<V, E>{(gbl_ptr<E>)(out + off), scale}, lds,
      (int)threadIdx.x, tw);
}

}  // namespace fyfft
