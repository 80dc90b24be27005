// Four-step passes for lengths beyond one workgroup's LDS, and the
// convolution pipelines (Rader, Bluestein) built on them.
//
// Length N = N1 * N2, input index n = n1*N2 + n2, output index
// k = k1 + N1*k2:
//   X[k1 + N1*k2] = sum_n2 w_N2^(n2*k2) * [ w_N^(n2*k1) * sum_n1 x[n1*N2 + n2] w_N1^(n1*k1) ]
// Column pass: C consecutive columns n2 per workgroup (coalesced rows of C
//   elements), length-N1 FFTs in LDS, twiddle w_N^(n2*k1) on the way out,
//   tmp[k1*N2 + n2].
// Row pass: C consecutive rows k1 (one contiguous span), length-N2 FFTs in
//   LDS, then either the transposed store X[k1 + N1*k2] (plain transform) or,
//   for a convolution, pointwise multiply by the kernel spectrum (stored in
//   the same [k1][k2] order), the inverse length-N2 FFT and the conjugate
//   twiddle w_N^(-n2*k1), back into tmp -- the first half of the inverse
//   four-step, so the spectrum never needs reordering.
// The inverse column pass then finishes the convolution and applies the
// Rader scatter or the Bluestein post-chirp.
//
// Twiddles w_N^m (m < N) come from two FP64-generated tables,
// w_N^m = hi[m >> lo_bits] * lo[m & (2^lo_bits - 1)], each about sqrt(N)
// entries, so they stay cache resident for any N.
#include "fourstep.hpp"

namespace fyfft {
using namespace dev;

// Compile-time sub-transforms (radices as planner.cpp radices_for, threads =
// ceil(n / points-per-thread) as the four-step geometry takes them): every
// mode for the power-of-two lengths; the plain modes only for the odd-radix
// factors of the reference table's four-step lengths (59049 = 243 x 243,
// 78125 = 125 x 625, FP64 15625 / 16807 / 19683), which live in
// kernels_fourstep_odd.hip.
struct FsFixed {
  int prec, n, threads, npass;
  int radix[4];
  bool plain_only;
};
static const FsFixed kFsFixed[] = {
    {kF32, 64, 4, 2, {8, 8}, false},       {kF32, 128, 8, 2, {16, 8}, false},
    {kF32, 256, 16, 2, {16, 16}, false},   {kF32, 512, 32, 3, {8, 8, 8}, false},
    {kF32, 1024, 64, 3, {16, 8, 8}, false}, {kF32, 2048, 128, 3, {16, 16, 8}, false},
    {kF32, 4096, 256, 3, {16, 16, 16}, false}, {kF32, 8192, 512, 4, {16, 8, 8, 8}, false},
    {kF32, 16384, 1024, 4, {16, 16, 8, 8}, false},
    {kF64, 64, 8, 2, {8, 8}, false},       {kF64, 128, 16, 2, {16, 8}, false},
    {kF64, 256, 32, 2, {16, 16}, false},   {kF64, 512, 64, 3, {8, 8, 8}, false},
    {kF64, 1024, 128, 3, {16, 8, 8}, false}, {kF64, 2048, 256, 3, {16, 16, 8}, false},
    {kF64, 4096, 512, 3, {16, 16, 16}, false}, {kF64, 8192, 1024, 4, {16, 8, 8, 8}, false},
    {kF32, 49, 4, 2, {7, 7}, true},        {kF32, 81, 6, 2, {9, 9}, true},
    {kF32, 125, 8, 3, {5, 5, 5}, true},    {kF32, 243, 16, 3, {9, 9, 3}, true},
    {kF32, 343, 22, 3, {7, 7, 7}, true},   {kF32, 625, 40, 4, {5, 5, 5, 5}, true},
    {kF64, 49, 7, 2, {7, 7}, true},        {kF64, 81, 11, 2, {9, 9}, true},
    {kF64, 125, 16, 3, {5, 5, 5}, true},   {kF64, 243, 31, 3, {9, 9, 3}, true},
    {kF64, 343, 43, 3, {7, 7, 7}, true},   {kF64, 625, 79, 4, {5, 5, 5, 5}, true},
};

int find_fs_fixed(int prec, const dev::PassDesc& pd, int n, int threads, bool all_modes) {
  for (int i = 0; i < (int)(sizeof(kFsFixed) / sizeof(kFsFixed[0])); ++i) {
    const FsFixed& e = kFsFixed[i];
    if (e.prec != prec || e.n != n || e.threads != threads || e.npass != pd.npass || (all_modes && e.plain_only))
      continue;
    bool same = true;
    for (int p = 0; p < pd.npass; ++p) same = same && e.radix[p] == pd.radix[p];
    if (same) return e.n;
  }
  return 0;
}

// The runtime-pass sub-transform: its Stockham table copied into LDS after
// the tile (tw_bytes > 0) or read from global memory.
template <typename V, int PPT>
static hipError_t dispatch_generic(bool col, int mode, int dir, const FsArgs<V>& a, size_t batch, size_t lds,
                                   size_t tw_bytes, int threads, hipStream_t s) {
  // composite radices: the plain modes of FP32 plans (their pass lists may
  // hold them); the convolution modes keep the base radices, whose bodies
  // measured slower with the composite passes' register budget
  constexpr bool CMP = kFsComposite<V, PPT>;
  const bool plain = mode == (col ? kFsPlainA : kFsPlainB);
  if (CMP && plain)
    return tw_bytes ? dispatch_fs_plain<V, GenericFft<V, PPT, true, CMP>>(col, dir, a, batch, lds + tw_bytes,
                                                                           threads, s)
                    : dispatch_fs_plain<V, GenericFft<V, PPT, false, CMP>>(col, dir, a, batch, lds, threads, s);
  if (tw_bytes)
    return dispatch_fs<V, GenericFft<V, PPT, true>>(col, mode, dir, a, batch, lds + tw_bytes, threads, s);
  return dispatch_fs<V, GenericFft<V, PPT>>(col, mode, dir, a, batch, lds, threads, s);
}

template <typename V, int PPT>
static hipError_t dispatch_len(int fixed_n, bool col, int mode, int dir, const FsArgs<V>& a,
                               size_t batch, size_t lds, size_t tw_bytes, int threads, hipStream_t s) {
  constexpr int P = sizeof(V) == 8 ? 16 : 8;  // points per thread of the fixed forms
  if (fixed_n & 1) {  // odd-radix forms: plain modes only
    if (mode == (col ? kFsPlainA : kFsPlainB))
      return launch_fs_odd(sizeof(V) == 16 ? kF64 : kF32, fixed_n, col, dir, &a, batch, lds, threads, s);
    return dispatch_generic<V, PPT>(col, mode, dir, a, batch, lds, tw_bytes, threads, s);
  }
  switch (fixed_n) {
    case 64: return dispatch_fs<V, FixedFft<V, 64, 64 / P, 8, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 128: return dispatch_fs<V, FixedFft<V, 128, 128 / P, 16, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 256: return dispatch_fs<V, FixedFft<V, 256, 256 / P, 16, 16>>(col, mode, dir, a, batch, lds, threads, s);
    case 512: return dispatch_fs<V, FixedFft<V, 512, 512 / P, 8, 8, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 1024: return dispatch_fs<V, FixedFft<V, 1024, 1024 / P, 16, 8, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 2048: return dispatch_fs<V, FixedFft<V, 2048, 2048 / P, 16, 16, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 4096: return dispatch_fs<V, FixedFft<V, 4096, 4096 / P, 16, 16, 16>>(col, mode, dir, a, batch, lds, threads, s);
    case 8192: return dispatch_fs<V, FixedFft<V, 8192, 8192 / P, 16, 8, 8, 8>>(col, mode, dir, a, batch, lds, threads, s);
    case 16384:
      if constexpr (sizeof(V) == 8)
        return dispatch_fs<V, FixedFft<V, 16384, 16384 / P, 16, 16, 8, 8>>(col, mode, dir, a, batch, lds, threads, s);
      break;
  }
  return dispatch_generic<V, PPT>(col, mode, dir, a, batch, lds, tw_bytes, threads, s);
}

hipError_t launch_fs_pass(int prec, bool col, int mode, int dir, const void* args, size_t batch,
                          const LdsGeom& g, int fixed_n, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const int threads = g.threads * g.per_block;
  // the twiddled column passes and the convolution row passes keep their
  // tables of roots after the tile (and after the Stockham table, which
  // dispatch_generic adds)
  const size_t elem = prec == kF64 ? 16 : 8;
  const size_t roots = col && fs_col_twiddled(mode) ? (size_t)fs_col_tw_row(g.ppt)
                       : !col && mode == kFsConvB ? (size_t)fs_row_tw_len(g.n)
                                                  : 0;
  const size_t lds = g.lds_bytes + (size_t)g.per_block * roots * elem;
  if (prec == kF64) {
    const auto& a = *static_cast<const FsArgs<double2>*>(args);
    return g.ppt <= 8 ? dispatch_len<double2, 8>(fixed_n, col, mode, dir, a, batch, lds, g.tw_bytes, threads, s)
                      : dispatch_len<double2, 16>(fixed_n, col, mode, dir, a, batch, lds, g.tw_bytes, threads, s);
  }
  const auto& a = *static_cast<const FsArgs<float2>*>(args);
  return g.ppt <= 16 ? dispatch_len<float2, 16>(fixed_n, col, mode, dir, a, batch, lds, g.tw_bytes, threads, s)
                     : dispatch_len<float2, 32>(fixed_n, col, mode, dir, a, batch, lds, g.tw_bytes, threads, s);
}

}  // namespace fyfft
