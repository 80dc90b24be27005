// Grouped ragged transforms: per task family, one persistent kernel works
// through that family's part of a device-side task table covering
// transforms of many different lengths.
//
// Families (GroupFamily): LDS-resident smooth, Rader and Bluestein
// transforms (up to `per_block` transforms of one length per task), and the
// four-step column pass, row pass and closing column pass of the plain,
// Rader and Bluestein four-step transforms.  Phase 0 holds the LDS families
// and the first column passes, phase 1 the row passes, phase 2 the closing
// column passes; the families of one phase are independent and run as
// parallel branches of the plan's graph.  Each family kernel inlines exactly
// one task body (the same device code as the batched kernels, bodies.hpp),
// so its registers are allocated for that body alone: no call frames, no
// scratch.  The plan bin-packs each family list onto the eight XCDs, keeping
// all tasks of one transform on one XCD, each XCD's list sorted by
// decreasing cost; workgroup b runs on XCD b mod 8 and takes every
// (gridDim/8)-th task of that XCD's list, so the long tasks start first.
//
// Element storage E is the compute type V (FP32, FP64), or 16-bit pairs
// widened to FP32 on load and rounded once on the final store (the four-step
// workspace stays FP32).
#pragma once
#include <type_traits>

#include "bodies.hpp"

namespace fyfft {
using namespace dev;

template <typename V>
struct TaskCtx {
  const long long* off;
  const long long* woff;
  V* ws;
  V* x0buf;
  V* lds;
  int tables;     // 0 forward tables, 1 inverse/adjoint tables
  int normalize;  // 1: scale by 1/n (normalised inverse)
};

template <int FAM> constexpr int family_mode() {
  return FAM == kGfColPlain ? kFsPlainA : FAM == kGfColRaderA ? kFsRaderA : FAM == kGfColBlueA ? kFsBlueA
       : FAM == kGfRowPlain ? kFsPlainB : FAM == kGfRowConv ? kFsConvB : FAM == kGfColRaderC ? kFsRaderC
       : kFsBlueC;
}

template <typename V, typename E, int FAM, int DIR, int PPT, bool TL>
__device__ __forceinline__ void run_family(const GLen<V>& gl, const GTask& t, const TaskCtx<V>& c, const E* in,
                                           E* out) {
  if constexpr (FAM == kGfLds || FAM == kGfRader || FAM == kGfBlue) {
    LdsBody<V> b = make_uniform(gl.lds);
    if (c.tables) {
      b.chirp = gl.chirp_inv;
      b.kern = gl.kern_inv;
    }
    b.scale = c.normalize ? (S<V>)(1.0 / (double)gl.n) : (S<V>)1;
    const int slot = threadIdx.x / b.threads, tid = threadIdx.x - slot * b.threads;
    const bool active = slot < t.count;
    const long long o = active ? c.off[t.first + slot] : 0;
    V* lds = c.lds + (active ? slot : 0) * b.stride;
    if constexpr (FAM == kGfRader) body_rader<V, GenericIo<V, PPT>>(b, in + o, out + o, lds, tid, active);
    else if constexpr (FAM == kGfBlue) body_blue<V, GenericIo<V, PPT>>(b, in + o, out + o, lds, tid, active);
    else body_lds_smooth<V, PPT, DIR>(b, in + o, out + o, lds, tid, active);
  } else {
    constexpr int MODE = family_mode<FAM>();
    constexpr bool col = FAM != kGfRowPlain && FAM != kGfRowConv;
    FsArgs<V> a = make_uniform(col ? gl.col : gl.row);
    if (c.tables) {
      a.chirp = gl.chirp_inv;
      a.kern = gl.kern_inv;
    }
    if constexpr (MODE == kFsRaderC) a.perm = gl.lds.perm_out;
    a.scale = c.normalize ? (S<V>)(1.0 / (double)gl.n) : (S<V>)1;
    const long long o = c.off[t.first], w = c.woff[t.first];
    V* x0 = c.x0buf + t.first;
    if constexpr (col)
      body_fs_col<V, GenericFft<V, PPT, TL, kFsComposite<V, PPT> && MODE == kFsPlainA>, DIR, MODE>(a, in + o, c.ws + w, out + o, x0, t.count, c.lds);
    else
      body_fs_row<V, GenericFft<V, PPT, TL, kFsComposite<V, PPT> && MODE == kFsPlainB>, DIR, MODE>(a, c.ws + w, out + o,
                                                    gl.kind == kGkFsRader ? x0 : nullptr, t.count, c.lds);
  }
}

// B threads per workgroup; at least four waves per SIMD so that several
// workgroups share a CU and one task's memory phase overlaps another's
// arithmetic.
template <typename V, typename E, int FAM, int DIR, int PPT, int B, bool TL>
__global__ __launch_bounds__(B, 4) void k_group_family(GroupedArgs<V> ga) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const TaskCtx<V> c{ga.off, ga.woff, ga.ws, ga.x0buf, reinterpret_cast<V*>(smem_raw), ga.tables, ga.normalize};
  const E* in = static_cast<const E*>(ga.in);
  E* out = static_cast<E*>(ga.out);
  const int x = blockIdx.x % kXcds, step = gridDim.x / kXcds;
  for (int ti = ga.xbase[x] + blockIdx.x / kXcds; ti < ga.xbase[x + 1]; ti += step) {
    const GTask t = ga.tasks[ti];
    __syncthreads();  // the previous task's LDS reads are complete
    run_family<V, E, FAM, DIR, PPT, TL>(ga.lens[t.len], t, c, in, out);
  }
}

// Launch (or, with prepare, set the dynamic-LDS limit of) one family kernel.
// The points-per-thread bound: lo (16 FP32 / 8 FP64) or, for the LDS
// families at 1024 threads, hi (twice that).  TL: the four-step families
// keep the sub-length's Stockham table in LDS after the tile.
template <typename V, typename E, int FAM, int DIR, int PPT, int B, bool TL>
static hipError_t group_family_op(bool prepare, const GroupedArgs<V>* a, unsigned grid, size_t lds, hipStream_t s) {
  auto k = k_group_family<V, E, FAM, DIR, PPT, B, TL>;
  if (prepare)
    return lds > 65536 ? hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                       : hipSuccess;
  hipLaunchKernelGGL(k, dim3(grid), dim3(B), lds, s, *a);
  return hipGetLastError();
}

template <typename V, typename E, int FAM, int DIR, int PPT, bool TL = false>
static hipError_t group_family_block(bool prepare, int block, const GroupedArgs<V>* a, unsigned grid, size_t lds,
                                     hipStream_t s) {
  constexpr int LO = sizeof(V) == 8 ? 16 : 8;
  if constexpr (PPT == LO) {
    if (block == 256) return group_family_op<V, E, FAM, DIR, PPT, 256, TL>(prepare, a, grid, lds, s);
    if (block == 512) return group_family_op<V, E, FAM, DIR, PPT, 512, TL>(prepare, a, grid, lds, s);
  }
  if (block == 1024) return group_family_op<V, E, FAM, DIR, PPT, 1024, TL>(prepare, a, grid, lds, s);
  return hipErrorInvalidValue;
}

// Four-step families: `global_table` (the table does not fit beside the
// tile) selects the form that reads the Stockham table from global memory.
template <typename V, typename E, int FAM, int DIR, int PPT>
static hipError_t group_fs_block(bool prepare, bool global_table, int block, const GroupedArgs<V>* a,
                                 unsigned grid, size_t lds, hipStream_t s) {
  return global_table ? group_family_block<V, E, FAM, DIR, PPT, false>(prepare, block, a, grid, lds, s)
                      : group_family_block<V, E, FAM, DIR, PPT, true>(prepare, block, a, grid, lds, s);
}

template <typename V, typename E>
static hipError_t group_family_dispatch(bool prepare, int fam, int dir, bool hi, int block, const void* args,
                                        unsigned grid, size_t lds, hipStream_t s) {
  constexpr int LO = sizeof(V) == 8 ? 16 : 8, HI = 2 * LO;
  const auto* a = static_cast<const GroupedArgs<V>*>(args);
  const bool fwd = dir < 0;
  switch (fam) {
    case kGfLds:
      if (hi) return fwd ? group_family_block<V, E, kGfLds, -1, HI>(prepare, block, a, grid, lds, s)
                         : group_family_block<V, E, kGfLds, +1, HI>(prepare, block, a, grid, lds, s);
      return fwd ? group_family_block<V, E, kGfLds, -1, LO>(prepare, block, a, grid, lds, s)
                 : group_family_block<V, E, kGfLds, +1, LO>(prepare, block, a, grid, lds, s);
    case kGfRader:
      return hi ? group_family_block<V, E, kGfRader, -1, HI>(prepare, block, a, grid, lds, s)
                : group_family_block<V, E, kGfRader, -1, LO>(prepare, block, a, grid, lds, s);
    case kGfBlue:
      return hi ? group_family_block<V, E, kGfBlue, -1, HI>(prepare, block, a, grid, lds, s)
                : group_family_block<V, E, kGfBlue, -1, LO>(prepare, block, a, grid, lds, s);
    default: break;
  }
  // four-step families: hi selects the global-table form
  switch (fam) {
    case kGfColPlain:
      return fwd ? group_fs_block<V, E, kGfColPlain, -1, LO>(prepare, hi, block, a, grid, lds, s)
                 : group_fs_block<V, E, kGfColPlain, +1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfRowPlain:
      return fwd ? group_fs_block<V, E, kGfRowPlain, -1, LO>(prepare, hi, block, a, grid, lds, s)
                 : group_fs_block<V, E, kGfRowPlain, +1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfColRaderA: return group_fs_block<V, E, kGfColRaderA, -1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfColBlueA: return group_fs_block<V, E, kGfColBlueA, -1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfRowConv: return group_fs_block<V, E, kGfRowConv, -1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfColRaderC: return group_fs_block<V, E, kGfColRaderC, +1, LO>(prepare, hi, block, a, grid, lds, s);
    case kGfColBlueC: return group_fs_block<V, E, kGfColBlueC, +1, LO>(prepare, hi, block, a, grid, lds, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fyfft
