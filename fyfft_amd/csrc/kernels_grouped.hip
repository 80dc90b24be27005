// Grouped ragged transforms: one persistent launch per phase works through a
// device-side task table covering transforms of many different lengths.
//
// Phase 0: every LDS-resident transform (smooth, Rader, Bluestein; up to
//          `per_block` transforms of one length per task) and the first
//          column pass of every four-step transform.
// Phase 1: the row pass of every four-step transform.
// Phase 2: the closing column pass of four-step Rader/Bluestein transforms.
// The plan bin-packs the tasks onto the eight XCDs, keeping all tasks of one
// transform on one XCD (so its cache lines are shared in one L2), each XCD's
// list sorted by decreasing cost.  Workgroup b runs on XCD b mod 8 and takes
// every (gridDim/8)-th task of that XCD's list, so the long tasks start
// first.  The task
// bodies are the same device code as the batched kernels (bodies.hpp); each
// runs out of line so that the register budget of one body does not add to
// the others'.
#include "bodies.hpp"

namespace fyfft {
using namespace dev;

template <typename V>
struct TaskCtx {
  const long long* off;
  const long long* woff;
  const V* in;
  V* out;
  V* ws;
  V* x0buf;
  V* lds;
  int tables;     // 0 forward tables, 1 inverse/adjoint tables
  int normalize;  // 1: scale by 1/n (normalised inverse)
};

template <typename V, int DIR, int PPT>
__device__ __noinline__ void task_lds(const GLen<V>& g, const GTask& t, const TaskCtx<V>& c) {
  LdsBody<V> b = make_uniform(g.lds);
  if (c.tables) {
    b.chirp = g.chirp_inv;
    b.kern = g.kern_inv;
  }
  b.scale = c.normalize ? (S<V>)(1.0 / (double)g.n) : (S<V>)1;
  const int slot = threadIdx.x / b.threads, tid = threadIdx.x - slot * b.threads;
  const bool active = slot < t.count;
  const long long o = active ? c.off[t.first + slot] : 0;
  V* lds = c.lds + (active ? slot : 0) * b.stride;
  if (g.kind == kGkRader) body_rader<V, GenericIo<V, PPT>>(b, c.in + o, c.out + o, lds, tid, active);
  else if (g.kind == kGkBlue) body_blue<V, GenericIo<V, PPT>>(b, c.in + o, c.out + o, lds, tid, active);
  else body_lds_smooth<V, PPT, DIR>(b, c.in + o, c.out + o, lds, tid, active);
}

template <typename V, int DIR, int PPT, int MODE>
__device__ __noinline__ void task_fs(const GLen<V>& g, const GTask& t, const TaskCtx<V>& c) {
  constexpr bool col = MODE != kFsPlainB && MODE != kFsConvB;
  FsArgs<V> a = make_uniform(col ? g.col : g.row);
  if (c.tables) {
    a.chirp = g.chirp_inv;
    a.kern = g.kern_inv;
  }
  if constexpr (MODE == kFsRaderC) a.perm = g.lds.perm_out;
  a.scale = c.normalize ? (S<V>)(1.0 / (double)g.n) : (S<V>)1;
  const long long o = c.off[t.first], w = c.woff[t.first];
  V* x0 = c.x0buf + t.first;
  if constexpr (col)
    body_fs_col<V, GenericFft<V, PPT>, DIR, MODE>(a, c.in + o, c.ws + w, c.out + o, x0, t.count, c.lds);
  else
    body_fs_row<V, GenericFft<V, PPT>, DIR, MODE>(a, c.ws + w, c.out + o, g.kind == kGkFsRader ? x0 : nullptr,
                                   t.count, c.lds);
}

template <typename V, int DIR, int PPT>
__device__ __forceinline__ void run_task(const GLen<V>& g, const GTask& t, const TaskCtx<V>& c) {
  switch (t.phase) {
    case 0:
      if (g.kind <= kGkBlue) task_lds<V, DIR, PPT>(g, t, c);
      else if (g.kind == kGkFsPlain) task_fs<V, DIR, PPT, kFsPlainA>(g, t, c);
      else if (g.kind == kGkFsRader) task_fs<V, -1, PPT, kFsRaderA>(g, t, c);
      else task_fs<V, -1, PPT, kFsBlueA>(g, t, c);
      break;
    case 1:
      if (g.kind == kGkFsPlain) task_fs<V, DIR, PPT, kFsPlainB>(g, t, c);
      else task_fs<V, -1, PPT, kFsConvB>(g, t, c);
      break;
    default:
      if (g.kind == kGkFsRader) task_fs<V, +1, PPT, kFsRaderC>(g, t, c);
      else task_fs<V, +1, PPT, kFsBlueC>(g, t, c);
      break;
  }
}

// B threads per workgroup; at least four waves per SIMD so that several
// workgroups share a CU and one task's memory phase overlaps another's
// arithmetic.
template <typename V, int DIR, int PLO, int PHI, int B>
__global__ __launch_bounds__(B, 4) void k_grouped(GroupedArgs<V> ga) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  TaskCtx<V> c{ga.off, ga.woff, ga.in, ga.out, ga.ws, ga.x0buf,
                    reinterpret_cast<V*>(smem_raw), ga.tables, ga.normalize};
  const int x = blockIdx.x % kXcds, step = gridDim.x / kXcds;
  for (int ti = ga.xbase[x] + blockIdx.x / kXcds; ti < ga.xbase[x + 1]; ti += step) {
    const GTask t = ga.tasks[ti];
    const GLen<V>& g = ga.lens[t.len];
    __syncthreads();  // the previous task's LDS reads are complete
    const int ppt = t.phase == 0 && g.kind <= kGkBlue ? g.lds_ppt
                    : (t.phase == 1 ? g.row_ppt : g.col_ppt);
    if (ppt <= PLO) run_task<V, DIR, PLO>(g, t, c);
    else run_task<V, DIR, PHI>(g, t, c);
  }
}

template <typename V, int PLO, int PHI, int B>
static hipError_t set_grouped_lds(size_t lds) {
  if (lds <= 65536) return hipSuccess;
  hipError_t e = hipFuncSetAttribute((const void*)k_grouped<V, -1, PLO, PHI, B>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_grouped<V, +1, PLO, PHI, B>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return e;
}

template <typename V, int PLO, int PHI>
static hipError_t go_grouped(int dir, const GroupedArgs<V>& a, unsigned grid, int block, size_t lds,
                             hipStream_t s) {
  // the LDS attribute is set once at plan time (prepare_grouped), so this
  // path is free of non-stream calls and can be captured into a graph
  auto pick = [&](auto k_fwd, auto k_inv) { return dir < 0 ? k_fwd : k_inv; };
  switch (block) {
    case 256:
      hipLaunchKernelGGL(pick(k_grouped<V, -1, PLO, PHI, 256>, k_grouped<V, +1, PLO, PHI, 256>),
                         dim3(grid), dim3(256), lds, s, a);
      break;
    case 512:
      hipLaunchKernelGGL(pick(k_grouped<V, -1, PLO, PHI, 512>, k_grouped<V, +1, PLO, PHI, 512>),
                         dim3(grid), dim3(512), lds, s, a);
      break;
    case 1024:
      hipLaunchKernelGGL(pick(k_grouped<V, -1, PLO, PHI, 1024>, k_grouped<V, +1, PLO, PHI, 1024>),
                         dim3(grid), dim3(1024), lds, s, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t prepare_grouped(int prec, const size_t lds[kGroupClasses]) {
  hipError_t e;
  if (prec == kF64) {
    e = set_grouped_lds<double2, 8, 16, 256>(lds[0]);
    if (e == hipSuccess) e = set_grouped_lds<double2, 8, 16, 512>(lds[1]);
    if (e == hipSuccess) e = set_grouped_lds<double2, 8, 16, 1024>(lds[2]);
  } else {
    e = set_grouped_lds<float2, 16, 32, 256>(lds[0]);
    if (e == hipSuccess) e = set_grouped_lds<float2, 16, 32, 512>(lds[1]);
    if (e == hipSuccess) e = set_grouped_lds<float2, 16, 32, 1024>(lds[2]);
  }
  return e;
}

hipError_t launch_grouped(int prec, int dir, const void* args, unsigned grid, int block, size_t lds,
                          hipStream_t s) {
  if (prec == kF64)
    return go_grouped<double2, 8, 16>(dir, *static_cast<const GroupedArgs<double2>*>(args), grid, block, lds, s);
  return go_grouped<float2, 16, 32>(dir, *static_cast<const GroupedArgs<float2>*>(args), grid, block, lds, s);
}

}  // namespace fyfft
