// C ABI (include/fyfft.h): plan creation, execution and queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <tuple>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fyfft.h"
#include "internal.hpp"

using namespace fyfft;

namespace {

constexpr uint32_t kMagic = 0xF1FF7A11u;
thread_local std::string g_last_error;

enum Kind { kTrivial = 0, kLds = 1, kFourStep = 2, kRaderLds = 3, kRaderFs = 4, kBlueLds = 5, kBlueFs = 6, kGrouped = 7, kMfma = 8,
            kReal = 9, kStrided = 10, kLowp = 11 };

// One strided pass: transforms of length n at (plane p, column c), element j
// at p*pdist + c*cdist + j*stride, separately for input and output.
struct Layout2 {
  long long stride = 1, cdist = 0, pdist = 0;
};
enum PassMethod { kPmRows = 0, kPmColumns = 1, kPmGather = 2, kPmColFour = 3 };
struct StridedPass {
  size_t n = 0, ncols = 1, planes = 1;
  Layout2 in, out;
  int method = kPmRows;
  fyfft_plan* sub = nullptr;  // contiguous plan of length n (rows, gather); owned by subs
  LdsGeom g;                  // columns
  dev::PassDesc pd{};
  void* tw = nullptr;
  int fixed_n = 0;
  size_t chunk = 0;           // gather: transforms per workspace chunk; colfour: columns
  // colfour: the column axis split as n = n1 * n2, two column passes
  size_t n1 = 0, n2 = 0;
  LdsGeom g2;
  dev::PassDesc pd2{};
  void *tw2 = nullptr, *tw_hi = nullptr, *tw_lo = nullptr;
  int fixed2 = 0, lo_bits = 0;
};

// One family launch of a grouped plan: the tasks [base, base + count) of the
// device task table, split per XCD at xbase, with its workgroup size class,
// dynamic LDS, grid, estimated cost, time measured at plan time (ms, 0 if
// not measured), and the phase and branch it is scheduled on.
struct GList {
  int fam = 0, cls = 0, hi = 0;
  int base = 0, count = 0;
  int xbase[dev::kXcds + 1] = {};
  size_t lds = 0;
  unsigned grid = 0;
  double cost = 0, ms = 0;
  double bytes = 0;  // HBM bytes the launch moves (data, workspace, kernel spectra)
  int phase = 0, branch = 0;
};
// Parallel branches the family launches of one phase are spread over
// (the box's hardware queues per process).
constexpr int kGroupBranches = 4;

// Forms of a real (R2C / C2R) plan.
enum RealForm { kRealFusedHalf = 0, kRealFusedFull = 1, kRealHalfPasses = 2, kRealFullPasses = 3 };

// Tuning time budget of FYFFT_GROUPED_TUNED plans (seconds of plan time).
constexpr double kGroupedTuneSeconds = 20.0;

// Internal plan form: TUNED restricted to what the runtime-pass kernels run
// (sub-plans of tuned grouped plans, which never launch compile-time kernels).
constexpr int kFormTunedGeneric = 16;
// Internal plan form: PLAIN for the sub-plans of grouped plans, whose family
// kernels run LDS-resident lengths at 32 points per thread better than as
// extra four-step tasks (ragged step 6.20 ms against 6.28).
constexpr int kFormGroupedPlain = 17;

// Workspace of multi-pass plans is capped; larger batches run in chunks.
constexpr size_t kWorkspaceCap = (size_t)1 << 30;
// Longest factor of a four-step split (one column/row FFT in LDS).
constexpr size_t kFactorCap = 4096;
static_assert(dev::kPadShift == kLdsPadShift, "planner and kernels must pad LDS alike");

fyfft_status from_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return FYFFT_SUCCESS;
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return FYFFT_NO_DEVICE;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return FYFFT_ALLOC_FAILED;
  return FYFFT_DEVICE_ERROR;
}

#define HIP_TRY(call)                                             \
  do {                                                            \
    hipError_t e_ = (call);                                       \
    if (e_ != hipSuccess) return from_hip(e_, #call);             \
  } while (0)

}  // namespace

// A four-step structure choice: the split n = n1 x (n / n1) and the number
// of sub-transforms per workgroup of the column and row passes.
struct FsChoice {
  size_t n1 = 0;
  int ccol = 0, crow = 0;
};

struct fyfft_plan {
  uint32_t magic = kMagic;
  size_t n = 0, batch = 0;
  int prec = kF32;
  int form = FYFFT_FORM_AUTO;
  bool grouped_sub = false;  // a sub-plan of a grouped plan (kFormGroupedPlain / kFormTunedGeneric)
  // tuning overrides for the structure being built (autotuner candidates)
  const FsChoice* force_fs = nullptr;
  size_t force_m = 0;
  int strategy = kDirect;
  int device = 0;
  Kind kind = kTrivial;
  LdsLaunch lds;        // kLds, and the conv length of kRaderLds / kBlueLds
  // four-step structure (kFourStep; conv length of kRaderFs / kBlueFs)
  size_t n1 = 0, n2 = 0;
  LdsGeom gcol, grow;
  dev::PassDesc pdcol{}, pdrow{};
  void *twcol = nullptr, *twrow = nullptr, *tw_hi = nullptr, *tw_lo = nullptr;
  int lo_bits = 0;
  int fixed_col = 0, fixed_row = 0;  // compile-time sub-transform lengths (0 = generic)
  // three-pass form (n3 > 0): n = n1 x n2 x n3, a second column pass over
  // the n2 x n3 rows (geometry gmid, twiddles of the row length n2 * n3)
  size_t n3 = 0;
  LdsGeom gmid;
  dev::PassDesc pdmid{};
  void *twmid = nullptr, *tw2_hi = nullptr, *tw2_lo = nullptr;
  int lo_bits2 = 0, fixed_mid = 0;
  size_t chunk = 0;     // transforms per workspace chunk
  void* tmp = nullptr;  // chunk * conv length elements
  void* x0buf = nullptr;
  // convolution plans
  size_t m = 0;                       // conv length (Rader N-1, Bluestein M)
  void* kern[2] = {nullptr, nullptr}; // [0] forward, [1] inverse/adjoint spectra / m
  void* chirp[2] = {nullptr, nullptr};
  int* perm_in = nullptr;
  int* perm_out = nullptr;
  int* pos_out = nullptr;  // LDS-resident Rader: inverse of perm_out (output position -> cyclic index)
  // four-step Rader: bucketed output permutation (chunk size rb_q; tables
  // of L entries) and the chunk x L staging buffer it runs through
  int rb_q = 0;
  int* rb_lout = nullptr;  // LDS index (closing column pass)
  int2* rb_out = nullptr;  // {zw position, offset in the chunk of y}
  void* zw = nullptr;
  // real plans: complex sub-plan subs[0] of length n/2 (half forms) or n
  int real_form = -1;
  // strided plans (many / multi-dimensional): passes run in order, the first
  // from in to out, the others in place on out
  std::vector<StridedPass> passes;
  void* lws = nullptr;   // gather workspace
  size_t in_ext = 0, out_ext = 0;  // elements spanned by the input / output layout
  // half-length tables of the spectral transforms (built on first use)
  mutable std::mutex half_mu;
  mutable int half_state = 0;       // 0 not built, 1 ready, -1 not available
  mutable LdsLaunch half;
  mutable void* half_rtw = nullptr;
  void* rtw = nullptr;   // exp(-2*pi*i*k/n), k <= n/2 (half forms)
  void* rws = nullptr;   // rchunk rows of the sub-plan's length (multi-pass forms)
  size_t rchunk = 0;
  // grouped plans
  unsigned gflags = 0;
  size_t count = 0, total = 0;               // transforms, sum of lengths
  std::vector<fyfft_plan*> subs;             // one plan per distinct length
  std::vector<int> sub_of;                   // transform -> index into subs
  std::vector<size_t> h_off;                 // transform -> packed element offset
  void *d_tasks = nullptr, *d_lens = nullptr, *d_off = nullptr, *d_woff = nullptr;
  void *gws = nullptr, *gx0 = nullptr;
  std::vector<GList> glists;                 // family launches, by phase then decreasing cost
  hipStream_t gbranch[kGroupBranches] = {};  // branch streams of multi-family phases
  hipEvent_t gfork[3] = {}, gjoin[3][kGroupBranches] = {};
  double gspan_ms = 0;  // sum over phases of the longest branch (measured weights; 0: estimates)
  void* gstage = nullptr;  // 16-bit per-transform baseline: FP32 copy of the longest transform
  mutable std::mutex graph_mu;
  mutable std::map<std::tuple<const void*, void*, int, hipStream_t>, hipGraphExec_t> graphs;
  std::vector<void*> allocs;
  size_t dev_bytes = 0;
  std::string desc;
  // staging for fyfft_execute_host (lazily allocated, reused): a device
  // buffer, a pinned host buffer of the same size (null if pinned memory
  // was refused: then pageable copies) and a stream of the plan's own
  mutable std::mutex host_mu;
  mutable void* d_io = nullptr;
  mutable void* h_io = nullptr;
  mutable size_t d_io_bytes = 0;
  mutable hipStream_t io_stream = nullptr;

  fyfft_status alloc(size_t bytes, void** p) {
    HIP_TRY(hipMalloc(p, bytes));
    allocs.push_back(*p);
    dev_bytes += bytes;
    return FYFFT_SUCCESS;
  }
  ~fyfft_plan() {
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (hipStream_t b : gbranch)
      if (b) (void)hipStreamDestroy(b);
    for (int ph = 0; ph < 3; ++ph) {
      if (gfork[ph]) (void)hipEventDestroy(gfork[ph]);
      for (hipEvent_t e : gjoin[ph])
        if (e) (void)hipEventDestroy(e);
    }
    for (fyfft_plan* sp : subs) delete sp;
    for (void* p : allocs) (void)hipFree(p);
    if (d_io) (void)hipFree(d_io);
    if (h_io) (void)hipHostFree(h_io);
    if (io_stream) (void)hipStreamDestroy(io_stream);
    magic = 0;
  }
};

static bool valid(const fyfft_plan* p) { return p && p->magic == kMagic; }

static size_t lds_cap_bytes(int device) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
    return 65536;
  return (size_t)v;
}

static std::string radix_string(const dev::PassDesc& pd) {
  std::string s;
  for (int i = 0; i < pd.npass; ++i) s += (i ? "x" : "") + std::to_string(pd.radix[i]);
  return s;
}

static int ilog2_exact(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// Pass list + FP64-generated Stockham twiddle table for a smooth length.
static fyfft_status make_pass_table(fyfft_plan* p, size_t n, dev::PassDesc* pd, void** tw,
                                    const std::vector<int>* radices = nullptr) {
  const std::vector<int> rad = radices ? *radices : radices_for(n);
  if (rad.empty() || (int)rad.size() > dev::kMaxPackedPasses) return FYFFT_UNSUPPORTED;
  pd->npass = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) pd->radix[i] = rad[i];
  fyfft_status st = p->alloc(std::max<size_t>(n, 2) * elem_bytes(p->prec), tw);
  if (st) return st;
  HIP_TRY(launch_gen_stockham_twiddles(p->prec, *tw, (int)n, *pd, nullptr));
  return FYFFT_SUCCESS;
}

// Plan an LDS-resident smooth transform of length n.
// short_threads, small_radix: lds_geometry's options for batched plans (not
// convolutions; small_radix for the C2C kernel only).
static fyfft_status plan_lds(fyfft_plan* p, size_t n, LdsLaunch* L, bool allow_fixed, bool short_threads,
                             bool small_radix = false) {
  if (!lds_geometry(n, elem_bytes(p->prec), lds_cap_bytes(p->device), &L->geom, short_threads, small_radix))
    return FYFFT_UNSUPPORTED;
  void* tw = nullptr;
  fyfft_status st = make_pass_table(p, n, &L->pd, &tw);
  if (st) return st;
  L->tw = tw;
  L->fixed = allow_fixed ? find_fixed_kernel(p->prec, L->pd, &L->geom) : -1;
  return FYFFT_SUCCESS;
}

// Geometry of a four-step pass whose LDS FFT has length len: up to 16
// sub-transforms (columns or rows) per workgroup, and at least two
// workgroups per CU where the LDS allows.
// LDS bytes of a copy of the Stockham table of a length (n - 1 entries).
static size_t table_lds_bytes(size_t n, size_t elem) { return (std::max<size_t>(n, 2) - 1) * elem; }

// FsArgs::inv_n of a four-step length: 1/n rounded to FP32 where the FP32
// kernels compute their twiddles (n <= 2^24: every remainder of the
// octant reduction is exact in FP32), 0 (table lookups) beyond.
static float fs_inv_n(size_t n) { return n <= ((size_t)1 << 24) ? (float)(1.0 / (double)n) : 0.0f; }

static bool fs_geometry(size_t len, size_t elem, size_t cap, LdsGeom* g, int force_c = 0,
                        bool transposed_rows = false) {
  if (!lds_geometry(len, elem, cap, g)) return false;
  // 16 sub-transforms per workgroup make 128-byte row segments for FP32
  // columns; fewer where 16 would leave one workgroup per CU: measured,
  // N = 2^20 (1024 x 1024) 8 columns at two workgroups per CU 1.26 ms
  // against 16 at one 1.36 ms for 128 transforms, FP64 2^18 1.76 vs 2.06.
  // But never below 8 where 8 fit at all: 32-byte segments cost more than
  // the occupancy (N = 2^21 = 1024 x 2048 x 128: rows of 4 per workgroup
  // 3.92 ms, of 8 at one workgroup per CU 2.18 ms).  The plain row pass
  // (transposed_rows) writes its output in segments of as many rows as a
  // workgroup holds and takes 32 where they fit half the LDS (256-byte
  // segments: N = 65536 x 1024 0.461 -> 0.423 ms).
  auto count = [&](size_t stride) {
    int c = 16;
    while (c > 1 && ((size_t)c * g->threads > 1024 || (size_t)c * stride * elem > cap / 2)) c >>= 1;
    if (c < 8 && (size_t)8 * g->threads <= 1024 && (size_t)8 * stride * elem <= cap) c = 8;
    if (transposed_rows && c == 16 && (size_t)32 * g->threads <= 1024 && (size_t)32 * stride * elem <= cap / 2)
      c = 32;
    return c;
  };
  int c = count((size_t)dev::lds_len((int)len, g->pad_shift));
  // Sub-lengths that are not multiples of 16 have no pass whose strides
  // fall on one bank: their tiles are stored unpadded where that keeps the
  // count per workgroup, which lets the runtime passes address a
  // butterfly's points as one base plus uniform steps (22-30 % fewer VALU
  // instructions per pass).  Measured: N = 75000 (200 x 375) 0.614 -> 0.599
  // ms; multiples of 16 run 5-60 % slower unpadded, and N = 78125 (125 x
  // 625), whose unpadded rows would fit 16 per workgroup at two workgroups
  // per CU instead of 8 at three, 2-4 % slower.
  if (len % 16 && count(len) == c) g->pad_shift = 30;
  const size_t stride = (size_t)dev::lds_len((int)len, g->pad_shift);
  if (force_c > 0) {  // an autotuner candidate: any power of two that fits
    if ((force_c & (force_c - 1)) || (size_t)force_c * g->threads > 1024) return false;
    c = force_c;
  }
  // the twiddled column passes and the convolution row passes add their
  // tables of roots (fs_col_twiddled, fs_row_tw_len)
  const size_t roots = (size_t)c * std::max(fs_col_tw_row(g->ppt), fs_row_tw_len((int)len)) * elem;
  if ((size_t)c * stride * elem + roots > cap) return false;
  g->per_block = c;
  g->lds_bytes = (size_t)c * stride * elem;
  // the runtime-pass sub-transform reads its Stockham table from a copy in
  // LDS after the tile where that fits (compile-time sub-transforms ignore it)
  const size_t tw = table_lds_bytes(len, elem);
  g->tw_bytes = g->lds_bytes + tw + roots <= cap ? tw : 0;
  return true;
}

// Split of a batched four-step plan.  The plain row pass writes its output
// transposed, in segments of as many rows as a workgroup holds, so it needs
// 16 rows per workgroup and the balanced split.  The convolution row pass
// of Rader/Bluestein plans reads and writes its rows in place, so those take
// the LDS-aware split (small column tiles, long rows): N = 1000003 at
// M = 2^21 = 256 x 8192 ran 0.53 ms against 0.58 ms for 1024 x 2048, while
// plain N = 2^21 at 256 x 8192 ran 3x slower than at 1024 x 2048.  Plain
// lengths that are not powers of two take the cost-model split (their
// runtime-pass sub-transforms vary widely in speed with the factors).
static bool fs_split(size_t n, size_t elem, bool conv, size_t* n1, size_t* n2) {
  return (conv && four_step_split_lds(n, elem, n1, n2)) || (!conv && four_step_split_model(n, kFactorCap, elem, n1, n2)) ||
         four_step_split(n, kFactorCap, n1, n2);
}

// Three-pass split n = a x b x c of a smooth length, each factor at most
// 512 (FP32) / 256 (FP64) points so that every pass holds 16 or more
// sub-transforms per workgroup within half a CU's LDS (128-byte segments):
// fewest passes, then the most balanced factors.
static bool three_pass_split(size_t n, size_t elem, size_t* a, size_t* b, size_t* c) {
  const size_t fmax = elem >= 16 ? 256 : 512;
  auto passes = [](size_t m) { return (double)radices_for(m).size(); };
  double best = 1e300;
  for (size_t x = 2; x <= fmax; ++x) {
    if (n % x || !is_smooth7(x)) continue;
    for (size_t y = 2; y <= fmax; ++y) {
      if ((n / x) % y || !is_smooth7(y)) continue;
      const size_t z = n / x / y;
      if (z < 2 || z > fmax) continue;
      const double cost = passes(x) + passes(y) + passes(z) +
                          0.1 * (std::fabs(std::log((double)x / (double)z)) + std::fabs(std::log((double)y / (double)z)));
      if (cost < best) {
        best = cost;
        *a = x;
        *b = y;
        *c = z;
      }
    }
  }
  return best < 1e300;
}

static bool fs_possible(size_t n, size_t elem, size_t cap) {
  size_t n1, n2;
  LdsGeom a, b;
  return fs_split(n, elem, false, &n1, &n2) && fs_geometry(n1, elem, cap, &a) && fs_geometry(n2, elem, cap, &b);
}

// Four-step structure for smooth n (tables only; workspace separately).
static fyfft_status plan_fourstep(fyfft_plan* p, size_t n, bool conv) {
  const size_t cap = lds_cap_bytes(p->device), elem = elem_bytes(p->prec);
  const FsChoice* f = p->force_fs;
  if (f) {
    if (f->n1 < 2 || n % f->n1 || n / f->n1 < 2 || !is_smooth7(f->n1) || !is_smooth7(n / f->n1))
      return FYFFT_UNSUPPORTED;
    p->n1 = f->n1;
    p->n2 = n / f->n1;
  }
  // two passes where the split and its tiles exist; the three-pass form
  // below otherwise (and for two-pass tiles narrower than 8 sub-transforms)
  const bool two = (f || fs_split(n, elem, conv, &p->n1, &p->n2)) &&
                   fs_geometry(p->n1, elem, cap, &p->gcol, f ? f->ccol : 0) &&
                   fs_geometry(p->n2, elem, cap, &p->grow, f ? f->crow : 0, !conv);
  if (!two && (f || conv || p->grouped_sub)) return FYFFT_UNSUPPORTED;
  // a compile-time sub-transform when one matches the planner's pass list;
  // otherwise the runtime-pass sub-transforms of plain FP32 transforms take
  // composite radices (convolutions keep the base radices)
  auto passes = [&](size_t len, const LdsGeom& g, int* fixed, dev::PassDesc* pd, void** tw) {
    std::vector<int> rad = radices_for(len);
    dev::PassDesc probe{};
    probe.npass = (int)std::min<size_t>(rad.size(), (size_t)dev::kMaxPasses);
    for (int i = 0; i < probe.npass; ++i) probe.radix[i] = rad[i];
    *fixed = find_fs_fixed(p->prec, probe, (int)len, g.threads, conv);
    if (!*fixed && !conv && p->prec == kF32 && g.ppt == 16) rad = radices_composite(len);
    return make_pass_table(p, len, pd, tw, &rad);
  };
  // very long lengths whose two-pass tiles hold fewer than 8 sub-transforms
  // (segments under 64 bytes: N = 2^24 = 4096 x 4096 ran 2 / 2 per
  // workgroup at 4.5 % of 8 TB/s) take three passes of short sub-transforms
  size_t a3 = 0, b3 = 0, c3 = 0;
  LdsGeom ga, gb, gc;
  if (!conv && !f && !p->grouped_sub && (!two || p->gcol.per_block < 8 || p->grow.per_block < 8) &&
      three_pass_split(n, elem, &a3, &b3, &c3) &&
      fs_geometry(a3, elem, cap, &ga) && fs_geometry(b3, elem, cap, &gb) && fs_geometry(c3, elem, cap, &gc, 0, true)) {
    p->n1 = a3;
    p->n2 = b3;
    p->n3 = c3;
    p->gcol = ga;
    p->gmid = gb;
    p->grow = gc;
  } else if (!two) {
    return FYFFT_UNSUPPORTED;
  }
  fyfft_status st = passes(p->n1, p->gcol, &p->fixed_col, &p->pdcol, &p->twcol);
  if (!st && p->n3) st = passes(p->n2, p->gmid, &p->fixed_mid, &p->pdmid, &p->twmid);
  if (!st) st = passes(p->n3 ? p->n3 : p->n2, p->grow, &p->fixed_row, &p->pdrow, &p->twrow);
  if (st) return st;
  // hi/lo root tables of length `len` (FP64, and beyond 2^24 points)
  auto roots = [&](size_t len, int* lo_bits, void** t_lo, void** t_hi) {
    int lg = 0;
    while (((size_t)1 << lg) < len) ++lg;
    *lo_bits = (lg + 1) / 2;
    const size_t lo = (size_t)1 << *lo_bits, hi = (len >> *lo_bits) + 1;
    fyfft_status e;
    if ((e = p->alloc(lo * elem, t_lo)) || (e = p->alloc(hi * elem, t_hi))) return e;
    HIP_TRY(launch_gen_roots(p->prec, *t_lo, (long long)lo, 1, (long long)len, nullptr));
    HIP_TRY(launch_gen_roots(p->prec, *t_hi, (long long)hi, (long long)lo, (long long)len, nullptr));
    return FYFFT_SUCCESS;
  };
  if ((st = roots(n, &p->lo_bits, &p->tw_lo, &p->tw_hi))) return st;
  if (p->n3 && (st = roots(p->n2 * p->n3, &p->lo_bits2, &p->tw2_lo, &p->tw2_hi))) return st;
  return FYFFT_SUCCESS;
}

static fyfft_status alloc_workspace(fyfft_plan* p, size_t len, bool x0) {
  const size_t per = len * elem_bytes(p->prec);
  p->chunk = std::max<size_t>(1, std::min(p->batch, kWorkspaceCap / per));
  fyfft_status st = p->alloc(p->chunk * per, &p->tmp);
  if (!st && x0) st = p->alloc(p->chunk * elem_bytes(p->prec), &p->x0buf);
  return st;
}

// Forward FFT of a length-len FP64 sequence in place with a temporary F64
// plan of this library (used to build kernel spectra from FP64 tables).

// ------------------------------------------------------------ autotuner
// FYFFT_FORM_TUNED plans time candidate configurations on the device at plan
// time and keep the fastest:
//   * LDS-resident smooth lengths: the compile-time kernel when one exists,
//     and the generic kernel over several radix orders (power-of-two split
//     and placement of the odd radices), LDS padding on/off and two
//     points-per-thread classes, each with its own FP64 twiddle table;
//   * four-step lengths (smooth N beyond one workgroup's LDS, and the
//     convolution length of four-step Rader plans): the split N = n1 x n2
//     over the plannable divisor pairs nearest the default one, and the
//     sub-transforms per workgroup of the column and row passes;
//   * Bluestein: the convolution length M among the cheapest candidates of
//     the length cost model.
// Choices are remembered as wisdom per (device, N, precision): the device
// is its architecture and compute-unit count, so wisdom tuned on one GPU
// model is never applied on another.
struct LdsChoice {
  bool fixed = false;
  int ppt = 16, ps = 4, tpb = 1;
  std::vector<int> rad;
};
enum WisdomFs { kWfSmooth = 0, kWfRader = 1 };
static std::mutex g_wisdom_mu;
// (device, N, prec, generic-only)
static std::map<std::tuple<std::string, size_t, int, int>, LdsChoice> g_wisdom;
static std::map<std::tuple<std::string, size_t, int, int>, FsChoice> g_wisdom_fs;      // (device, N, prec, kind)
static std::map<std::tuple<std::string, size_t, int>, size_t> g_wisdom_m;             // Bluestein M

// "gfx950 256": the architecture (without target features) and CU count of
// a device, the identity wisdom is keyed by.
static std::string device_identity(int dev) {
  static std::mutex mu;
  static std::map<int, std::string> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  std::string id = "unknown 0";
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    id = arch + " " + std::to_string(prop.multiProcessorCount);
  }
  cache[dev] = id;
  return id;
}

static std::vector<std::vector<int>> radix_orders(size_t n) {
  const std::vector<int> base = radices_for(n);
  std::vector<int> pow2, odd;
  int bits = 0;
  for (int r : base) {
    if (r & (r - 1)) odd.push_back(r);
    else { pow2.push_back(r); int b = 0; while ((1 << b) < r) ++b; bits += b; }
  }
  std::vector<std::vector<int>> p2s = {pow2};
  if (bits > 0) {
    std::vector<int> greedy;
    int left = bits;
    while (left > 0) { const int b = std::min(4, left); greedy.push_back(1 << b); left -= b; }
    p2s.push_back(greedy);
    std::vector<int> rev(greedy.rbegin(), greedy.rend());
    p2s.push_back(rev);
  }
  std::vector<std::vector<int>> odds = {odd};
  if (odd.size() > 1) odds.push_back(std::vector<int>(odd.rbegin(), odd.rend()));
  std::vector<std::vector<int>> out;
  for (const auto& a : p2s)
    for (const auto& o : odds)
      for (int odd_first = 0; odd_first < 2; ++odd_first) {
        if (odd_first && (o.empty() || a.empty())) continue;
        std::vector<int> r = odd_first ? o : a;
        const auto& tail = odd_first ? a : o;
        r.insert(r.end(), tail.begin(), tail.end());
        if (std::find(out.begin(), out.end(), r) == out.end()) out.push_back(r);
      }
  return out;
}

// Geometry + table for one candidate; the table is a plain hipMalloc owned by
// the caller (*tw_alloc).
static fyfft_status make_candidate(fyfft_plan* p, size_t n, const LdsChoice& c, LdsLaunch* L,
                                   void** tw_alloc) {
  const size_t elem = elem_bytes(p->prec), cap = lds_cap_bytes(p->device);
  *tw_alloc = nullptr;
  if (c.rad.empty() || (int)c.rad.size() > dev::kMaxPackedPasses) return FYFFT_UNSUPPORTED;
  L->pd.npass = (int)c.rad.size();
  for (size_t i = 0; i < c.rad.size(); ++i) L->pd.radix[i] = c.rad[i];
  LdsGeom g;
  g.n = (int)n;
  g.ppt = c.ppt;
  g.threads = (int)((n + c.ppt - 1) / c.ppt);
  g.pad_shift = c.ps;
  g.per_block = c.tpb;
  const size_t stride = (size_t)dev::lds_len((int)n, c.ps);
  g.lds_bytes = stride * c.tpb * elem;
  if (g.threads > 1024 || g.threads * c.tpb > 1024 || g.lds_bytes > cap) return FYFFT_UNSUPPORTED;
  L->geom = g;
  L->fixed = -1;
  if (c.fixed) {
    L->fixed = find_fixed_kernel(p->prec, L->pd, &L->geom);
    if (L->fixed < 0) return FYFFT_UNSUPPORTED;
  }
  HIP_TRY(hipMalloc(tw_alloc, std::max<size_t>(n, 2) * elem));
  HIP_TRY(launch_gen_stockham_twiddles(p->prec, *tw_alloc, (int)n, L->pd, nullptr));
  L->tw = *tw_alloc;
  return FYFFT_SUCCESS;
}

static std::vector<LdsChoice> lds_candidates(fyfft_plan* p, size_t n) {
  std::vector<LdsChoice> out;
  LdsGeom g;
  if (!lds_geometry(n, elem_bytes(p->prec), lds_cap_bytes(p->device), &g, true)) return out;
  const bool f64 = p->prec == kF64;
  const int lo = f64 ? 8 : 16;  // the alternative: twice or half the points per thread
  const int alt_ppt = g.ppt == lo ? 2 * lo : (g.ppt < lo || g.ppt == 2 * lo ? lo : 2 * lo);
  auto tpb_for = [&](int ppt, int ps) {
    const int T = (int)((n + ppt - 1) / ppt);
    const size_t per = (size_t)dev::lds_len((int)n, ps) * elem_bytes(p->prec);
    int t = std::max(1, 256 / T);
    t = (int)std::min<size_t>(t, lds_cap_bytes(p->device) / per);
    return std::max(1, std::min(t, 1024 / T));
  };
  const auto orders = radix_orders(n);
  {
    LdsChoice c;
    c.fixed = true;
    c.rad = orders[0];
    c.ppt = g.ppt;
    c.ps = g.pad_shift;
    c.tpb = 1;
    out.push_back(c);  // kept only if a fixed kernel exists (checked in make_candidate)
  }
  for (size_t o = 0; o < orders.size() && out.size() < 16; ++o)
    for (int ps : {g.pad_shift, 30}) {
      LdsChoice c;
      c.rad = orders[o];
      c.ppt = g.ppt;
      c.ps = ps;
      c.tpb = tpb_for(c.ppt, ps);
      out.push_back(c);
      if (g.pad_shift == 30) break;
    }
  LdsChoice c;
  c.rad = orders[0];
  c.ppt = alt_ppt;
  c.ps = g.pad_shift;
  if ((n + alt_ppt - 1) / alt_ppt <= 1024) {
    c.tpb = tpb_for(alt_ppt, c.ps);
    out.push_back(c);
  }
  return out;
}

// Device buffers, stream and events of one tuning session: in/out of
// `elems` elements (the input zeroed: timing does not depend on the values).
struct TuneBench {
  void *in = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipSuccess;
  TuneBench(size_t bytes) {
    err = hipMalloc(&in, bytes);
    if (err == hipSuccess) err = hipMalloc(&out, bytes);
    if (err == hipSuccess) err = hipMemset(in, 0, bytes);
    if (err == hipSuccess) err = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (err == hipSuccess) err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
  }
  ~TuneBench() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
  }
  // ms per call of f(stream): the best of three groups of three calls,
  // after two warm-up calls
  template <typename F> float time(F f) {
    for (int i = 0; i < 2 && err == hipSuccess; ++i) err = f(st);
    float best = 1e30f;
    for (int g = 0; g < 3 && err == hipSuccess; ++g) {
      err = hipEventRecord(e0, st);
      for (int i = 0; i < 3 && err == hipSuccess; ++i) err = f(st);
      if (err == hipSuccess) err = hipEventRecord(e1, st);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 1e30f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err == hipSuccess) best = std::min(best, ms / 3);
    }
    return err == hipSuccess ? best : 1e30f;
  }
};

// Transforms per tuning launch: the plan's own batch, up to 1 GiB of data.
// Smaller samples misled: 256 MB fits the 256 MB Infinity Cache, and the
// structures fastest there were 5-10 % slower at the full batch (N = 2^16 x
// 1024, 2^20 x 64).
static size_t tune_batch(size_t batch, size_t n, size_t elem) {
  return std::max<size_t>(1, std::min<size_t>(batch, ((size_t)1 << 30) / (n * elem)));
}

// generic_only: candidates the runtime-pass kernels can run (sub-plans of
// grouped plans, which never launch the compile-time kernels)
static fyfft_status plan_lds_tuned(fyfft_plan* p, size_t n, bool generic_only = false) {
  const auto key = std::make_tuple(device_identity(p->device), n, p->prec, generic_only ? 1 : 0);
  LdsChoice best;
  bool have = false;
  {
    std::lock_guard<std::mutex> lock(g_wisdom_mu);
    auto it = g_wisdom.find(key);
    if (it != g_wisdom.end()) { best = it->second; have = true; }
  }
  if (!have) {
    const size_t elem = elem_bytes(p->prec);
    const size_t b = tune_batch(p->batch, n, elem);
    TuneBench tb(b * n * elem);
    float best_ms = 1e30f;
    for (const LdsChoice& c : lds_candidates(p, n)) {
      if (tb.err != hipSuccess) break;
      if (generic_only && c.fixed) continue;
      LdsLaunch L;
      void* tw = nullptr;
      if (make_candidate(p, n, c, &L, &tw) == FYFFT_SUCCESS) {
        const float ms =
            tb.time([&](hipStream_t st) { return launch_lds(p->prec, kForward, L, tb.in, tb.out, b, n, 1.0, st); });
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
          have = true;
        }
      }
      if (tw) (void)hipFree(tw);
    }
    if (tb.err != hipSuccess) return from_hip(tb.err, "autotune");
    if (!have) return FYFFT_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_wisdom_mu);
    g_wisdom[key] = best;
  }
  void* tw = nullptr;
  fyfft_status st2 = make_candidate(p, n, best, &p->lds, &tw);
  if (tw) {
    p->allocs.push_back(tw);
    p->dev_bytes += std::max<size_t>(n, 2) * elem_bytes(p->prec);
  }
  if (!st2 && generic_only) p->lds.fixed = -1;
  return st2;
}

// Four-step structure candidates for length L: the default split, the
// plannable divisor pairs nearest to it, and for each split the default,
// halved and (where threads and LDS allow) doubled sub-transforms per
// workgroup of both passes (doubled column tiles read 256-byte row segments,
// which straddle fewer cache lines when the row length is not a multiple of
// the line).
static std::vector<FsChoice> fs_candidates(const fyfft_plan* p, size_t L, bool conv) {
  const size_t elem = elem_bytes(p->prec), cap = lds_cap_bytes(p->device);
  std::vector<FsChoice> out;
  size_t d1 = 0, d2 = 0;
  if (!fs_split(L, elem, conv, &d1, &d2)) return out;
  std::vector<size_t> splits = {d1};
  std::vector<std::pair<double, size_t>> others;
  for (size_t d = 2; d * 2 <= L; ++d) {
    if (L % d || d == d1) continue;
    LdsGeom a, b;
    if (!is_smooth7(d) || !is_smooth7(L / d) || !fs_geometry(d, elem, cap, &a) || !fs_geometry(L / d, elem, cap, &b))
      continue;
    others.push_back({std::fabs(std::log((double)d / (double)d1)), d});
  }
  std::sort(others.begin(), others.end());
  for (size_t i = 0; i < others.size() && splits.size() < 4; ++i) splits.push_back(others[i].second);
  for (size_t n1 : splits) {
    LdsGeom a, b;
    if (!fs_geometry(n1, elem, cap, &a) || !fs_geometry(L / n1, elem, cap, &b)) continue;
    auto fits = [&](const LdsGeom& q, size_t len, int c) {
      return c >= 1 && (size_t)c * q.threads <= 1024 &&
             (size_t)c * (size_t)dev::lds_len((int)len, q.pad_shift) * elem <= cap;
    };
    for (int ca : {a.per_block, a.per_block / 2, a.per_block * 2})
      for (int cb : {b.per_block, b.per_block / 2, b.per_block * 2}) {
        if (!fits(a, n1, ca) || !fits(b, L / n1, cb)) continue;
        FsChoice c;
        c.n1 = n1;
        c.ccol = ca;
        c.crow = cb;
        out.push_back(c);
      }
  }
  return out;
}

static fyfft_status create_plan_impl(fyfft_plan**, size_t, size_t, fyfft_precision, int, int force = -1,
                                     const FsChoice* force_fs = nullptr, size_t force_m = 0);

static fyfft_precision api_precision(int prec) {
  return prec == kF64 ? FYFFT_F64 : (prec == kBF16 ? FYFFT_BF16 : (prec == kF16 ? FYFFT_F16 : FYFFT_F32));
}

// The fastest four-step structure for the plan's length (kind: a smooth
// length, or the convolution of a Rader plan), from wisdom or by timing a
// candidate plan per structure on the plan's batch (up to 1 GiB).
static fyfft_status tune_fourstep(fyfft_plan* p, size_t L, int kind, FsChoice* best) {
  const auto key = std::make_tuple(device_identity(p->device), p->n, p->prec, kind);
  {
    std::lock_guard<std::mutex> lock(g_wisdom_mu);
    auto it = g_wisdom_fs.find(key);
    if (it != g_wisdom_fs.end()) {
      *best = it->second;
      return FYFFT_SUCCESS;
    }
  }
  const std::vector<FsChoice> cands = fs_candidates(p, L, kind == kWfRader);
  if (cands.empty()) return FYFFT_UNSUPPORTED;
  const size_t elem = elem_bytes(p->prec), b = tune_batch(p->batch, p->n, elem);
  TuneBench tb(b * p->n * elem);
  float best_ms = 1e30f;
  bool have = false;
  for (const FsChoice& c : cands) {
    if (tb.err != hipSuccess) break;
    fyfft_plan* q = nullptr;
    if (create_plan_impl(&q, p->n, b, api_precision(p->prec), FYFFT_FORM_AUTO, kind == kWfRader ? kRader : -1, &c) !=
        FYFFT_SUCCESS)
      continue;
    const float ms = tb.time([&](hipStream_t st) {
      return fyfft_execute(q, tb.in, tb.out, FYFFT_FORWARD, st) == FYFFT_SUCCESS ? hipSuccess : hipErrorUnknown;
    });
    fyfft_plan_destroy(q);
    if (ms < best_ms) {
      best_ms = ms;
      *best = c;
      have = true;
    }
  }
  if (tb.err != hipSuccess) return from_hip(tb.err, "autotune four-step");
  if (!have) return FYFFT_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  g_wisdom_fs[key] = *best;
  return FYFFT_SUCCESS;
}

static double conv_length_cost(size_t c, size_t elem, size_t cap, int prec);

// Bluestein convolution lengths worth timing: the cheapest few of the
// length cost model in [2N-1, 1.25 (2N-1)] (the model's own choice first).
static std::vector<size_t> bluestein_candidates(const fyfft_plan* p, size_t n, size_t count) {
  const size_t elem = elem_bytes(p->prec), cap = lds_cap_bytes(p->device);
  std::vector<std::pair<double, size_t>> c;
  const size_t lo = 2 * n - 1, hi = lo + lo / 4;
  for (size_t m = lo; m <= hi; ++m) {
    const double cost = conv_length_cost(m, elem, cap, p->prec);
    if (cost > 0) c.push_back({cost, m});
  }
  std::sort(c.begin(), c.end());
  std::vector<size_t> out;
  for (size_t i = 0; i < c.size() && out.size() < count; ++i) out.push_back(c[i].second);
  return out;
}

static fyfft_status tune_bluestein(fyfft_plan* p, size_t* best_m) {
  const auto key = std::make_tuple(device_identity(p->device), p->n, p->prec);
  {
    std::lock_guard<std::mutex> lock(g_wisdom_mu);
    auto it = g_wisdom_m.find(key);
    if (it != g_wisdom_m.end()) {
      *best_m = it->second;
      return FYFFT_SUCCESS;
    }
  }
  const std::vector<size_t> cands = bluestein_candidates(p, p->n, 4);
  if (cands.size() < 2) {
    *best_m = 0;  // nothing to choose: the cost model's length
    return FYFFT_SUCCESS;
  }
  const size_t elem = elem_bytes(p->prec), b = tune_batch(p->batch, p->n, elem);
  TuneBench tb(b * p->n * elem);
  float best_ms = 1e30f;
  for (size_t m : cands) {
    if (tb.err != hipSuccess) break;
    fyfft_plan* q = nullptr;
    if (create_plan_impl(&q, p->n, b, api_precision(p->prec), FYFFT_FORM_AUTO, kBluestein, nullptr, m) !=
        FYFFT_SUCCESS)
      continue;
    const float ms = tb.time([&](hipStream_t st) {
      return fyfft_execute(q, tb.in, tb.out, FYFFT_FORWARD, st) == FYFFT_SUCCESS ? hipSuccess : hipErrorUnknown;
    });
    fyfft_plan_destroy(q);
    if (ms < best_ms) {
      best_ms = ms;
      *best_m = m;
    }
  }
  if (tb.err != hipSuccess) return from_hip(tb.err, "autotune bluestein");
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  g_wisdom_m[key] = *best_m;
  return FYFFT_SUCCESS;
}

static std::string wisdom_text() {
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  std::map<std::string, std::string> per_dev;
  for (const auto& kv : g_wisdom) {
    const LdsChoice& c = kv.second;
    std::string ln = std::string(std::get<3>(kv.first) ? "ldsg " : "lds ") + std::to_string(std::get<1>(kv.first)) +
                     " " + std::to_string(std::get<2>(kv.first)) +
                     " " + std::to_string(c.fixed ? 1 : 0) + " " + std::to_string(c.ppt) + " " +
                     std::to_string(c.ps) + " " + std::to_string(c.tpb) + " ";
    for (size_t i = 0; i < c.rad.size(); ++i) ln += (i ? "x" : "") + std::to_string(c.rad[i]);
    per_dev[std::get<0>(kv.first)] += ln + "\n";
  }
  for (const auto& kv : g_wisdom_fs)
    per_dev[std::get<0>(kv.first)] += "fs " + std::to_string(std::get<1>(kv.first)) + " " +
                                      std::to_string(std::get<2>(kv.first)) + " " +
                                      std::to_string(std::get<3>(kv.first)) + " " + std::to_string(kv.second.n1) +
                                      " " + std::to_string(kv.second.ccol) + " " + std::to_string(kv.second.crow) +
                                      "\n";
  for (const auto& kv : g_wisdom_m)
    per_dev[std::get<0>(kv.first)] += "bluestein " + std::to_string(std::get<1>(kv.first)) + " " +
                                      std::to_string(std::get<2>(kv.first)) + " " + std::to_string(kv.second) +
                                      "\n";
  std::string s = "fyfft-wisdom 2\n";
  for (const auto& kv : per_dev) s += "device " + kv.first + "\n" + kv.second;
  return s;
}

static fyfft_status wisdom_bad(const std::string& why) {
  g_last_error = "fyfft_wisdom_import: " + why;
  return FYFFT_INVALID_ARGUMENT;
}

static fyfft_status wisdom_parse(const char* text) {
  std::string t(text);
  size_t pos = t.find('\n');
  if (t.compare(0, 15, "fyfft-wisdom 2\n") != 0 || pos == std::string::npos)
    return wisdom_bad("expected the header line \"fyfft-wisdom 2\"");
  // wisdom is applied only on the GPU model it was tuned on: a device line
  // that does not match the current device (when one is visible) is refused
  int cur = -1;
  const std::string here = hipGetDevice(&cur) == hipSuccess ? device_identity(cur) : "";
  (void)hipGetLastError();
  decltype(g_wisdom) lds;
  decltype(g_wisdom_fs) fs;
  decltype(g_wisdom_m) bm;
  std::string dev;
  size_t start = pos + 1;
  while (start < t.size()) {
    size_t end = t.find('\n', start);
    if (end == std::string::npos) end = t.size();
    const std::string line = t.substr(start, end - start);
    start = end + 1;
    if (line.empty()) continue;
    char arch[64] = {0};
    int cus = 0;
    if (line.compare(0, 7, "device ") == 0) {
      if (std::sscanf(line.c_str(), "device %63s %d", arch, &cus) != 2 || cus <= 0)
        return wisdom_bad("malformed device line: " + line);
      dev = std::string(arch) + " " + std::to_string(cus);
      if (!here.empty() && dev != here)
        return wisdom_bad("tuned on " + dev + ", the current device is " + here);
      continue;
    }
    if (dev.empty()) return wisdom_bad("entry before any device line: " + line);
    unsigned long long n = 0;
    int prec = 0;
    if (line.compare(0, 3, "fs ") == 0) {
      unsigned long long n1 = 0;
      int kind = 0, ccol = 0, crow = 0;
      if (std::sscanf(line.c_str(), "fs %llu %d %d %llu %d %d", &n, &prec, &kind, &n1, &ccol, &crow) != 6 ||
          (prec != kF32 && prec != kF64) || (kind != kWfSmooth && kind != kWfRader) || n1 < 2 ||
          n > FYFFT_MAX_LENGTH || ccol < 1 || crow < 1 || ccol > 1024 || crow > 1024 || (ccol & (ccol - 1)) ||
          (crow & (crow - 1)))
        return wisdom_bad("malformed four-step entry: " + line);
      const unsigned long long L = kind == kWfRader ? n - 1 : n;
      if (L % n1 || !is_smooth7((size_t)n1) || !is_smooth7((size_t)(L / n1)))
        return wisdom_bad("split does not divide the length: " + line);
      FsChoice c;
      c.n1 = (size_t)n1;
      c.ccol = ccol;
      c.crow = crow;
      fs[std::make_tuple(dev, (size_t)n, prec, kind)] = c;
      continue;
    }
    if (line.compare(0, 10, "bluestein ") == 0) {
      unsigned long long m = 0;
      if (std::sscanf(line.c_str(), "bluestein %llu %d %llu", &n, &prec, &m) != 3 || (prec != kF32 && prec != kF64) ||
          n < 2 || n > FYFFT_MAX_LENGTH || m < 2 * n - 1 || m > 8 * n || !is_smooth7((size_t)m))
        return wisdom_bad("malformed Bluestein entry: " + line);
      bm[std::make_tuple(dev, (size_t)n, prec)] = (size_t)m;
      continue;
    }
    int fixed = 0, ppt = 0, ps = 0, tpb = 0;
    char rad[256] = {0};
    const bool gen = line.compare(0, 5, "ldsg ") == 0;
    if (std::sscanf(line.c_str() + (gen ? 5 : 4), "%llu %d %d %d %d %d %255s", &n, &prec, &fixed, &ppt, &ps, &tpb,
                    rad) != 7 ||
        (!gen && line.compare(0, 4, "lds ") != 0) || (gen && fixed))
      return wisdom_bad("malformed entry: " + line);
    LdsChoice c;
    c.fixed = fixed != 0;
    c.ppt = ppt;
    c.ps = ps;
    c.tpb = tpb;
    size_t prod = 1;
    for (char* tok = std::strtok(rad, "x"); tok; tok = std::strtok(nullptr, "x")) {
      const int r = std::atoi(tok);
      if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 9 && r != 16)
        return wisdom_bad("radix outside the codelets: " + line);
      c.rad.push_back(r);
      prod *= (size_t)r;
    }
    // the planner's points per thread: lo or 2*lo, or what its geometry for n
    // takes (n spread over 1024 threads, or fewer points for short lengths)
    const unsigned long long lo = prec == kF64 ? 8 : 16;
    LdsGeom geo;
    const bool ppt_ok = (unsigned long long)ppt == lo || (unsigned long long)ppt == 2 * lo ||
                        (n >= 1 && n <= (1ULL << 20) &&
                         lds_geometry((size_t)n, prec == kF64 ? 16 : 8, (size_t)1 << 40, &geo, true) && geo.ppt == ppt);
    if (prod != n || (prec != kF32 && prec != kF64) || !ppt_ok || (ps != 30 && (ps < 2 || ps > 8)) ||
        tpb < 1 || tpb > 1024 || c.rad.empty() || (int)c.rad.size() > dev::kMaxPackedPasses)
      return wisdom_bad("invalid LDS entry: " + line);
    lds[std::make_tuple(dev, (size_t)n, prec, gen ? 1 : 0)] = c;
  }
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  for (auto& kv : lds) g_wisdom[kv.first] = kv.second;
  for (auto& kv : fs) g_wisdom_fs[kv.first] = kv.second;
  for (auto& kv : bm) g_wisdom_m[kv.first] = kv.second;
  return FYFFT_SUCCESS;
}


static fyfft_status fft_f64_inplace(double2* buf, size_t len) {
  fyfft_plan* sub = nullptr;
  fyfft_status st = create_plan_impl(&sub, len, 1, FYFFT_F64, FYFFT_FORM_AUTO);
  if (st) return st;
  st = fyfft_execute(sub, buf, buf, FYFFT_FORWARD, nullptr);
  if (!st) {
    hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) st = from_hip(e, "kernel spectrum FFT");
  }
  fyfft_plan_destroy(sub);
  return st;
}

// Lengths that fit one workgroup's LDS only beyond the low points-per-thread
// bound (FP32 16384 < n <= 20480, FP64 8192 < n <= 10240) and have no
// compile-time kernel take the four-step form in batched plans: the
// runtime-pass kernel at 1024 threads and 32 points per thread keeps call
// frames in scratch and one workgroup per CU (FP32 batches of 2^26 points:
// N = 16800 1.11 -> 0.71 ms, 18432 0.80 -> 0.54 ms; Rader N = 18433
// 2.01 -> 1.47 ms, 17011 2.76 -> 1.75 ms).  Grouped sub-plans keep them in
// LDS (see kFormGroupedPlain).
static bool lds_hi_to_fourstep(fyfft_plan* p, size_t n, const LdsGeom& g, bool conv = false) {
  if (p->grouped_sub) return false;
  const size_t elem = elem_bytes(p->prec);
  const int lo = elem >= 16 ? 8 : 16;
  if (g.ppt <= lo) return false;
  std::vector<int> rad = radices_for(n);
  dev::PassDesc pd{};
  pd.npass = (int)std::min<size_t>(rad.size(), (size_t)dev::kMaxPasses);
  for (int i = 0; i < pd.npass; ++i) pd.radix[i] = rad[i];
  LdsGeom q = g;
  if (conv ? find_conv_fixed(p->prec, pd, &q) >= 0 : find_fixed_kernel(p->prec, pd, &q) >= 0) return false;
  return fs_possible(n, elem, lds_cap_bytes(p->device));
}

// Convolution structure of length m (LDS-fused or four-step) for a Rader or
// Bluestein plan.
static fyfft_status plan_conv_structure(fyfft_plan* p, size_t m, bool rader) {
  p->m = m;
  LdsGeom g;
  if (lds_geometry(m, elem_bytes(p->prec), lds_cap_bytes(p->device), &g) && !lds_hi_to_fourstep(p, m, g, true)) {
    p->kind = rader ? kRaderLds : kBlueLds;
    fyfft_status st = plan_lds(p, m, &p->lds, false, false);
    if (!st) p->lds.fixed = find_conv_fixed(p->prec, p->lds.pd, &p->lds.geom);
    return st;
  }
  p->kind = rader ? kRaderFs : kBlueFs;
  FsChoice tuned_fs;
  fyfft_status st = FYFFT_SUCCESS;
  if (rader && !p->force_fs && (p->form == FYFFT_FORM_TUNED || p->form == kFormTunedGeneric)) {
    st = tune_fourstep(p, m, kWfRader, &tuned_fs);
    if (!st) p->force_fs = &tuned_fs;
  }
  if (!st) st = plan_fourstep(p, m, true);
  if (p->force_fs == &tuned_fs) p->force_fs = nullptr;
  if (!st) st = alloc_workspace(p, m, rader);
  return st;
}

// Kernel spectrum K = FFT_m(b) / m from an FP64 sequence b (consumed),
// stored as the plan precision, in [k1][k2] order for four-step plans.
static fyfft_status store_spectrum(fyfft_plan* p, double2* b, int slot) {
  fyfft_status st = fft_f64_inplace(b, p->m);
  if (st) return st;
  if ((st = p->alloc(p->m * elem_bytes(p->prec), &p->kern[slot]))) return st;
  const bool fs = p->kind == kRaderFs || p->kind == kBlueFs;
  HIP_TRY(launch_convert(p->prec, p->kern[slot], b, (long long)p->m, 1.0 / (double)p->m,
                         (long long)p->n1, (long long)p->n2, fs, nullptr));
  return FYFFT_SUCCESS;
}

// Tables of the bucketed Rader output permutation (kernels_rader.hip) for
// a four-step plan L = N1 x N2 with C columns per column-pass workgroup:
// conv index n = k1*N2 + n2 belongs to column group b = n2 / C and its
// result goes to y[pout[n]], in chunk (pout[n] - 1) / Q of y.  Group b's
// results are sorted by destination chunk (lout: the LDS index to read,
// c*stride + k1 + pad, for zw position b*N1*C + i); chunk k's elements in
// group order (out: {zw position, offset in the chunk}).
static fyfft_status plan_rader_buckets(fyfft_plan* p, const std::vector<int>& pout) {
  const size_t L = p->m, N1 = p->n1, N2 = p->n2;
  const int C = p->gcol.per_block, ps = p->gcol.pad_shift;
  const int stride = dev::lds_len((int)N1, ps);
  const size_t Q = 4096, K = (L + Q - 1) / Q, G = (N2 + C - 1) / C;
  std::vector<int> lout(L), dest(L);
  std::vector<int2> out2(L);
  std::vector<std::vector<size_t>> out_lo(G, std::vector<size_t>(K + 1, 0));
  for (size_t b = 0; b < G; ++b) {
    const size_t c0 = b * (size_t)C, ncols = std::min((size_t)C, N2 - c0), start = c0 * N1;
    std::vector<size_t> cnt(K + 1, 0);
    for (size_t k1 = 0; k1 < N1; ++k1)
      for (size_t c = 0; c < ncols; ++c) ++cnt[((size_t)pout[k1 * N2 + c0 + c] - 1) / Q + 1];
    for (size_t k = 0; k < K; ++k) cnt[k + 1] += cnt[k];
    out_lo[b] = cnt;
    std::vector<size_t> fill = cnt;
    for (size_t k1 = 0; k1 < N1; ++k1)
      for (size_t c = 0; c < ncols; ++c) {
        const size_t d = (size_t)pout[k1 * N2 + c0 + c], i = fill[(d - 1) / Q]++;
        lout[start + i] = (int)(c * (size_t)stride + k1 + (k1 >> ps));
        dest[start + i] = (int)d;
      }
  }
  for (size_t k = 0, pos = 0; k < K; ++k)
    for (size_t b = 0; b < G; ++b) {
      const size_t start = b * (size_t)C * N1;
      for (size_t i = out_lo[b][k]; i < out_lo[b][k + 1]; ++i, ++pos)
        out2[pos] = int2{(int)(start + i), (int)((size_t)dest[start + i] - 1 - k * Q)};
    }
  void* d = nullptr;
  fyfft_status st;
  if ((st = p->alloc(L * sizeof(int), &d))) return st;
  p->rb_lout = (int*)d;
  HIP_TRY(hipMemcpy(d, lout.data(), L * sizeof(int), hipMemcpyHostToDevice));
  if ((st = p->alloc(L * sizeof(int2), &d))) return st;
  p->rb_out = (int2*)d;
  HIP_TRY(hipMemcpy(d, out2.data(), L * sizeof(int2), hipMemcpyHostToDevice));
  if ((st = p->alloc(p->chunk * L * elem_bytes(p->prec), &p->zw))) return st;
  p->rb_q = (int)Q;
  return FYFFT_SUCCESS;
}

static fyfft_status plan_rader(fyfft_plan* p) {
  const size_t n = p->n, L = n - 1;
  fyfft_status st = plan_conv_structure(p, L, true);
  if (st) return st;
  const uint64_t g = primitive_root(n), ginv = power_mod(g, n - 2, n);
  std::vector<int> pin(L), pout(L);
  uint64_t a = 1, b = 1;
  for (size_t i = 0; i < L; ++i) {
    pout[i] = (int)a;
    pin[i] = (int)b;
    a = a * g % n;
    b = b * ginv % n;
  }
  void* d = nullptr;
  if ((st = p->alloc(L * sizeof(int), &d))) return st;
  p->perm_in = (int*)d;
  if ((st = p->alloc(L * sizeof(int), &d))) return st;
  p->perm_out = (int*)d;
  HIP_TRY(hipMemcpy(p->perm_in, pin.data(), L * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->perm_out, pout.data(), L * sizeof(int), hipMemcpyHostToDevice));
  if (p->kind == kRaderLds) {
    std::vector<int> pos(L);
    for (size_t i = 0; i < L; ++i) pos[pout[i] - 1] = (int)i;
    if ((st = p->alloc(L * sizeof(int), &d))) return st;
    p->pos_out = (int*)d;
    HIP_TRY(hipMemcpy(p->pos_out, pos.data(), L * sizeof(int), hipMemcpyHostToDevice));
  }
  double2* bb = nullptr;
  HIP_TRY(hipMalloc(&bb, L * sizeof(double2)));
  for (int slot = 0; slot < 2 && !st; ++slot) {
    hipError_t e = launch_gen_rader_b(bb, p->perm_out, (long long)n, (long long)L, slot ? +1 : -1, nullptr);
    st = e != hipSuccess ? from_hip(e, "rader kernel") : store_spectrum(p, bb, slot);
  }
  (void)hipFree(bb);
  if (!st && p->kind == kRaderFs) st = plan_rader_buckets(p, pout);
  if (!st)
    p->desc = std::string(p->kind == kRaderLds ? "rader-lds" : "rader-fourstep") +
              (p->kind == kRaderFs && (p->form == FYFFT_FORM_TUNED || p->form == kFormTunedGeneric) ? "(tuned)" : "") +
              " N=" +
              std::to_string(n) + " L=" + std::to_string(L) + " g=" + std::to_string(g) +
              (p->rb_q ? " bucketed-permutation Q=" + std::to_string(p->rb_q) : "") +
              (p->kind == kRaderLds && p->lds.fixed >= 0 ? " fixed" : "");
  return st;
}

// Estimated cost of a Bluestein convolution of length c (0: not plannable):
// c times its pass count (forward and inverse FFT of c, whether in LDS or as
// four-step column/row passes), times 0.6 for a compile-time LDS kernel
// (measured 1.4-1.8x faster: N = 10007 at M = 20480 0.33 -> 0.18 ms for
// 1024, N = 8191 at M = 16384 0.24 -> 0.16 ms for 2048), times 1.6 for the
// four-step's extra HBM round trips, times 0.7 when both four-step factors
// have compile-time sub-transforms.  Measured on MI355X against the smallest smooth length:
// N = 10007 at M = 20480 (4 passes) instead of 20160 (6): 0.367 -> 0.335 ms
// for 1024 transforms; N = 1000003 at M = 2^21 instead of 2000376 =
// 1372 x 1458: 1.29 -> 0.63 ms for 16.
static double conv_length_cost(size_t c, size_t elem, size_t cap, int prec) {
  if (!is_smooth7(c)) return 0;
  LdsGeom g;
  if (lds_geometry(c, elem, cap, &g)) {
    const std::vector<int> r = radices_for(c);
    dev::PassDesc pd{};
    pd.npass = (int)std::min<size_t>(r.size(), (size_t)dev::kMaxPasses);
    for (int i = 0; i < pd.npass; ++i) pd.radix[i] = r[i];
    return (double)c * (double)r.size() * (find_conv_fixed(prec, pd, &g) >= 0 ? 0.6 : 1.0);
  }
  size_t n1, n2;
  LdsGeom a, b;
  if (!fs_split(c, elem, true, &n1, &n2) || !fs_geometry(n1, elem, cap, &a) || !fs_geometry(n2, elem, cap, &b))
    return 0;
  auto fixed = [&](size_t len, const LdsGeom& q) {
    const std::vector<int> r = radices_for(len);
    dev::PassDesc pd{};
    pd.npass = (int)r.size();
    for (size_t i = 0; i < r.size() && i < (size_t)dev::kMaxPasses; ++i) pd.radix[i] = r[i];
    return find_fs_fixed(prec, pd, (int)len, q.threads, true) != 0;
  };
  const double passes = (double)(radices_for(n1).size() + radices_for(n2).size());
  return (double)c * passes * 1.6 * (fixed(n1, a) && fixed(n2, b) ? 0.7 : 1.0);
}

static fyfft_status plan_bluestein(fyfft_plan* p) {
  const size_t n = p->n, elem = elem_bytes(p->prec), cap = lds_cap_bytes(p->device);
  // the cheapest plannable smooth length in [2N-1, 1.25 (2N-1)], else the
  // first plannable one above
  size_t m = 0;
  if (p->force_m) {
    if (p->force_m < 2 * n - 1 || conv_length_cost(p->force_m, elem, cap, p->prec) <= 0) return FYFFT_UNSUPPORTED;
    m = p->force_m;
  } else if (p->form == FYFFT_FORM_TUNED || p->form == kFormTunedGeneric) {
    fyfft_status st = tune_bluestein(p, &m);
    if (st) return st;
  }
  double best = 0;
  const size_t lo = 2 * n - 1, hi = lo + lo / 4;
  for (size_t c = lo; c <= hi && !m; ++c) {
    const double cost = conv_length_cost(c, elem, cap, p->prec);
    if (cost > 0 && (best == 0 || cost < best)) best = cost;
  }
  for (size_t c = lo; c <= hi && !m; ++c)
    if (best > 0 && conv_length_cost(c, elem, cap, p->prec) == best) m = c;
  for (size_t c = hi + 1; !m && c < 4 * lo + 64; ++c)
    if (conv_length_cost(c, elem, cap, p->prec) > 0) m = c;
  if (!m) return FYFFT_UNSUPPORTED;
  fyfft_status st = plan_conv_structure(p, m, false);
  if (st) return st;
  double2* c = nullptr;
  double2* bb = nullptr;
  HIP_TRY(hipMalloc(&c, n * sizeof(double2)));
  hipError_t e = hipMalloc(&bb, m * sizeof(double2));
  if (e != hipSuccess) { (void)hipFree(c); return from_hip(e, "bluestein tables"); }
  // chirp[0] = exp(-i pi n^2/N) (forward), chirp[1] = exp(+i pi n^2/N)
  for (int slot = 0; slot < 2 && !st; ++slot) {
    if ((st = p->alloc(n * elem, &p->chirp[slot]))) break;
    e = launch_gen_chirp(c, (long long)n, (long long)n, slot ? +1 : -1, nullptr);
    if (e == hipSuccess) e = launch_convert(p->prec, p->chirp[slot], c, (long long)n, 1.0, 0, 0, false, nullptr);
    if (e != hipSuccess) st = from_hip(e, "bluestein chirp");
  }
  // c now holds exp(+i pi n^2/N); b (wrapped) and conj(b) give the spectra
  for (int slot = 0; slot < 2 && !st; ++slot) {
    e = launch_gen_bluestein_b(bb, c, (long long)n, (long long)m, slot == 1, nullptr);
    st = e != hipSuccess ? from_hip(e, "bluestein kernel") : store_spectrum(p, bb, slot);
  }
  (void)hipFree(c);
  (void)hipFree(bb);
  if (!st)
    p->desc = std::string(p->kind == kBlueLds ? "bluestein-lds" : "bluestein-fourstep") +
              (p->form == FYFFT_FORM_TUNED || p->form == kFormTunedGeneric ? "(tuned)" : "") + " N=" +
              std::to_string(n) + " M=" + std::to_string(m) +
              (p->kind == kBlueLds && p->lds.fixed >= 0 ? " fixed" : "");
  return st;
}

// ------------------------------------------------------------ execution
// Rows of n elements that are not a whole number of 128-byte lines: the
// four-step row passes stage them through LDS, and strided segments of such
// rows (column-pass reads, transposed row-pass stores) go through the caches
// instead of streaming, since they share their first and last lines with
// the neighbouring segments (ragged grouped step 6.81 -> 6.70 ms).
static bool rows_misaligned(size_t n, size_t elem) { return (n * elem) % 128 != 0; }

template <typename V>
static void fill_pass(FsArgs<V>& a, const LdsGeom& g, size_t len, const dev::PassDesc& pd,
                      const void* tw) {
  a.threads = g.threads;
  a.log2c = ilog2_exact(g.per_block);
  a.ps = g.pad_shift;
  a.stride = dev::lds_len((int)len, g.pad_shift);
  a.pd = pd;
  a.tw = (const V*)tw;
}

// Device memory a multi-pass execution runs through: `chunk` transforms'
// worth of convolution-length workspace (tmp), Rader x[0] slots and the
// bucketed-permutation staging (zw).  The plan's own, or carved from the
// library's static scratch by fyfft_execute_noalloc.
struct Workspace {
  void* tmp = nullptr;
  void* x0buf = nullptr;
  void* zw = nullptr;
  size_t chunk = 0;
};

static Workspace plan_workspace(const fyfft_plan* p) {
  Workspace w;
  w.tmp = p->tmp;
  w.x0buf = p->x0buf;
  w.zw = p->zw;
  w.chunk = p->chunk;
  return w;
}

// Three-pass form of a plain length n = A x B x C (n3 = C): pass 1 is the
// four-step column pass over A (x -> out, twiddles of n), pass 2 treats
// each of the A rows of length B*C as a B x C four-step and runs its column
// pass over B (out -> workspace, twiddles of B*C), writing row kb of row k1
// at (kb*A + k1)*C, and pass 3 runs the length-C rows of that workspace with
// the transposed store, X[q + A*B*kc] for row q = kb*A + k1.
template <typename V>
static fyfft_status exec_fourstep3(const fyfft_plan* p, const V* in, V* out, int dir, double scale, size_t rows,
                                   size_t dist, const Workspace& ws, hipStream_t s) {
  const size_t n = p->n, A = p->n1, B = p->n2, C = p->n3, M = B * C, elem = elem_bytes(p->prec);
  for (size_t b0 = 0; b0 < rows; b0 += ws.chunk) {
    const size_t cb = std::min(ws.chunk, rows - b0);
    FsArgs<V> a;
    a.in = in + b0 * dist;
    a.tmp = out + b0 * dist;
    a.in_dist = a.tmp_dist = (long long)dist;
    a.n1 = (int)A;
    a.n2 = (int)M;
    a.tw_hi = (const V*)p->tw_hi;
    a.tw_lo = (const V*)p->tw_lo;
    a.lo_bits = p->lo_bits;
    a.inv_n = fs_inv_n(n);
    fill_pass(a, p->gcol, A, p->pdcol, p->twcol);
    a.col_cached = rows_misaligned(M, elem);
    HIP_TRY(launch_fs_pass(p->prec, true, kFsPlainA, dir, &a, cb, p->gcol, p->fixed_col, s));
    FsArgs<V> m;
    m.in = out + b0 * dist;
    m.tmp = (V*)ws.tmp;
    m.group = (int)A;
    m.in_dist = (long long)M;
    m.in_gdist = (long long)dist;
    m.tmp_dist = (long long)C;
    m.tmp_gdist = (long long)n;
    m.tmp_row = (int)(A * C);
    m.n1 = (int)B;
    m.n2 = (int)C;
    m.tw_hi = (const V*)p->tw2_hi;
    m.tw_lo = (const V*)p->tw2_lo;
    m.lo_bits = p->lo_bits2;
    m.inv_n = fs_inv_n(M);
    fill_pass(m, p->gmid, B, p->pdmid, p->twmid);
    m.col_cached = rows_misaligned(C, elem);
    HIP_TRY(launch_fs_pass(p->prec, true, kFsPlainA, dir, &m, cb * A, p->gmid, p->fixed_mid, s));
    FsArgs<V> r;
    r.tmp = (V*)ws.tmp;
    r.tmp_dist = (long long)n;
    r.out = out + b0 * dist;
    r.out_dist = (long long)dist;
    r.n1 = (int)(A * B);
    r.n2 = (int)C;
    r.scale = (dev::S<V>)scale;
    fill_pass(r, p->grow, C, p->pdrow, p->twrow);
    r.stage_rows = rows_misaligned(C, elem);
    r.out_cached = rows_misaligned(A * B, elem);
    HIP_TRY(launch_fs_pass(p->prec, false, kFsPlainB, dir, &r, cb, p->grow, p->fixed_row, s));
  }
  return FYFFT_SUCCESS;
}

template <typename V>
static fyfft_status exec_fourstep(const fyfft_plan* p, const V* in, V* out, int dir, bool inv_tables,
                                  double scale, size_t rows, size_t dist, const Workspace& ws, hipStream_t s) {
  const size_t n = p->n, L = (p->kind == kFourStep) ? n : p->m;
  for (size_t b0 = 0; b0 < rows; b0 += ws.chunk) {
    const size_t cb = std::min(ws.chunk, rows - b0);
    FsArgs<V> a;
    a.in = in + b0 * dist;
    a.out = out + b0 * dist;
    a.tmp = (V*)ws.tmp;
    a.in_dist = (long long)dist;
    a.out_dist = (long long)dist;
    a.tmp_dist = (long long)L;
    a.n1 = (int)p->n1;
    a.n2 = (int)p->n2;
    a.tw_hi = (const V*)p->tw_hi;
    a.tw_lo = (const V*)p->tw_lo;
    a.lo_bits = p->lo_bits;
    a.inv_n = fs_inv_n(p->n1 * p->n2);
    a.n_orig = (int)n;
    a.scale = (dev::S<V>)scale;
    const int slot = inv_tables ? 1 : 0;
    a.kern = (const V*)p->kern[slot];
    a.chirp = (const V*)p->chirp[slot];
    a.x0buf = (V*)ws.x0buf;
    int colA, colC = -1, rowB;
    // Rader: the input permutation as an element gather in the column pass
    // (measured faster than gathering it through chunked runs as well), the
    // output permutation through buckets
    const bool buckets = p->kind == kRaderFs && p->rb_q > 0;
    if (p->kind == kFourStep) { colA = kFsPlainA; rowB = kFsPlainB; }
    else if (p->kind == kRaderFs) {
      colA = kFsRaderA;
      rowB = kFsConvB;
      colC = buckets ? kFsRaderC2 : kFsRaderC;
    } else { colA = kFsBlueA; rowB = kFsConvB; colC = kFsBlueC; }
    if (buckets) {
      a.zw = (V*)ws.zw;
      a.zw_dist = (long long)L;
    }

    fill_pass(a, p->gcol, p->n1, p->pdcol, p->twcol);
    a.perm = p->perm_in;
    a.col_cached = rows_misaligned(p->n2, elem_bytes(p->prec));
    a.out_cached = rows_misaligned(p->n1, elem_bytes(p->prec));
    HIP_TRY(launch_fs_pass(p->prec, true, colA, dir, &a, cb, p->gcol, p->fixed_col, s));
    fill_pass(a, p->grow, p->n2, p->pdrow, p->twrow);
    a.stage_rows = rows_misaligned(p->n2, elem_bytes(p->prec));
    HIP_TRY(launch_fs_pass(p->prec, false, rowB, dir, &a, cb, p->grow, p->fixed_row, s));
    a.stage_rows = 0;
    if (colC >= 0) {
      fill_pass(a, p->gcol, p->n1, p->pdcol, p->twcol);
      a.perm = p->perm_out;
      a.lidx = p->rb_lout;
      HIP_TRY(launch_fs_pass(p->prec, true, colC, +1, &a, cb, p->gcol, p->fixed_col, s));
    }
    if (buckets)
      HIP_TRY(launch_rader_bucket(p->prec, ws.zw, a.out, (long long)cb, (long long)L, p->rb_q, p->rb_out,
                                  (long long)L, (long long)dist, s));
  }
  return FYFFT_SUCCESS;
}

template <typename V>
static fyfft_status exec_conv_lds(const fyfft_plan* p, const V* in, V* out, bool inv_tables,
                                  double scale, size_t rows, size_t dist, hipStream_t s) {
  ConvArgs<V> a;
  a.in = in;
  a.out = out;
  a.batch = (long long)rows;
  a.dist = (long long)dist;
  a.n_orig = (int)p->n;
  a.m = (int)p->m;
  a.threads = p->lds.geom.threads;
  a.per_block = p->lds.geom.per_block;
  a.ps = p->lds.geom.pad_shift;
  a.stride = dev::lds_len((int)p->m, a.ps);
  a.pd = p->lds.pd;
  a.tw = (const V*)p->lds.tw;
  a.perm_in = p->perm_in;
  a.perm_out = p->perm_out;
  a.pos_out = p->pos_out;
  const int slot = inv_tables ? 1 : 0;
  a.chirp = (const V*)p->chirp[slot];
  a.kern = (const V*)p->kern[slot];
  a.scale = (dev::S<V>)scale;
  HIP_TRY(launch_conv_lds(p->prec, p->kind == kRaderLds, &a, p->lds.geom, p->lds.fixed, s));
  return FYFFT_SUCCESS;
}

template <typename V>
static fyfft_status exec_spectral(const fyfft_plan* p, bool inverse, const void* src, void* dst,
                                  size_t k, double scale0, double scale, bool mode_major, hipStream_t s) {
  SpecArgs<V> a;
  a.rdist = mode_major ? 1 : (long long)k;
  a.kdist = mode_major ? (long long)p->batch : 1;
  if (inverse) {
    a.cplx_in = (const V*)src;
    a.real_out = (dev::S<V>*)dst;
  } else {
    a.real_in = (const dev::S<V>*)src;
    a.cplx_out = (V*)dst;
  }
  const LdsGeom& g = p->lds.geom;
  a.rows = (long long)p->batch;
  a.n = (int)p->n;
  a.k = (int)k;
  a.threads = g.threads;
  a.per_block = g.per_block;
  a.ps = g.pad_shift;
  a.stride = dev::lds_len((int)p->n, g.pad_shift);
  a.pd = p->lds.pd;
  a.tw = (const V*)p->lds.tw;
  a.a0 = (dev::S<V>)scale0;
  a.a = (dev::S<V>)scale;
  HIP_TRY(launch_spectral(p->prec, inverse, &a, g, s));
  return FYFFT_SUCCESS;
}

// Tables of the half-length spectral forms: the pass list, Stockham table
// and geometry of length n/2 (compile-time kernel geometry where one
// exists) and exp(-2 pi i k / n) for k <= n/2.  Built once per plan.
static bool spectral_half_tables(const fyfft_plan* cp) {
  std::lock_guard<std::mutex> lock(cp->half_mu);
  if (cp->half_state) return cp->half_state > 0;
  fyfft_plan* p = const_cast<fyfft_plan*>(cp);
  const size_t n = p->n, h = n / 2;
  cp->half_state = -1;
  if (n % 2 || h < 2 || !is_smooth7(h)) return false;
  if (plan_lds(p, h, &p->half, true, true)) return false;
  if (p->alloc((h + 1) * elem_bytes(p->prec), &p->half_rtw)) return false;
  if (launch_gen_roots(p->prec, p->half_rtw, (long long)(h + 1), 1, (long long)n, nullptr) != hipSuccess) return false;
  if (hipDeviceSynchronize() != hipSuccess) return false;
  cp->half_state = 1;
  return true;
}

template <typename V>
static fyfft_status exec_spectral_half(const fyfft_plan* p, bool inverse, const void* src, void* dst,
                                       size_t k, double scale0, double scale, bool mode_major, hipStream_t s) {
  SpecArgs<V> a;
  a.rdist = mode_major ? 1 : (long long)k;
  a.kdist = mode_major ? (long long)p->batch : 1;
  if (inverse) {
    a.cplx_in = (const V*)src;
    a.real_out = (dev::S<V>*)dst;
  } else {
    a.real_in = (const dev::S<V>*)src;
    a.cplx_out = (V*)dst;
  }
  const LdsGeom& g = p->half.geom;
  const size_t h = p->n / 2;
  a.rows = (long long)p->batch;
  a.n = (int)p->n;
  a.k = (int)k;
  a.threads = g.threads;
  a.per_block = g.per_block;
  a.ps = g.pad_shift;
  a.stride = dev::lds_len((int)h, g.pad_shift);
  a.pd = p->half.pd;
  a.tw = (const V*)p->half.tw;
  a.rtw = (const V*)p->half_rtw;
  if (inverse && k > 0) {
    // the modes array's extent: under 2 GiB the inverse reads it through a
    // buffer resource whose out-of-range reads return zero
    const unsigned long long span =
        ((unsigned long long)(a.rows - 1) * a.rdist + (unsigned long long)(k - 1) * a.kdist + 1) * sizeof(V);
    a.mode_bytes = span < (1ull << 31) && scale != 0 ? (unsigned)span : 0u;
  }
  a.a0 = (dev::S<V>)scale0;
  a.a = (dev::S<V>)scale;
  HIP_TRY(launch_spectral_half(p->prec, inverse, &a, g, p->half.fixed >= 0 ? (int)h : 0, s));
  return FYFFT_SUCCESS;
}

static fyfft_status spectral(const fyfft_plan* p, bool inverse, const void* src, void* dst,
                             size_t k, double scale0, double scale, int layout, void* stream) {
  if (layout != FYFFT_MODES_ROW_MAJOR && layout != FYFFT_MODES_MODE_MAJOR) return FYFFT_INVALID_ARGUMENT;
  const bool mm = layout == FYFFT_MODES_MODE_MAJOR;
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!src || !dst || k < 1 || k > p->n) return FYFFT_INVALID_ARGUMENT;
  if (p->kind != kLds && p->kind != kFourStep) return FYFFT_UNSUPPORTED;
  const hipStream_t s = (hipStream_t)stream;
  if (k <= p->n / 2 + 1 && spectral_half_tables(p))
    return p->prec == kF64 ? exec_spectral_half<double2>(p, inverse, src, dst, k, scale0, scale, mm, s)
                           : exec_spectral_half<float2>(p, inverse, src, dst, k, scale0, scale, mm, s);
  if (p->kind != kLds) return FYFFT_UNSUPPORTED;
  if (p->prec == kF64)
    return exec_spectral<double2>(p, inverse, src, dst, k, scale0, scale, mm, s);
  return exec_spectral<float2>(p, inverse, src, dst, k, scale0, scale, mm, s);
}

// force: -1 the library's strategy; kBluestein plans a Bluestein convolution
// for any N > 1; kRader a Rader convolution for primes whose N - 1 is
// 7-smooth (other N keep the library's choice); the other strategies keep
// the library's choice (the caller checks eligibility).
static fyfft_status create_plan_impl(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision,
                                     int form, int force, const FsChoice* force_fs, size_t force_m) {
  if (!out) return FYFFT_INVALID_ARGUMENT;
  *out = nullptr;
  if (batch == 0) return FYFFT_INVALID_ARGUMENT;
  if (precision != FYFFT_F32 && precision != FYFFT_F64 && precision != FYFFT_BF16 &&
      precision != FYFFT_F16)
    return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  if (form != FYFFT_FORM_AUTO && form != FYFFT_FORM_PLAIN && form != FYFFT_FORM_TUNED && form != FYFFT_FORM_MFMA &&
      form != kFormTunedGeneric && form != kFormGroupedPlain)
    return FYFFT_INVALID_ARGUMENT;
  const bool grouped_sub = form == kFormGroupedPlain || form == kFormTunedGeneric;
  if (form == kFormGroupedPlain) form = FYFFT_FORM_PLAIN;
  const bool low = precision == FYFFT_BF16 || precision == FYFFT_F16;
  const bool tuned = form == FYFFT_FORM_TUNED || form == kFormTunedGeneric;
  if (form == FYFFT_FORM_MFMA && !low) return FYFFT_INVALID_ARGUMENT;
  int r0 = 1;
  dev::PassDesc lowp_pd{};
  // 16-bit storage: FP32 arithmetic in the compile-time kernels unless the
  // MFMA form is asked for (or the length has no compile-time kernel)
  const bool lowp = low && form != FYFFT_FORM_MFMA && lowp_fixed_supported(n, &lowp_pd);
  if (low && !lowp && !mfma16_supported(n, &r0)) return FYFFT_UNSUPPORTED;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  fyfft_plan* p = new (std::nothrow) fyfft_plan();
  if (!p) return FYFFT_ALLOC_FAILED;
  p->n = n;
  p->batch = batch;
  p->prec = precision == FYFFT_F64 ? kF64 : (precision == FYFFT_BF16 ? kBF16 : (precision == FYFFT_F16 ? kF16 : kF32));
  p->device = dev;
  p->form = form;
  p->grouped_sub = grouped_sub;
  p->force_fs = force_fs;
  p->force_m = force_m;
  FsChoice tuned_fs;
  p->strategy = strategy_for(n);
  fyfft_status st = FYFFT_SUCCESS;
  if (lowp) {
    p->kind = kLowp;
    p->lds.pd = lowp_pd;
    void* tw = nullptr;
    st = p->alloc(n * sizeof(float2), &tw);
    if (!st) {
      const hipError_t e = launch_gen_stockham_twiddles(kF32, tw, (int)n, lowp_pd, nullptr);
      if (e != hipSuccess) st = from_hip(e, "twiddles");
    }
    p->lds.tw = tw;
    p->lds.geom.n = (int)n;
    p->desc = std::string("lowp-fixed ") + (precision == FYFFT_BF16 ? "bf16" : "fp16") + " N=" +
              std::to_string(n) + " radices=" + radix_string(lowp_pd) + " FP32 arithmetic, 16-bit storage";
  } else if (low) {
    // FP32 Stockham table for the pass list [R0,] 16, ..., 16
    p->kind = kMfma;
    dev::PassDesc& pd = p->lds.pd;
    pd.npass = 0;
    if (r0 > 1) pd.radix[pd.npass++] = r0;
    for (size_t m = r0; m < n; m *= 16) pd.radix[pd.npass++] = 16;
    void* tw = nullptr;
    st = p->alloc(n * sizeof(float2), &tw);
    if (!st) {
      const hipError_t e = launch_gen_stockham_twiddles(kF32, tw, (int)n, pd, nullptr);
      if (e != hipSuccess) st = from_hip(e, "twiddles");
    }
    p->lds.tw = tw;
    p->lds.geom.n = (int)n;
    p->desc = std::string("mfma16 ") + (precision == FYFFT_BF16 ? "bf16" : "fp16") + " N=" +
              std::to_string(n) + " radices=" + radix_string(pd) + " split-hi/lo fp16 MFMA";
  } else if (n == 1) {
    p->kind = kTrivial;
    p->desc = "trivial N=1";
  } else if (force == kBluestein) {
    p->strategy = kBluestein;
    st = plan_bluestein(p);
  } else if (force == kRader && n >= 3 && is_prime(n) && is_smooth7(n - 1)) {
    p->strategy = kRader;
    st = plan_rader(p);
  } else if (is_smooth7(n)) {
    LdsGeom g;
    if (lds_geometry(n, elem_bytes(p->prec), lds_cap_bytes(dev), &g) && !lds_hi_to_fourstep(p, n, g)) {
      p->kind = kLds;
      st = tuned ? plan_lds_tuned(p, n, form == kFormTunedGeneric)
                 : plan_lds(p, n, &p->lds, form == FYFFT_FORM_AUTO, true, true);
      if (!st) {
        const LdsGeom& q = p->lds.geom;
        p->desc = std::string(p->lds.fixed >= 0 ? "lds-fixed" : "lds-generic") +
                  (tuned ? "(tuned)" : "") + " N=" +
                  std::to_string(n) + " radices=" + radix_string(p->lds.pd) + " T=" + std::to_string(q.threads) +
                  (fixed_kernel_split_w(p->lds.fixed) > 0
                       ? " split-exchange W=" + std::to_string(fixed_kernel_split_w(p->lds.fixed))
                       : " TPB=" + std::to_string(q.per_block) + " pad=" + std::to_string(q.pad_shift));
      }
    } else {
      p->kind = kFourStep;
      if (tuned && !p->force_fs) {
        st = tune_fourstep(p, n, kWfSmooth, &tuned_fs);
        if (!st) p->force_fs = &tuned_fs;
        else if (st == FYFFT_UNSUPPORTED) st = FYFFT_SUCCESS;  // no two-pass candidate: the three-pass form
      }
      if (!st) st = plan_fourstep(p, n, false);
      if (!st) st = alloc_workspace(p, n, false);
      if (!st)
        p->desc = p->n3 ? "fourstep N=" + std::to_string(n) + " = " + std::to_string(p->n1) + "(" +
                              radix_string(p->pdcol) + ") x " + std::to_string(p->n2) + "(" + radix_string(p->pdmid) +
                              ") x " + std::to_string(p->n3) + "(" + radix_string(p->pdrow) + ") three-pass C=" +
                              std::to_string(p->gcol.per_block) + "/" + std::to_string(p->gmid.per_block) + "/" +
                              std::to_string(p->grow.per_block) + " chunk=" + std::to_string(p->chunk)
                        : std::string(tuned ? "fourstep(tuned)" : "fourstep") + " N=" + std::to_string(n) + " = " +
                              std::to_string(p->n1) + "(" + radix_string(p->pdcol) + ") x " + std::to_string(p->n2) +
                              "(" + radix_string(p->pdrow) + ") C=" + std::to_string(p->gcol.per_block) + "/" +
                              std::to_string(p->grow.per_block) + " chunk=" + std::to_string(p->chunk);
    }
  } else if (p->strategy == kRader) {
    st = plan_rader(p);
  } else {
    st = plan_bluestein(p);
  }
  p->force_fs = nullptr;  // pointed at this function's candidate
  if (!st) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) st = from_hip(e, "plan table generation");
  }
  if (st) {
    delete p;
    return st;
  }
  *out = p;
  return FYFFT_SUCCESS;
}


// ------------------------------------------------------------ grouped plans
// Workgroup-size class of a grouped task: the smallest class whose thread
// count and LDS budget (cap * threads / 1024) hold it.
static size_t class_lds_cap(size_t cap, int k) { return cap * (size_t)kGroupClassThreads[k] / 1024; }

// Class and column count of a four-step sub-pass.  A workgroup's row
// segments are C elements wide; a strided FP32 copy with 128-byte segments
// ran at 60-65 % of HBM against 33-47 % with 64-byte ones, so 16 columns are
// taken in the smallest class that holds them (the two smaller classes
// only: a 1024-thread class holds one workgroup per CU), then 8; below
// that, the largest class takes whatever fits.
// The tile is followed by the LDS copy of the sub-length's Stockham table
// (*tl) when some class holds a one-column tile and the table; otherwise
// the table is read from global memory.
static void grouped_fs_class(const LdsGeom& q, size_t len, size_t elem, size_t cap, size_t roots, int* cls,
                             int* log2c, bool* tl) {
  // roots: bytes of root tables per sub-transform (twiddled column passes,
  // convolution row passes)
  const size_t per = (size_t)dev::lds_len((int)len, q.pad_shift) * elem + roots;
  const size_t twb = table_lds_bytes(len, elem);
  *tl = per + twb <= class_lds_cap(cap, kGroupClasses - 1);
  const size_t extra = *tl ? twb : 0;
  auto fits = [&](int c, int k) {
    return (size_t)c * q.threads <= (size_t)kGroupClassThreads[k] && c * per + extra <= class_lds_cap(cap, k);
  };
  for (int want : {16, 8})
    for (int k = 0; k < kGroupClasses; ++k) {
      if (want == 16 && k == kGroupClasses - 1) continue;
      if (fits(want, k)) {
        // more columns while the class's threads and LDS budget allow: the
        // persistent workgroups of a class otherwise leave threads idle
        // (N1 = 125: 16 columns of 8 threads use half of a 256-thread class)
        int c = want;
        while (c < 64 && fits(2 * c, k)) c *= 2;
        *cls = k;
        *log2c = ilog2_exact(c);
        return;
      }
    }
  int c = 16;
  while (c > 1 && !fits(c, kGroupClasses - 1)) c >>= 1;
  *cls = kGroupClasses - 1;
  *log2c = ilog2_exact(c);
}

// Sub-transforms per workgroup (a multiple of gran, up to 64) of a grouped
// four-step pass: the count that fills the most threads of a class within
// its LDS budget, the smallest class reaching 90 %; unchanged (the power of
// two of grouped_fs_class) where no class fits gran of them.  A convolution
// row of 300 points (19 threads) takes 13 rows in the 256-thread class (96 %
// of its threads) instead of 16 in the 512-thread class (59 %).
static void grouped_sub_count(const LdsGeom& q, size_t len, size_t elem, size_t cap, size_t roots, bool tl,
                              int gran, int* cls, int* nsub) {
  const size_t per = (size_t)dev::lds_len((int)len, q.pad_shift) * elem + roots;
  const size_t extra = tl ? table_lds_bytes(len, elem) : 0;
  double best = 0;
  for (int k = 0; k < kGroupClasses; ++k) {
    if ((size_t)gran * q.threads > (size_t)kGroupClassThreads[k] || gran * per + extra > class_lds_cap(cap, k))
      continue;
    int r = (int)std::min<size_t>({64, (size_t)kGroupClassThreads[k] / q.threads,
                                   (class_lds_cap(cap, k) - extra) / per});
    r -= r % gran;
    const double use = (double)r * q.threads / kGroupClassThreads[k];
    if (use > best + 1e-9) {
      best = use;
      *cls = k;
      *nsub = r;
    }
    if (best >= 0.9) break;
  }
}

template <typename V>
struct GLenPlan {
  GLen<V> g;
  int per_block = 1;                   // LDS kinds: transforms per task
  int lds_cls = 0, col_cls = 0, row_cls = 0;
  bool col_tl = false, row_tl = false;  // four-step: Stockham table in LDS after the tile
};

template <typename V>
static GLenPlan<V> make_glen(const fyfft_plan* s, size_t count_len, size_t cap, size_t io_elem) {
  GLenPlan<V> r;
  GLen<V>& g = r.g;
  g.n = (int)s->n;
  auto fill_lds = [&](size_t m) {
    const LdsGeom& q = s->lds.geom;
    g.lds.n = (int)s->n;
    g.lds.m = (int)m;
    g.lds.threads = q.threads;
    g.lds.ps = q.pad_shift;
    g.lds.stride = dev::lds_len((int)m, q.pad_shift);
    g.lds.pd = s->lds.pd;
    g.lds.tw = (const V*)s->lds.tw;
    g.lds_ppt = q.ppt;
    const size_t per = (size_t)g.lds.stride * sizeof(V);
    int k = 0;
    while (k < kGroupClasses - 1 &&
           (q.threads > kGroupClassThreads[k] || per > class_lds_cap(cap, k))) ++k;
    const size_t pb = std::min<size_t>(kGroupClassThreads[k] / q.threads, class_lds_cap(cap, k) / per);
    r.lds_cls = k;
    r.per_block = (int)std::max<size_t>(1, std::min(pb, count_len));
  };
  auto fill_fs = [&](FsArgs<V>& a, const LdsGeom& q, size_t len, const dev::PassDesc& pd, const void* tw,
                     int* cls, bool* tl) {
    a.n1 = (int)s->n1;
    a.n2 = (int)s->n2;
    // root tables per sub-transform: a row of roots per column (the family
    // kernels run 16 FP32 / 8 FP64 points per thread), per convolution row
    const bool conv_row = &a == &g.row && g.kind != kGkFsPlain;
    const size_t roots = (&a == &g.col ? fs_col_tw_row(sizeof(V) == 8 ? 16 : 8)
                          : conv_row ? fs_row_tw_len((int)len) : 0) * sizeof(V);
    grouped_fs_class(q, len, sizeof(V), cap, roots, cls, &a.log2c, tl);
    // then the count filling the class best: any count for convolution rows
    // (read and written in place), multiples of 16 for columns and plain
    // rows, whose segments stay whole 128-byte lines (ragged step, 20000
    // lengths: powers of two 5.48 ms, multiples of 16 5.26, of 8 5.28, of 4
    // 5.37, any 5.57)
    grouped_sub_count(q, len, sizeof(V), cap, roots, *tl, conv_row ? 1 : 16, cls, &a.nsub);
    a.threads = q.threads;
    a.ps = q.pad_shift;
    a.stride = dev::lds_len((int)len, q.pad_shift);
    a.pd = pd;
    a.tw = (const V*)tw;
    a.tw_hi = (const V*)s->tw_hi;
    a.tw_lo = (const V*)s->tw_lo;
    a.lo_bits = s->lo_bits;
    a.inv_n = fs_inv_n(s->n1 * s->n2);
    // the Rader closing scatter of a grouped plan sweeps its destinations in
    // windows of 2^14 elements (ragged step 5.37 -> 5.27 ms; the batched
    // plans' bucketed permutation ran as fast as windows of 2^15, 2.60 vs
    // 2.56 ms, and keeps the two-pass form)
    a.win_bits = s->kind == kRaderFs ? 14 : 0;
    a.n_orig = (int)s->n;
    a.perm = s->perm_in;
    a.chirp = (const V*)s->chirp[0];
    a.kern = (const V*)s->kern[0];
  };
  g.lds.perm_in = s->perm_in;
  g.lds.perm_out = s->perm_out;
  g.lds.pos_out = s->pos_out;
  g.lds.chirp = (const V*)s->chirp[0];
  g.lds.kern = (const V*)s->kern[0];
  g.chirp_inv = (const V*)s->chirp[1];
  g.kern_inv = (const V*)s->kern[1];
  switch (s->kind) {
    case kTrivial:
      g.kind = kGkLds;
      g.lds.n = g.lds.m = 1;
      g.lds.threads = 1;
      g.lds.ps = 30;
      g.lds.stride = 1;
      g.lds.pd.npass = 0;
      r.lds_cls = 0;
      r.per_block = (int)std::max<size_t>(1, std::min<size_t>(kGroupClassThreads[0], count_len));
      break;
    case kLds: g.kind = kGkLds; fill_lds(s->n); break;
    case kRaderLds: g.kind = kGkRader; fill_lds(s->m); break;
    case kBlueLds: g.kind = kGkBlue; fill_lds(s->m); break;
    default:
      g.kind = s->kind == kFourStep ? kGkFsPlain : (s->kind == kRaderFs ? kGkFsRader : kGkFsBlue);
      fill_fs(g.col, s->gcol, s->n1, s->pdcol, s->twcol, &r.col_cls, &r.col_tl);
      fill_fs(g.row, s->grow, s->n2, s->pdrow, s->twrow, &r.row_cls, &r.row_tl);
      g.row.stage_rows = rows_misaligned(s->n2, sizeof(V));
      g.col.col_cached = rows_misaligned(s->n2, io_elem);
      g.row.out_cached = rows_misaligned(s->n1, io_elem);
      g.col_ppt = s->gcol.ppt;
      g.row_ppt = s->grow.ppt;
      break;
  }
  return r;
}

template <typename V>
static GroupedArgs<V> grouped_args(const fyfft_plan* p, const void* in, void* out, int direction) {
  GroupedArgs<V> a;
  a.lens = (const GLen<V>*)p->d_lens;
  a.off = (const long long*)p->d_off;
  a.woff = (const long long*)p->d_woff;
  a.in = in;
  a.out = out;
  a.ws = (V*)p->gws;
  a.x0buf = (V*)p->gx0;
  a.tables = direction != FYFFT_FORWARD;
  a.normalize = direction == FYFFT_INVERSE;
  return a;
}

template <typename V>
static hipError_t launch_list(const fyfft_plan* p, const GList& L, GroupedArgs<V>& a, int dir, hipStream_t st) {
  a.tasks = (const GTask*)p->d_tasks + L.base;
  for (int x = 0; x <= dev::kXcds; ++x) a.xbase[x] = L.xbase[x];
  return group_family(p->prec, false, L.fam, dir, L.hi != 0, kGroupClassThreads[L.cls], &a, L.grid, L.lds, st);
}

// Times every family launch of a grouped plan on its own (forward, on a
// zeroed scratch copy of the data, one warm-up round then the best of two),
// in phase order so each reads what the earlier phases wrote.  Leaves
// GList::ms at 0 when anything fails (scheduling then uses the estimates).
template <typename V>
static void time_grouped_lists(fyfft_plan* p) {
  const size_t bytes = std::max<size_t>(1, p->total) * elem_bytes(p->prec);
  void* buf = nullptr;
  if (hipMalloc(&buf, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
            hipMemsetAsync(buf, 0, bytes, st) == hipSuccess;
  GroupedArgs<V> a = grouped_args<V>(p, buf, buf, FYFFT_FORWARD);
  std::vector<size_t> order(p->glists.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
    return family_phase(p->glists[x].fam) < family_phase(p->glists[y].fam);
  });
  std::vector<double> best(order.size(), 1e30);
  for (int rep = 0; ok && rep < 3; ++rep)
    for (size_t i : order) {
      float ms = 0;
      ok = ok && hipEventRecord(e0, st) == hipSuccess && launch_list(p, p->glists[i], a, -1, st) == hipSuccess &&
           hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
           hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      if (ok && rep > 0) best[i] = std::min(best[i], (double)ms);
    }
  if (ok)
    for (size_t i = 0; i < order.size(); ++i) p->glists[i].ms = std::max(best[i], 1e-3);
  else
    (void)hipGetLastError();
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(buf);
}

// Places the family launches on phases and branches: four-step passes keep
// their phase (column passes, row passes, closing passes), longest first
// onto the least loaded branch; single-pass families (LDS-resident plain,
// Rader and Bluestein transforms) depend on no other launch, so each goes,
// longest first, to the phase and branch where it lengthens the sum of the
// phase spans least (the least loaded such branch on a tie).  Weights are
// the plan-time measurements when every launch has one, else the estimates.
// The launches are then ordered by phase, longest first.
static void schedule_branches(fyfft_plan* p) {
  bool measured = !p->glists.empty();
  for (const GList& L : p->glists) measured = measured && L.ms > 0;
  auto w = [&](const GList& L) { return measured ? L.ms : L.cost; };
  const int nb = p->gbranch[0] ? kGroupBranches : 1;
  double load[3][kGroupBranches] = {};
  bool bound_phase[3] = {false, false, false};
  std::vector<GList*> bound, free_lists;
  for (GList& L : p->glists) {
    L.phase = family_phase(L.fam);
    (L.fam <= kGfBlue ? free_lists : bound).push_back(&L);
  }
  auto heavier = [&](const GList* x, const GList* y) { return w(*x) > w(*y); };
  std::stable_sort(bound.begin(), bound.end(), heavier);
  std::stable_sort(free_lists.begin(), free_lists.end(), heavier);
  for (GList* L : bound) {
    const int b = (int)(std::min_element(load[L->phase], load[L->phase] + nb) - load[L->phase]);
    load[L->phase][b] += w(*L);
    L->branch = b;
    bound_phase[L->phase] = true;
  }
  if (!bound_phase[1] && !bound_phase[2]) bound_phase[0] = true;
  for (GList* L : free_lists) {
    int bp = 0, bb = 0;
    double best_d = 1e300, best_l = 1e300;
    for (int ph = 0; ph < 3; ++ph) {
      if (!bound_phase[ph]) continue;
      const double span = *std::max_element(load[ph], load[ph] + nb);
      for (int b = 0; b < nb; ++b) {
        const double l = load[ph][b] + w(*L), d = std::max(span, l) - span;
        if (d < best_d || (d == best_d && l < best_l)) best_d = d, best_l = l, bp = ph, bb = b;
      }
    }
    load[bp][bb] += w(*L);
    L->phase = bp;
    L->branch = bb;
  }
  std::stable_sort(p->glists.begin(), p->glists.end(), [&](const GList& x, const GList& y) {
    return x.phase != y.phase ? x.phase < y.phase : w(x) > w(y);
  });
  p->gspan_ms = 0;
  if (measured)
    for (int ph = 0; ph < 3; ++ph) p->gspan_ms += *std::max_element(load[ph], load[ph] + nb);
}

template <typename V>
static fyfft_status build_grouped(fyfft_plan* p, const std::map<size_t, std::vector<size_t>>& by_len) {
  const size_t cap = lds_cap_bytes(p->device), elem = sizeof(V);
  const double eio = (double)elem_bytes(p->prec);  // storage bytes per element of the data
  constexpr int kLo = sizeof(V) == 8 ? 16 : 8;  // points-per-thread bound of the lo family kernels
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess || ncu <= 0) ncu = 256;
  std::vector<GLen<V>> lens;
  std::vector<long long> off, woff;
  struct T { GTask t; double cost; double bytes; };
  // (family, class, hi) -> tasks, LDS
  std::map<std::tuple<int, int, int>, std::vector<T>> lists;
  std::map<std::tuple<int, int, int>, size_t> list_lds;
  auto add = [&](int fam, int cls, int hi, const T& t, size_t lds) {
    const auto key = std::make_tuple(fam, cls, hi);
    lists[key].push_back(t);
    list_lds[key] = std::max(list_lds[key], lds);
  };
  long long wsize = 0;
  int li = 0;
  for (const auto& kv : by_len) {
    const fyfft_plan* s = p->subs[li];
    const std::vector<size_t>& idx = kv.second;
    const GLenPlan<V> lp = make_glen<V>(s, idx.size(), cap, elem_bytes(p->prec));
    lens.push_back(lp.g);
    const GLen<V>& g = lens.back();
    const int first = (int)off.size();
    for (size_t i : idx) {
      off.push_back((long long)p->h_off[i]);
      const bool fs = g.kind >= kGkFsPlain;
      woff.push_back(fs ? wsize : 0);
      if (fs) wsize += (long long)(g.col.n1) * g.col.n2;
    }
    const double lg = std::log2((double)std::max<size_t>(2, kv.first));
    if (g.kind <= kGkBlue) {
      const size_t m = (size_t)g.lds.m;
      const double conv = g.kind == kGkLds ? 1.0 : 2.0;
      const int fam = g.kind == kGkLds ? kGfLds : (g.kind == kGkRader ? kGfRader : kGfBlue);
      const int hi = g.lds_ppt > kLo, k = lp.lds_cls, pb = lp.per_block;
      if (hi && kGroupClassThreads[k] != 1024) return FYFFT_UNSUPPORTED;
      for (size_t i = 0; i < idx.size(); i += pb) {
        const int c = (int)std::min<size_t>(pb, idx.size() - i);
        const double by = (double)c * (2.0 * g.n * eio + (g.kind == kGkLds ? 0.0 : (double)m * elem));
        add(fam, k, hi, T{GTask{li, 0, first + (int)i, c}, conv * c * m * lg, by}, (size_t)pb * g.lds.stride * elem);
      }
    } else {
      if (g.col_ppt > kLo || g.row_ppt > kLo) return FYFFT_UNSUPPORTED;
      const int C = g.col.nsub ? g.col.nsub : 1 << g.col.log2c, R = g.row.nsub ? g.row.nsub : 1 << g.row.log2c;
      const double conv = g.kind == kGkFsPlain ? 1.0 : 2.0;
      const int kc = lp.col_cls, kr = lp.row_cls;
      const int famA = g.kind == kGkFsPlain ? kGfColPlain : (g.kind == kGkFsRader ? kGfColRaderA : kGfColBlueA);
      const int famB = g.kind == kGkFsPlain ? kGfRowPlain : kGfRowConv;
      const int famC = g.kind == kGkFsRader ? kGfColRaderC : kGfColBlueC;
      // hi = 1: the global-table form of the family (no room for the table)
      const int chi = lp.col_tl ? 0 : 1, rhi = lp.row_tl ? 0 : 1;
      const size_t col_lds = (size_t)C * g.col.stride * elem + (lp.col_tl ? table_lds_bytes(g.col.n1, elem) : 0);
      const size_t colA_lds = col_lds + (size_t)C * fs_col_tw_row(kLo) * elem;
      const size_t row_lds = (size_t)R * g.row.stride * elem + (lp.row_tl ? table_lds_bytes(g.col.n2, elem) : 0) +
                             (g.kind != kGkFsPlain ? (size_t)R * fs_row_tw_len(g.col.n2) * elem : 0);
      for (size_t i = 0; i < idx.size(); ++i) {
        for (int c0 = 0; c0 < g.col.n2; c0 += C) {
          const double frac = (double)g.n / ((double)g.col.n1 * g.col.n2);  // data share of the conv length
          add(famA, kc, chi, T{GTask{li, 0, first + (int)i, c0}, (double)C * g.col.n1 * lg,
                               (double)C * g.col.n1 * (frac * eio + elem)}, colA_lds);
          if (g.kind != kGkFsPlain)
            add(famC, kc, chi, T{GTask{li, 2, first + (int)i, c0}, (double)C * g.col.n1 * lg,
                                 (double)C * g.col.n1 * (frac * eio + elem)}, col_lds);
        }
        for (int r0 = 0; r0 < g.col.n1; r0 += R)
          add(famB, kr, rhi, T{GTask{li, 1, first + (int)i, r0}, conv * R * g.col.n2 * lg,
                             (double)R * g.col.n2 * (g.kind == kGkFsPlain ? elem + eio : 3.0 * elem)}, row_lds);
      }
    }
    ++li;
  }
  // Bin-pack each family list onto the XCDs: all tasks of one transform
  // (four-step groups) stay together, greedy by decreasing group cost onto
  // the least loaded XCD; within an XCD, tasks by decreasing cost.
  std::vector<GTask> tasks;
  p->glists.clear();
  for (const auto& kv : lists) {
    const std::vector<T>& list = kv.second;
    GList L;
    L.fam = std::get<0>(kv.first);
    L.cls = std::get<1>(kv.first);
    L.hi = std::get<2>(kv.first);
    std::map<std::pair<int, int>, std::vector<T>> groups;  // (len, transform or task) -> tasks
    for (size_t i = 0; i < list.size(); ++i) {
      const T& t = list[i];
      const bool fs = lens[t.t.len].kind >= kGkFsPlain;
      groups[{t.t.len, fs ? t.t.first : -1 - (int)i}].push_back(t);
    }
    std::vector<std::pair<double, std::vector<T>*>> order;
    for (auto& g : groups) {
      double c = 0;
      for (const T& t : g.second) c += t.cost, L.bytes += t.bytes;
      order.push_back({c, &g.second});
      L.cost += c;
    }
    std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
    std::vector<T> per_x[dev::kXcds];
    double load[dev::kXcds] = {0};
    for (auto& g : order) {
      const int x = (int)(std::min_element(load, load + dev::kXcds) - load);
      load[x] += g.first;
      per_x[x].insert(per_x[x].end(), g.second->begin(), g.second->end());
    }
    L.base = (int)tasks.size();
    L.count = (int)list.size();
    for (int x = 0; x < dev::kXcds; ++x) {
      std::stable_sort(per_x[x].begin(), per_x[x].end(), [](const T& u, const T& v) { return u.cost > v.cost; });
      L.xbase[x] = (int)tasks.size() - L.base;
      for (const T& t : per_x[x]) tasks.push_back(t.t);
    }
    L.xbase[dev::kXcds] = (int)tasks.size() - L.base;
    L.lds = std::max<size_t>(list_lds[kv.first], 16);
    // resident workgroups per CU: 1024 / class threads (the class LDS budget
    // keeps them within the CU); the grid is a multiple of the XCD count
    const size_t per_cu = 1024 / (size_t)kGroupClassThreads[L.cls];
    size_t grid = std::min<size_t>(list.size(), (size_t)ncu * per_cu);
    grid = std::max<size_t>(dev::kXcds, (grid + dev::kXcds - 1) / dev::kXcds * dev::kXcds);
    L.grid = (unsigned)grid;
    p->glists.push_back(L);
  }
  std::stable_sort(p->glists.begin(), p->glists.end(), [](const GList& x, const GList& y) {
    const int px = family_phase(x.fam), py = family_phase(y.fam);
    return px != py ? px < py : x.cost > y.cost;
  });
  fyfft_status st;
  if ((st = p->alloc(std::max<size_t>(1, tasks.size()) * sizeof(GTask), &p->d_tasks)) ||
      (st = p->alloc(lens.size() * sizeof(GLen<V>), &p->d_lens)) ||
      (st = p->alloc(off.size() * sizeof(long long), &p->d_off)) ||
      (st = p->alloc(woff.size() * sizeof(long long), &p->d_woff)) ||
      (st = p->alloc(std::max<long long>(1, wsize) * elem, &p->gws)) ||
      (st = p->alloc(p->count * elem, &p->gx0)))
    return st;
  HIP_TRY(hipMemcpy(p->d_tasks, tasks.data(), tasks.size() * sizeof(GTask), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->d_lens, lens.data(), lens.size() * sizeof(GLen<V>), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->d_off, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(p->d_woff, woff.data(), woff.size() * sizeof(long long), hipMemcpyHostToDevice));
  for (const GList& L : p->glists)
    for (int dir : {-1, +1})
      HIP_TRY(group_family(p->prec, true, L.fam, dir, L.hi != 0, kGroupClassThreads[L.cls], nullptr, 0, L.lds,
                           nullptr));
  // branch streams and fork/join events when there are several launches
  // (default priority: streams of other priorities have hung these GPUs)
  if (p->glists.size() > 1 && !(p->gflags & FYFFT_GROUPED_PER_TRANSFORM)) {
    for (hipStream_t& b : p->gbranch) HIP_TRY(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
    for (int ph = 0; ph < 3; ++ph) {
      HIP_TRY(hipEventCreateWithFlags(&p->gfork[ph], hipEventDisableTiming));
      for (hipEvent_t& e : p->gjoin[ph]) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
  }
  if (p->gbranch[0]) time_grouped_lists<V>(p);
  schedule_branches(p);
  if (const char* rep = std::getenv("FYFFT_GROUPED_REPORT"); rep && *rep == '1') {
    // one line per family launch: its plan-time measurement alone on the GPU
    for (const GList& L : p->glists)
      std::fprintf(stderr, "grouped-list fam=%d cls=%d hi=%d tasks=%d grid=%u lds=%zu phase=%d branch=%d "
                   "ms=%.4f MB=%.1f GB/s=%.0f\n", L.fam, kGroupClassThreads[L.cls], L.hi, L.count, L.grid, L.lds,
                   L.phase, L.branch, L.ms, L.bytes / 1e6, L.ms > 0 ? L.bytes / L.ms / 1e6 : 0.0);
  }
  size_t ntask[3] = {0, 0, 0};
  for (const GList& L : p->glists) ntask[family_phase(L.fam)] += (size_t)L.count;
  static const char* kStore[] = {"f32", "f64", "bf16", "fp16"};
  p->desc = "grouped count=" + std::to_string(p->count) + " lengths=" + std::to_string(lens.size()) +
            " total=" + std::to_string(p->total) + " storage=" + kStore[p->prec] + " tasks=" +
            std::to_string(ntask[0]) + "/" + std::to_string(ntask[1]) + "/" + std::to_string(ntask[2]) +
            " family-launches=" + std::to_string(p->glists.size()) + " workspace=" + std::to_string(wsize);
  if (p->gspan_ms > 0) {
    int per_phase[3] = {0, 0, 0};
    for (const GList& L : p->glists) ++per_phase[L.phase];
    char buf[96];
    std::snprintf(buf, sizeof(buf), " schedule=measured phases=%d/%d/%d est-ms=%.3f", per_phase[0], per_phase[1],
                  per_phase[2], p->gspan_ms);
    p->desc += buf;
  }
  return FYFFT_SUCCESS;
}

template <typename V>
static fyfft_status launch_grouped_phases(const fyfft_plan* p, const void* in, void* out, int direction,
                                          hipStream_t s) {
  GroupedArgs<V> a = grouped_args<V>(p, in, out, direction);
  const int dir = direction == FYFFT_FORWARD ? -1 : +1;
  for (int ph = 0; ph < 3; ++ph) {
    bool used[kGroupBranches] = {};
    int nused = 0;
    for (const GList& L : p->glists)
      if (L.phase == ph && !used[L.branch]) used[L.branch] = true, ++nused;
    if (nused == 0) continue;
    if (nused == 1 || !p->gbranch[0]) {
      for (const GList& L : p->glists)
        if (L.phase == ph) HIP_TRY(launch_list(p, L, a, dir, s));
      continue;
    }
    // fork: the phase's launches wait for the previous phase and run on
    // their scheduled branches (longest first within a branch); the stream
    // waits for every branch before the next phase
    HIP_TRY(hipEventRecord(p->gfork[ph], s));
    for (int b = 0; b < kGroupBranches; ++b)
      if (used[b]) HIP_TRY(hipStreamWaitEvent(p->gbranch[b], p->gfork[ph], 0));
    for (const GList& L : p->glists)
      if (L.phase == ph) HIP_TRY(launch_list(p, L, a, dir, p->gbranch[L.branch]));
    for (int b = 0; b < kGroupBranches; ++b)
      if (used[b]) {
        HIP_TRY(hipEventRecord(p->gjoin[ph][b], p->gbranch[b]));
        HIP_TRY(hipStreamWaitEvent(s, p->gjoin[ph][b], 0));
      }
  }
  return FYFFT_SUCCESS;
}

static fyfft_status exec_grouped(const fyfft_plan* p, const void* in, void* out, int direction, hipStream_t s) {
  const size_t elem = elem_bytes(p->prec);
  const bool low = p->prec == kBF16 || p->prec == kF16;
  if (p->gflags & FYFFT_GROUPED_PER_TRANSFORM) {
    // the baseline: one launch of the plain kernels per transform; 16-bit
    // data is widened into an FP32 copy around each one
    for (size_t i = 0; i < p->count; ++i) {
      const fyfft_plan* sp = p->subs[p->sub_of[i]];
      const char* x = (const char*)in + p->h_off[i] * elem;
      char* y = (char*)out + p->h_off[i] * elem;
      if (low) {
        HIP_TRY(launch_convert16(p->prec == kBF16, true, x, p->gstage, (long long)sp->n, s));
        fyfft_status st = fyfft_execute(sp, p->gstage, p->gstage, direction, s);
        if (st) return st;
        HIP_TRY(launch_convert16(p->prec == kBF16, false, p->gstage, y, (long long)sp->n, s));
      } else {
        const fyfft_status st = fyfft_execute(sp, x, y, direction, s);
        if (st) return st;
      }
    }
    return FYFFT_SUCCESS;
  }
  auto run = [&]() {
    return p->prec == kF64 ? launch_grouped_phases<double2>(p, in, out, direction, s)
                           : launch_grouped_phases<float2>(p, in, out, direction, s);
  };
  if (!(p->gflags & FYFFT_GROUPED_GRAPH) || s == nullptr) return run();
  std::lock_guard<std::mutex> lock(p->graph_mu);
  const auto key = std::make_tuple(in, out, direction, s);
  auto it = p->graphs.find(key);
  if (it == p->graphs.end()) {
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const fyfft_status st = run();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (st) {
      if (graph) (void)hipGraphDestroy(graph);
      return st;
    }
    if (e != hipSuccess) return from_hip(e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) return from_hip(e2, "hipGraphInstantiate");
    it = p->graphs.emplace(key, exec).first;
  }
  HIP_TRY(hipGraphLaunch(it->second, s));
  return FYFFT_SUCCESS;
}

// Execution over the first `rows` (<= batch) transforms of a plan.
static fyfft_status exec_real(const fyfft_plan* p, const void* in, void* out, int direction, size_t rows,
                              hipStream_t s);

static fyfft_status exec_strided(const fyfft_plan* p, const void* in, void* out, int direction, hipStream_t s);

// dist: elements between consecutive transforms (0: the length n).  Only the
// plain batched kinds accept dist != n.
static fyfft_status execute_rows(const fyfft_plan* p, const void* in, void* out, int direction, size_t rows,
                                 hipStream_t s, size_t dist = 0, const Workspace* wsp = nullptr) {
  if (dist == 0) dist = p->n;
  if (dist != p->n && p->kind != kTrivial && p->kind != kLds && p->kind != kFourStep && p->kind != kRaderFs &&
      p->kind != kBlueFs && p->kind != kRaderLds && p->kind != kBlueLds)
    return FYFFT_UNSUPPORTED;
  if (!in || !out) return FYFFT_INVALID_ARGUMENT;
  if (direction != FYFFT_FORWARD && direction != FYFFT_INVERSE && direction != FYFFT_ADJOINT)
    return FYFFT_INVALID_ARGUMENT;
  const double scale = direction == FYFFT_INVERSE ? 1.0 / (double)p->n : 1.0;
  const int dir = direction == FYFFT_FORWARD ? kForward : (direction == FYFFT_INVERSE ? kInverse : kAdjoint);
  const bool inv_tables = direction != FYFFT_FORWARD;
  switch (p->kind) {
    case kTrivial:
      if (in != out) {
        const size_t e = elem_bytes(p->prec);
        HIP_TRY(hipMemcpy2DAsync(out, dist * e, in, dist * e, e, rows, hipMemcpyDeviceToDevice, s));
      }
      return FYFFT_SUCCESS;
    case kLds:
      HIP_TRY(launch_lds(p->prec, dir, p->lds, in, out, rows, dist, scale, s));
      return FYFFT_SUCCESS;
    case kFourStep:
    case kRaderFs:
    case kBlueFs: {
      const int d = dir == kForward ? -1 : +1;
      const Workspace ws = wsp ? *wsp : plan_workspace(p);
      if (p->kind == kFourStep && p->n3) {
        if (p->prec == kF64)
          return exec_fourstep3<double2>(p, (const double2*)in, (double2*)out, d, scale, rows, dist, ws, s);
        return exec_fourstep3<float2>(p, (const float2*)in, (float2*)out, d, scale, rows, dist, ws, s);
      }
      if (p->prec == kF64)
        return exec_fourstep<double2>(p, (const double2*)in, (double2*)out, d, inv_tables, scale, rows, dist, ws, s);
      return exec_fourstep<float2>(p, (const float2*)in, (float2*)out, d, inv_tables, scale, rows, dist, ws, s);
    }
    case kRaderLds:
    case kBlueLds:
      if (p->prec == kF64)
        return exec_conv_lds<double2>(p, (const double2*)in, (double2*)out, inv_tables, scale, rows, dist, s);
      return exec_conv_lds<float2>(p, (const float2*)in, (float2*)out, inv_tables, scale, rows, dist, s);
    case kGrouped:
      return exec_grouped(p, in, out, direction, s);
    case kMfma:
      HIP_TRY(launch_mfma16(p->prec == kBF16, dir == kForward ? -1 : +1, p->n, in, out, rows,
                            p->lds.tw, scale, s));
      return FYFFT_SUCCESS;
    case kLowp:
      HIP_TRY(launch_lowp_fixed(p->prec == kBF16, dir == kForward ? -1 : +1, p->n, in, out, rows,
                                p->lds.tw, scale, s));
      return FYFFT_SUCCESS;
    case kReal:
      return exec_real(p, in, out, direction, rows, s);
    case kStrided:
      return exec_strided(p, in, out, direction, s);
  }
  return FYFFT_INVALID_PLAN;
}

// ------------------------------------------------------------ real plans
template <typename V>
static fyfft_status exec_real_fused(const fyfft_plan* p, const void* in, void* out, bool r2c, size_t rows,
                                    hipStream_t s) {
  const fyfft_plan* sp = p->subs[0];
  const bool half = p->real_form == kRealFusedHalf;
  const size_t len = half ? p->n / 2 : p->n;
  RealArgs<V> a;
  a.in = in;
  a.out = out;
  a.rows = (long long)rows;
  a.n = (int)p->n;
  const LdsGeom& g = sp->lds.geom;
  a.threads = g.threads;
  a.per_block = g.per_block;
  a.ps = g.pad_shift;
  a.stride = dev::lds_len((int)len, g.pad_shift);
  a.pd = sp->lds.pd;
  a.tw = (const V*)sp->lds.tw;
  a.rtw = (const V*)p->rtw;
  a.scale = (dev::S<V>)(r2c ? 1.0 : 1.0 / (double)len);
  const int mode = half ? (r2c ? kRealR2cHalf : kRealC2rHalf) : (r2c ? kRealR2cFull : kRealC2rFull);
  HIP_TRY(launch_real_lds(p->prec, mode, &a, g, sp->lds.fixed >= 0 ? (int)len : 0, s));
  return FYFFT_SUCCESS;
}

static fyfft_status exec_real(const fyfft_plan* p, const void* in, void* out, int direction, size_t rows,
                              hipStream_t s) {
  if (direction == FYFFT_ADJOINT) {
    g_last_error = "real plans run FYFFT_FORWARD (R2C) and FYFFT_INVERSE (C2R) only";
    return FYFFT_UNSUPPORTED;
  }
  const bool r2c = direction == FYFFT_FORWARD;
  if (p->real_form == kRealFusedHalf || p->real_form == kRealFusedFull)
    return p->prec == kF64 ? exec_real_fused<double2>(p, in, out, r2c, rows, s)
                           : exec_real_fused<float2>(p, in, out, r2c, rows, s);
  const fyfft_plan* sp = p->subs[0];
  const size_t n = p->n, h = n / 2, hk = n / 2 + 1;
  const size_t cplx = elem_bytes(p->prec), real = cplx / 2;
  const bool half = p->real_form == kRealHalfPasses;
  for (size_t b0 = 0; b0 < rows; b0 += p->rchunk) {
    const long long cb = (long long)std::min(p->rchunk, rows - b0);
    if (r2c) {
      const char* x = (const char*)in + b0 * n * real;
      char* X = (char*)out + b0 * hk * cplx;
      if (half) {
        fyfft_status st = execute_rows(sp, x, p->rws, FYFFT_FORWARD, (size_t)cb, s);
        if (st) return st;
        HIP_TRY(launch_real_pass(p->prec, kRealOpPost, p->rws, X, cb, (int)n, p->rtw, 1.0, s));
      } else {
        HIP_TRY(launch_real_pass(p->prec, kRealOpExpand, x, p->rws, cb, (int)n, nullptr, 1.0, s));
        fyfft_status st = execute_rows(sp, p->rws, p->rws, FYFFT_FORWARD, (size_t)cb, s);
        if (st) return st;
        HIP_TRY(launch_real_pass(p->prec, kRealOpTruncate, p->rws, X, cb, (int)n, nullptr, 1.0, s));
      }
    } else {
      const char* X = (const char*)in + b0 * hk * cplx;
      char* y = (char*)out + b0 * n * real;
      if (half) {
        HIP_TRY(launch_real_pass(p->prec, kRealOpPre, X, p->rws, cb, (int)n, p->rtw, 1.0, s));
        fyfft_status st = execute_rows(sp, p->rws, y, FYFFT_INVERSE, (size_t)cb, s);
        if (st) return st;
      } else {
        HIP_TRY(launch_real_pass(p->prec, kRealOpHermExpand, X, p->rws, cb, (int)n, nullptr, 1.0, s));
        fyfft_status st = execute_rows(sp, p->rws, p->rws, FYFFT_INVERSE, (size_t)cb, s);
        if (st) return st;
        HIP_TRY(launch_real_pass(p->prec, kRealOpRealPart, p->rws, y, cb, (int)n, nullptr, 1.0, s));
      }
    }
    (void)h;
  }
  return FYFFT_SUCCESS;
}

static fyfft_status create_real_impl(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision) {
  if (!out) return FYFFT_INVALID_ARGUMENT;
  *out = nullptr;
  if (batch == 0) return FYFFT_INVALID_ARGUMENT;
  if (precision == FYFFT_BF16 || precision == FYFFT_F16) return FYFFT_UNSUPPORTED;
  if (precision != FYFFT_F32 && precision != FYFFT_F64) return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  fyfft_plan* p = new (std::nothrow) fyfft_plan();
  if (!p) return FYFFT_ALLOC_FAILED;
  p->kind = kReal;
  p->n = n;
  p->batch = batch;
  p->prec = precision == FYFFT_F64 ? kF64 : kF32;
  p->device = dev;
  p->strategy = strategy_for(n);
  const size_t elem = elem_bytes(p->prec), cap = lds_cap_bytes(dev), h = n / 2;
  LdsGeom g;
  // (the complex sub-plan takes the four-step form where a batched plan of
  // its length would, see lds_hi_to_fourstep)
  if (n % 2 == 0 && h >= 2 && is_smooth7(h) && lds_geometry(h, elem, cap, &g) && !lds_hi_to_fourstep(p, h, g))
    p->real_form = kRealFusedHalf;
  else if (n % 2 == 0 && h >= 2 && is_smooth7(h) && fs_possible(h, elem, cap)) p->real_form = kRealHalfPasses;
  else if (n >= 2 && is_smooth7(n) && lds_geometry(n, elem, cap, &g) && !lds_hi_to_fourstep(p, n, g))
    p->real_form = kRealFusedFull;
  else p->real_form = kRealFullPasses;
  const bool half = p->real_form == kRealFusedHalf || p->real_form == kRealHalfPasses;
  const bool fused = p->real_form == kRealFusedHalf || p->real_form == kRealFusedFull;
  const size_t len = half ? h : n;
  p->rchunk = fused ? batch : std::min(batch, std::max<size_t>(1, kWorkspaceCap / (len * elem)));
  fyfft_plan* sp = nullptr;
  fyfft_status st = create_plan_impl(&sp, len, fused ? 1 : p->rchunk, precision, FYFFT_FORM_AUTO);
  if (!st) {
    p->subs.push_back(sp);
    if (fused && sp->kind != kLds) st = FYFFT_UNSUPPORTED;
    // the real kernels have no lean small-radix form: default geometry
    if (!st && fused && sp->lds.geom.small_radix && !lds_geometry(len, elem, cap, &sp->lds.geom, true, false))
      st = FYFFT_UNSUPPORTED;
  }
  if (!st && half) {
    st = p->alloc((h + 1) * elem, &p->rtw);
    if (!st) {
      const hipError_t e = launch_gen_roots(p->prec, p->rtw, (long long)(h + 1), 1, (long long)n, nullptr);
      if (e != hipSuccess) st = from_hip(e, "real twiddles");
    }
  }
  if (!st && !fused) st = p->alloc(p->rchunk * len * elem, &p->rws);
  if (!st) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) st = from_hip(e, "real plan tables");
  }
  if (st) {
    delete p;
    return st;
  }
  static const char* kForms[] = {"fused half-length LDS", "fused full-length LDS",
                                 "half-length transform + pre/post pass", "full-length transform + expand pass"};
  // the fused forms run their own kernels on the sub-plan's geometry and
  // tables: describe those rather than the sub-plan's C2C kernel
  const LdsGeom& q = sp->lds.geom;
  const std::string sub =
      fused ? std::string(sp->lds.fixed >= 0 ? "compile-time" : "runtime") + " passes N=" + std::to_string(len) +
                  " radices=" + radix_string(sp->lds.pd) + " T=" + std::to_string(q.threads) +
                  " TPB=" + std::to_string(q.per_block) + " pad=" + std::to_string(q.pad_shift)
            : sp->desc;
  p->desc = "real N=" + std::to_string(n) + " form=" + kForms[p->real_form] + " [" + sub + "]";
  *out = p;
  return FYFFT_SUCCESS;
}

// ------------------------------------------------------------ strided plans
// Plans one pass: rows when each transform is contiguous and the transforms
// sit at one distance; columns when transforms are adjacent (cdist 1) and
// fit one workgroup's LDS; otherwise gather into a contiguous workspace,
// transform there and scatter back.
static fyfft_status plan_strided_pass(fyfft_plan* p, StridedPass& sp, fyfft_precision precision,
                                      size_t* ws_elems) {
  const size_t elem = elem_bytes(p->prec), cap = lds_cap_bytes(p->device);
  const size_t count = sp.ncols * sp.planes;
  const bool rows = sp.in.stride == 1 && sp.out.stride == 1 && sp.in.cdist == sp.out.cdist &&
                    sp.in.cdist >= (long long)sp.n &&
                    (sp.planes == 1 || (sp.in.pdist == (long long)sp.ncols * sp.in.cdist &&
                                        sp.out.pdist == (long long)sp.ncols * sp.out.cdist));
  // the column kernel needs >= 64-byte row segments (8 FP32 / 4 FP64
  // columns per workgroup); longer columns go through the tiled transpose
  // into contiguous rows and back
  const bool cols = sp.in.cdist == 1 && sp.out.cdist == 1 && sp.n >= 2 && is_smooth7(sp.n) &&
                    fs_geometry(sp.n, elem, cap, &sp.g) && (size_t)sp.g.per_block * elem >= 64;
  fyfft_plan* sub = nullptr;
  fyfft_status st;
  if (rows) {
    sp.method = kPmRows;
    st = create_plan_impl(&sub, sp.n, count, precision, FYFFT_FORM_AUTO);
    if (st) return st;
    p->subs.push_back(sub);
    sp.sub = sub;
    return FYFFT_SUCCESS;
  }
  if (cols) {
    sp.method = kPmColumns;
    st = make_pass_table(p, sp.n, &sp.pd, &sp.tw);
    if (st) return st;
    sp.fixed_n = find_fs_fixed(p->prec, sp.pd, (int)sp.n, sp.g.threads, true);
    return FYFFT_SUCCESS;
  }
  // Adjacent transforms too long for one column tile: two column passes over
  // n = n1 x n2 (lengths n1 then n2, the twiddle w_n^(n2*k1) between them)
  // through a workspace of `chunk` columns, each pass reading and writing
  // row segments of adjacent columns -- two round trips instead of the
  // gather's three.
  size_t n1 = 0, n2 = 0;
  if (sp.in.cdist == 1 && sp.out.cdist == 1 && sp.ncols >= 16 && is_smooth7(sp.n) &&
      four_step_split(sp.n, kFactorCap, &n1, &n2) && fs_geometry(n1, elem, cap, &sp.g) &&
      fs_geometry(n2, elem, cap, &sp.g2) && (size_t)sp.g.per_block * elem >= 64 &&
      (size_t)sp.g2.per_block * elem >= 64) {
    sp.method = kPmColFour;
    sp.n1 = n1;
    sp.n2 = n2;
    st = make_pass_table(p, n1, &sp.pd, &sp.tw);
    if (!st) st = make_pass_table(p, n2, &sp.pd2, &sp.tw2);
    if (st) return st;
    sp.fixed_n = find_fs_fixed(p->prec, sp.pd, (int)n1, sp.g.threads, true);
    sp.fixed2 = find_fs_fixed(p->prec, sp.pd2, (int)n2, sp.g2.threads, true);
    int lg = 0;
    while (((size_t)1 << lg) < sp.n) ++lg;
    sp.lo_bits = (lg + 1) / 2;
    const size_t lo = (size_t)1 << sp.lo_bits, hi = (sp.n >> sp.lo_bits) + 1;
    if ((st = p->alloc(lo * elem, &sp.tw_lo)) || (st = p->alloc(hi * elem, &sp.tw_hi))) return st;
    HIP_TRY(launch_gen_roots(p->prec, sp.tw_lo, (long long)lo, 1, (long long)sp.n, nullptr));
    HIP_TRY(launch_gen_roots(p->prec, sp.tw_hi, (long long)hi, (long long)lo, (long long)sp.n, nullptr));
    size_t cw = std::max<size_t>(16, kWorkspaceCap / (sp.n * elem)) & ~(size_t)15;
    sp.chunk = std::min(sp.ncols, cw);
    *ws_elems = std::max(*ws_elems, sp.chunk * sp.n);
    return FYFFT_SUCCESS;
  }
  sp.method = kPmGather;
  sp.chunk = std::min(count, std::max<size_t>(1, kWorkspaceCap / (sp.n * elem)));
  st = create_plan_impl(&sub, sp.n, sp.chunk, precision, FYFFT_FORM_AUTO);
  if (st) return st;
  p->subs.push_back(sub);
  sp.sub = sub;
  *ws_elems = std::max(*ws_elems, sp.chunk * sp.n);
  return FYFFT_SUCCESS;
}

template <typename V>
static fyfft_status exec_columns(const fyfft_plan* p, const StridedPass& sp, const void* in, void* out,
                                 int direction, hipStream_t s) {
  ColArgs<V> a;
  a.in = (const V*)in;
  a.out = (V*)out;
  a.istride = sp.in.stride;
  a.ostride = sp.out.stride;
  a.ipdist = sp.in.pdist;
  a.opdist = sp.out.pdist;
  a.n = (int)sp.n;
  a.ncols = (int)sp.ncols;
  a.log2c = ilog2_exact(sp.g.per_block);
  a.threads = sp.g.threads;
  a.ps = sp.g.pad_shift;
  a.stride = dev::lds_len((int)sp.n, sp.g.pad_shift);
  a.pd = sp.pd;
  a.tw = (const V*)sp.tw;
  a.scale = (dev::S<V>)(direction == FYFFT_INVERSE ? 1.0 / (double)sp.n : 1.0);
  HIP_TRY(launch_columns(p->prec, direction == FYFFT_FORWARD ? -1 : +1, &a, sp.planes, sp.g, sp.fixed_n, s));
  return FYFFT_SUCCESS;
}

// The two column passes of a long column axis, per plane and column chunk:
// pass 1 transforms columns of length n1 (rows n1*n2*stride apart, inner
// plane n2 at offset n2*stride) into ws[(k1*n2 + n2)*cw + c] with the twiddle
// w_n^(n2*k1); pass 2 transforms length n2 along ws and writes
// X[(k1 + n1*k2)*stride + c].
template <typename V>
static fyfft_status exec_colfour(const fyfft_plan* p, const StridedPass& sp, const void* in, void* out,
                                 int direction, hipStream_t s) {
  const int dir = direction == FYFFT_FORWARD ? -1 : +1;
  V* ws = (V*)p->lws;
  const long long n1 = (long long)sp.n1, n2 = (long long)sp.n2;
  for (size_t pl = 0; pl < sp.planes; ++pl)
    for (size_t c0 = 0; c0 < sp.ncols; c0 += sp.chunk) {
      const long long cw = (long long)std::min(sp.chunk, sp.ncols - c0);
      ColArgs<V> a;
      a.in = (const V*)in + (long long)pl * sp.in.pdist + (long long)c0;
      a.out = ws;
      a.istride = n2 * sp.in.stride;
      a.ipdist2 = sp.in.stride;
      a.ostride = n2 * cw;
      a.opdist2 = cw;
      a.inner = (int)n2;
      a.n = (int)n1;
      a.ncols = (int)cw;
      a.log2c = ilog2_exact(sp.g.per_block);
      a.threads = sp.g.threads;
      a.ps = sp.g.pad_shift;
      a.stride = dev::lds_len((int)n1, sp.g.pad_shift);
      a.pd = sp.pd;
      a.tw = (const V*)sp.tw;
      a.tw_hi = (const V*)sp.tw_hi;
      a.tw_lo = (const V*)sp.tw_lo;
      a.lo_bits = sp.lo_bits;
      HIP_TRY(launch_columns(p->prec, dir, &a, (size_t)n2, sp.g, sp.fixed_n, s));
      ColArgs<V> b;
      b.in = ws;
      b.out = (V*)out + (long long)pl * sp.out.pdist + (long long)c0;
      b.istride = cw;
      b.ipdist2 = n2 * cw;
      b.ostride = n1 * sp.out.stride;
      b.opdist2 = sp.out.stride;
      b.inner = (int)n1;
      b.n = (int)n2;
      b.ncols = (int)cw;
      b.log2c = ilog2_exact(sp.g2.per_block);
      b.threads = sp.g2.threads;
      b.ps = sp.g2.pad_shift;
      b.stride = dev::lds_len((int)n2, sp.g2.pad_shift);
      b.pd = sp.pd2;
      b.tw = (const V*)sp.tw2;
      b.scale = (dev::S<V>)(direction == FYFFT_INVERSE ? 1.0 / (double)sp.n : 1.0);
      HIP_TRY(launch_columns(p->prec, dir, &b, (size_t)n1, sp.g2, sp.fixed2, s));
    }
  return FYFFT_SUCCESS;
}

static fyfft_status exec_strided(const fyfft_plan* p, const void* in, void* out, int direction, hipStream_t s) {
  for (size_t i = 0; i < p->passes.size(); ++i) {
    const StridedPass& sp = p->passes[i];
    const void* src = i == 0 ? in : out;
    fyfft_status st = FYFFT_SUCCESS;
    if (sp.method == kPmRows) {
      st = execute_rows(sp.sub, src, out, direction, sp.ncols * sp.planes, s, (size_t)sp.in.cdist);
    } else if (sp.method == kPmColumns) {
      st = p->prec == kF64 ? exec_columns<double2>(p, sp, src, out, direction, s)
                           : exec_columns<float2>(p, sp, src, out, direction, s);
    } else if (sp.method == kPmColFour) {
      st = p->prec == kF64 ? exec_colfour<double2>(p, sp, src, out, direction, s)
                           : exec_colfour<float2>(p, sp, src, out, direction, s);
    } else {
      const size_t count = sp.ncols * sp.planes;
      for (size_t t0 = 0; t0 < count && !st; t0 += sp.chunk) {
        const long long cnt = (long long)std::min(sp.chunk, count - t0);
        HIP_TRY(launch_gather(p->prec, false, src, p->lws, (long long)sp.n, sp.in.stride, sp.in.cdist,
                              (long long)sp.ncols, sp.in.pdist, (long long)t0, cnt, s));
        st = execute_rows(sp.sub, p->lws, p->lws, direction, (size_t)cnt, s);
        if (!st)
          HIP_TRY(launch_gather(p->prec, true, out, p->lws, (long long)sp.n, sp.out.stride, sp.out.cdist,
                                (long long)sp.ncols, sp.out.pdist, (long long)t0, cnt, s));
      }
    }
    if (st) return st;
  }
  return FYFFT_SUCCESS;
}

static const char* method_name(int m) {
  return m == kPmRows ? "rows" : (m == kPmColumns ? "columns" : (m == kPmColFour ? "colfour" : "gather"));
}

static fyfft_status finish_strided(fyfft_plan* p, size_t ws_elems, const std::string& head) {
  fyfft_status st = FYFFT_SUCCESS;
  if (ws_elems) st = p->alloc(ws_elems * elem_bytes(p->prec), &p->lws);
  if (!st) {
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) st = from_hip(e, "strided plan tables");
  }
  if (st) return st;
  p->desc = head;
  for (const StridedPass& sp : p->passes)
    p->desc += std::string(" | n=") + std::to_string(sp.n) + " " + method_name(sp.method) +
               (sp.sub ? " [" + sp.sub->desc + "]"
                : sp.method == kPmColFour
                    ? " [" + std::to_string(sp.n1) + "(" + radix_string(sp.pd) + ") x " + std::to_string(sp.n2) + "(" +
                          radix_string(sp.pd2) + ") C=" + std::to_string(sp.g.per_block) + "/" +
                          std::to_string(sp.g2.per_block) + " chunk=" + std::to_string(sp.chunk) + "]"
                    : " [C=" + std::to_string(sp.g.per_block) + " " + radix_string(sp.pd) + "]");
  return FYFFT_SUCCESS;
}

static fyfft_status new_strided_plan(fyfft_plan** out, fyfft_precision precision, fyfft_plan** pp) {
  if (!out) return FYFFT_INVALID_ARGUMENT;
  *out = nullptr;
  if (precision == FYFFT_BF16 || precision == FYFFT_F16) return FYFFT_UNSUPPORTED;
  if (precision != FYFFT_F32 && precision != FYFFT_F64) return FYFFT_INVALID_ARGUMENT;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  fyfft_plan* p = new (std::nothrow) fyfft_plan();
  if (!p) return FYFFT_ALLOC_FAILED;
  p->kind = kStrided;
  p->prec = precision == FYFFT_F64 ? kF64 : kF32;
  p->device = dev;
  *pp = p;
  return FYFFT_SUCCESS;
}

static fyfft_status create_many_impl(fyfft_plan** out, size_t n, size_t howmany, const fyfft_layout* L,
                                     fyfft_precision precision) {
  if (!out || !L || howmany == 0) return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  if (L->istride == 0 || L->ostride == 0 || (howmany > 1 && (L->idist == 0 || L->odist == 0)))
    return FYFFT_INVALID_ARGUMENT;
  fyfft_plan* p = nullptr;
  fyfft_status st = new_strided_plan(out, precision, &p);
  if (st) return st;
  p->n = n;
  p->batch = howmany;
  p->strategy = strategy_for(n);
  StridedPass sp;
  sp.n = n;
  sp.ncols = howmany;
  sp.planes = 1;
  sp.in.stride = (long long)L->istride;
  sp.in.cdist = (long long)(howmany > 1 ? L->idist : n * L->istride);
  sp.out.stride = (long long)L->ostride;
  sp.out.cdist = (long long)(howmany > 1 ? L->odist : n * L->ostride);
  sp.in.pdist = sp.in.cdist * (long long)howmany;
  sp.out.pdist = sp.out.cdist * (long long)howmany;
  p->in_ext = (size_t)((long long)(howmany - 1) * sp.in.cdist + (long long)(n - 1) * sp.in.stride + 1);
  p->out_ext = (size_t)((long long)(howmany - 1) * sp.out.cdist + (long long)(n - 1) * sp.out.stride + 1);
  size_t ws = 0;
  p->passes.push_back(sp);
  st = plan_strided_pass(p, p->passes.back(), precision, &ws);
  if (!st)
    st = finish_strided(p, ws, "many N=" + std::to_string(n) + " howmany=" + std::to_string(howmany) +
                                   " in(" + std::to_string(L->istride) + "," + std::to_string(L->idist) +
                                   ") out(" + std::to_string(L->ostride) + "," + std::to_string(L->odist) + ")");
  if (st) {
    delete p;
    return st;
  }
  *out = p;
  return FYFFT_SUCCESS;
}

static fyfft_status create_nd_impl(fyfft_plan** out, int rank, const size_t* dims, size_t batch,
                                   fyfft_precision precision) {
  if (!out || !dims || rank < 1 || rank > 3 || batch == 0) return FYFFT_INVALID_ARGUMENT;
  size_t vol = 1;
  for (int d = 0; d < rank; ++d) {
    if (dims[d] < 1 || dims[d] > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
    vol *= dims[d];
  }
  if (vol > ((size_t)1 << 40) / batch) return FYFFT_INVALID_LENGTH;
  fyfft_plan* p = nullptr;
  fyfft_status st = new_strided_plan(out, precision, &p);
  if (st) return st;
  p->n = vol;
  p->batch = batch;
  p->strategy = strategy_for(dims[rank - 1]);
  p->in_ext = p->out_ext = vol * batch;
  size_t ws = 0;
  std::string head = "nd dims=";
  for (int d = 0; d < rank; ++d) head += (d ? "x" : "") + std::to_string(dims[d]);
  head += " batch=" + std::to_string(batch);
  // innermost axis first (contiguous rows), then the strided axes
  for (int d = rank - 1; d >= 0 && !st; --d) {
    if (dims[d] == 1) continue;
    size_t inner = 1, outer = batch;
    for (int e = d + 1; e < rank; ++e) inner *= dims[e];
    for (int e = 0; e < d; ++e) outer *= dims[e];
    StridedPass sp;
    sp.n = dims[d];
    if (inner == 1) {  // contiguous rows
      sp.ncols = 1;
      sp.planes = outer;
      sp.in.stride = 1;
      sp.in.cdist = (long long)dims[d];
      sp.in.pdist = (long long)dims[d];
    } else {
      sp.ncols = inner;
      sp.planes = outer;
      sp.in.stride = (long long)inner;
      sp.in.cdist = 1;
      sp.in.pdist = (long long)(inner * dims[d]);
    }
    sp.out = sp.in;
    p->passes.push_back(sp);
    st = plan_strided_pass(p, p->passes.back(), precision, &ws);
  }
  if (!st && p->passes.empty()) {  // every dimension 1: a copy
    StridedPass sp;
    sp.n = 1;
    sp.planes = batch;
    sp.in.cdist = sp.in.pdist = 1;
    sp.out = sp.in;
    p->passes.push_back(sp);
    st = plan_strided_pass(p, p->passes.back(), precision, &ws);
  }
  if (!st) st = finish_strided(p, ws, head);
  if (st) {
    delete p;
    return st;
  }
  *out = p;
  return FYFFT_SUCCESS;
}

// ------------------------------------------------------------ no-allocation execution
// Workspace of fyfft_execute_noalloc: a file-scope device array, one copy per
// GPU at its own address, held from module load until the process ends.  In
// host code the array's name is only a stand-in for the symbol: its device
// address is taken once per device with hipGetSymbolAddress (that device
// current) and only that address is handed to kernels.  Every call on one
// GPU shares the array, so calls that overlap in time on one GPU race.
constexpr size_t kScratchBytes = FYFFT_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kScratchDevices] = {};

static fyfft_status scratch_of_current_device(int dev, void** ptr) {
  if (dev < 0 || dev >= kScratchDevices) {
    g_last_error = "fyfft_execute_noalloc: device index beyond the supported 64";
    return FYFFT_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return from_hip(e, "hipGetSymbolAddress(scratch)");
    if (!a) {
      g_last_error = "hipGetSymbolAddress(scratch) returned a null address";
      return FYFFT_DEVICE_ERROR;
    }
    g_scratch_ptr[dev] = a;
  }
  *ptr = g_scratch_ptr[dev];
  return FYFFT_SUCCESS;
}

// Carves a four-step plan's workspace out of the scratch array: as many
// transforms per chunk as the array holds (tmp and, for bucketed Rader
// plans, zw: one convolution length each per transform; Rader x[0] slots),
// each region 256-byte aligned.  None of it needs clearing: every pass
// writes what the next one reads.
static fyfft_status scratch_workspace(const fyfft_plan* p, void* base, Workspace* w) {
  const size_t elem = elem_bytes(p->prec), L = p->kind == kFourStep ? p->n : p->m;
  const bool x0 = p->kind == kRaderFs, zw = p->kind == kRaderFs && p->rb_q > 0;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  auto need = [&](size_t c) { return up(c * L * elem) + (zw ? up(c * L * elem) : 0) + (x0 ? up(c * elem) : 0); };
  const size_t per = L * elem * (zw ? 2 : 1) + (x0 ? elem : 0);
  size_t chunk = std::min(p->batch, kScratchBytes / per);
  while (chunk > 0 && need(chunk) > kScratchBytes) --chunk;
  if (chunk == 0) {
    g_last_error = "fyfft_execute_noalloc: one transform needs " + std::to_string(need(1)) +
                   " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                   "-byte static scratch (use fyfft_execute)";
    return FYFFT_INVALID_LENGTH;
  }
  char* b = (char*)base;
  w->chunk = chunk;
  w->tmp = b;
  b += up(chunk * L * elem);
  w->zw = zw ? b : nullptr;
  if (zw) b += up(chunk * L * elem);
  w->x0buf = x0 ? b : nullptr;
  return FYFFT_SUCCESS;
}

extern "C" {

const char* fyfft_status_string(fyfft_status s) {
  switch (s) {
    case FYFFT_SUCCESS: return "success";
    case FYFFT_INVALID_ARGUMENT: return "invalid argument";
    case FYFFT_INVALID_LENGTH: return "invalid length";
    case FYFFT_UNSUPPORTED: return "unsupported";
    case FYFFT_ALLOC_FAILED: return "allocation failed";
    case FYFFT_DEVICE_ERROR: return "device error";
    case FYFFT_INVALID_PLAN: return "invalid plan";
    case FYFFT_NO_DEVICE: return "no device";
  }
  return "unknown status";
}

const char* fyfft_last_error(void) { return g_last_error.c_str(); }

fyfft_status fyfft_length_strategy(size_t n, int* strategy) {
  if (!strategy) return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  *strategy = strategy_for(n);
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_length_radices(size_t n, int* radices, int max, int* count) {
  if (!count || (max > 0 && !radices) || max < 0) return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  if (!is_smooth7(n)) return FYFFT_UNSUPPORTED;
  const std::vector<int> r = radices_for(n);
  *count = (int)r.size();
  for (int i = 0; i < (int)r.size() && i < max; ++i) radices[i] = r[i];
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_create(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision) {
  return create_plan_impl(out, n, batch, precision, FYFFT_FORM_AUTO);
}

fyfft_status fyfft_plan_create_ex(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision,
                                  const fyfft_options* options) {
  int form = FYFFT_FORM_AUTO;
  if (options) {
    for (int r : options->reserved)
      if (r != 0) return FYFFT_INVALID_ARGUMENT;
    form = options->form;
    if (form < FYFFT_FORM_AUTO || form > FYFFT_FORM_MFMA) return FYFFT_INVALID_ARGUMENT;
  }
  return create_plan_impl(out, n, batch, precision, form);
}

fyfft_status fyfft_plan_create_strategy(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision,
                                       int strategy) {
  if (strategy < FYFFT_KERNEL_DIRECT || strategy > FYFFT_DFT) return FYFFT_INVALID_ARGUMENT;
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  if ((strategy == FYFFT_KERNEL_DIRECT && !in_reference_registry(n)) ||
      (strategy == FYFFT_MIX_RADIX && !is_smooth7(n))) {
    g_last_error = "fyfft_plan_create_strategy: N=" + std::to_string(n) + " is not " +
                   (strategy == FYFFT_KERNEL_DIRECT ? "a directly supported length" : "2/3/5/7-smooth");
    return FYFFT_UNSUPPORTED;
  }
  if (precision == FYFFT_BF16 || precision == FYFFT_F16) {
    if (strategy == FYFFT_RADER || strategy == FYFFT_BLUESTEIN) return FYFFT_UNSUPPORTED;
  }
  return create_plan_impl(out, n, batch, precision, FYFFT_FORM_AUTO, strategy);
}

fyfft_status fyfft_strategy_processable(int strategy, size_t n, int* processable) {
  if (!processable) return FYFFT_INVALID_ARGUMENT;
  switch (strategy) {
    case FYFFT_KERNEL_DIRECT: *processable = in_reference_registry(n); break;
    case FYFFT_MIX_RADIX: *processable = n >= 1 && is_smooth7(n); break;
    // a prime outside the registry whose 2N - 3 has a registry length above it
    case FYFFT_RADER: *processable = !in_reference_registry(n) && is_prime(n) && 2 * n - 3 <= 78125; break;
    case FYFFT_BLUESTEIN: *processable = n <= 39062; break;
    case FYFFT_DFT: *processable = n >= 1; break;
    default: return FYFFT_INVALID_ARGUMENT;
  }
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_create_real(fyfft_plan** out, size_t n, size_t batch, fyfft_precision precision) {
  return create_real_impl(out, n, batch, precision);
}

fyfft_status fyfft_plan_create_many(fyfft_plan** out, size_t n, size_t howmany, const fyfft_layout* layout,
                                    fyfft_precision precision) {
  return create_many_impl(out, n, howmany, layout, precision);
}

fyfft_status fyfft_plan_create_nd(fyfft_plan** out, int rank, const size_t* dims, size_t batch,
                                  fyfft_precision precision) {
  return create_nd_impl(out, rank, dims, batch, precision);
}

fyfft_status fyfft_wisdom_export(char* buf, size_t len, size_t* needed) {
  const std::string t = wisdom_text();
  if (needed) *needed = t.size() + 1;
  if (!buf) return needed ? FYFFT_SUCCESS : FYFFT_INVALID_ARGUMENT;
  if (len < t.size() + 1) return FYFFT_INVALID_ARGUMENT;
  std::memcpy(buf, t.c_str(), t.size() + 1);
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_wisdom_import(const char* text) {
  if (!text) return FYFFT_INVALID_ARGUMENT;
  return wisdom_parse(text);
}

fyfft_status fyfft_wisdom_forget(void) {
  std::lock_guard<std::mutex> lock(g_wisdom_mu);
  g_wisdom.clear();
  g_wisdom_fs.clear();
  g_wisdom_m.clear();
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_create_grouped(fyfft_plan** out, const size_t* lengths, size_t count,
                                       fyfft_precision precision, unsigned flags) {
  if (!out) return FYFFT_INVALID_ARGUMENT;
  *out = nullptr;
  if (!lengths || count == 0 || count > (size_t)INT32_MAX / 2) return FYFFT_INVALID_ARGUMENT;
  if (flags & ~(unsigned)(FYFFT_GROUPED_GRAPH | FYFFT_GROUPED_PER_TRANSFORM | FYFFT_GROUPED_TUNED))
    return FYFFT_INVALID_ARGUMENT;
  if (precision != FYFFT_F32 && precision != FYFFT_F64 && precision != FYFFT_BF16 && precision != FYFFT_F16)
    return FYFFT_INVALID_ARGUMENT;
  for (size_t i = 0; i < count; ++i)
    if (lengths[i] < 1 || lengths[i] > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  fyfft_plan* p = new (std::nothrow) fyfft_plan();
  if (!p) return FYFFT_ALLOC_FAILED;
  p->kind = kGrouped;
  // storage precision; 16-bit plans compute in FP32 with FP32 sub-plans
  p->prec = precision == FYFFT_F64 ? kF64 : (precision == FYFFT_BF16 ? kBF16 : (precision == FYFFT_F16 ? kF16 : kF32));
  const fyfft_precision sub_prec = precision == FYFFT_F64 ? FYFFT_F64 : FYFFT_F32;
  p->device = dev;
  p->gflags = flags;
  p->count = count;
  p->batch = 1;
  p->h_off.resize(count);
  p->sub_of.resize(count);
  std::map<size_t, std::vector<size_t>> by_len;
  for (size_t i = 0; i < count; ++i) {
    p->h_off[i] = p->total;
    p->total += lengths[i];
    by_len[lengths[i]].push_back(i);
  }
  p->n = p->total;
  fyfft_status st = FYFFT_SUCCESS;
  // FYFFT_GROUPED_TUNED: the lengths that carry the most elements are tuned
  // first (radix order, padding and points per thread of LDS lengths; split
  // and tile of four-step and Rader convolution lengths; Bluestein's
  // convolution length), each timed on a batch of its own length, until the
  // tuning time budget is spent; wisdom makes later plans of the same
  // lengths tune instantly.  Other lengths keep the planner's defaults.
  std::vector<size_t> tune_order;
  for (const auto& kv : by_len) tune_order.push_back(kv.first);
  std::stable_sort(tune_order.begin(), tune_order.end(), [&](size_t a, size_t b) {
    return a * by_len.at(a).size() > b * by_len.at(b).size();
  });
  std::map<size_t, int> tune_form;
  if (flags & FYFFT_GROUPED_TUNED) {
    const auto t0 = std::chrono::steady_clock::now();
    for (size_t len : tune_order) {
      if (std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() > kGroupedTuneSeconds) break;
      tune_form[len] = kFormTunedGeneric;
      fyfft_plan* probe = nullptr;  // tunes (or finds wisdom for) this length once
      const size_t cnt = by_len.at(len).size();
      if (create_plan_impl(&probe, len, std::max<size_t>(cnt, 64), sub_prec, kFormTunedGeneric) == FYFFT_SUCCESS)
        fyfft_plan_destroy(probe);
      else
        tune_form[len] = kFormGroupedPlain;
    }
  }
  for (const auto& kv : by_len) {
    fyfft_plan* sp = nullptr;
    auto tf = tune_form.find(kv.first);
    st = create_plan_impl(&sp, kv.first, 1, sub_prec, tf == tune_form.end() ? kFormGroupedPlain : tf->second);
    if (st) break;
    for (size_t i : kv.second) p->sub_of[i] = (int)p->subs.size();
    p->subs.push_back(sp);
  }
  if (!st) st = p->prec == kF64 ? build_grouped<double2>(p, by_len) : build_grouped<float2>(p, by_len);
  if (!st && (p->prec == kBF16 || p->prec == kF16) && (flags & FYFFT_GROUPED_PER_TRANSFORM))
    st = p->alloc(by_len.rbegin()->first * sizeof(float2), &p->gstage);
  if (st) {
    delete p;
    return st;
  }
  *out = p;
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_elements(const fyfft_plan* p, size_t* elements) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!elements) return FYFFT_INVALID_ARGUMENT;
  *elements = p->kind == kGrouped ? p->total : (p->kind == kStrided ? p->in_ext : p->n * p->batch);
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_destroy(fyfft_plan* p) {
  if (!p) return FYFFT_SUCCESS;
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  delete p;
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_execute(const fyfft_plan* p, const void* in, void* out, int direction, void* stream) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  return execute_rows(p, in, out, direction, p->batch, (hipStream_t)stream);
}

fyfft_status fyfft_execute_noalloc(const fyfft_plan* p, const void* in, void* out, int direction, void* stream) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  const hipStream_t s = (hipStream_t)stream;
  switch (p->kind) {
    case kTrivial:
    case kLds:
    case kRaderLds:
    case kBlueLds:
    case kMfma:
    case kLowp:  // single-kernel plans: no workspace at all
      return execute_rows(p, in, out, direction, p->batch, s);
    case kFourStep:
    case kRaderFs:
    case kBlueFs: {
      int dev = 0;
      HIP_TRY(hipGetDevice(&dev));
      if (dev != p->device) {
        g_last_error = "fyfft_execute_noalloc: the plan belongs to device " + std::to_string(p->device) +
                       ", the current device is " + std::to_string(dev);
        return FYFFT_INVALID_ARGUMENT;
      }
      void* base = nullptr;
      fyfft_status st = scratch_of_current_device(dev, &base);
      Workspace w;
      if (!st) st = scratch_workspace(p, base, &w);
      if (st) return st;
      return execute_rows(p, in, out, direction, p->batch, s, 0, &w);
    }
    default:
      g_last_error = "fyfft_execute_noalloc: grouped, real, strided and multi-dimensional plans are not served";
      return FYFFT_UNSUPPORTED;
  }
}

fyfft_status fyfft_rfft_modes(const fyfft_plan* p, const void* x, void* modes, size_t k,
                              double scale0, double scale, void* stream) {
  return spectral(p, false, x, modes, k, scale0, scale, FYFFT_MODES_ROW_MAJOR, stream);
}

fyfft_status fyfft_irfft_modes(const fyfft_plan* p, const void* modes, void* y, size_t k,
                               double scale0, double scale, void* stream) {
  return spectral(p, true, modes, y, k, scale0, scale, FYFFT_MODES_ROW_MAJOR, stream);
}

fyfft_status fyfft_rfft_modes_ex(const fyfft_plan* p, const void* x, void* modes, size_t k, double scale0,
                                 double scale, int layout, void* stream) {
  return spectral(p, false, x, modes, k, scale0, scale, layout, stream);
}

fyfft_status fyfft_irfft_modes_ex(const fyfft_plan* p, const void* modes, void* y, size_t k, double scale0,
                                  double scale, int layout, void* stream) {
  return spectral(p, true, modes, y, k, scale0, scale, layout, stream);
}

fyfft_status fyfft_mix_modes(int kind, fyfft_precision precision, const void* p, const void* q, void* out,
                             size_t k, size_t rows, size_t cin, size_t cout, void* stream) {
  if (kind != FYFFT_MIX_FORWARD && kind != FYFFT_MIX_ADJOINT_X && kind != FYFFT_MIX_GRAD_W) {
    g_last_error = "fyfft_mix_modes: unknown kind";
    return FYFFT_INVALID_ARGUMENT;
  }
  if (precision != FYFFT_F32 && precision != FYFFT_F64) {
    g_last_error = "fyfft_mix_modes: FYFFT_F32 (complex float) or FYFFT_F64 (complex double) operands only";
    return FYFFT_UNSUPPORTED;
  }
  if (k == 0 || rows == 0 || cin == 0 || cout == 0) return FYFFT_SUCCESS;
  if (!p || !q || !out) return FYFFT_INVALID_ARGUMENT;
  if (k > 65535 || rows > (size_t)1 << 30 || cin > (size_t)1 << 30 || cout > (size_t)1 << 30) {
    g_last_error = "fyfft_mix_modes: k must be <= 65535 and rows, cin, cout < 2^30";
    return FYFFT_INVALID_ARGUMENT;
  }
  HIP_TRY(launch_mix_modes(precision == FYFFT_F64 ? kF64 : kF32, kind, p, q, out, (long long)k, (long long)rows,
                           (long long)cin, (long long)cout, (hipStream_t)stream));
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_execute_host(const fyfft_plan* p, const void* in, void* out, int direction) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!in || !out) return FYFFT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(p->host_mu);
  const size_t elem = elem_bytes(p->prec);
  size_t in_bytes, out_bytes;
  if (p->kind == kReal) {
    const size_t real = p->batch * p->n * (elem / 2), half = p->batch * (p->n / 2 + 1) * elem;
    in_bytes = direction == FYFFT_FORWARD ? real : half;
    out_bytes = direction == FYFFT_FORWARD ? half : real;
  } else if (p->kind == kStrided) {
    in_bytes = p->in_ext * elem;
    out_bytes = p->out_ext * elem;
  } else {
    in_bytes = out_bytes = (p->kind == kGrouped ? p->total : p->n * p->batch) * elem;
  }
  // in place for batched complex plans; real and strided plans stage input
  // and output apart (the output starts as zeros where a layout has gaps)
  const bool apart = p->kind == kReal || p->kind == kStrided;
  const size_t out_off = apart ? (in_bytes + 255) / 256 * 256 : 0;
  const size_t bytes = std::max(in_bytes, out_off + out_bytes);
  if (!p->io_stream) HIP_TRY(hipStreamCreateWithFlags(&p->io_stream, hipStreamNonBlocking));
  if (p->d_io_bytes < bytes) {
    if (p->d_io) (void)hipFree(p->d_io);
    if (p->h_io) (void)hipHostFree(p->h_io);
    p->d_io = p->h_io = nullptr;
    p->d_io_bytes = 0;
    HIP_TRY(hipMalloc(&p->d_io, bytes));
    // pinned staging: the copies run as DMA on the plan's stream and only
    // the final synchronisation waits (pageable copies if it is refused)
    if (hipHostMalloc(&p->h_io, bytes, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p->h_io = nullptr;
    }
    p->d_io_bytes = bytes;
  }
  const hipStream_t s = p->io_stream;
  void* dout = (char*)p->d_io + out_off;
  char* h = (char*)p->h_io;
  if (h) {
    std::memcpy(h, in, in_bytes);
    if (apart) std::memcpy(h + out_off, out, out_bytes);
    HIP_TRY(hipMemcpyAsync(p->d_io, h, apart ? out_off + out_bytes : in_bytes, hipMemcpyHostToDevice, s));
  } else {
    HIP_TRY(hipMemcpyAsync(p->d_io, in, in_bytes, hipMemcpyHostToDevice, s));
    if (apart) HIP_TRY(hipMemcpyAsync(dout, out, out_bytes, hipMemcpyHostToDevice, s));
  }
  fyfft_status st = fyfft_execute(p, p->d_io, dout, direction, s);
  if (st) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  HIP_TRY(hipMemcpyAsync(h ? (void*)(h + out_off) : out, dout, out_bytes, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (h) std::memcpy(out, h + out_off, out_bytes);
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_strategy(const fyfft_plan* p, int* strategy) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!strategy) return FYFFT_INVALID_ARGUMENT;
  *strategy = p->strategy;
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_workspace_bytes(const fyfft_plan* p, size_t* bytes) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!bytes) return FYFFT_INVALID_ARGUMENT;
  *bytes = p->dev_bytes;
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_plan_describe(const fyfft_plan* p, char* buf, size_t len) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!buf || len == 0) return FYFFT_INVALID_ARGUMENT;
  std::snprintf(buf, len, "%s", p->desc.c_str());
  return FYFFT_SUCCESS;
}

}  // extern "C"
