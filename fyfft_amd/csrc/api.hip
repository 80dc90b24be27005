// C ABI (include/fyfft.h): plan creation, execution and queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
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

enum Kind { kTrivial = 0, kLds = 1 };

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

struct fyfft_plan {
  uint32_t magic = kMagic;
  size_t n = 0, batch = 0;
  int prec = kF32;
  int strategy = kDirect;
  int device = 0;
  Kind kind = kTrivial;
  LdsLaunch lds;
  std::vector<void*> allocs;
  size_t dev_bytes = 0;
  std::string desc;
  // staging for fyfft_execute_host (lazily allocated, reused)
  mutable std::mutex host_mu;
  mutable void* d_io = nullptr;
  mutable size_t d_io_bytes = 0;

  fyfft_status alloc(size_t bytes, void** p) {
    HIP_TRY(hipMalloc(p, bytes));
    allocs.push_back(*p);
    dev_bytes += bytes;
    return FYFFT_SUCCESS;
  }
  ~fyfft_plan() {
    for (void* p : allocs) (void)hipFree(p);
    if (d_io) (void)hipFree(d_io);
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

// Plan an LDS-resident smooth transform of length n; fills L and allocates
// and generates its twiddle table.
static fyfft_status plan_lds(fyfft_plan* p, size_t n, LdsLaunch* L) {
  const size_t cap = lds_cap_bytes(p->device);
  if (!lds_geometry(n, elem_bytes(p->prec), cap, &L->geom)) return FYFFT_UNSUPPORTED;
  const std::vector<int> rad = radices_for(n);
  if (rad.empty() || (int)rad.size() > dev::kMaxPasses) return FYFFT_UNSUPPORTED;
  L->pd.npass = (int)rad.size();
  for (size_t i = 0; i < rad.size(); ++i) L->pd.radix[i] = rad[i];
  L->fixed = find_fixed_kernel(p->prec, L->pd, &L->geom);
  void* tw = nullptr;
  fyfft_status st = p->alloc(std::max<size_t>(n, 2) * elem_bytes(p->prec), &tw);
  if (st) return st;
  HIP_TRY(launch_gen_stockham_twiddles(p->prec, tw, (int)n, L->pd, nullptr));
  L->tw = tw;
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
  if (!out) return FYFFT_INVALID_ARGUMENT;
  *out = nullptr;
  if (batch == 0) return FYFFT_INVALID_ARGUMENT;
  if (precision != FYFFT_F32 && precision != FYFFT_F64) {
    if (precision == FYFFT_BF16 || precision == FYFFT_F16) return FYFFT_UNSUPPORTED;
    return FYFFT_INVALID_ARGUMENT;
  }
  if (n < 1 || n > FYFFT_MAX_LENGTH) return FYFFT_INVALID_LENGTH;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  fyfft_plan* p = new (std::nothrow) fyfft_plan();
  if (!p) return FYFFT_ALLOC_FAILED;
  p->n = n;
  p->batch = batch;
  p->prec = precision == FYFFT_F64 ? kF64 : kF32;
  p->device = dev;
  p->strategy = strategy_for(n);
  fyfft_status st = FYFFT_SUCCESS;
  if (n == 1) {
    p->kind = kTrivial;
    p->desc = "trivial N=1";
  } else if (is_smooth7(n)) {
    p->kind = kLds;
    st = plan_lds(p, n, &p->lds);
    if (!st) {
      const LdsGeom& g = p->lds.geom;
      p->desc = std::string(p->lds.fixed >= 0 ? "lds-fixed" : "lds-generic") + " N=" +
                std::to_string(n) + " radices=" + radix_string(p->lds.pd) +
                " T=" + std::to_string(g.threads) + " TPB=" + std::to_string(g.per_block) +
                " pad=" + std::to_string(g.pad_shift);
    }
  } else {
    st = FYFFT_UNSUPPORTED;
  }
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

fyfft_status fyfft_plan_destroy(fyfft_plan* p) {
  if (!p) return FYFFT_SUCCESS;
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  delete p;
  return FYFFT_SUCCESS;
}

fyfft_status fyfft_execute(const fyfft_plan* p, const void* in, void* out, int direction, void* stream) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!in || !out) return FYFFT_INVALID_ARGUMENT;
  if (direction != FYFFT_FORWARD && direction != FYFFT_INVERSE && direction != FYFFT_ADJOINT)
    return FYFFT_INVALID_ARGUMENT;
  hipStream_t s = (hipStream_t)stream;
  const double scale = direction == FYFFT_INVERSE ? 1.0 / (double)p->n : 1.0;
  const int dir = direction == FYFFT_FORWARD ? kForward : (direction == FYFFT_INVERSE ? kInverse : kAdjoint);
  switch (p->kind) {
    case kTrivial:
      if (in != out)
        HIP_TRY(hipMemcpyAsync(out, in, p->batch * elem_bytes(p->prec), hipMemcpyDeviceToDevice, s));
      return FYFFT_SUCCESS;
    case kLds:
      HIP_TRY(launch_lds(p->prec, dir, p->lds, in, out, p->batch, p->n, scale, s));
      return FYFFT_SUCCESS;
  }
  return FYFFT_INVALID_PLAN;
}

fyfft_status fyfft_execute_host(const fyfft_plan* p, const void* in, void* out, int direction) {
  if (!valid(p)) return FYFFT_INVALID_PLAN;
  if (!in || !out) return FYFFT_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(p->host_mu);
  const size_t bytes = p->n * p->batch * elem_bytes(p->prec);
  if (p->d_io_bytes < bytes) {
    if (p->d_io) (void)hipFree(p->d_io);
    p->d_io = nullptr;
    p->d_io_bytes = 0;
    HIP_TRY(hipMalloc(&p->d_io, bytes));
    p->d_io_bytes = bytes;
  }
  HIP_TRY(hipMemcpy(p->d_io, in, bytes, hipMemcpyHostToDevice));
  fyfft_status st = fyfft_execute(p, p->d_io, p->d_io, direction, nullptr);
  if (st) return st;
  HIP_TRY(hipMemcpy(out, p->d_io, bytes, hipMemcpyDeviceToHost));
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
