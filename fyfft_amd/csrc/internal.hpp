// Declarations shared between the C ABI (api.hip) and the kernel
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_device.hpp"
#include "planner.hpp"

namespace fyfft {

enum Prec { kF32 = 0, kF64 = 1 };
enum Dir { kForward = 0, kInverse = 1, kAdjoint = 2 };

inline size_t elem_bytes(int prec) { return prec == kF64 ? 16 : 8; }

// Everything one LDS-resident launch needs.
struct LdsLaunch {
  LdsGeom geom;
  dev::PassDesc pd;
  const void* tw = nullptr;  // Stockham twiddle table (N-1 entries)
  int fixed = -1;            // index of a compile-time specialised kernel, -1 = generic
};

// --- kernels_lds.hip
// Batched LDS-resident transform: `batch` transforms of length N, transform
// b at in + b*dist (elements).  scale multiplies the output.
hipError_t launch_lds(int prec, int dir, const LdsLaunch& L, const void* in, void* out,
                      size_t batch, size_t dist, double scale, hipStream_t s);
// Index of a compile-time specialised kernel for (prec, radices), or -1; on
// a hit the geometry is replaced by the kernel's own.
int find_fixed_kernel(int prec, const dev::PassDesc& pd, LdsGeom* g);

// --- kernels_tables.hip
// Stockham twiddles for pass list pd (N-1 entries) in FP64, stored as prec.
hipError_t launch_gen_stockham_twiddles(int prec, void* tw, int n, const dev::PassDesc& pd,
                                        hipStream_t s);

}  // namespace fyfft
