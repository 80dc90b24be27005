// Host planner: strategy choice, radix factorisation, launch geometry and
// number theory.  Pure C++ (no HIP), so the C ABI's planning queries work
// without a GPU.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace fyfft {

// strategy codes, same values as fyfft_strategy in include/fyfft.h
enum Strategy { kDirect = 0, kMix = 1, kRader = 2, kBluestein = 3, kDft = 4 };

bool in_reference_registry(size_t n);  // the reference's 38 direct lengths
bool is_smooth7(size_t n);
bool is_prime(size_t n);
uint64_t power_mod(uint64_t base, uint64_t e, uint64_t mod);
uint64_t primitive_root(uint64_t p);  // smallest generator of (Z/pZ)*, p prime
int strategy_for(size_t n);

// Radix passes (each in {2,3,4,5,7,8,9,16}) for a smooth length: powers of
// two spread over ceil(log2/4) passes of at most 16, then 7s, 5s, the threes
// paired as 9s and a last 3 if their count is odd.  Empty if n is not
// 7-smooth or n < 2.
std::vector<int> radices_for(size_t n);

// Pass list of a runtime-pass four-step sub-transform (FP32, 16 points per
// thread): like radices_for, but a pass may also have one of the composite
// radices 6, 10, 12, 14, 15, 20, 21, 25 (two levels in registers).  Chosen
// by dynamic programming over the exponents of 2, 3, 5 and 7 to minimise a
// cost of one unit per pass (an LDS round trip and barrier) plus 0.3 per
// unit of idle lanes in that pass at 16 points per thread.  Empty if n is
// not 7-smooth or n < 2.
std::vector<int> radices_composite(size_t n);

// LDS padding of the planner's geometries: index i is stored at
// i + (i >> kLdsPadShift) (the device code's dev::kPadShift, checked equal
// where both are visible).
constexpr int kLdsPadShift = 4;

// Geometry of one LDS-resident transform group.
struct LdsGeom {
  int n = 0;         // transform length
  int ppt = 16;      // max points per thread (kernel template)
  int threads = 0;   // threads per transform (T)
  int per_block = 1; // transforms per workgroup (TPB)
  int pad_shift = 4; // LDS index i -> i + (i >> pad_shift)
  size_t lds_bytes = 0;
  bool small_radix = false;  // FP32, no radix-16 pass: the leaner 8-points-per-thread kernel
  // four-step tiles: bytes of the LDS copy of the sub-length's Stockham
  // table after the tile (runtime-pass sub-transforms), 0 when it does not fit
  size_t tw_bytes = 0;
};
// Fits n (elements of `elem_bytes`) in one workgroup's LDS of `lds_cap`
// bytes; returns false if it does not fit even unpadded.  short_threads
// (batched LDS plans): FP32 lengths below 512 take n/8 threads instead of
// n/16 (with n/16, fewer than 32, passes of few butterflies leave most lanes
// waiting; measured 1.5-3x slower).  Four-step sub-transforms keep n/16,
// which their compile-time forms assume.  small_radix (batched C2C plans
// without a radix-16 pass; FP32 128 <= n < 2048, FP64 64 <= n < 400): n/8
// (FP32) or n/4 (FP64) threads and the leaner runtime-pass kernel sized for
// radix 9 at most (FP32 81 VGPRs against 116; FP64 116 VGPRs and no
// call-frame scratch against 128 + 244 B): measured 4-10 % faster for FP32
// 168...1000 and 15-34 % for FP64 100...200, slower for longer lengths and
// for the real-transform kernels, which keep the default.
bool lds_geometry(size_t n, size_t elem_bytes, size_t lds_cap, LdsGeom* g, bool short_threads = false,
                  bool small_radix = false);
size_t lds_capacity_elems(size_t elem_bytes, size_t lds_cap);  // unpadded

// Four-step split n = n1 * n2 with both factors smooth and LDS-resident
// (n1 <= n2, as balanced as possible).  False if none exists.
bool four_step_split(size_t n, size_t cap_elems, size_t* n1, size_t* n2);
// Split of a smooth length that is not a power of two by a cost model of the
// runtime-pass sub-transforms: the passes of both factors (radices_for),
// weighted by how much of an LDS round trip each radix uses (3, 5, 7 and 9
// heavier), plus a penalty for each factor whose elements do not fill whole
// 128-byte lines (the column pass's workspace rows and the row pass's
// transposed output then write partial lines), and a small imbalance term;
// factors within a ratio of 8 of each other.  Measured against every split
// of eight such lengths (24000...120000) it picked the fastest or one
// within 10 % (N = 96000: 300 x 320 0.95 ms -> 240 x 400 0.60 ms).
bool four_step_split_model(size_t n, size_t cap_elems, size_t elem_bytes, size_t* n1, size_t* n2);
// The split batched four-step plans prefer: the largest n1 <= sqrt(n) whose
// 16-column tiles fit 40 KB of LDS, with n2 <= 16384 (FP32) / 8192 (FP64).
bool four_step_split_lds(size_t n, size_t elem_bytes, size_t* n1, size_t* n2);

}  // namespace fyfft
