// Plain four-step column and row passes with compile-time odd-radix
// sub-transforms: the factors of the reference table's four-step lengths
// (59049 = 243 x 243, 78125 = 125 x 625, and in FP64 15625 = 125 x 125,
// 16807 = 49 x 343, 19683 = 81 x 243).  Threads per sub-transform are
// ceil(n / 16) (FP32) or ceil(n / 8) (FP64), the four-step geometry's.
// A translation unit of its own so that it compiles beside the power-of-two
// forms.
#include "fourstep.hpp"

namespace fyfft {

namespace {

template <typename V, typename F>
hipError_t plain_fs(bool col, int dir, const FsArgs<V>& a, size_t batch, size_t lds, int threads, hipStream_t s) {
  if (col)
    return dir < 0 ? go_col<V, F, -1, kFsPlainA>(a, batch, lds, threads, s)
                   : go_col<V, F, +1, kFsPlainA>(a, batch, lds, threads, s);
  return dir < 0 ? go_row<V, F, -1, kFsPlainB>(a, batch, lds, threads, s)
                 : go_row<V, F, +1, kFsPlainB>(a, batch, lds, threads, s);
}

template <typename V>
hipError_t odd_len(int n, bool col, int dir, const FsArgs<V>& a, size_t batch, size_t lds, int threads,
                   hipStream_t s) {
  constexpr int P = sizeof(V) == 8 ? 16 : 8;
  switch (n) {
    case 49: return plain_fs<V, FixedFft<V, 49, cdiv(49, P), 7, 7>>(col, dir, a, batch, lds, threads, s);
    case 81: return plain_fs<V, FixedFft<V, 81, cdiv(81, P), 9, 9>>(col, dir, a, batch, lds, threads, s);
    case 125: return plain_fs<V, FixedFft<V, 125, cdiv(125, P), 5, 5, 5>>(col, dir, a, batch, lds, threads, s);
    case 243: return plain_fs<V, FixedFft<V, 243, cdiv(243, P), 9, 9, 3>>(col, dir, a, batch, lds, threads, s);
    case 343: return plain_fs<V, FixedFft<V, 343, cdiv(343, P), 7, 7, 7>>(col, dir, a, batch, lds, threads, s);
    case 625: return plain_fs<V, FixedFft<V, 625, cdiv(625, P), 5, 5, 5, 5>>(col, dir, a, batch, lds, threads, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_fs_odd(int prec, int n, bool col, int dir, const void* args, size_t batch, size_t lds,
                         int threads, hipStream_t s) {
  if (prec == kF64) return odd_len(n, col, dir, *static_cast<const FsArgs<double2>*>(args), batch, lds, threads, s);
  return odd_len(n, col, dir, *static_cast<const FsArgs<float2>*>(args), batch, lds, threads, s);
}

}  // namespace fyfft
