// Grouped ragged transforms, FP32 and FP64 storage: the family kernels of
// grouped.hpp instantiated for complex float and double.
#include "grouped.hpp"

namespace fyfft {

hipError_t group_family(int storage, bool prepare, int fam, int dir, bool hi, int block, const void* args,
                        unsigned grid, size_t lds, hipStream_t s) {
  switch (storage) {
    case kF32: return group_family_dispatch<float2, float2>(prepare, fam, dir, hi, block, args, grid, lds, s);
    case kF64: return group_family_dispatch<double2, double2>(prepare, fam, dir, hi, block, args, grid, lds, s);
    case kBF16: return group_family_lowp(true, prepare, fam, dir, hi, block, args, grid, lds, s);
    case kF16: return group_family_lowp(false, prepare, fam, dir, hi, block, args, grid, lds, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fyfft
