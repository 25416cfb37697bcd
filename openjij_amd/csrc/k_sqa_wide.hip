// k_sqa_wide.hip -- Trotter kernel instantiations for 2049-4096 spins (64
// slots of 64 spins per slice).  A slice's fields take 64 registers (FP32)
// or 128 (FP64) per lane, so only the 256-thread variant with one slice pair
// per wave (FP64) or up to two (FP32) fits the 512 registers of a wave:
// at most 8 slices (FP64) or 16 (FP32) per replica.
#include "ojmi_sqa.h"

namespace ojmi {

hipError_t launch_sqa_wide(int ftype, const DenseArgs& a, int PPW, int W, hipStream_t st) {
  if (W > 4) return hipErrorInvalidValue;
  dim3 grid(a.R), block(64 * W);
  const size_t lds = (size_t)a.M * 64 * sizeof(uint64_t);
  if (ftype == FT_F64 && PPW == 1) {
    hipLaunchKernelGGL((k_sqa_wave<double, 64, 1, 256, 0>), grid, block, lds, st, a);
  } else if (ftype == FT_F32 && PPW == 1) {
    hipLaunchKernelGGL((k_sqa_wave<float, 64, 1, 256, 0>), grid, block, lds, st, a);
  } else if (ftype == FT_F32 && PPW == 2) {
    hipLaunchKernelGGL((k_sqa_wave<float, 64, 2, 256, 0>), grid, block, lds, st, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ojmi
