// k_sqa.hip -- Suzuki-Trotter dense sweeps (ojmi_sqa.h): dispatch by field
// type and slot count.
#include "ojmi_sqa.h"

namespace ojmi {

hipError_t launch_sqa_wave(int ftype, const DenseArgs& a, int NS, int PPW, int W,
                           hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (NS == 64) return launch_sqa_wide(ftype, a, PPW, W, st);
  const bool large = NS >= 16;
  if (ftype == FT_F64)
    return large ? launch_sqa_f64_large(a, NS, PPW, W, st) : launch_sqa_f64_small(a, NS, PPW, W, st);
  if (ftype == FT_F32)
    return large ? launch_sqa_f32_large(a, NS, PPW, W, st) : launch_sqa_f32_small(a, NS, PPW, W, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
