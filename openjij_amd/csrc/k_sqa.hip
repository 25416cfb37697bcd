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

int sqa_lds_slots(int ftype, int NS, int n, int M, size_t budget) {
  // the widest of 4, 2, 1 slots per row whose copy (built where the largest
  // system of NS slots fits 160 KiB) fits beside the slices' spin words
  if (ftype != FT_F32 && ftype != FT_F64) return 0;
  const size_t fb = ftype == FT_F64 ? 8 : 4;
  const size_t spin_words = (size_t)M * (size_t)NS * 8;
  for (int sl : {4, 2, 1}) {
    const bool built = sl <= NS && (size_t)sl * 64 * NS * 64 * fb <= 160 * 1024;
    if (built && spin_words + (size_t)sl * (size_t)n * 64 * fb <= budget) return sl;
  }
  return 0;
}

}  // namespace ojmi
