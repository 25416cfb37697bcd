// k_sqa_f32_large.hip -- Trotter kernel instantiations, float fields, 513-2048 spins.
#include "ojmi_sqa.h"

namespace ojmi {

hipError_t launch_sqa_f32_large(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st) {
  switch (NS) {
    case 16: return launch_sqa_ns<float, 16>(a, PPW, W, st);
    case 32: return launch_sqa_ns<float, 32>(a, PPW, W, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ojmi
