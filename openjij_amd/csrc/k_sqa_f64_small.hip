// k_sqa_f64_small.hip -- Trotter kernel instantiations, double fields, up to 512 spins.
#include "ojmi_sqa.h"

namespace ojmi {

hipError_t launch_sqa_f64_small(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st) {
  switch (NS) {
    case 1: return launch_sqa_ns<double, 1>(a, PPW, W, st);
    case 2: return launch_sqa_ns<double, 2>(a, PPW, W, st);
    case 4: return launch_sqa_ns<double, 4>(a, PPW, W, st);
    case 8: return launch_sqa_ns<double, 8>(a, PPW, W, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ojmi
