// k_colown_i32.hip -- column-owned dense sweep, int32 couplings / int32 fields.
#include "ojmi_colown.h"

namespace ojmi {

hipError_t launch_colown_i32(const SharedArgs& a, int NS, int G, hipStream_t st) {
  return launch_colown_ft<int32_t, int32_t>(a, NS, G, st);
}

}  // namespace ojmi
