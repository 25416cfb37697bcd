// k_colown_i8.hip -- column-owned dense sweep, int8 couplings / int32 fields.
#include "ojmi_colown.h"

namespace ojmi {

hipError_t launch_colown_i8(const SharedArgs& a, int NS, int G, hipStream_t st) {
  return launch_colown_ft<int32_t, int8_t>(a, NS, G, st);
}

}  // namespace ojmi
