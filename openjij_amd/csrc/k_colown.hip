// k_colown.hip -- column-owned dense sweep (ojmi_colown.h): shape rules,
// LDS size and dispatch by type; FP32 instantiations.
#include "ojmi_colown.h"

namespace ojmi {

size_t colown_lds_bytes(int n, int NS, int G, int jbytes, int fbytes) {
  (void)NS;
  // tile, block fields [G][64], flip and sign words [G], spin bytes [G][64],
  // init table of per-column spin bytes
  return (size_t)64 * 64 * jbytes + (size_t)G * 64 * fbytes + 2 * 8 * (size_t)G +
         (size_t)G * 64 + 4 * (size_t)((n + 3) / 4) + 16;
}

hipError_t launch_colown_f32(const SharedArgs& a, int NS, int G, hipStream_t st) {
  return launch_colown_ft<float, float>(a, NS, G, st);
}

bool colown_supported(int NS, int G) { return (G == 4 || G == 8) && NS >= 8 && NS % G == 0; }

hipError_t launch_dense_colown(int ftype, int jtype, const SharedArgs& a, int NS, int G,
                               hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (!colown_supported(NS, G)) return hipErrorInvalidValue;
  if (ftype == FT_F32 && jtype == JT_F32) return launch_colown_f32(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_colown_i8(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_colown_i32(a, NS, G, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
