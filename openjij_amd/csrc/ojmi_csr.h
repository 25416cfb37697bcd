// ojmi_csr.h -- one pass of a level-order CSR sweep: the neighbour lists of
// 64 level-order positions, loaded one pass ahead by the CSR kernels
// (k_csr.hip, k_sqa_csr.hip).
#pragma once
#include "ojmi_kernels.h"

namespace ojmi {

// Neighbour lists are padded to DMAX entries (zero couplings on column 0), so
// every pass loads the same number of values and the lists of the next pass
// can be in flight while the current one is processed (DMAX <= 16; wider
// lists are loaded at the start of their own pass).  A zero coupling adds +-0
// to the field, which changes no decision.
template <typename FT, int DMAX>
struct CsrPass {
  int i;        // spin index of this lane's position
  bool valid;   // position inside the level
  FT hv;
  int col[DMAX];
  FT val[DMAX];
};

// lp: the level pointers (global memory, or a copy in LDS); lv is
// wave-uniform, and so is the level end read from it.
template <typename FT, int DMAX>
__device__ __forceinline__ void csr_load_pass(const CsrArgs& a, const int32_t* lp, int lv,
                                              int base, int lane, CsrPass<FT, DMAX>& P) {
  const int b1 = __builtin_amdgcn_readfirstlane(lp[__builtin_amdgcn_readfirstlane(lv) + 1]);
  const int pos = base + lane;
  P.valid = pos < b1;
  const int pc = P.valid ? pos : b1 - 1;  // levels are never empty
  const int n = a.n;
  P.i = a.order[pc];
  P.hv = reinterpret_cast<const FT*>(a.h)[pc];
  const FT* ev = reinterpret_cast<const FT*>(a.val);
#pragma unroll
  for (int d = 0; d < DMAX; ++d) {
    P.col[d] = a.ell_col[(int64_t)d * n + pc];
    P.val[d] = ev[(int64_t)d * n + pc];
  }
}

}  // namespace ojmi
