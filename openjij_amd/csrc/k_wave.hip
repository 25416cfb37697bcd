// k_wave.hip -- classical dense sweeps, one wave64 per replica of one problem
// (the anneal itself is in ojmi_anneal.h).
#include "ojmi_anneal.h"

namespace ojmi {

// One wave per replica of one problem.
template <typename FT, typename JT, int NS>
__global__ __launch_bounds__(256) void k_dense_wave(DenseArgs a) {
  typedef typename ThreshT<FT>::type TT;
  const int lane = threadIdx.x & 63;
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (r >= a.R) return;
  const FT* F_in = a.F_in ? reinterpret_cast<const FT*>(a.F_in) + (int64_t)r * (64 * NS) : nullptr;
  FT* F_out = a.F_out ? reinterpret_cast<FT*>(a.F_out) + (int64_t)r * (64 * NS) : nullptr;
  anneal_replica<FT, JT, NS>(reinterpret_cast<const char*>(a.J), reinterpret_cast<const FT*>(a.h),
                             a.spins + (int64_t)r * a.n, a.n, F_in, F_out,
                             (uint32_t)(a.rep_offset + r), 0u, a.seed, a.t0, a.T,
                             p0_ptr<TT>(a.p0_d, a.p0_f), a.rule, lane, a.stats);
}

template <typename FT, typename JT>
static hipError_t launch_dense_ft(const DenseArgs& a, int NS, hipStream_t st) {
  const int wpb = 4;
  dim3 grid((a.R + wpb - 1) / wpb), block(64 * wpb);
  switch (NS) {
#define OJMI_CASE(ns)                                                          \
  case ns:                                                                     \
    hipLaunchKernelGGL((k_dense_wave<FT, JT, ns>), grid, block, 0, st, a);     \
    break;
    OJMI_CASE(1)
    OJMI_CASE(2)
    OJMI_CASE(4)
    OJMI_CASE(8)
    OJMI_CASE(16)
    OJMI_CASE(32)
    OJMI_CASE(64)
#undef OJMI_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dense_wave(int ftype, int jtype, const DenseArgs& a, int NS, hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_dense_ft<double, double>(a, NS, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_dense_ft<float, float>(a, NS, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_dense_ft<int32_t, int8_t>(a, NS, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_dense_ft<int32_t, int32_t>(a, NS, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
