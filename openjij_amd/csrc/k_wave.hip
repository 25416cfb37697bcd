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

// The same anneal with the coupling rows in LDS: the workgroup copies the
// n rows (row_bytes each) once, then every accepted flip reads its row from
// LDS instead of L2.  For systems whose rows fit (int8 rows of up to 256
// spins: 64 KiB), where the flip-to-flip chain of a replica waits on that
// read.
template <typename FT, typename JT, int NS>
__global__ __launch_bounds__(256) void k_dense_wave_lds(DenseArgs a) {
  typedef typename ThreshT<FT>::type TT;
  extern __shared__ __align__(16) char sJ[];
  constexpr int64_t RB = row_bytes<JT, NS>();
  const int64_t words = (int64_t)a.n * RB / 16;  // RB is a multiple of 16
  const uint4* src = reinterpret_cast<const uint4*>(a.J);
  uint4* dst = reinterpret_cast<uint4*>(sJ);
  for (int64_t e = threadIdx.x; e < words; e += blockDim.x) dst[e] = src[e];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (r >= a.R) return;  // after the only barrier
  const FT* F_in = a.F_in ? reinterpret_cast<const FT*>(a.F_in) + (int64_t)r * (64 * NS) : nullptr;
  FT* F_out = a.F_out ? reinterpret_cast<FT*>(a.F_out) + (int64_t)r * (64 * NS) : nullptr;
  anneal_replica<FT, JT, NS>(sJ, reinterpret_cast<const FT*>(a.h), a.spins + (int64_t)r * a.n,
                             a.n, F_in, F_out, (uint32_t)(a.rep_offset + r), 0u, a.seed, a.t0,
                             a.T, p0_ptr<TT>(a.p0_d, a.p0_f), a.rule, lane, a.stats);
}

size_t wave_lds_bytes(int n, int NS, int jbytes) {
  const int64_t rb = jbytes == 1 ? (int64_t)((NS + 3) / 4) * 256 : (int64_t)NS * 64 * jbytes;
  return (size_t)n * (size_t)rb;
}

template <typename FT, typename JT, int NS>
static hipError_t launch_wave_lds_t(const DenseArgs& a, int wpb, hipStream_t st) {
  const size_t lds = wave_lds_bytes(a.n, NS, (int)sizeof(JT));
  auto kern = k_dense_wave_lds<FT, JT, NS>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((a.R + wpb - 1) / wpb), block(64 * wpb);
  hipLaunchKernelGGL(kern, grid, block, lds, st, a);
  return hipGetLastError();
}

template <typename FT, typename JT>
static hipError_t launch_wave_lds_ft(const DenseArgs& a, int NS, int wpb, hipStream_t st) {
  switch (NS) {
    case 1: return launch_wave_lds_t<FT, JT, 1>(a, wpb, st);
    case 2: return launch_wave_lds_t<FT, JT, 2>(a, wpb, st);
    case 4: return launch_wave_lds_t<FT, JT, 4>(a, wpb, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_dense_wave_lds(int ftype, int jtype, const DenseArgs& a, int NS, int wpb,
                                 hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (wpb < 1 || wpb > 4) return hipErrorInvalidValue;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_wave_lds_ft<double, double>(a, NS, wpb, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_wave_lds_ft<float, float>(a, NS, wpb, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_wave_lds_ft<int32_t, int8_t>(a, NS, wpb, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_wave_lds_ft<int32_t, int32_t>(a, NS, wpb, st);
  return hipErrorInvalidValue;
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
