// k_energy.hip -- exact energies of returned states for batches of problems
// (the single-problem energies run on the matrix cores, k_spin_gemm.hip).
#include "ojmi_kernels.h"

namespace ojmi {

// ---------------------------------------------------------------------------
// Energies of a batch of problems, one wave per (problem, replica):
//   E = sum_i s_i (h_i + sum_{j>i} J_ij s_j)
// read from the sweep's own coupling copy (row stride rb[k] bytes; int8 rows
// in the packed layout, whose element offset does not depend on the row
// length).  Integer problems accumulate in int64 (exact); real ones in FP64,
// per lane in index order and then across lanes in a fixed butterfly order.
template <typename JT>
__device__ __forceinline__ JT batch_coupling(const char* row, int j) {
  if constexpr (sizeof(JT) == 1)
    return (JT)row[(((j >> 8) * 64 + (j & 63)) << 2) + ((j >> 6) & 3)];
  else
    return reinterpret_cast<const JT*>(row)[j];
}

template <typename JT, typename HT, typename AT>
__global__ __launch_bounds__(256) void k_batch_energy(BatchArgs b, const int64_t* __restrict__ rb,
                                                     int K, double* __restrict__ E) {
  const int lane = threadIdx.x & 63;
  const int g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= K * b.R) return;
  const int k = g / b.R, r = g - k * b.R;
  const int n = b.n[k];
  const char* J = reinterpret_cast<const char*>(b.J) + b.j_off[k];
  const HT* h = reinterpret_cast<const HT*>(b.h) + b.h_off[k];
  const int8_t* sp = b.spins + b.s_off[k] + (int64_t)r * n;
  const int64_t stride = rb[k];
  AT part = 0;
  for (int i = 0; i < n; ++i) {
    const char* row = J + (int64_t)i * stride;
    AT acc = 0;
    for (int j = i + 1 + lane; j < n; j += 64) acc += (AT)batch_coupling<JT>(row, j) * (AT)sp[j];
    if (lane == 0) acc += (AT)h[i];
    part += (AT)sp[i] * acc;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
  if (lane == 0) E[g] = (double)part;
}

hipError_t launch_batch_energy(int ftype, int jtype, const BatchArgs& b, const int64_t* rb, int K,
                               double* E, hipStream_t st) {
  if (K <= 0 || b.R <= 0) return hipSuccess;
  dim3 grid((unsigned)(((int64_t)K * b.R + 3) / 4)), block(256);
  if (ftype == FT_I32 && jtype == JT_I8)
    hipLaunchKernelGGL((k_batch_energy<int8_t, int32_t, long long>), grid, block, 0, st, b, rb, K, E);
  else if (ftype == FT_I32 && jtype == JT_I32)
    hipLaunchKernelGGL((k_batch_energy<int32_t, int32_t, long long>), grid, block, 0, st, b, rb, K, E);
  else if (ftype == FT_F32 && jtype == JT_F32)
    hipLaunchKernelGGL((k_batch_energy<float, float, double>), grid, block, 0, st, b, rb, K, E);
  else if (ftype == FT_F64 && jtype == JT_F64)
    hipLaunchKernelGGL((k_batch_energy<double, double, double>), grid, block, 0, st, b, rb, K, E);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace ojmi
