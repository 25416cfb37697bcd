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

// ---------------------------------------------------------------------------
// Energies of sparse systems from the CSR kernel's own level-order ELL arrays
// (no dense copy): one workgroup of 256 per configuration c,
//   E_c = sum_pos s_i (h_i + sum_{d < deg, j > i} v[d][pos] s_j),  i = order[pos],
// each pair counted once (the neighbour lists hold both directions).  Thread t
// sums positions t, t + 256, ... in order in FP64, then a fixed butterfly and
// the four wave sums in order: deterministic, and exact for integer couplings.
__global__ __launch_bounds__(256) void k_csr_energy(const int8_t* __restrict__ spins, int n,
                                                   const int32_t* __restrict__ order,
                                                   const int32_t* __restrict__ deg,
                                                   const int32_t* __restrict__ ell,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ hpos,
                                                   double* __restrict__ E) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  const int8_t* s = spins + (int64_t)c * n;
  double part = 0.0;
  for (int pos = threadIdx.x; pos < n; pos += 256) {
    const int i = order[pos];
    double acc = hpos[pos];
    const int dg = deg[pos];
    for (int d = 0; d < dg; ++d) {
      const int j = ell[(int64_t)d * n + pos];
      if (j > i) acc += val[(int64_t)d * n + pos] * (double)s[j];
    }
    part += (double)s[i] * acc;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
  __syncthreads();
  if (threadIdx.x == 0) E[c] = ((red[0] + red[1]) + red[2]) + red[3];
}

hipError_t launch_csr_energy(const int8_t* spins, int n, int C, const int32_t* order,
                             const int32_t* deg, const int32_t* ell, const double* val,
                             const double* hpos, double* E, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_csr_energy, dim3(C), dim3(256), 0, st, spins, n, order, deg, ell, val,
                     hpos, E);
  return hipGetLastError();
}

}  // namespace ojmi
