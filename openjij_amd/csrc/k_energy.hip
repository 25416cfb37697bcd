// k_energy.hip -- FP64 field evaluation and energies for returned states.
//
//   H[c][j] = sum_i s_c,i J64[i][j]          (tiled FP64 GEMM, VALU)
//   E[c]    = sum_j s_c,j (H[c][j]/2 + h_j)  (= sum_{i<j} J s s + sum h s,
//                                             include/openjij/graph/dense.hpp:176-192)
//   drift[c] = max_j |F32[c][j] - (H[c][j] + h_j)|   (when F32 is given)
//
// This is the only FP64 work of the mixed-precision path.  The reduction of
// each configuration runs in a fixed order, so energies are deterministic.
#include "ojmi_kernels.h"

namespace ojmi {

constexpr int EB_M = 64, EB_N = 64, EB_K = 16;

// grid: (Npad/64, ceil(C/64)), block 256.  Each thread computes 4x4 outputs.
__global__ __launch_bounds__(256) void k_gemm_spin_f64(const double* __restrict__ J,
                                                      const int8_t* __restrict__ S, int n,
                                                      int Npad, int C,
                                                      double* __restrict__ H) {
  __shared__ double sS[EB_K][EB_M + 1];
  __shared__ double sJ[EB_K][EB_N];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.y * EB_M, j0 = blockIdx.x * EB_N;
  const int tr = (tid >> 4) * 4, tc = (tid & 15) * 4;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;

  for (int k0 = 0; k0 < n; k0 += EB_K) {
    // S tile: 64 configs x 16 spins (1024 entries, 4 per thread)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int cc = e >> 4, kk = e & 15;
      const int c = c0 + cc, k = k0 + kk;
      double v = 0.0;
      if (c < C && k < n) v = (double)S[(int64_t)c * n + k];
      sS[kk][cc] = v;
    }
    // J tile: 16 rows x 64 columns
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;
      const int kk = e >> 6, jj = e & 63;
      const int k = k0 + kk;
      sJ[kk][jj] = (k < n) ? J[(int64_t)k * Npad + j0 + jj] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < EB_K; ++kk) {
      double sv[4], jv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) sv[a] = sS[kk][tr + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) jv[b] = sJ[kk][tc + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_fma(sv[a], jv[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int c = c0 + tr + a;
    if (c >= C) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) H[(int64_t)c * Npad + j0 + tc + b] = acc[a][b];
  }
}

// One wave per configuration: energy and drift in a fixed reduction order.
__global__ __launch_bounds__(256) void k_energy_reduce(const double* __restrict__ H,
                                                      const double* __restrict__ h,
                                                      const int8_t* __restrict__ S,
                                                      const float* __restrict__ F32, int n,
                                                      int Npad, int C, double* __restrict__ E,
                                                      double* __restrict__ drift) {
  const int lane = threadIdx.x & 63;
  const int c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (c >= C) return;
  double e = 0.0, d = 0.0;
  for (int j = lane; j < n; j += 64) {
    const double hj = h[j];
    const double Hj = H[(int64_t)c * Npad + j];
    const double s = (double)S[(int64_t)c * n + j];
    e = __builtin_fma(s, __builtin_fma(0.5, Hj, hj), e);
    if (F32) {
      const double diff = __builtin_fabs((double)F32[(int64_t)c * Npad + j] - (Hj + hj));
      d = diff > d ? diff : d;
    }
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    e += __shfl_xor(e, off);
    const double od = __shfl_xor(d, off);
    d = od > d ? od : d;
  }
  if (lane == 0) {
    E[c] = e;
    if (drift) drift[c] = d;
  }
}

hipError_t launch_energy_f64(const double* J64, const double* h64, const int8_t* spins, int n,
                             int Npad, int C, double* H, double* E, const float* F32,
                             double* drift, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  dim3 g1(Npad / EB_N, (C + EB_M - 1) / EB_M);
  hipLaunchKernelGGL(k_gemm_spin_f64, g1, dim3(256), 0, st, J64, spins, n, Npad, C, H);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_reduce, dim3((C + 3) / 4), dim3(256), 0, st, H, h64, spins, F32,
                     n, Npad, C, E, drift);
  return hipGetLastError();
}

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
