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

}  // namespace ojmi
