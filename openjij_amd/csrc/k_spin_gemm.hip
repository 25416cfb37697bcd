// k_spin_gemm.hip -- the dense contractions of the path on the matrix cores
// (gfx950 MFMA): local fields H = S J of every replica at once, and the FP64
// energies s^T J s of returned configurations.
//
//   fields (reset_dE, include/openjij/system/classical_ising.hpp:87-90, for R
//   replicas at once):
//     F[c][j] = h_j + sum_i S[c][i] J[i][j]
//       * bf16 hi/lo: J = J_hi + J_lo with J_hi = bf16(J), J_lo = bf16(J - J_hi)
//         (|J - J_hi - J_lo| <= 2^-18 |J|), spins +-1 exact in bf16, both
//         halves accumulated into one FP32 accumulator
//         (v_mfma_f32_16x16x32_bf16): the FP32 path's initial fields and its
//         periodic refresh;
//       * int8: J and spins int8, int32 accumulation (v_mfma_i32_16x16x64_i8):
//         exact, so the result equals the sequential accumulation bit for bit.
//   energies (Dense::energy, include/openjij/graph/dense.hpp:176-192):
//     H = S J64 in FP64 (v_mfma_f64_16x16x4_f64), fused with
//     E[c] = sum_j s_cj (H_cj / 2 + h_j) and the drift check
//     max_j |F32[c][j] - (H_cj + h_j)|, both as per-tile partials reduced in a
//     fixed order (deterministic).
//
// Operands: S is the spins zero-padded to [C][Kp] (k_pack_spins), J the
// zero-padded Npad x Npad natural-layout couplings with Kp = Npad.  MFMA lane
// layout (16x16 tiles): lane l supplies A[l % 16][k-group of l / 16] and
// B[k-group of l / 16][l % 16], and receives D[4 (l / 16) + j][l % 16] in its
// j-th accumulator (bf16 and int8 forms; the FP64 form holds D[4 j + l / 16]
// [l % 16] there, as a lane-layout probe on gfx950 showed).  A dot product does not depend on how the K index is
// split over lanes as long as A and B use the same split, so each lane loads
// 8 (bf16) or 16 (int8) consecutive k of one spin row and, through J's
// symmetry (B[k][col] = J[col][k]), of one coupling row: 16-byte loads.
//
// Tiling: a workgroup is 4 waves stacked along the configurations (64 each)
// over the same 64 columns, so the coupling tile a wave loads is the one its
// neighbours load (L1 hits); grid = (Npad / 64, ceil(C / 256)).
#include "ojmi_kernels.h"

#include <algorithm>
#include <mutex>

namespace ojmi {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef double f64x4_t __attribute__((ext_vector_type(4)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

constexpr int GW = 4;    // waves per workgroup (along the configurations)
constexpr int WM = 64;   // configurations per wave
constexpr int TN = 64;   // columns per workgroup

__device__ __forceinline__ uint16_t bf16_of_spin(int8_t s) {
  // +1 -> 0x3F80, -1 -> 0xBF80, 0 -> 0
  return s == 0 ? (uint16_t)0 : (uint16_t)(0x3F80u | (s < 0 ? 0x8000u : 0u));
}

// ---------------------------------------------------------------------------
// S8[c][k] = spins[c][k] (k < n) else 0, and the same as bf16 bits in Sbf.
__global__ __launch_bounds__(256) void k_pack_spins(const int8_t* __restrict__ spins, int n,
                                                   int C, int Kp, int8_t* __restrict__ S8,
                                                   uint16_t* __restrict__ Sbf) {
  const int64_t total = (int64_t)C * Kp;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(e / Kp), k = (int)(e - (int64_t)c * Kp);
    const int8_t v = (k < n) ? spins[(int64_t)c * n + k] : (int8_t)0;
    if (S8) S8[e] = v;
    if (Sbf) Sbf[e] = bf16_of_spin(v);
  }
}

// ---------------------------------------------------------------------------
// F = h + S (J_hi + J_lo), FP32 accumulation.
__global__ __launch_bounds__(256) void k_fields_bf16x2(const uint16_t* __restrict__ S,
                                                      const uint16_t* __restrict__ Jhi,
                                                      const uint16_t* __restrict__ Jlo,
                                                      const float* __restrict__ h, int C, int Kp,
                                                      int Npad, float* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int n0 = blockIdx.x * TN;
  const int m0 = (blockIdx.y * GW + w) * WM;
  if (m0 >= C) return;  // whole wave; the kernel has no barriers
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const uint16_t* ap[4];
  const uint16_t* bh[4];
  const uint16_t* bl[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int r = min(m0 + 16 * t + lr, C - 1);
    ap[t] = S + (int64_t)r * Kp + lk;
    const int col = n0 + 16 * t + lr;
    bh[t] = Jhi + (int64_t)col * Npad + lk;
    bl[t] = Jlo + (int64_t)col * Npad + lk;
  }
  f32x4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f32x4_t){0.f, 0.f, 0.f, 0.f};

  auto load = [&](int k0, u32x4_t (&A)[4], u32x4_t (&BH)[4], u32x4_t (&BL)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      A[t] = *reinterpret_cast<const u32x4_t*>(ap[t] + k0);
      BH[t] = *reinterpret_cast<const u32x4_t*>(bh[t] + k0);
      BL[t] = *reinterpret_cast<const u32x4_t*>(bl[t] + k0);
    }
  };
  auto mma = [&](const u32x4_t (&A)[4], const u32x4_t (&BH)[4], const u32x4_t (&BL)[4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const bf16x8_t av = __builtin_bit_cast(bf16x8_t, A[a]);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            av, __builtin_bit_cast(bf16x8_t, BH[b]), acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            av, __builtin_bit_cast(bf16x8_t, BL[b]), acc[a][b], 0, 0, 0);
      }
    }
  };
  // two register stages: the next K step's operands load while this one runs
  u32x4_t A0[4], H0[4], L0[4], A1[4], H1[4], L1[4];
  load(0, A0, H0, L0);
  int k0 = 0;
  for (; k0 + 64 <= Kp; k0 += 64) {
    load(k0 + 32, A1, H1, L1);
    mma(A0, H0, L0);
    if (k0 + 64 < Kp) load(k0 + 64, A0, H0, L0);
    mma(A1, H1, L1);
  }
  if (k0 < Kp) mma(A0, H0, L0);  // Kp is a multiple of 32; one step left

#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int col = n0 + 16 * b + lr;
    const float hv = h[col];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = m0 + 16 * a + 4 * (lane >> 4) + j;
        if (c < C) F[(int64_t)c * Npad + col] = acc[a][b][j] + hv;
      }
  }
}

// ---------------------------------------------------------------------------
// F = h + S J8, exact int32.
__global__ __launch_bounds__(256) void k_fields_i8(const int8_t* __restrict__ S,
                                                  const int8_t* __restrict__ J8,
                                                  const int32_t* __restrict__ h, int C, int Kp,
                                                  int Npad, int32_t* __restrict__ F) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int n0 = blockIdx.x * TN;
  const int m0 = (blockIdx.y * GW + w) * WM;
  if (m0 >= C) return;
  const int lr = lane & 15, lk = (lane >> 4) * 16;
  const int8_t* ap[4];
  const int8_t* bp[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    ap[t] = S + (int64_t)min(m0 + 16 * t + lr, C - 1) * Kp + lk;
    bp[t] = J8 + (int64_t)(n0 + 16 * t + lr) * Npad + lk;
  }
  i32x4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (i32x4_t){0, 0, 0, 0};
  auto load = [&](int k0, u32x4_t (&A)[4], u32x4_t (&B)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      A[t] = *reinterpret_cast<const u32x4_t*>(ap[t] + k0);
      B[t] = *reinterpret_cast<const u32x4_t*>(bp[t] + k0);
    }
  };
  auto mma = [&](const u32x4_t (&A)[4], const u32x4_t (&B)[4]) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i32x4_t, A[a]),
                                                          __builtin_bit_cast(i32x4_t, B[b]),
                                                          acc[a][b], 0, 0, 0);
  };
  u32x4_t A0[4], B0[4], A1[4], B1[4];
  load(0, A0, B0);
  int k0 = 0;
  for (; k0 + 128 <= Kp; k0 += 128) {
    load(k0 + 64, A1, B1);
    mma(A0, B0);
    if (k0 + 128 < Kp) load(k0 + 128, A0, B0);
    mma(A1, B1);
  }
  if (k0 < Kp) mma(A0, B0);  // Kp is a multiple of 64
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int col = n0 + 16 * b + lr;
    const int32_t hv = h[col];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = m0 + 16 * a + 4 * (lane >> 4) + j;
        if (c < C) F[(int64_t)c * Npad + col] = acc[a][b][j] + hv;
      }
  }
}

// ---------------------------------------------------------------------------
// FP64 energies and drift.  K is split so that MFMA step kk of a group of 16
// covers k = k0 + 4 (l / 16) + kk: lane l then reads one dword of 4 spins and,
// per step, element [k][col] of a natural coupling row (16 lanes read 16
// consecutive doubles of one row).
__global__ __launch_bounds__(256) void k_energy_f64_mfma(const int8_t* __restrict__ S,
                                                        const double* __restrict__ J,
                                                        const double* __restrict__ h, int C,
                                                        int Kp, int Npad,
                                                        const float* __restrict__ F32,
                                                        double* __restrict__ part_e,
                                                        double* __restrict__ part_d) {
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int n0 = blockIdx.x * TN;
  const int m0 = (blockIdx.y * GW + w) * WM;
  if (m0 >= C) return;
  const int lr = lane & 15, lq = lane >> 4;
  const int8_t* ap[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) ap[t] = S + (int64_t)min(m0 + 16 * t + lr, C - 1) * Kp + 4 * lq;
  f64x4_t acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = (f64x4_t){0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < Kp; k0 += 16) {
    uint32_t sw[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) sw[a] = *reinterpret_cast<const uint32_t*>(ap[a] + k0);
    double bv[4][4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const double* row = J + (int64_t)(k0 + 4 * lq + kk) * Npad + n0 + lr;
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[kk][b] = row[16 * b];
    }
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double av = (double)(int8_t)(sw[a] >> (8 * kk));
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[kk][b], acc[a][b], 0, 0, 0);
      }
    }
  }
  // epilogue: the FP64 16x16 result places D[4 j + l / 16][l % 16] in
  // register j of lane l (rows interleaved over the lane groups, unlike the
  // FP32 / int32 layouts above), so row c = m0 + 16 a + 4 j + lq, column
  // n0 + 16 b + lr
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = m0 + 16 * a + 4 * j + lq;
      const int cc = min(c, C - 1);
      double e = 0.0, d = 0.0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int col = n0 + 16 * b + lr;
        const double H = acc[a][b][j];
        const double s = (double)S[(int64_t)cc * Kp + col];
        e = __builtin_fma(s, __builtin_fma(0.5, H, h[col]), e);
        if (F32) {
          const double diff = __builtin_fabs((double)F32[(int64_t)cc * Npad + col] - (H + h[col]));
          d = diff > d ? diff : d;
        }
      }
      // fixed-order reduction over the 16 lanes of this row group
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        e += __shfl_xor(e, off);
        const double od = __shfl_xor(d, off);
        d = od > d ? od : d;
      }
      if (lr == 0 && c < C) {
        part_e[(int64_t)c * gridDim.x + blockIdx.x] = e;
        part_d[(int64_t)c * gridDim.x + blockIdx.x] = d;
      }
    }
}

__global__ __launch_bounds__(256) void k_energy_parts(const double* __restrict__ part_e,
                                                     const double* __restrict__ part_d, int C,
                                                     int nb, double* __restrict__ E,
                                                     double* __restrict__ drift) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  double e = 0.0, d = 0.0;
  for (int b = 0; b < nb; ++b) {
    e += part_e[(int64_t)c * nb + b];
    const double x = part_d[(int64_t)c * nb + b];
    d = x > d ? x : d;
  }
  E[c] = e;
  if (drift) drift[c] = d;
}

}  // namespace

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free energies (ojmi_energies_static): the
// padded spins and the per-tile partials live in this file-scope device
// array, one copy per GPU, held until the process ends.  In host code the
// name `scratch` is only a stand-in; kernels get the device address looked up
// by hipGetSymbolAddress on each device.
static __device__ __align__(256) unsigned char scratch[STATIC_SCRATCH_BYTES];

hipError_t static_scratch(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[MAX_DEVICES] = {};
  if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = 0;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

hipError_t launch_pack_spins(const int8_t* spins, int n, int C, int Kp, int8_t* S8,
                             uint16_t* Sbf, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  const int64_t total = (int64_t)C * Kp;
  const int blocks = (int)std::min<int64_t>(4096, (total + 255) / 256);
  hipLaunchKernelGGL(k_pack_spins, dim3(blocks), dim3(256), 0, st, spins, n, C, Kp, S8, Sbf);
  return hipGetLastError();
}

hipError_t launch_fields_bf16x2(const uint16_t* Sbf, const uint16_t* Jhi, const uint16_t* Jlo,
                                const float* h, int C, int Npad, float* F, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  if (Npad % TN) return hipErrorInvalidValue;
  dim3 grid(Npad / TN, (C + GW * WM - 1) / (GW * WM));
  hipLaunchKernelGGL(k_fields_bf16x2, grid, dim3(64 * GW), 0, st, Sbf, Jhi, Jlo, h, C, Npad, Npad,
                     F);
  return hipGetLastError();
}

hipError_t launch_fields_i8(const int8_t* S8, const int8_t* J8, const int32_t* h, int C, int Npad,
                            int32_t* F, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  if (Npad % TN) return hipErrorInvalidValue;
  dim3 grid(Npad / TN, (C + GW * WM - 1) / (GW * WM));
  hipLaunchKernelGGL(k_fields_i8, grid, dim3(64 * GW), 0, st, S8, J8, h, C, Npad, Npad, F);
  return hipGetLastError();
}

size_t energy_part_count(int C, int Npad) { return (size_t)C * (size_t)(Npad / TN); }

hipError_t launch_energy_f64(const int8_t* S8, const double* J64, const double* h64, int C,
                             int Npad, const float* F32, double* part_e, double* part_d,
                             double* E, double* drift, hipStream_t st) {
  if (C <= 0) return hipSuccess;
  if (Npad % TN) return hipErrorInvalidValue;
  dim3 grid(Npad / TN, (C + GW * WM - 1) / (GW * WM));
  hipLaunchKernelGGL(k_energy_f64_mfma, grid, dim3(64 * GW), 0, st, S8, J64, h64, C, Npad, Npad,
                     F32, part_e, part_d);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_energy_parts, dim3((C + 255) / 256), dim3(256), 0, st, part_e, part_d, C,
                     (int)grid.x, E, drift);
  return hipGetLastError();
}

}  // namespace ojmi
