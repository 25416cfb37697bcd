// ojmi_wave.h -- register-resident local fields for the wave-per-replica
// dense kernels.
//
// Layout: a replica (or a Trotter slice of one) is owned by one wave64.  Lane l
// holds the local fields F_j of the spins j = 64*m + l for slots m = 0..NS-1,
// and the spin signs as bit m of a 64-bit word (bit set = spin -1).  The
// fields are kept in clang vectors of at most 1024 bits, which the compiler
// keeps in VGPRs; a wave-uniform dynamic slot index (the slot being decided)
// becomes s_set_gpr_idx indirect addressing instead of scratch memory.
//
// Device coupling layout (row i of the padded Npad x Npad matrix, Npad = 64*NS):
//   fp64 / fp32 / int32 : element j at position j (natural order)
//   int8               : element (m, l) at byte ((m>>2)*64 + l)*4 + (m&3), so
//                        that one dword load per lane fetches 4 slots and one
//                        wave instruction reads 256 contiguous bytes.
#pragma once
#include "ojmi_device.h"

namespace ojmi {

template <typename FT, int NS>
struct FieldVec {
  static constexpr int EPV = (sizeof(FT) == 8) ? 16 : 32;
  static constexpr int VL0 = NS < EPV ? NS : EPV;
  static constexpr int VL = VL0 < 2 ? 2 : VL0;
  static constexpr int NV = (NS + VL - 1) / VL;
  typedef FT vec_t __attribute__((ext_vector_type(VL)));
  vec_t v[NV];

  __device__ __forceinline__ FT get(int m) const { return v[m / VL][m % VL]; }
  __device__ __forceinline__ void set(int m, FT x) { v[m / VL][m % VL] = x; }
  // k is wave-uniform and in [0, NS).  readfirstlane makes that explicit to
  // the compiler so the access becomes s_set_gpr_idx indirect addressing on a
  // scalar index (without it, index arithmetic placed in VGPRs turns into a
  // waterfall loop).  For 32-bit fields an unsigned split keeps the vector /
  // element arithmetic to a shift and a mask; for FP64 fields that form makes
  // the compiler move the whole vector to scratch memory, so FP64 keeps the
  // signed split.
  __device__ __forceinline__ FT dyn(int k) const {
    if constexpr (NV == 1) {
      return v[0][__builtin_amdgcn_readfirstlane(k)];
    } else if constexpr (sizeof(FT) == 8) {
      const int ku = __builtin_amdgcn_readfirstlane(k);
      const int q = ku / VL, e = ku - q * VL;
      FT r = v[0][e];
#pragma unroll
      for (int qq = 1; qq < NV; ++qq)
        if (q == qq) r = v[qq][e];
      return r;
    } else {
      const unsigned ku = (unsigned)__builtin_amdgcn_readfirstlane(k);
      const unsigned q = ku / (unsigned)VL, e = ku % (unsigned)VL;
      FT r = v[0][e];
#pragma unroll
      for (int qq = 1; qq < NV; ++qq)
        if (q == (unsigned)qq) r = v[qq][e];
      return r;
    }
  }
};

// Number of bytes between consecutive rows of the device coupling matrix.
template <typename JT, int NS>
__host__ __device__ constexpr int64_t row_bytes() {
  if constexpr (sizeof(JT) == 1) {
    return (int64_t)((NS + 3) / 4) * 256;
  } else {
    return (int64_t)NS * 64 * (int64_t)sizeof(JT);
  }
}
// The same for a slot count known at run time.
template <typename JT>
__host__ __device__ inline int64_t row_bytes_rt(int ns) {
  if constexpr (sizeof(JT) == 1) {
    return (int64_t)((ns + 3) / 4) * 256;
  } else {
    return (int64_t)ns * 64 * (int64_t)sizeof(JT);
  }
}

// Value of element (slot m, lane) of a coupling row.  ns (wave-uniform,
// <= NS): slots the row holds; slots m >= ns read as 0 and are not loaded.
template <typename JT, typename FT, int NS>
struct RowLoad {
  FT x[NS];
  __device__ __forceinline__ void load(const char* __restrict__ row, int lane, int ns = NS) {
    if constexpr (sizeof(JT) == 1) {
      constexpr int NQ = (NS + 3) / 4;
      uint32_t w[NQ];
      const uint32_t* p = reinterpret_cast<const uint32_t*>(row);
#pragma unroll
      for (int q = 0; q < NQ; ++q) w[q] = (4 * q < ns) ? p[q * 64 + lane] : 0u;
#pragma unroll
      for (int m = 0; m < NS; ++m) {
        const int32_t b = ((int32_t)(w[m >> 2] << (24 - 8 * (m & 3)))) >> 24;
        x[m] = (FT)b;
      }
    } else {
      const JT* p = reinterpret_cast<const JT*>(row);
#pragma unroll
      for (int m = 0; m < NS; ++m) x[m] = (m < ns) ? (FT)p[m * 64 + lane] : (FT)0;
    }
  }
};

template <typename FT>
__device__ __forceinline__ FT fma_ft(FT a, FT b, FT c);
template <>
__device__ __forceinline__ double fma_ft<double>(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_ft<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
template <>
__device__ __forceinline__ int32_t fma_ft<int32_t>(int32_t a, int32_t b, int32_t c) {
  return a * b + c;
}

// F += c * row  (c = +-2, so c * J is exact and each field sees one rounding).
// FP64 fields of 64 slots fill 128 VGPRs; a whole FP64 row in flight would take
// the other 128, so such rows are applied in two halves of 32 slots.
// ns (wave-uniform, <= NS): slots the row holds; fields of the other slots
// (columns past the problem) receive zero updates.
template <typename JT, typename FT, int NS>
__device__ __forceinline__ void row_axpy(FieldVec<FT, NS>& F, const char* __restrict__ row,
                                         int lane, FT c, int ns = NS) {
  constexpr bool halves = sizeof(FT) == 8 && sizeof(JT) == 8 && NS > 32;
  if constexpr (halves) {
    const JT* p = reinterpret_cast<const JT*>(row);
#pragma unroll
    for (int m0 = 0; m0 < NS; m0 += 32) {
      if (m0 >= ns) break;
      FT x[32];
#pragma unroll
      for (int m = 0; m < 32; ++m) x[m] = (m0 + m < ns) ? (FT)p[(m0 + m) * 64 + lane] : (FT)0;
#pragma unroll
      for (int m = 0; m < 32; ++m) F.set(m0 + m, fma_ft<FT>(c, x[m], F.get(m0 + m)));
    }
  } else {
    RowLoad<JT, FT, NS> r;
    r.load(row, lane, ns);
    // slots past ns read as 0: their fields (columns past the problem) stay as they are
#pragma unroll
    for (int m = 0; m < NS; ++m) F.set(m, fma_ft<FT>(c, r.x[m], F.get(m)));
  }
}

// F = h + sum_k s_k J[k][.], k = 0..n-1 in order (one rounding per term).
// `down` holds the lane's spin bits; the wave-uniform spin of k comes from a
// ballot of slot k>>6.
// ns (wave-uniform, <= NS): the problem's slot count (rows of ns slots).
template <typename JT, typename FT, int NS>
__device__ __forceinline__ void init_fields_seq(FieldVec<FT, NS>& F, const char* __restrict__ J,
                                                const FT* __restrict__ h, uint64_t down,
                                                int n, int lane, int ns = NS) {
#pragma unroll
  for (int m = 0; m < NS; ++m) F.set(m, (m < ns) ? h[m * 64 + lane] : (FT)0);
  const int64_t RB = (ns == NS) ? row_bytes<JT, NS>() : row_bytes_rt<JT>(ns);
  for (int q = 0; q < ns; ++q) {
    const uint64_t bal = ballot64((down >> q) & 1ull);
    const int kend = min(64, n - 64 * q);
    for (int l2 = 0; l2 < kend; ++l2) {
      const int k = 64 * q + l2;
      const FT c = ((bal >> l2) & 1ull) ? (FT)(-1) : (FT)1;
      row_axpy<JT, FT, NS>(F, J + (int64_t)k * RB, lane, c, ns);
    }
  }
}

}  // namespace ojmi
