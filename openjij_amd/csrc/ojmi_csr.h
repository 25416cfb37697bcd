// ojmi_csr.h -- one pass of a level-order CSR sweep: the neighbour lists of
// 64 level-order positions, loaded one pass ahead by the CSR kernels
// (k_csr.hip, k_sqa_csr.hip).
#pragma once
#include "ojmi_kernels.h"

#include <type_traits>

namespace ojmi {

// Neighbour lists are padded to DMAX entries (zero couplings on column 0), so
// every pass loads the same number of values and the lists of the next pass
// can be in flight while the current one is processed (DMAX <= 16 in the
// wave-per-replica kernel, 32 in the workgroup form; wider lists are loaded
// at the start of their own pass).  A zero coupling adds +-0
// to the field, which changes no decision.
// V8 (integer couplings within int8 and integer fields h, any field type):
// the values of a pass are DMAX / 4 words of four packed int8, and the field
// is formed by dot4 instructions on the gathered neighbour spins packed the
// same way, then converted to the field type.  The sum of such integers is
// exact in int32, and so is every partial sum of the per-neighbour adds in
// FP32 (|h| <= 2^20, so below 2^24) and FP64: the same field, bit for bit.
// C16 (spin indices below 65536): the columns are DMAX / 2 words of two
// 16-bit indices, loaded from the position records below.  Both forms cut
// the vector memory work of a pass, which bounds the workgroup form (texture
// units 60-70% busy with 18 dword loads per pass).
template <typename FT, int DMAX, bool V8 = false, bool C16 = false>
struct CsrPass {
  int i;        // spin index of this lane's position
  bool valid;   // position inside the level
  FT hv;
  int32_t h8;   // V8: the field h as an integer
  int col[C16 ? 1 : DMAX];
  uint32_t cw[C16 ? DMAX / 2 : 1];
  FT val[V8 ? 1 : DMAX];
  uint32_t v8[V8 ? DMAX / 4 : 1];
};

template <typename FT, int DMAX, bool V8, bool C16>
__device__ __forceinline__ int csr_col(const CsrPass<FT, DMAX, V8, C16>& P, int d) {
  if constexpr (C16) return (int)((P.cw[d >> 1] >> (16 * (d & 1))) & 0xffffu);
  else return P.col[d];
}

// Position records of the workgroup form (spin indices below 65536): the
// words of position pos are 16-byte chunks k at rec[k * n + pos], so each
// chunk of a pass is one coalesced 16-byte load per lane.  Words: the
// 16-bit column pairs [0, DMAX/2), the values (int8 quadruples if V8, else
// FT), the spin index, the field h (int32 if V8, else FT); zero padding to a
// multiple of 4.
template <typename FT, int DMAX, bool V8>
struct CsrRec {
  static constexpr int kCols = 0;
  static constexpr int kVals = DMAX / 2;
  static constexpr int kIdx = kVals + (V8 ? DMAX / 4 : DMAX * (int)sizeof(FT) / 4);
  static constexpr int kH = kIdx + 1;
  static constexpr int kWords = (kH + (V8 ? 1 : (int)sizeof(FT) / 4) + 3) / 4 * 4;
};

// The level-order arrays as buffer resources: a pass's loads then share one
// per-lane byte offset (its position) and take the ELL row d as a scalar
// offset, so no per-load address arithmetic runs on the vector unit.
struct CsrBufs {
  __amdgpu_buffer_rsrc_t order, h, col, val;
};

template <typename FT>
__device__ __forceinline__ CsrBufs csr_bufs(const CsrArgs& a) {
  constexpr int flags = 0x00020000;
  const int n = a.n, fb = (int)sizeof(FT);
  CsrBufs b;
  b.order = __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(a.order), (short)0, n * 4, flags);
  b.h = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.h), (short)0, n * fb, flags);
  b.col = __builtin_amdgcn_make_buffer_rsrc(const_cast<int32_t*>(a.ell_col), (short)0, a.D * n * 4,
                                            flags);
  b.val = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(a.val), (short)0, a.D * n * fb, flags);
  return b;
}

template <typename FT>
__device__ __forceinline__ FT csr_ld(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(FT) == 8) {
    return __builtin_bit_cast(FT, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(FT, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  }
}

// lp: the level pointers (global memory, or a copy in LDS); lv is
// wave-uniform, and so is the level end read from it.
// EMPTY: base may lie past the level's end (the workgroup form's waves
// beyond a short level); such a pass issues no loads (EA3D L = 16,
// R = 1024: 72.4 -> 67.2 ms; the check costs the wave-per-replica kernel,
// which has no empty passes, 7%).
// C16: the pass loads its per-position records (CsrRec), 16 bytes per load.
// BUF: word loads through the buffer resources B (the Trotter kernel's form
// without records: EA3D L = 16, M = 8, R = 1024 4.7 ms against 5.4 with
// plain loads); else plain global loads (the classical workgroup form, with
// a barrier per level, took 111 ms against 88 on EA3D L = 16 with buffer
// loads).
// csr_load_pass_at: the same with the level's end b1 (wave-uniform) given.
template <typename FT, int DMAX, bool BUF, bool V8 = false, bool EMPTY = false, bool C16 = false>
__device__ __forceinline__ void csr_load_pass_at(const CsrArgs& a, const CsrBufs& B, int b1,
                                                 int base, int lane,
                                                 CsrPass<FT, DMAX, V8, C16>& P) {
  static_assert(!(BUF && (V8 || C16)), "packed arrays are loaded by the global-load form");
  if (EMPTY && base >= b1) {  // an empty pass: no loads
    P.valid = false;
    return;
  }
  const int pos = base + lane;
  P.valid = pos < b1;
  const int pc = P.valid ? pos : b1 - 1;  // levels are never empty
  if constexpr (BUF) {
    const int n = __builtin_amdgcn_readfirstlane(a.n);
    P.i = csr_ld<int32_t>(B.order, pc * 4, 0);
    P.hv = csr_ld<FT>(B.h, pc * (int)sizeof(FT), 0);
#pragma unroll
    for (int d = 0; d < DMAX; ++d) {
      P.col[d] = csr_ld<int32_t>(B.col, pc * 4, d * n * 4);
      P.val[d] = csr_ld<FT>(B.val, pc * (int)sizeof(FT), d * n * (int)sizeof(FT));
    }
  } else {
    const int n = a.n;
    if constexpr (C16) {
      typedef CsrRec<FT, DMAX, V8> L;
      uint32_t wd[L::kWords];
#pragma unroll
      for (int k = 0; k < L::kWords / 4; ++k) {
        const uint4 v = a.rec[(int64_t)k * n + pc];
        wd[4 * k] = v.x;
        wd[4 * k + 1] = v.y;
        wd[4 * k + 2] = v.z;
        wd[4 * k + 3] = v.w;
      }
#pragma unroll
      for (int q = 0; q < DMAX / 2; ++q) P.cw[q] = wd[L::kCols + q];
      P.i = (int)wd[L::kIdx];
      if constexpr (V8) {
        P.h8 = (int32_t)wd[L::kH];
#pragma unroll
        for (int q = 0; q < DMAX / 4; ++q) P.v8[q] = wd[L::kVals + q];
      } else if constexpr (sizeof(FT) == 8) {
        P.hv = __builtin_bit_cast(FT, (uint64_t)wd[L::kH] | ((uint64_t)wd[L::kH + 1] << 32));
#pragma unroll
        for (int d = 0; d < DMAX; ++d)
          P.val[d] = __builtin_bit_cast(
              FT, (uint64_t)wd[L::kVals + 2 * d] | ((uint64_t)wd[L::kVals + 2 * d + 1] << 32));
      } else {
        P.hv = __builtin_bit_cast(FT, wd[L::kH]);
#pragma unroll
        for (int d = 0; d < DMAX; ++d) P.val[d] = __builtin_bit_cast(FT, wd[L::kVals + d]);
      }
    } else {
      static_assert(!V8 || std::is_same<FT, int32_t>::value,
                    "packed word arrays (val8) hold INT32 systems");
      P.i = a.order[pc];
      if constexpr (V8) P.h8 = reinterpret_cast<const int32_t*>(a.h)[pc];
      else P.hv = reinterpret_cast<const FT*>(a.h)[pc];
#pragma unroll
      for (int d = 0; d < DMAX; ++d) P.col[d] = a.ell_col[(int64_t)d * n + pc];
      if constexpr (V8) {
#pragma unroll
        for (int q = 0; q < DMAX / 4; ++q) P.v8[q] = a.val8[(int64_t)q * n + pc];
      } else {
        const FT* ev = reinterpret_cast<const FT*>(a.val);
#pragma unroll
        for (int d = 0; d < DMAX; ++d) P.val[d] = ev[(int64_t)d * n + pc];
      }
    }
  }
}

template <typename FT, int DMAX, bool BUF, bool V8 = false, bool EMPTY = false, bool C16 = false>
__device__ __forceinline__ void csr_load_pass(const CsrArgs& a, const CsrBufs& B,
                                              const int32_t* lp, int lv, int base, int lane,
                                              CsrPass<FT, DMAX, V8, C16>& P) {
  const int b1 = __builtin_amdgcn_readfirstlane(lp[__builtin_amdgcn_readfirstlane(lv) + 1]);
  csr_load_pass_at<FT, DMAX, BUF, V8, EMPTY, C16>(a, B, b1, base, lane, P);
}

// The field of a pass's spin from the slice's int8 spins s (stored order).
template <typename FT, int DMAX, bool V8, bool C16>
__device__ __forceinline__ FT csr_field(const CsrPass<FT, DMAX, V8, C16>& P, const int8_t* s) {
  if constexpr (V8) {
    int32_t acc = P.h8;
#pragma unroll
    for (int q = 0; q < DMAX / 4; ++q) {
      const uint32_t w = (uint32_t)(uint8_t)s[csr_col(P, 4 * q)] |
                         ((uint32_t)(uint8_t)s[csr_col(P, 4 * q + 1)] << 8) |
                         ((uint32_t)(uint8_t)s[csr_col(P, 4 * q + 2)] << 16) |
                         ((uint32_t)(uint8_t)s[csr_col(P, 4 * q + 3)] << 24);
      acc = __builtin_amdgcn_sdot4((int)w, (int)P.v8[q], acc, false);
    }
    return (FT)acc;
  } else {
    FT f = P.hv;
#pragma unroll
    for (int d = 0; d < DMAX; ++d) f = (s[csr_col(P, d)] < 0) ? f - P.val[d] : f + P.val[d];
    return f;
  }
}

}  // namespace ojmi
