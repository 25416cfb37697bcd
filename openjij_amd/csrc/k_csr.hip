// k_csr.hip -- CSR single-spin-flip sweeps, one wave64 per replica, spins in
// level order.
//
// The reference visits spins 0..n-1 in order
// (include/openjij/updater/single_spin_flip.hpp:78 with a CSRSparse row at
// :84-86).  Spin i only reads the current state of its neighbours; those with
// a smaller index were already visited this sweep, those with a larger index
// still hold last sweep's state.  With level(i) = 1 + max level of its
// smaller-index neighbours (0 if none), two spins of one level are never
// neighbours, and every smaller-index neighbour of a spin sits in an earlier
// level while every larger-index neighbour sits in a later one.  Visiting the
// levels in order, all spins of a level at once, therefore reads exactly the
// states the sequential sweep reads: the trajectory is the sequential one.
// (EA3D, L = 16: 46 levels of ~89 spins.)
//
// Fields are recomputed per attempt in stored CSR order,
//   F_i = h_i + sum_e val_e s_col(e),
// from spins kept as int8 in LDS; the Philox words of one sweep are generated
// four per lane up front into LDS (counter (i>>2, t, r, 0), word i&3), or, in
// the workgroup form, the thresholds computed from them.
#include "ojmi_kernels.h"
#include "ojmi_tuned.h"
#include "ojmi_device.h"
#include "ojmi_csr.h"

#include <type_traits>

namespace ojmi {

template <typename FT>
struct CsrThresh {
  typedef double type;
};
template <>
struct CsrThresh<float> {
  typedef float type;
};

// Thresholds of the workgroup form: INT32 fields hold ceil(theta) as int32
// (int_threshold), half the LDS of a double and an integer compare.
// FP64 fields over int8 integer couplings (V8) have integer-valued dE too,
// so they take the INT32 thresholds as well (the same double theta).
template <typename FT, bool V8>
struct WgThresh {
  typedef typename CsrThresh<FT>::type type;
};
template <bool V8>
struct WgThresh<int32_t, V8> {
  typedef int32_t type;
};
template <>
struct WgThresh<double, true> {
  typedef int32_t type;
};

// REC: the pass loads the per-position records (rec; V8: int8 quadruples);
// else word arrays (V8: INT32 systems' val8).
template <typename FT, int DMAX, bool V8, bool REC>
__global__ __launch_bounds__(256) void k_csr_level(CsrArgs a) {
  typedef typename CsrThresh<FT>::type TT;
  constexpr bool PREFETCH = DMAX <= 16;
  extern __shared__ char smem[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + wv);
  if (r >= a.R) return;
  const int n = a.n;
  const int ng = (n + 3) >> 2;
  // per-wave LDS: words [4*ng] uint32, spins [4*ng] int8
  const size_t per_wave = (size_t)ng * 16 + (size_t)ng * 4;
  uint32_t* s_w = reinterpret_cast<uint32_t*>(smem + per_wave * wv);
  int8_t* s_s = reinterpret_cast<int8_t*>(smem + per_wave * wv + (size_t)ng * 16);
  int8_t* sp = a.spins + (int64_t)r * n;
  for (int i = lane; i < n; i += 64) s_s[i] = sp[i] < 0 ? (int8_t)-1 : (int8_t)1;

  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);

  const CsrBufs CB{};  // unused: plain global loads
  CsrPass<FT, DMAX, V8, REC> cur;
  if constexpr (PREFETCH)
    csr_load_pass<FT, DMAX, false, V8, false, REC>(a, CB, a.lvl_ptr, 0, 0, lane, cur);
  uint64_t nflip = 0;  // wave-uniform
  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    TT nb;
    if constexpr (sizeof(TT) == 8) nb = sweep_scalar(a.p0_d[t]);
    else nb = sweep_scalar(a.p0_f[t]);
    for (int g = lane; g < ng; g += 64) {
      uint32_t w0, w1, w2, w3;
      philox4x32_10((uint32_t)g, tg, rep, 0u, slo, shi, w0, w1, w2, w3);
      s_w[4 * g + 0] = w0;
      s_w[4 * g + 1] = w1;
      s_w[4 * g + 2] = w2;
      s_w[4 * g + 3] = w3;
    }
    int lv = 0, base = 0;
    while (true) {
      // the pass after this one (the first pass of the next sweep after the last)
      int nlv = lv, nbase = base + 64;
      if (nbase >= a.lvl_ptr[lv + 1]) {
        nlv = lv + 1;
        nbase = a.lvl_ptr[lv + 1];
      }
      const bool last = nlv >= a.nlev;
      if (last) {
        nlv = 0;
        nbase = 0;
      }
      CsrPass<FT, DMAX, V8, REC> nxt;
      if constexpr (PREFETCH) {
        csr_load_pass<FT, DMAX, false, V8, false, REC>(a, CB, a.lvl_ptr, nlv, nbase, lane, nxt);
      } else {
        csr_load_pass<FT, DMAX, false, V8, false, REC>(a, CB, a.lvl_ptr, lv, base, lane, cur);
      }
      bool flipped = false;
      if (cur.valid) {
        const FT f = csr_field(cur, s_s);
        const int i = cur.i;
        const int8_t si = s_s[i];
        const FT dE = (si < 0) ? (FT)2 * f : (FT)(-2) * f;
        const TT th = accept_log<TT>(s_w[i], a.rule) * nb;
        flipped = (TT)dE < th;
        if (flipped) s_s[i] = (int8_t)-si;
      }
      nflip += (uint64_t)__builtin_popcountll(ballot64(flipped));
      if constexpr (PREFETCH) cur = nxt;
      if (last) break;
      lv = nlv;
      base = nbase;
    }
  }
  if (a.stats && lane == 0) atomicAdd(a.stats, (unsigned long long)nflip);
  for (int i = lane; i < n; i += 64) sp[i] = s_s[i];
}

// Workgroup form: W waves anneal one replica, sharing its spins and the
// sweep's acceptance thresholds in LDS (computed from the Philox words before
// the first level, so the per-level chain holds no logarithm).  Within a level wave w takes the passes at
// level start + 64*(w + W*j); a barrier ends every level (only LDS writes are
// waited for, so the prefetched neighbour lists of the next pass stay in
// flight).  Spins of one level are never neighbours, so the order of the
// passes inside a level does not matter.
template <typename FT, int DMAX, int W, bool V8>
__global__ __launch_bounds__(64 * W) void k_csr_wg(CsrArgs a) {
  typedef typename WgThresh<FT, V8>::type TT;
  typedef typename CsrThresh<FT>::type NB;  // the sweep's scale and exact thresholds
  static_assert(DMAX <= 16 || (DMAX == 32 && sizeof(FT) == 4),
                "the workgroup form holds two passes' records in registers");
  extern __shared__ char smem[];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = blockIdx.x;
  if (r >= a.R) return;  // whole workgroup
  const int n = a.n;
  const int ng = (n + 3) >> 2;
  // thresholds of the sweep's attempts, theta_i = log-factor(u_i) * p0[t],
  // computed for all spins before the levels so that no level waits on a
  // logarithm; then the spins
  TT* s_t = reinterpret_cast<TT*>(smem);                                          // [4*ng]
  int8_t* s_s = reinterpret_cast<int8_t*>(smem + (size_t)ng * 4 * sizeof(TT));    // [4*ng]
  // level pointers [nlev + 1]: read every pass, and global reads of them would
  // be vector loads with a full wait (the barriers' memory clobbers keep the
  // compiler from scalar loads)
  int32_t* s_lp = reinterpret_cast<int32_t*>(smem + (size_t)ng * 4 * (sizeof(TT) + 1));
  int8_t* sp = a.spins + (int64_t)r * n;
  for (int i = tid; i < n; i += 64 * W) s_s[i] = sp[i] < 0 ? (int8_t)-1 : (int8_t)1;
  for (int q = tid; q <= a.nlev; q += 64 * W) s_lp[q] = a.lvl_ptr[q];
  __syncthreads();

  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  auto bar = []() {
    __builtin_amdgcn_s_waitcnt((0 & 15) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
    __asm__ __volatile__("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ __volatile__("" ::: "memory");
  };

  const CsrBufs CB{};  // unused: plain global loads
  const int nlev = a.nlev;
  const int end0 = __builtin_amdgcn_readfirstlane(s_lp[1]);  // end of level 0
  CsrPass<FT, DMAX, V8, true> cur;
  csr_load_pass_at<FT, DMAX, false, V8, true, true>(a, CB, end0, 64 * w, lane, cur);
  uint64_t nflip = 0;  // wave-uniform
  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    NB nb;
    if constexpr (sizeof(NB) == 8) nb = sweep_scalar(a.p0_d[t]);
    else nb = sweep_scalar(a.p0_f[t]);
    auto thresh = [&](uint32_t u) -> TT {
      if constexpr (std::is_same<TT, int32_t>::value) return int_threshold(u, a.rule, nb);
      else return accept_log<TT>(u, a.rule) * nb;
    };
    for (int g = tid; g < ng; g += 64 * W) {
      uint32_t w0, w1, w2, w3;
      philox4x32_10((uint32_t)g, tg, rep, 0u, slo, shi, w0, w1, w2, w3);
      s_t[4 * g + 0] = thresh(w0);
      s_t[4 * g + 1] = thresh(w1);
      s_t[4 * g + 2] = thresh(w2);
      s_t[4 * g + 3] = thresh(w3);
    }
    bar();  // thresholds of this sweep and spins of the last are visible
    int lv = 0, base = 64 * w;
    while (true) {
      // the ends of this level and the next in one LDS trip
      const int lq = __builtin_amdgcn_readfirstlane(lv);
      const int e1v = s_lp[lq + 1], e2v = s_lp[min(lq + 2, nlev)];
      const int b1 = __builtin_amdgcn_readfirstlane(e1v);
      const int b2 = __builtin_amdgcn_readfirstlane(e2v);
      int nlv = lv, nbase = base + 64 * W, nend = b1;
      const bool end_of_level = nbase >= b1;
      bool last = false;
      if (end_of_level) {
        nlv = lv + 1;
        if (nlv >= nlev) {
          last = true;
          nlv = 0;
          nbase = 64 * w;
          nend = end0;
        } else {
          nbase = b1 + 64 * w;
          nend = b2;
        }
      }
      CsrPass<FT, DMAX, V8, true> nxt;
      csr_load_pass_at<FT, DMAX, false, V8, true, true>(a, CB, nend, nbase, lane, nxt);
      bool flipped = false;
      if (cur.valid) {
        // own spin and threshold read with the neighbour gathers: one LDS trip
        const int i = cur.i;
        const int8_t si = s_s[i];
        const TT th = s_t[i];
        const FT f = csr_field(cur, s_s);
        const FT dE = (si < 0) ? (FT)2 * f : (FT)(-2) * f;
        flipped = (TT)dE < th;
        if (flipped) s_s[i] = (int8_t)-si;
      }
      nflip += (uint64_t)__builtin_popcountll(ballot64(flipped));
      cur = nxt;
      if (end_of_level) bar();  // this level's flips are visible to every wave
      if (last) break;
      lv = nlv;
      base = nbase;
    }
  }
  if (a.stats && lane == 0) atomicAdd(a.stats, (unsigned long long)nflip);
  __syncthreads();
  for (int i = tid; i < n; i += 64 * W) sp[i] = s_s[i];
}

// ELL width the kernels are instantiated for: the smallest of 8, 16, 32, 64
// that holds the maximum degree D.
int csr_padded_width(int D) { return D <= 8 ? 8 : (D <= 16 ? 16 : (D <= 32 ? 32 : 64)); }

template <typename FT, bool V8, bool REC>
static hipError_t launch_csr_ft(const CsrArgs& a, dim3 grid, dim3 block, size_t lds,
                                hipStream_t st) {
  switch (a.D) {
    case 8: hipLaunchKernelGGL((k_csr_level<FT, 8, V8, REC>), grid, block, lds, st, a); break;
    case 16: hipLaunchKernelGGL((k_csr_level<FT, 16, V8, REC>), grid, block, lds, st, a); break;
    case 32: hipLaunchKernelGGL((k_csr_level<FT, 32, V8, REC>), grid, block, lds, st, a); break;
    case 64: hipLaunchKernelGGL((k_csr_level<FT, 64, V8, REC>), grid, block, lds, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// LDS of the workgroup form: thresholds [4*ng] of `tbytes` each, the int8
// spins [4*ng] and the level pointers [nlev + 1] (tbytes =
// sizeof(WgThresh<FT>::type): 4 for FP32 and INT32 fields, 8 for FP64 fields).
static size_t csr_wg_lds_bytes(int n, int nlev, size_t tbytes) {
  return (size_t)((n + 3) >> 2) * (4 * tbytes + 4) + 4 * ((size_t)nlev + 1);
}
constexpr size_t CSR_WG_LDS_MAX = 160 * 1024;  // LDS of one gfx950 workgroup

template <typename FT, int DMAX, int W, bool V8>
static hipError_t launch_csr_wg_d(const CsrArgs& a, hipStream_t st) {
  const size_t lds = csr_wg_lds_bytes(a.n, a.nlev, sizeof(typename WgThresh<FT, V8>::type));
  if (lds > CSR_WG_LDS_MAX) return hipErrorInvalidValue;
  auto kern = k_csr_wg<FT, DMAX, W, V8>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(a.R), dim3(64 * W), lds, st, a);
  return hipGetLastError();
}

template <typename FT, int W, bool V8>
static hipError_t launch_csr_wg_t(const CsrArgs& a, hipStream_t st) {
  if (a.D == 8) return launch_csr_wg_d<FT, 8, W, V8>(a, st);
  if (a.D == 16) return launch_csr_wg_d<FT, 16, W, V8>(a, st);
  if constexpr (sizeof(FT) == 4) {
    if (a.D == 32) return launch_csr_wg_d<FT, 32, W, V8>(a, st);
  }
  return hipErrorInvalidValue;
}

template <typename FT, bool V8 = false>
static hipError_t launch_csr_wg_ft(const CsrArgs& a, int W, hipStream_t st) {
  if (W == 2) return launch_csr_wg_t<FT, 2, V8>(a, st);
  if (W == 3) return launch_csr_wg_t<FT, 3, V8>(a, st);
  if (W == 4) return launch_csr_wg_t<FT, 4, V8>(a, st);
  if (W == 8) return launch_csr_wg_t<FT, 8, V8>(a, st);
  return hipErrorInvalidValue;
}

// Waves per replica for the workgroup form (1 = the wave-per-replica kernel):
// several waves split a level's passes, so each SIMD holds more waves and each
// wave fewer serial passes.  `requested` (1, 2, 3, 4 or 8) overrides the choice;
// 0 chooses by R.  Graphs of degree > 32 (> 16 with FP64 fields: two passes'
// records must fit the registers), or whose thresholds, spins and level
// pointers exceed one workgroup's LDS, take the wave-per-replica kernel.
int csr_waves_per_replica(int64_t R, int D, int requested, int n, int nlev, int ftype) {
  if (D > 32 || (D > 16 && ftype == FT_F64)) return 1;
  const size_t tb = ftype == FT_F64 ? 8 : 4;
  if (csr_wg_lds_bytes(n, nlev, tb) > CSR_WG_LDS_MAX) return 1;
  if (requested > 0) return requested;
  // the autotuner's table (openjij_amd/tune.py, ojmi_tuned.h)
  for (const tuned::CsrShape& e : tuned::kCsr)
    if (e.ftype == ftype && R <= e.r_max) return e.waves;
  return ftype == FT_F64 ? 1 : 4;
}

hipError_t launch_csr_level(int ftype, const CsrArgs& a, int wpb, int waves, hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  const int W = csr_waves_per_replica(a.R, a.D, waves, a.n, a.nlev, ftype);
  if (W > 1) {
    if (!a.rec) return hipErrorInvalidValue;  // position records: built for every n it takes
    if (a.pack8) {
      if (ftype == FT_F64) return launch_csr_wg_ft<double, true>(a, W, st);
      if (ftype == FT_F32) return launch_csr_wg_ft<float, true>(a, W, st);
      if (ftype == FT_I32) return launch_csr_wg_ft<int32_t, true>(a, W, st);
    } else {
      if (ftype == FT_F64) return launch_csr_wg_ft<double>(a, W, st);
      if (ftype == FT_F32) return launch_csr_wg_ft<float>(a, W, st);
      if (ftype == FT_I32) return launch_csr_wg_ft<int32_t>(a, W, st);
    }
    return hipErrorInvalidValue;
  }
  const int ng = (a.n + 3) >> 2;
  const size_t lds = (size_t)wpb * ((size_t)ng * 20);
  dim3 grid((a.R + wpb - 1) / wpb), block(64 * wpb);
  if (a.rec && a.pack8) {
    if (ftype == FT_F64) return launch_csr_ft<double, true, true>(a, grid, block, lds, st);
    if (ftype == FT_F32) return launch_csr_ft<float, true, true>(a, grid, block, lds, st);
    if (ftype == FT_I32) return launch_csr_ft<int32_t, true, true>(a, grid, block, lds, st);
  } else if (a.rec) {
    if (ftype == FT_F64) return launch_csr_ft<double, false, true>(a, grid, block, lds, st);
    if (ftype == FT_F32) return launch_csr_ft<float, false, true>(a, grid, block, lds, st);
    if (ftype == FT_I32) return launch_csr_ft<int32_t, false, true>(a, grid, block, lds, st);
  } else {
    if (ftype == FT_F64) return launch_csr_ft<double, false, false>(a, grid, block, lds, st);
    if (ftype == FT_F32) return launch_csr_ft<float, false, false>(a, grid, block, lds, st);
    if (ftype == FT_I32 && a.val8)
      return launch_csr_ft<int32_t, true, false>(a, grid, block, lds, st);
    if (ftype == FT_I32) return launch_csr_ft<int32_t, false, false>(a, grid, block, lds, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace ojmi
