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
// four per lane up front into LDS (counter (i>>2, t, r, 0), word i&3).
#include "ojmi_kernels.h"
#include "ojmi_device.h"

namespace ojmi {

template <typename FT>
struct CsrThresh {
  typedef double type;
};
template <>
struct CsrThresh<float> {
  typedef float type;
};

template <typename FT>
__global__ __launch_bounds__(256) void k_csr_level(CsrArgs a) {
  typedef typename CsrThresh<FT>::type TT;
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

  const FT* h = reinterpret_cast<const FT*>(a.h);
  const FT* ev = reinterpret_cast<const FT*>(a.val);
  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  const int D = a.D;

  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    TT nb;
    if constexpr (sizeof(TT) == 8) nb = a.p0_d[t];
    else nb = a.p0_f[t];
    for (int g = lane; g < ng; g += 64) {
      uint32_t w0, w1, w2, w3;
      philox4x32_10((uint32_t)g, tg, rep, 0u, slo, shi, w0, w1, w2, w3);
      s_w[4 * g + 0] = w0;
      s_w[4 * g + 1] = w1;
      s_w[4 * g + 2] = w2;
      s_w[4 * g + 3] = w3;
    }
    for (int lv = 0; lv < a.nlev; ++lv) {
      const int b0 = a.lvl_ptr[lv], b1 = a.lvl_ptr[lv + 1];
      for (int base = b0; base < b1; base += 64) {
        const int pos = base + lane;
        if (pos < b1) {
          const int i = a.order[pos];
          const int dg = a.deg[pos];
          FT f = h[pos];
          for (int d = 0; d < dg; ++d) {
            const int j = a.ell_col[(int64_t)d * n + pos];
            const FT v = ev[(int64_t)d * n + pos];
            f = (s_s[j] < 0) ? f - v : f + v;
          }
          const int8_t si = s_s[i];
          const FT dE = (si < 0) ? (FT)2 * f : (FT)(-2) * f;
          const TT th = log_uniform<TT>(s_w[i]) * nb;
          if ((TT)dE < th) s_s[i] = (int8_t)-si;
        }
      }
    }
  }
  for (int i = lane; i < n; i += 64) sp[i] = s_s[i];
}

hipError_t launch_csr_level(int ftype, const CsrArgs& a, int wpb, hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  const int ng = (a.n + 3) >> 2;
  const size_t lds = (size_t)wpb * ((size_t)ng * 20);
  dim3 grid((a.R + wpb - 1) / wpb), block(64 * wpb);
  if (ftype == FT_F64)
    hipLaunchKernelGGL(k_csr_level<double>, grid, block, lds, st, a);
  else if (ftype == FT_F32)
    hipLaunchKernelGGL(k_csr_level<float>, grid, block, lds, st, a);
  else if (ftype == FT_I32)
    hipLaunchKernelGGL(k_csr_level<int32_t>, grid, block, lds, st, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace ojmi
