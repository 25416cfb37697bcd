// k_lds.hip -- dense single-spin-flip sweeps for systems too large for
// register-resident fields (N > 4096): one workgroup per replica, the fields
// and spins of the replica in LDS.
//
// Semantics are those of the reference sweep
// (include/openjij/updater/single_spin_flip.hpp:67-95) and of k_dense_wave:
// spins visited in index order, Philox threshold test, every field receiving
// one-rounding updates c * J[u][j] in flip order, so trajectories equal the
// oracle replay bit for bit.  Per 64-spin block:
//   A. wave 0 decides the block by rejection skipping, using the block's
//      fields from LDS and the diagonal tile column J[64kb + q][64kb + lane]
//      in registers (in-block flips change only in-block fields);
//   B. all threads apply the flipped rows, in ascending order, to their
//      columns: a chunk of fields is held in registers while the rows stream
//      from memory, then written back to LDS.
// Couplings: natural rows of Npad = 64 * ceil(n / 64) elements.
#include "ojmi_kernels.h"
#include "ojmi_wave.h"

#include <type_traits>

namespace ojmi {

namespace {

template <typename FT>
struct LdsThresh {
  typedef double type;
};
template <>
struct LdsThresh<float> {
  typedef float type;
};

constexpr int LDS_THREADS = 256;
constexpr int LDS_CH = 8;  // fields per thread per pass of the row update

template <typename TT>
__device__ __forceinline__ const TT* lds_p0(const DenseArgs& a);
template <>
__device__ __forceinline__ const double* lds_p0<double>(const DenseArgs& a) {
  return a.p0_d;
}
template <>
__device__ __forceinline__ const float* lds_p0<float>(const DenseArgs& a) {
  return a.p0_f;
}

template <typename FT>
__device__ __forceinline__ FT plus_minus2(bool plus) {
  return plus ? (FT)2 : (FT)(-2);
}

}  // namespace

template <typename FT, typename JT>
__global__ __launch_bounds__(LDS_THREADS) void k_dense_lds(DenseArgs a, int Npad) {
  typedef typename LdsThresh<FT>::type TT;
  extern __shared__ __align__(16) char smem[];
  FT* sF = reinterpret_cast<FT*>(smem);                          // [Npad]
  int8_t* sS = reinterpret_cast<int8_t*>(sF + Npad);             // [Npad] spins +-1
  uint64_t* sW = reinterpret_cast<uint64_t*>(smem + (size_t)Npad * sizeof(FT) +
                                             (((size_t)Npad + 7) & ~(size_t)7));  // [2]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int r = blockIdx.x;
  if (r >= a.R) return;  // whole workgroup
  const int n = a.n;
  const JT* J = reinterpret_cast<const JT*>(a.J);
  const FT* h = reinterpret_cast<const FT*>(a.h);
  int8_t* sp = a.spins + (int64_t)r * n;

  const FT* F_in = a.F_in ? reinterpret_cast<const FT*>(a.F_in) + (int64_t)r * Npad : nullptr;
  for (int j = tid; j < Npad; j += LDS_THREADS) {
    sS[j] = (j < n) ? sp[j] : (int8_t)1;
    sF[j] = F_in ? F_in[j] : h[j];
  }
  __syncthreads();
  // F = h + sum_k s_k J[k][.], k ascending, one rounding per term (unless the
  // caller passed the fields)
  for (int c0 = 0; c0 < (F_in ? 0 : Npad); c0 += LDS_THREADS * LDS_CH) {
    FT f[LDS_CH];
#pragma unroll
    for (int m = 0; m < LDS_CH; ++m) {
      const int col = c0 + tid + LDS_THREADS * m;
      f[m] = (col < Npad) ? sF[col] : (FT)0;
    }
    for (int k = 0; k < n; ++k) {
      const FT c = (sS[k] < 0) ? (FT)(-1) : (FT)1;
      const JT* row = J + (int64_t)k * Npad;
#pragma unroll
      for (int m = 0; m < LDS_CH; ++m) {
        const int col = c0 + tid + LDS_THREADS * m;
        if (col < Npad) f[m] = fma_ft<FT>(c, (FT)row[col], f[m]);
      }
    }
#pragma unroll
    for (int m = 0; m < LDS_CH; ++m) {
      const int col = c0 + tid + LDS_THREADS * m;
      if (col < Npad) sF[col] = f[m];
    }
  }
  __syncthreads();

  const int NB = (n + 63) >> 6;  // 64-spin blocks; NB * 64 <= Npad
  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  const TT* p0 = lds_p0<TT>(a);

  uint64_t nflip = 0;  // workgroup-uniform; one row per flip (one replica per workgroup)
  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    const TT nb = sweep_scalar(p0[t]);
    for (int kb = 0; kb < NB; ++kb) {
      // ---- A: wave 0 decides block kb
      if (w == 0) {
        const int i = 64 * kb + lane;
        uint32_t wd[4];
        {
          const int Q = kb >> 2;
          const uint32_t c0 = (uint32_t)(16 * (4 * Q + (lane & 3)) + (lane >> 2));
          philox4x32_10(c0, tg, rep, 0u, slo, shi, wd[0], wd[1], wd[2], wd[3]);
          quad_transpose(wd[0], wd[1], wd[2], wd[3], lane);
        }
        const int jj = kb & 3;
        const uint32_t wj = jj == 0 ? wd[0] : (jj == 1 ? wd[1] : (jj == 2 ? wd[2] : wd[3]));
        const TT th = (i < n) ? accept_log<TT>(wj, a.rule) * nb : neg_inf<TT>();
        FieldVec<FT, 64> Dc;  // Dc[q] = J[64kb + q][64kb + lane]
        const JT* tile = J + (int64_t)(64 * kb) * Npad + 64 * kb + lane;
#pragma unroll
        for (int q = 0; q < 64; ++q) Dc.set(q, (FT)tile[(int64_t)q * Npad]);
        const bool dn = sS[i] < 0;
        const uint64_t sb = ballot64(dn);
        const FT s2 = dn ? (FT)2 : (FT)(-2);
        FT fk = sF[i];
        uint64_t live = ballot64(i < n);
        uint64_t flips = 0;
        while (true) {
          const FT dE = s2 * fk;
          const uint64_t cand = ballot64((TT)dE < th) & live;
          if (cand == 0) break;
          const uint64_t low = cand & (~cand + 1ull);
          live &= ~(cand ^ (cand - 1ull));
          flips |= low;
          const int i0 = __builtin_ctzll(cand);
          fk = fma_ft<FT>(plus_minus2<FT>((sb & low) != 0ull), Dc.dyn(i0), fk);
        }
        if ((flips >> lane) & 1ull) sS[i] = (int8_t)(-sS[i]);
        if (lane == 0) {
          sW[0] = flips;
          sW[1] = sb;
        }
      }
      __syncthreads();
      const uint64_t flips = sW[0], sb = sW[1];
      nflip += (uint64_t)__builtin_popcountll(flips);
      // ---- B: rows of the flipped spins, ascending, onto every column
      if (flips) {
        const JT* Jb = J + (int64_t)(64 * kb) * Npad;
        for (int c0 = 0; c0 < Npad; c0 += LDS_THREADS * LDS_CH) {
          FT f[LDS_CH];
#pragma unroll
          for (int m = 0; m < LDS_CH; ++m) {
            const int col = c0 + tid + LDS_THREADS * m;
            f[m] = (col < Npad) ? sF[col] : (FT)0;
          }
          uint64_t U = flips;
          while (U) {
            const int u = __builtin_ctzll(U);
            U &= U - 1;
            const FT c = plus_minus2<FT>((sb >> u) & 1ull);
            const JT* row = Jb + (int64_t)u * Npad;
#pragma unroll
            for (int m = 0; m < LDS_CH; ++m) {
              const int col = c0 + tid + LDS_THREADS * m;
              if (col < Npad) f[m] = fma_ft<FT>(c, (FT)row[col], f[m]);
            }
          }
#pragma unroll
          for (int m = 0; m < LDS_CH; ++m) {
            const int col = c0 + tid + LDS_THREADS * m;
            if (col < Npad) sF[col] = f[m];
          }
        }
      }
      __syncthreads();
    }
  }

  if (a.stats && tid == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nflip);
  }
  for (int j = tid; j < n; j += LDS_THREADS) sp[j] = sS[j];
  if (a.F_out) {
    FT* fo = reinterpret_cast<FT*>(a.F_out) + (int64_t)r * Npad;
    for (int j = tid; j < Npad; j += LDS_THREADS) fo[j] = sF[j];
  }
}

size_t lds_kernel_bytes(int Npad, int fbytes) {
  return (size_t)Npad * fbytes + (((size_t)Npad + 7) & ~(size_t)7) + 16;
}

template <typename FT, typename JT>
static hipError_t launch_lds_t(const DenseArgs& a, int Npad, hipStream_t st) {
  const size_t lds = lds_kernel_bytes(Npad, (int)sizeof(FT));
  auto kern = k_dense_lds<FT, JT>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(a.R), dim3(LDS_THREADS), lds, st, a, Npad);
  return hipGetLastError();
}

hipError_t launch_dense_lds(int ftype, int jtype, const DenseArgs& a, int Npad, hipStream_t st) {
  if (a.R <= 0 || a.T < 0 || Npad % 64 != 0 || a.n > Npad) return hipErrorInvalidValue;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_lds_t<double, double>(a, Npad, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_lds_t<float, float>(a, Npad, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_lds_t<int32_t, int8_t>(a, Npad, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_lds_t<int32_t, int32_t>(a, Npad, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
