// k_shared.hip -- batched dense sweep in which the G replicas of a workgroup
// share every coupling row they accept.
//
// Same sweep semantics as k_dense_wave (reference order, Philox threshold
// test; include/openjij/updater/single_spin_flip.hpp:67-95), reorganised per
// 64-spin block k:
//   A. decide: each wave (one replica) runs its block-k decisions on its own
//      copy of the block's field F[k], using only the diagonal tile
//      J[block k][block k] staged once in LDS (in-block flips change only
//      in-block fields through that tile);
//   B. update: the union of the G replicas' flipped rows is staged in LDS,
//      each row fetched from memory once for the workgroup, and every wave adds
//      c_u * J[u][.] to all of its fields (slot k included, in the same order
//      as in A, so F[k] ends bit-identical to A's copy).
// Each field therefore receives the same sequence of one-rounding updates as
// in the per-replica sweep, and the trajectory stays bit-identical to the
// oracle replay.  Rows of spins no replica flipped are never read.
//
// Workgroups are dispatched round-robin over the 8 XCDs; those that share an
// XCD share its L2.  With `xcd_window` > 0 a workgroup waits (bounded spin)
// before block k until its XCD peers are within `xcd_window` blocks, so rows
// fetched by one are still in L2 for the others.  The wait only shapes cache
// reuse: after the spin bound it proceeds regardless.
#include "ojmi_kernels.h"
#include "ojmi_wave.h"

namespace ojmi {

template <typename FT>
struct SharedThresh {
  typedef double type;
};
template <>
struct SharedThresh<float> {
  typedef float type;
};

template <typename JT, typename FT>
__device__ __forceinline__ void stage_vec4(FT* dst, const JT* src) {
  if constexpr (sizeof(JT) == 4 && sizeof(FT) == 4) {
    *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
  } else if constexpr (sizeof(JT) == 8) {
    // 4 doubles
    const double2 a = reinterpret_cast<const double2*>(src)[0];
    const double2 b = reinterpret_cast<const double2*>(src)[1];
    dst[0] = (FT)a.x;
    dst[1] = (FT)a.y;
    dst[2] = (FT)b.x;
    dst[3] = (FT)b.y;
  } else {
    const uint32_t w = *reinterpret_cast<const uint32_t*>(src);
    dst[0] = (FT)(int8_t)(w & 0xff);
    dst[1] = (FT)(int8_t)((w >> 8) & 0xff);
    dst[2] = (FT)(int8_t)((w >> 16) & 0xff);
    dst[3] = (FT)(int8_t)(w >> 24);
  }
}

// Stage rows rows[0..nr) of the natural-layout matrix (row length Npad) into
// LDS slots 0..nr-1, 4 elements per thread per step.
template <typename JT, typename FT>
__device__ __forceinline__ void stage_rows(FT* ring, const JT* __restrict__ J, int Npad,
                                           const int* rows, int nr, int tid, int nthreads) {
  const int per_row = Npad >> 2;
  const int total = nr * per_row;
  for (int e = tid; e < total; e += nthreads) {
    const int slot = e / per_row;
    const int c4 = (e - slot * per_row) << 2;
    stage_vec4<JT, FT>(ring + (int64_t)slot * Npad + c4, J + (int64_t)rows[slot] * Npad + c4);
  }
}

__device__ __forceinline__ uint32_t peer_min_progress(const uint32_t* prog, int b, int nblocks,
                                                      int lane) {
  const int xcd = b & 7;
  uint32_t mn = 0xffffffffu;
  for (int p = xcd + 8 * lane; p < nblocks; p += 8 * 64) {
    const uint32_t v = __hip_atomic_load(prog + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    mn = v < mn ? v : mn;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t o = __shfl_xor(mn, off);
    mn = o < mn ? o : mn;
  }
  return mn;
}

template <typename FT, typename JT, int NS, int G>
__global__ __launch_bounds__(64 * G) void k_dense_shared(SharedArgs a) {
  typedef typename SharedThresh<FT>::type TT;
  constexpr int Npad = 64 * NS;
  extern __shared__ __align__(16) char smem[];
  FT* sD = reinterpret_cast<FT*>(smem);                         // 64 x 64
  FT* ring = sD + 64 * 64;                                      // RB x Npad
  uint64_t* sFlip = reinterpret_cast<uint64_t*>(ring + (int64_t)a.ring_rows * Npad);  // [G]
  uint64_t* sSign = sFlip + G;                                  // [G]
  int* sRows = reinterpret_cast<int*>(sSign + G);               // [ring_rows]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x;
  const int r = b * G + w;
  const bool active = r < a.R;  // wave-uniform
  const int n = a.n;
  const JT* J = reinterpret_cast<const JT*>(a.J);
  const int RB = a.ring_rows;

  uint64_t down = 0;
  if (active) {
    const int8_t* sp = a.spins + (int64_t)r * n;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int j = 64 * m + lane;
      if (j < n && sp[j] < 0) down |= 1ull << m;
    }
  }
  FieldVec<FT, NS> F;
  if (a.F_in) {
    const FT* fi = reinterpret_cast<const FT*>(a.F_in) + (int64_t)(active ? r : 0) * Npad;
#pragma unroll
    for (int m = 0; m < NS; ++m) F.set(m, fi[64 * m + lane]);
  } else {
    // F = h + sum_k s_k J[k][.] in k order, rows shared through the ring
    const FT* h = reinterpret_cast<const FT*>(a.h);
#pragma unroll
    for (int m = 0; m < NS; ++m) F.set(m, h[m * 64 + lane]);
    for (int k0 = 0; k0 < n; k0 += RB) {
      const int nr = min(RB, n - k0);
      if (tid < nr) sRows[tid] = k0 + tid;
      __syncthreads();
      stage_rows<JT, FT>(ring, J, Npad, sRows, nr, tid, 64 * G);
      __syncthreads();
      for (int j = 0; j < nr; ++j) {
        const int k = k0 + j;
        const uint64_t bal = ballot64((down >> (k >> 6)) & 1ull);
        const FT c = ((bal >> (k & 63)) & 1ull) ? (FT)(-1) : (FT)1;
        const FT* row = ring + (int64_t)j * Npad;
#pragma unroll
        for (int m = 0; m < NS; ++m) F.set(m, fma_ft<FT>(c, row[m * 64 + lane], F.get(m)));
      }
      __syncthreads();
    }
  }

  const uint32_t rep = (uint32_t)(a.rep_offset + (active ? r : 0));
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  constexpr int NQ = (NS + 3) / 4;
  const int nblocks = gridDim.x;
  uint32_t done = 0;  // blocks completed by this workgroup

  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    TT nb;
    if constexpr (sizeof(TT) == 8) nb = a.p0_d[t];
    else nb = a.p0_f[t];
    for (int Q = 0; Q < NQ; ++Q) {
      uint32_t wd[4];
      {
        const uint32_t c0 = (uint32_t)(16 * (4 * Q + (lane & 3)) + (lane >> 2));
        philox4x32_10(c0, tg, rep, 0u, slo, shi, wd[0], wd[1], wd[2], wd[3]);
        quad_transpose(wd[0], wd[1], wd[2], wd[3], lane);
      }
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int k = 4 * Q + jj;
        if (k >= NS) break;
        // soft XCD pacing (performance only)
        if (a.xcd_window > 0 && w == 0 && done > (uint32_t)a.xcd_window) {
          const uint32_t need = done - (uint32_t)a.xcd_window;
          for (int spin = 0; spin < a.xcd_spin; ++spin) {
            if (peer_min_progress(a.progress, b, nblocks, lane) >= need) break;
            __builtin_amdgcn_s_sleep(8);
          }
        }
        // stage the diagonal tile D[i][l] = J[64k+i][64k+l]
        for (int e = tid; e < 64 * 16; e += 64 * G) {
          const int i = e >> 4, c4 = (e & 15) << 2;
          stage_vec4<JT, FT>(sD + i * 64 + c4, J + (int64_t)(64 * k + i) * Npad + 64 * k + c4);
        }
        __syncthreads();
        // ---- A: in-block decisions
        const int i = 64 * k + lane;
        uint64_t flips = 0, sb0 = 0;
        if (active) {
          const TT th = (i < n) ? log_uniform<TT>(wd[jj]) * nb : neg_inf<TT>();
          uint64_t live = ballot64(i < n);
          uint64_t sb = ballot64((down >> k) & 1ull);
          sb0 = sb;
          FT fk = F.dyn(k);
          while (true) {
            const bool dn = (sb >> lane) & 1ull;
            const FT dE = dn ? (FT)2 * fk : (FT)(-2) * fk;
            const uint64_t cand = ballot64((TT)dE < th) & live;
            if (cand == 0) break;
            const int i0 = __builtin_ctzll(cand);
            live &= lanes_above(i0);
            const FT c = ((sb >> i0) & 1ull) ? (FT)2 : (FT)(-2);
            sb ^= 1ull << i0;
            flips |= 1ull << i0;
            fk = fma_ft<FT>(c, sD[i0 * 64 + lane], fk);
          }
          if ((flips >> lane) & 1ull) down ^= 1ull << k;
        }
        if (lane == 0) {
          sFlip[w] = flips;
          sSign[w] = sb0;
        }
        __syncthreads();
        uint64_t U = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) U |= sFlip[g];
        // ---- B: union rows through the ring, in increasing spin order
        while (U) {
          int nr = 0;
          uint64_t Ut = U;
          while (Ut && nr < RB) {
            const int u = __builtin_ctzll(Ut);
            Ut &= Ut - 1;
            if (tid == 0) sRows[nr] = 64 * k + u;
            ++nr;
          }
          __syncthreads();
          stage_rows<JT, FT>(ring, J, Npad, sRows, nr, tid, 64 * G);
          __syncthreads();
          if (active) {
            uint64_t Uw = U;
            for (int j = 0; j < nr; ++j) {
              const int u = __builtin_ctzll(Uw);
              Uw &= Uw - 1;
              if ((flips >> u) & 1ull) {
                const FT c = ((sb0 >> u) & 1ull) ? (FT)2 : (FT)(-2);
                const FT* row = ring + (int64_t)j * Npad;
#pragma unroll
                for (int m = 0; m < NS; ++m)
                  F.set(m, fma_ft<FT>(c, row[m * 64 + lane], F.get(m)));
              }
            }
          }
          U = Ut;
          __syncthreads();
        }
        ++done;
        if (a.xcd_window > 0 && tid == 0)
          __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }

  if (active) {
    int8_t* sp = a.spins + (int64_t)r * n;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int j = 64 * m + lane;
      if (j < n) sp[j] = ((down >> m) & 1ull) ? (int8_t)-1 : (int8_t)1;
    }
    if (a.F_out) {
      FT* fo = reinterpret_cast<FT*>(a.F_out) + (int64_t)r * Npad;
#pragma unroll
      for (int m = 0; m < NS; ++m) fo[64 * m + lane] = F.get(m);
    }
  }
}

size_t shared_lds_bytes(int NS, int G, int ring_rows, int fbytes) {
  return (size_t)fbytes * (64 * 64 + (size_t)ring_rows * 64 * NS) + 16 * (size_t)G +
         4 * (size_t)ring_rows + 16;
}

template <typename FT, typename JT, int NS, int G>
static hipError_t launch_shared_t(const SharedArgs& a, hipStream_t st) {
  const size_t lds = shared_lds_bytes(NS, G, a.ring_rows, (int)sizeof(FT));
  auto kern = k_dense_shared<FT, JT, NS, G>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((a.R + G - 1) / G), block(64 * G);
  hipLaunchKernelGGL(kern, grid, block, lds, st, a);
  return hipGetLastError();
}

template <typename FT, typename JT, int G>
static hipError_t launch_shared_g(const SharedArgs& a, int NS, hipStream_t st) {
  switch (NS) {
    case 1: return launch_shared_t<FT, JT, 1, G>(a, st);
    case 2: return launch_shared_t<FT, JT, 2, G>(a, st);
    case 4: return launch_shared_t<FT, JT, 4, G>(a, st);
    case 8: return launch_shared_t<FT, JT, 8, G>(a, st);
    case 16: return launch_shared_t<FT, JT, 16, G>(a, st);
    case 32: return launch_shared_t<FT, JT, 32, G>(a, st);
    case 64: return launch_shared_t<FT, JT, 64, G>(a, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename FT, typename JT>
static hipError_t launch_shared_ft(const SharedArgs& a, int NS, int G, hipStream_t st) {
  switch (G) {
    case 4: return launch_shared_g<FT, JT, 4>(a, NS, st);
    case 8: return launch_shared_g<FT, JT, 8>(a, NS, st);
    case 16: return launch_shared_g<FT, JT, 16>(a, NS, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_dense_shared(int ftype, int jtype, const SharedArgs& a, int NS, int G,
                               hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (ftype == FT_F32 && jtype == JT_F32) return launch_shared_ft<float, float>(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_shared_ft<int32_t, int8_t>(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_shared_ft<int32_t, int32_t>(a, NS, G, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
