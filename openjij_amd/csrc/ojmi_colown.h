// ojmi_colown.h -- batched dense sweep with column-owned fields: the G waves of
// a workgroup anneal G replicas, and wave w holds the local fields of ALL G
// replicas for its own chunk of CS = NS/G slots (64*CS columns).
//
// Semantics are those of k_dense_wave / the reference sweep
// (include/openjij/updater/single_spin_flip.hpp:67-95): spins visited in index
// order, Philox threshold test, each field receiving the same one-rounding
// updates in the same order, so trajectories equal the oracle replay bit for
// bit.  Per 64-spin block k (owned by wave ow = k / CS):
//   0. the owner publishes the block's fields and spin words of every replica
//      in LDS;
//   A. wave g decides the block for replica g (rejection skipping against the
//      diagonal tile J[block k][block k], held column-per-lane in VGPRs);
//   B. every wave streams the rows of the union of flipped spins, each row
//      read once per workgroup (its own 64*CS-column chunk, straight from
//      L2 into registers), and applies it to every replica that flipped it.
// Two barriers per block; no row is staged in LDS and each coupling is read
// once per workgroup.  XCD pacing as in k_shared.hip (performance only,
// bounded).
#pragma once
#include "ojmi_kernels.h"
#include "ojmi_wave.h"

#include <type_traits>

namespace ojmi {

template <typename FT>
struct ColThresh {
  typedef double type;
};
template <>
struct ColThresh<float> {
  typedef float type;
};

namespace {

// Diagonal tile D[i][l] = J[64k+i][64k+l] (natural-layout rows of Npad),
// staged through registers: wave w loads rows 4q..4q+3 for q = w, w+G, ...,
// four consecutive columns per lane, and later stores them to LDS in the same
// order.  Plain vector loads keep the count of loads in flight exact, so row
// loads issued after them can be waited for individually.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <typename JT>
struct TileVec {
  typedef typename std::conditional<sizeof(JT) == 1, uint32_t, u32x4>::type type;
};

template <typename JT, int NS, int G>
__device__ __forceinline__ void load_tile(typename TileVec<JT>::type (&st)[16 / G],
                                          const char* __restrict__ J, int k, int w, int lane) {
  constexpr int EB = (int)sizeof(JT);
  constexpr int64_t RB = (int64_t)64 * NS * EB;
  typedef typename TileVec<JT>::type V;
#pragma unroll
  for (int j = 0; j < 16 / G; ++j) {
    const int q = w + G * j;
    const int i = 4 * q + (lane >> 4);
    st[j] = *reinterpret_cast<const V*>(J + (int64_t)(64 * k + i) * RB +
                                         (64 * k + 4 * (lane & 15)) * EB);
  }
}

template <typename JT, int G>
__device__ __forceinline__ void store_tile(const typename TileVec<JT>::type (&st)[16 / G],
                                           char* dst, int w, int lane) {
  constexpr int EB = (int)sizeof(JT);
  typedef typename TileVec<JT>::type V;
#pragma unroll
  for (int j = 0; j < 16 / G; ++j) {
    const int q = w + G * j;
    *reinterpret_cast<V*>(dst + q * (4 * 64 * EB) + lane * (4 * EB)) = st[j];
  }
}

__device__ __forceinline__ void bar() {
  __builtin_amdgcn_s_waitcnt((0 & 15) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
  __asm__ __volatile__("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ __volatile__("" ::: "memory");
}

__device__ __forceinline__ uint32_t peer_min(const uint32_t* prog, int b, int nblocks, int lane) {
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

// A wave-uniform 64-bit value read from LDS, moved to scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// w[j] for a wave-uniform j, without indexing a register array (which the
// compiler would place in scratch memory when the block loop is not unrolled).
__device__ __forceinline__ uint32_t pick4(const uint32_t (&w)[4], int j) {
  return j == 0 ? w[0] : (j == 1 ? w[1] : (j == 2 ? w[2] : w[3]));
}

// s_waitcnt: at most `vm` vector memory loads still in flight (gfx9 encoding:
// vmcnt in bits 3:0 and 15:14, expcnt and lgkmcnt left at "no wait").
template <int vm>
__device__ __forceinline__ void wait_vm() {
  static_assert(vm >= 0 && vm < 64, "vmcnt is 6 bits");
  __builtin_amdgcn_s_waitcnt((vm & 15) | (((vm >> 4) & 3) << 14) | (7 << 4) | (15 << 8));
}

// +2 or -2 as a wave-uniform (scalar register) value.
template <typename FT>
__device__ __forceinline__ FT coef2(bool plus) {
  if constexpr (std::is_floating_point<FT>::value) {
    const uint32_t b = __builtin_amdgcn_readfirstlane(plus ? 0x40000000u : 0xC0000000u);
    return __builtin_bit_cast(FT, b);
  } else {
    return (FT)(int)__builtin_amdgcn_readfirstlane(plus ? 2 : -2);
  }
}

// Column layout of a wave's chunk: each lane loads VEC consecutive columns
// (16 bytes where the chunk allows), so slot s of lane l is column
//   col0 + 64*VEC*(s / VEC) + VEC*l + s % VEC.
template <typename JT, int CS>
__host__ __device__ constexpr int vec_cols() {
  return (int)(16 / sizeof(JT)) < CS ? (int)(16 / sizeof(JT)) : CS;
}
template <typename JT, int V>
struct JVec {
  typedef JT type __attribute__((ext_vector_type(V)));
};
template <typename JT>
struct JVec<JT, 1> {
  typedef JT type;
};

template <int VEC>
__device__ __forceinline__ int slot_col(int col0, int s, int lane) {
  return col0 + 64 * VEC * (s / VEC) + VEC * lane + (s % VEC);
}

// x[s] = element (slot s, lane) of one coupling row chunk starting at p
template <typename JT, typename FT, int CS>
__device__ __forceinline__ void load_chunk(FT (&x)[CS], const JT* __restrict__ p, int lane) {
  constexpr int VEC = vec_cols<JT, CS>();
  typedef typename JVec<JT, VEC>::type V;
#pragma unroll
  for (int q = 0; q < CS / VEC; ++q) {
    const V v = *reinterpret_cast<const V*>(p + 64 * VEC * q + VEC * lane);
    if constexpr (VEC == 1) {
      x[q] = (FT)v;
    } else {
#pragma unroll
      for (int e = 0; e < VEC; ++e) x[VEC * q + e] = (FT)v[e];
    }
  }
}

}  // namespace

// RULE: acceptance rule (RULE_METROPOLIS / RULE_HEAT_BATH) fixed at compile
// time, so the Metropolis decision code has no rule branch.
template <typename FT, typename JT, int NS, int G, int RULE>
__global__ __launch_bounds__(64 * G) __attribute__((amdgpu_waves_per_eu(2)))
void k_dense_colown(SharedArgs a) {
  typedef typename ColThresh<FT>::type TT;
  constexpr int CS = NS / G;
  static_assert(CS >= 1 && CS * G == NS, "NS must be a multiple of G");
  constexpr int VEC = vec_cols<JT, CS>();  // columns per lane per load
  static_assert(CS % VEC == 0 && 64 % VEC == 0, "chunk layout");
  constexpr int Npad = 64 * NS;
  constexpr int TILE_BYTES = 64 * 64 * (int)sizeof(JT);
  constexpr int PF = (CS >= 16) ? 2 : ((CS >= 4) ? 32 / CS : 8);  // rows per prefetch group

  if (a.sel && *a.sel != a.sel_id) return;  // another kernel runs this segment
  extern __shared__ __align__(16) char smem[];
  char* sD = smem;                                                   // 1 tile
  FT* sF = reinterpret_cast<FT*>(smem + TILE_BYTES);                 // [G][64]
  uint64_t* sFlip = reinterpret_cast<uint64_t*>(sF + G * 64);        // [G]
  uint64_t* sSign = sFlip + G;                                       // [G]
  uint8_t* sDn = reinterpret_cast<uint8_t*>(sSign + G);              // [G][64]
  uint32_t* sTab = reinterpret_cast<uint32_t*>(sDn + G * 64);        // [(n+3)/4]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x;
  const int n = a.n;
  const int rbase = b * G;
  const JT* J = reinterpret_cast<const JT*>(a.J);
  const int col0 = 64 * CS * w;  // first column of this wave's chunk

  // spins of the chunk for every replica: bit g*CS+s
  uint64_t down = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (rbase + g < a.R) {
      const int8_t* sp = a.spins + (int64_t)(rbase + g) * n;
#pragma unroll
      for (int s = 0; s < CS; ++s) {
        const int col = slot_col<VEC>(col0, s, lane);
        if (col < n && sp[col] < 0) down |= 1ull << (g * CS + s);
      }
    }
  }

  // Fields as NS separate values, every index static: a replica's row update
  // under its flip test then writes its registers in place (a vector type
  // would tie the NS values to one register tuple, which the conditional
  // updates would copy around).
  FT F[NS];
  if (a.F_in) {
    const FT* fi = reinterpret_cast<const FT*>(a.F_in);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int rg = (rbase + g < a.R) ? rbase + g : 0;
#pragma unroll
      for (int s = 0; s < CS; ++s)
        F[g * CS + s] = fi[(int64_t)rg * Npad + slot_col<VEC>(col0, s, lane)];
    }
  } else {
    // F = h + sum_k s_k J[k][.], k in order, for every replica; the spins of
    // column k for all replicas come from a byte-per-column LDS table.
    const int nw = (n + 3) >> 2;
    for (int e = tid; e < nw; e += 64 * G) sTab[e] = 0;
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int s = 0; s < CS; ++s) {
        const int col = slot_col<VEC>(col0, s, lane);
        if ((down >> (g * CS + s)) & 1ull)
          atomicOr(&sTab[col >> 2], 1u << ((col & 3) * 8 + g));
      }
    __syncthreads();
    const FT* h = reinterpret_cast<const FT*>(a.h);
#pragma unroll
    for (int s = 0; s < CS; ++s) {
      const FT hv = h[slot_col<VEC>(col0, s, lane)];
#pragma unroll
      for (int g = 0; g < G; ++g) F[g * CS + s] = hv;
    }
    for (int k = 0; k < n; ++k) {
      FT x[CS];
      load_chunk<JT, FT, CS>(x, J + (int64_t)k * Npad + col0, lane);
      const uint32_t bits = (sTab[k >> 2] >> ((k & 3) * 8)) & 0xffu;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const FT c = ((bits >> g) & 1u) ? (FT)(-1) : (FT)1;
#pragma unroll
        for (int s = 0; s < CS; ++s) F[g * CS + s] = fma_ft<FT>(c, x[s], F[g * CS + s]);
      }
    }
    __syncthreads();
  }

  const int r = rbase + w;  // replica decided by this wave
  const bool active = r < a.R;
  const uint32_t rep = (uint32_t)(a.rep_offset + (active ? r : 0));
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  constexpr int NQ = (NS + 3) / 4;
  const int nblocks = gridDim.x;
  uint32_t done = 0;
  uint64_t nflip = 0, nrow = 0;  // workgroup-uniform counters
  typename TileVec<JT>::type tile[16 / G];

  if (a.T > 0) {
    load_tile<JT, NS, G>(tile, reinterpret_cast<const char*>(J), 0, w, lane);
    store_tile<JT, G>(tile, sD, w, lane);  // visible after the first barrier
  }

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
        const int ow = k / CS, sk = k - ow * CS;
        // block k = columns 64*sk + t of the owner's chunk: slots VEC*qb + e
        // of lanes lb .. lb + 64/VEC - 1, column t = VEC*(lane - lb) + e
        const int qb = sk / VEC;
        const int lb = (64 / VEC) * (sk % VEC);
        const bool holds = lane >= lb && lane < lb + 64 / VEC;
        // ---- 0: the owner publishes block k's fields and spin states
        if (w == ow && holds) {
          // qb is wave-uniform: scalar branches pick the static slots
          FT pub[G][VEC];
#pragma unroll
          for (int g = 0; g < G; ++g)
#pragma unroll
            for (int e = 0; e < VEC; ++e) pub[g][e] = F[g * CS + e];
#pragma unroll
          for (int cq = 1; cq < CS / VEC; ++cq) {
            if (cq != qb) continue;
#pragma unroll
            for (int g = 0; g < G; ++g)
#pragma unroll
              for (int e = 0; e < VEC; ++e) pub[g][e] = F[g * CS + VEC * cq + e];
          }
          FT* pf = sF + VEC * (lane - lb);
          uint8_t* pd = sDn + VEC * (lane - lb);
#pragma unroll
          for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              pf[g * 64 + e] = pub[g][e];
              pd[g * 64 + e] = (uint8_t)((down >> (g * CS + VEC * qb + e)) & 1ull);
            }
          }
        }
        bar();
        // ---- A: wave w decides block k for replica r
        uint64_t flips = 0, sb0 = 0;
        if (active) {
          const JT* D = reinterpret_cast<const JT*>(sD);
          FieldVec<FT, 64> Dc;  // Dc[i] = D[i][lane]
#pragma unroll
          for (int q = 0; q < 64; ++q) Dc.set(q, (FT)D[q * 64 + lane]);
          __builtin_amdgcn_sched_barrier(0);
          const int i = 64 * k + lane;
          const TT th = (i < n) ? accept_log<TT>(pick4(wd, jj), RULE) * nb : neg_inf<TT>();
          uint64_t live = ballot64(i < n);
          // A lane is never tested again once it has flipped, so within the
          // block every live lane keeps its sign: dE = s2 * F.
          const bool dn = sDn[w * 64 + lane] != 0;
          const uint64_t sb = ballot64(dn);
          sb0 = sb;
          const FT s2 = dn ? (FT)2 : (FT)(-2);
          FT fk = sF[w * 64 + lane];
          // the chain runs on the scalar unit: the next flip i0 and the lanes
          // <= i0 (done) from one subtraction, the coefficient from lane i0
          uint64_t cand = ballot64((TT)(s2 * fk) < th) & live;
          while (cand != 0) {
            const uint64_t below = cand ^ (cand - 1ull);
            const int i0 = __builtin_ctzll(cand);
            live &= ~below;
            flips |= cand & below;
            fk = fma_ft<FT>(lane_value(s2, i0), Dc.dyn(i0), fk);
            cand = ballot64((TT)(s2 * fk) < th) & live;
          }
        }
        if (lane == 0) {
          sFlip[w] = flips;
          sSign[w] = sb0;
        }
        bar();
        uint64_t fl[G], sg[G];
        uint64_t U = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          fl[g] = uniform64(sFlip[g]);
          sg[g] = uniform64(sSign[g]);
          U |= fl[g];
          nflip += (uint64_t)__builtin_popcountll(fl[g]);
        }
        nrow += (uint64_t)__builtin_popcountll(U);
        if (w == ow && holds) {
          constexpr uint64_t m = (VEC == 64) ? ~0ull : ((1ull << VEC) - 1);
#pragma unroll
          for (int g = 0; g < G; ++g)
            down ^= ((fl[g] >> (VEC * (lane - lb))) & m) << (g * CS + VEC * qb);
        }
        // soft XCD pacing (performance only, bounded)
        if (a.xcd_window > 0 && w == 0 && done > (uint32_t)a.xcd_window) {
          const uint32_t need = done - (uint32_t)a.xcd_window;
          for (int spin = 0; spin < a.xcd_spin; ++spin) {
            if (peer_min(a.progress, b, nblocks, lane) >= need) break;
            __builtin_amdgcn_s_sleep(4);
          }
        }
        // next block's tile: loaded now, stored to LDS after the row updates
        // (every wave has its decisions' tile column in registers by now)
        const int kn = (k + 1 < NS) ? k + 1 : 0;
        load_tile<JT, NS, G>(tile, reinterpret_cast<const char*>(J), kn, w, lane);
        // ---- B: union rows, ascending, PF rows per load group, two groups in
        // flight.  Every group issues exactly PF row loads (a drained mask loads
        // row 0 of the block as a dummy), so the number of loads in flight is
        // the same on every path and each apply waits only for its own group.
        const JT* Jk = J + (int64_t)(64 * k) * Npad + col0;
        uint64_t Ug = U;  // rows not yet loaded
        const int ng = (__builtin_popcountll(U) + PF - 1) / PF;
        FT xa[PF][CS], xb[PF][CS];
        int ua[PF], ub[PF];
        auto load_group = [&](FT (&x)[PF][CS], int (&uu)[PF]) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int u = Ug ? __builtin_ctzll(Ug) : -1;
            Ug &= Ug - 1;
            uu[p] = u;
            load_chunk<JT, FT, CS>(x[p], Jk + (int64_t)(u < 0 ? 0 : u) * Npad, lane);
          }
        };
        auto apply_group = [&](FT (&x)[PF][CS], int (&uu)[PF]) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int u = uu[p];
            if (u < 0) continue;
#pragma unroll
            for (int g = 0; g < G; ++g) {
              if ((fl[g] >> u) & 1ull) {
                const FT c = coef2<FT>((sg[g] >> u) & 1ull);
#pragma unroll
                for (int s = 0; s < CS; ++s)
                  F[g * CS + s] = fma_ft<FT>(c, x[p][s], F[g * CS + s]);
              }
            }
          }
        };
        if (ng > 0) {
          load_group(xa, ua);
          int j = 0;
          // each apply waits once for its own group (the other group's
          // PF * CS loads may stay in flight)
          for (; j + 2 < ng; j += 2) {
            load_group(xb, ub);
            __builtin_amdgcn_sched_barrier(0);
            wait_vm<PF * CS>();
            apply_group(xa, ua);
            load_group(xa, ua);
            __builtin_amdgcn_sched_barrier(0);
            wait_vm<PF * CS>();
            apply_group(xb, ub);
          }
          if (j + 1 < ng) {
            load_group(xb, ub);
            __builtin_amdgcn_sched_barrier(0);
            wait_vm<PF * CS>();
            apply_group(xa, ua);
            wait_vm<0>();
            apply_group(xb, ub);
          } else {
            wait_vm<0>();
            apply_group(xa, ua);
          }
        }
        store_tile<JT, G>(tile, sD, w, lane);
        ++done;
        if (a.xcd_window > 0 && tid == 0)
          __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }

  if (a.stats && tid == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nrow);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (rbase + g >= a.R) continue;
    int8_t* sp = a.spins + (int64_t)(rbase + g) * n;
#pragma unroll
    for (int s = 0; s < CS; ++s) {
      const int col = slot_col<VEC>(col0, s, lane);
      if (col < n) sp[col] = ((down >> (g * CS + s)) & 1ull) ? (int8_t)-1 : (int8_t)1;
    }
    if (a.F_out) {
      FT* fo = reinterpret_cast<FT*>(a.F_out) + (int64_t)(rbase + g) * Npad;
#pragma unroll
      for (int s = 0; s < CS; ++s) fo[slot_col<VEC>(col0, s, lane)] = F[g * CS + s];
    }
  }
}

template <typename FT, typename JT, int NS, int G, int RULE>
inline hipError_t launch_colown_t(const SharedArgs& a, hipStream_t st) {
  const size_t lds = colown_lds_bytes(a.n, NS, G, (int)sizeof(JT), (int)sizeof(FT));
  auto kern = k_dense_colown<FT, JT, NS, G, RULE>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((a.R + G - 1) / G), block(64 * G);
  SharedArgs b = a;
  if (b.xcd_window > 0 && !all_resident(reinterpret_cast<const void*>(kern), (int)grid.x,
                                         (int)block.x, lds))
    b.xcd_window = 0;  // pacing waits on peers that may not have started
  hipLaunchKernelGGL(kern, grid, block, lds, st, b);
  return hipGetLastError();
}

template <typename FT, typename JT, int G, int RULE>
inline hipError_t launch_colown_g(const SharedArgs& a, int NS, hipStream_t st) {
  switch (NS) {
    case 8: return launch_colown_t<FT, JT, 8, G, RULE>(a, st);
    case 16: return launch_colown_t<FT, JT, 16, G, RULE>(a, st);
    case 32: return launch_colown_t<FT, JT, 32, G, RULE>(a, st);
    case 64: return launch_colown_t<FT, JT, 64, G, RULE>(a, st);
    default: return hipErrorInvalidValue;
  }
}

// Heat-bath kernels are instantiated for G = 4 only (a replica's trajectory
// does not depend on G; callers pass G = 4 for such runs).
template <typename FT, typename JT>
inline hipError_t launch_colown_ft(const SharedArgs& a, int NS, int G, hipStream_t st) {
  if (a.rule == RULE_HEAT_BATH)
    return G == 4 ? launch_colown_g<FT, JT, 4, RULE_HEAT_BATH>(a, NS, st) : hipErrorInvalidValue;
  if (a.rule != RULE_METROPOLIS) return hipErrorInvalidValue;
  switch (G) {
    case 4: return launch_colown_g<FT, JT, 4, RULE_METROPOLIS>(a, NS, st);
    case 8: return launch_colown_g<FT, JT, 8, RULE_METROPOLIS>(a, NS, st);
    default: return hipErrorInvalidValue;
  }
}

// One translation unit per (field, coupling) type pair, so that the
// instantiations compile in parallel.
hipError_t launch_colown_f32(const SharedArgs& a, int NS, int G, hipStream_t st);
hipError_t launch_colown_i8(const SharedArgs& a, int NS, int G, hipStream_t st);
hipError_t launch_colown_i32(const SharedArgs& a, int NS, int G, hipStream_t st);

}  // namespace ojmi
