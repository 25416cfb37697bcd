// k_colown.hip -- batched dense sweep with column-owned fields: the G waves of
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
#include "ojmi_kernels.h"
#include "ojmi_wave.h"

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

template <int BYTES>
__device__ __forceinline__ void g2lds_c(const char* gsrc, char* ldst_wave) {
  auto* l = (__attribute__((address_space(3))) void*)ldst_wave;
  if constexpr (BYTES == 16) {
    __builtin_amdgcn_global_load_lds((const void*)gsrc, l, 16, 0, 0);
  } else {
    static_assert(BYTES == 4, "lane width must be 4 or 16 bytes");
    __builtin_amdgcn_global_load_lds((const void*)gsrc, l, 4, 0, 0);
  }
}

// Diagonal tile D[i][l] = J[64k+i][64k+l], natural-layout rows of Npad.
template <typename JT, int NS, int G>
__device__ __forceinline__ void issue_tile_c(char* dst, const char* __restrict__ J, int k, int w,
                                             int lane) {
  constexpr int EB = (int)sizeof(JT);
  constexpr int64_t RB = (int64_t)64 * NS * EB;
  for (int q = w; q < 16; q += G) {
    const int i = 4 * q + (lane >> 4);
    const char* src = J + (int64_t)(64 * k + i) * RB + (64 * k + 4 * (lane & 15)) * EB;
    g2lds_c<4 * EB>(src, dst + q * (4 * 64 * EB));
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

// x[s] = element (chunk slot s, lane) of one coupling row
template <typename JT, typename FT, int CS>
__device__ __forceinline__ void load_chunk(FT (&x)[CS], const JT* __restrict__ p, int lane) {
#pragma unroll
  for (int s = 0; s < CS; ++s) x[s] = (FT)p[64 * s + lane];
}

}  // namespace

template <typename FT, typename JT, int NS, int G>
__global__ __launch_bounds__(64 * G) void k_dense_colown(SharedArgs a) {
  typedef typename ColThresh<FT>::type TT;
  constexpr int CS = NS / G;
  static_assert(CS >= 1 && CS * G == NS, "NS must be a multiple of G");
  constexpr int Npad = 64 * NS;
  constexpr int TILE_BYTES = 64 * 64 * (int)sizeof(JT);
  constexpr int PF = (CS >= 8) ? 2 : ((CS >= 4) ? 4 : 8);  // rows per prefetch group

  extern __shared__ __align__(16) char smem[];
  char* sD = smem;                                                   // 2 tiles
  FT* sF = reinterpret_cast<FT*>(smem + 2 * TILE_BYTES);             // [G][64]
  uint64_t* sS = reinterpret_cast<uint64_t*>(sF + G * 64);           // [G]
  uint64_t* sFlip = sS + G;                                          // [G]
  uint64_t* sSign = sFlip + G;                                       // [G]
  uint32_t* sTab = reinterpret_cast<uint32_t*>(sSign + G);           // [(n+3)/4]

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
        const int col = col0 + 64 * s + lane;
        if (col < n && sp[col] < 0) down |= 1ull << (g * CS + s);
      }
    }
  }

  FieldVec<FT, NS> F;
  if (a.F_in) {
    const FT* fi = reinterpret_cast<const FT*>(a.F_in);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int rg = (rbase + g < a.R) ? rbase + g : 0;
#pragma unroll
      for (int s = 0; s < CS; ++s) F.set(g * CS + s, fi[(int64_t)rg * Npad + col0 + 64 * s + lane]);
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
        const int col = col0 + 64 * s + lane;
        if ((down >> (g * CS + s)) & 1ull)
          atomicOr(&sTab[col >> 2], 1u << ((col & 3) * 8 + g));
      }
    __syncthreads();
    const FT* h = reinterpret_cast<const FT*>(a.h);
#pragma unroll
    for (int s = 0; s < CS; ++s) {
      const FT hv = h[col0 + 64 * s + lane];
#pragma unroll
      for (int g = 0; g < G; ++g) F.set(g * CS + s, hv);
    }
    for (int k = 0; k < n; ++k) {
      FT x[CS];
      load_chunk<JT, FT, CS>(x, J + (int64_t)k * Npad + col0, lane);
      const uint32_t bits = (sTab[k >> 2] >> ((k & 3) * 8)) & 0xffu;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const FT c = ((bits >> g) & 1u) ? (FT)(-1) : (FT)1;
#pragma unroll
        for (int s = 0; s < CS; ++s) F.set(g * CS + s, fma_ft<FT>(c, x[s], F.get(g * CS + s)));
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
  int tb = 0;

  if (a.T > 0) {
    issue_tile_c<JT, NS, G>(sD, reinterpret_cast<const char*>(J), 0, w, lane);
    __builtin_amdgcn_s_waitcnt(0);
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
        // ---- 0: the owner publishes block k's fields and spin words
        if (w == ow) {
#pragma unroll
          for (int g = 0; g < G; ++g) {
            sF[g * 64 + lane] = F.dyn(g * CS + sk);
            const uint64_t bal = ballot64((down >> (g * CS + sk)) & 1ull);
            if (lane == 0) sS[g] = bal;
          }
        }
        bar();
        // ---- A: wave w decides block k for replica r
        uint64_t flips = 0, sb0 = 0;
        if (active) {
          const JT* D = reinterpret_cast<const JT*>(sD + tb * TILE_BYTES);
          FieldVec<FT, 64> Dc;  // Dc[i] = D[i][lane]
#pragma unroll
          for (int q = 0; q < 64; ++q) Dc.set(q, (FT)D[q * 64 + lane]);
          __builtin_amdgcn_sched_barrier(0);
          const int i = 64 * k + lane;
          const TT th = (i < n) ? log_uniform<TT>(wd[jj]) * nb : neg_inf<TT>();
          uint64_t live = ballot64(i < n);
          uint64_t sb = sS[w];
          sb0 = sb;
          FT fk = sF[w * 64 + lane];
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
            fk = fma_ft<FT>(c, Dc.dyn(i0), fk);
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
          fl[g] = sFlip[g];
          sg[g] = sSign[g];
          U |= fl[g];
        }
        if (w == ow) {
#pragma unroll
          for (int g = 0; g < G; ++g) down ^= ((fl[g] >> lane) & 1ull) << (g * CS + sk);
        }
        // soft XCD pacing (performance only, bounded)
        if (a.xcd_window > 0 && w == 0 && done > (uint32_t)a.xcd_window) {
          const uint32_t need = done - (uint32_t)a.xcd_window;
          for (int spin = 0; spin < a.xcd_spin; ++spin) {
            if (peer_min(a.progress, b, nblocks, lane) >= need) break;
            __builtin_amdgcn_s_sleep(4);
          }
        }
        // next block's tile into the other buffer
        const int kn = (k + 1 < NS) ? k + 1 : 0;
        issue_tile_c<JT, NS, G>(sD + (tb ^ 1) * TILE_BYTES, reinterpret_cast<const char*>(J), kn,
                                w, lane);
        // ---- B: union rows, ascending, PF rows per load group, two groups in flight
        const JT* Jk = J + (int64_t)(64 * k) * Npad + col0;
        uint64_t Ug = U;  // rows not yet loaded
        uint64_t Uc = U;  // rows not yet applied
        FT xa[PF][CS], xb[PF][CS];
        int ua[PF], ub[PF];
        auto load_group = [&](FT (&x)[PF][CS], int (&uu)[PF]) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            const int u = Ug ? __builtin_ctzll(Ug) : -1;
            if (Ug) Ug &= Ug - 1;
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
                const FT c = ((sg[g] >> u) & 1ull) ? (FT)2 : (FT)(-2);
#pragma unroll
                for (int s = 0; s < CS; ++s)
                  F.set(g * CS + s, fma_ft<FT>(c, x[p][s], F.get(g * CS + s)));
              }
            }
          }
        };
        if (Uc) load_group(xa, ua);
        while (Uc) {
          if (Ug) load_group(xb, ub);
          __builtin_amdgcn_sched_barrier(0);
          apply_group(xa, ua);
#pragma unroll
          for (int p = 0; p < PF && Uc; ++p) Uc &= Uc - 1;
          if (!Uc) break;
          if (Ug) load_group(xa, ua);
          __builtin_amdgcn_sched_barrier(0);
          apply_group(xb, ub);
#pragma unroll
          for (int p = 0; p < PF && Uc; ++p) Uc &= Uc - 1;
        }
        // the next tile must have landed before the next block's decisions
        __builtin_amdgcn_s_waitcnt(0);
        tb ^= 1;
        ++done;
        if (a.xcd_window > 0 && tid == 0)
          __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }

#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (rbase + g >= a.R) continue;
    int8_t* sp = a.spins + (int64_t)(rbase + g) * n;
#pragma unroll
    for (int s = 0; s < CS; ++s) {
      const int col = col0 + 64 * s + lane;
      if (col < n) sp[col] = ((down >> (g * CS + s)) & 1ull) ? (int8_t)-1 : (int8_t)1;
    }
    if (a.F_out) {
      FT* fo = reinterpret_cast<FT*>(a.F_out) + (int64_t)(rbase + g) * Npad;
#pragma unroll
      for (int s = 0; s < CS; ++s) fo[col0 + 64 * s + lane] = F.get(g * CS + s);
    }
  }
}

size_t colown_lds_bytes(int n, int NS, int G, int jbytes, int fbytes) {
  (void)NS;
  return (size_t)2 * 64 * 64 * jbytes + (size_t)G * 64 * fbytes + 3 * 8 * (size_t)G +
         4 * (size_t)((n + 3) / 4) + 16;
}

template <typename FT, typename JT, int NS, int G>
static hipError_t launch_colown_t(const SharedArgs& a, hipStream_t st) {
  const size_t lds = colown_lds_bytes(a.n, NS, G, (int)sizeof(JT), (int)sizeof(FT));
  auto kern = k_dense_colown<FT, JT, NS, G>;
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

template <typename FT, typename JT, int G>
static hipError_t launch_colown_g(const SharedArgs& a, int NS, hipStream_t st) {
  switch (NS) {
    case 8: return launch_colown_t<FT, JT, 8, G>(a, st);
    case 16: return launch_colown_t<FT, JT, 16, G>(a, st);
    case 32: return launch_colown_t<FT, JT, 32, G>(a, st);
    case 64: return launch_colown_t<FT, JT, 64, G>(a, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename FT, typename JT>
static hipError_t launch_colown_ft(const SharedArgs& a, int NS, int G, hipStream_t st) {
  switch (G) {
    case 4: return launch_colown_g<FT, JT, 4>(a, NS, st);
    case 8: return launch_colown_g<FT, JT, 8>(a, NS, st);
    default: return hipErrorInvalidValue;
  }
}

bool colown_supported(int NS, int G) { return (G == 4 || G == 8) && NS >= 8 && NS % G == 0; }

hipError_t launch_dense_colown(int ftype, int jtype, const SharedArgs& a, int NS, int G,
                               hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (!colown_supported(NS, G)) return hipErrorInvalidValue;
  if (ftype == FT_F32 && jtype == JT_F32) return launch_colown_ft<float, float>(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_colown_ft<int32_t, int8_t>(a, NS, G, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_colown_ft<int32_t, int32_t>(a, NS, G, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
