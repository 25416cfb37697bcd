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

// ---------------------------------------------------------------------------
// Asynchronous global -> LDS staging (gfx950 global_load_lds): each lane moves
// BYTES bytes; one wave instruction fills 64*BYTES contiguous LDS bytes.
template <int BYTES>
__device__ __forceinline__ void g2lds(const char* gsrc, char* ldst_wave) {
  auto* l = (__attribute__((address_space(3))) void*)ldst_wave;
  if constexpr (BYTES == 16) {
    __builtin_amdgcn_global_load_lds((const void*)gsrc, l, 16, 0, 0);
  } else if constexpr (BYTES == 4) {
    __builtin_amdgcn_global_load_lds((const void*)gsrc, l, 4, 0, 0);
  } else {
    static_assert(BYTES == 1, "lane width must be 1, 4 or 16 bytes");
    __builtin_amdgcn_global_load_lds((const void*)gsrc, l, 1, 0, 0);
  }
}

__device__ __forceinline__ void wait_all_loads() { __builtin_amdgcn_s_waitcnt(0); }

// Workgroup barrier without the workgroup-scope memory fence of __syncthreads:
// that fence would also drain the in-flight global->LDS prefetches of later
// rounds.  Ordering is explicit instead: this wave's LDS accesses are complete
// (lgkmcnt 0) and the awaited prefetch rounds were waited for by wait_vmcnt.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt((0 & 15) | (3 << 14) | (7 << 4) | (0 << 8));  // lgkmcnt(0)
  __asm__ __volatile__("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ __volatile__("" ::: "memory");
}

// Wait until at most n (wave-uniform, 0..63) vector memory operations of this
// wave are outstanding; loads complete in issue order, so the oldest ones are
// then done.  s_waitcnt takes an immediate, hence the table.
#define OJMI_VMCNT_IMM(n) (((n)&15) | ((((n) >> 4) & 3) << 14) | (7 << 4) | (15 << 8))
__device__ __forceinline__ void wait_vmcnt(int n) {
  switch (n) {
#define OJMI_W(n) \
  case n:         \
    __builtin_amdgcn_s_waitcnt(OJMI_VMCNT_IMM(n)); \
    break;
#define OJMI_W8(b) OJMI_W(b) OJMI_W(b + 1) OJMI_W(b + 2) OJMI_W(b + 3) OJMI_W(b + 4) \
  OJMI_W(b + 5) OJMI_W(b + 6) OJMI_W(b + 7)
    OJMI_W8(0) OJMI_W8(8) OJMI_W8(16) OJMI_W8(24) OJMI_W8(32) OJMI_W8(40) OJMI_W8(48)
    OJMI_W8(56)
#undef OJMI_W8
#undef OJMI_W
    default:
      __builtin_amdgcn_s_waitcnt(0);
  }
}

// Row copies: per-lane width so that 64 lanes tile a row exactly.
template <typename JT, int NS>
struct RowCopy {
  static constexpr int ROW_BYTES = 64 * NS * (int)sizeof(JT);
  static constexpr int LANE_BYTES = (ROW_BYTES % 1024 == 0) ? 16 : ((ROW_BYTES % 256 == 0) ? 4 : 1);
  static constexpr int CHUNK = 64 * LANE_BYTES;
  static constexpr int CPR = ROW_BYTES / CHUNK;  // wave instructions per row
};

// Issue the copies of one round's rows into LDS slots 0.. of `dst`.  The rows
// are base + (set bits of `mask`, ascending) when `mask` is given, else
// base, base+1, ..., base+nr-1.  Chunk c = slot*CPR + piece goes to wave c % G.
template <typename JT, int NS, int G>
__device__ __forceinline__ void issue_rows(char* dst, const char* __restrict__ J, uint64_t mask,
                                           int base, int nr, int w, int lane) {
  typedef RowCopy<JT, NS> RC;
  for (int slot = 0; slot < nr; ++slot) {
    int row;
    if (mask) {
      row = base + __builtin_ctzll(mask);
      mask &= mask - 1;
    } else {
      row = base + slot;
    }
    int piece = w - ((slot * RC::CPR) % G);
    if (piece < 0) piece += G;
    for (; piece < RC::CPR; piece += G) {
      const int off = piece * RC::CHUNK;
      g2lds<RC::LANE_BYTES>(J + (int64_t)row * RC::ROW_BYTES + off + lane * RC::LANE_BYTES,
                            dst + slot * RC::ROW_BYTES + off);
    }
  }
}

// Load instructions wave w issues for a round of nr rows (chunks c = slot*CPR
// + piece with c % G == w, as issue_rows assigns them).
template <typename JT, int NS, int G>
__device__ __forceinline__ int round_loads(int nr, int w) {
  typedef RowCopy<JT, NS> RC;
  const int total = nr * RC::CPR;
  return total > w ? (total - w + G - 1) / G : 0;
}

template <int G>
__device__ __forceinline__ int tile_loads(int w) {
  return 16 > w ? (16 - w + G - 1) / G : 0;
}

// Diagonal tile D[i][l] = J[64k+i][64k+l] (64 x 64 elements): 16 wave
// instructions of 4 tile rows each, 4 elements per lane.
template <typename JT, int NS, int G>
__device__ __forceinline__ void issue_tile(char* dst, const char* __restrict__ J, int k, int w,
                                           int lane) {
  typedef RowCopy<JT, NS> RC;
  constexpr int EB = (int)sizeof(JT);
  for (int q = w; q < 16; q += G) {
    const int i = 4 * q + (lane >> 4);
    const char* src = J + (int64_t)(64 * k + i) * RC::ROW_BYTES + (64 * k + 4 * (lane & 15)) * EB;
    g2lds<4 * EB>(src, dst + q * (4 * 64 * EB));
  }
}

// Set bits [j*S, j*S + S) (in ascending order) of U: the rows of round j.
__device__ __forceinline__ uint64_t round_mask(uint64_t U, int j, int S) {
  for (int skip = j * S; skip > 0 && U; --skip) U &= U - 1;
  uint64_t out = 0;
  for (int k = 0; k < S && U; ++k) {
    const uint64_t low = U & (~U + 1);
    out |= low;
    U ^= low;
  }
  return out;
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

// F += c * row for a row staged in LDS (natural layout).  All of the lane's
// elements are read first and a scheduling barrier keeps the reads batched,
// so their LDS latencies overlap instead of being paid one by one.
template <typename FT, typename JT, int NS>
__device__ __forceinline__ void lds_row_axpy(FieldVec<FT, NS>& F, const JT* row, int lane, FT c) {
  FT x[NS];
#pragma unroll
  for (int m = 0; m < NS; ++m) x[m] = (FT)row[m * 64 + lane];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int m = 0; m < NS; ++m) F.set(m, fma_ft<FT>(c, x[m], F.get(m)));
}

// Row pipeline: the rows of a stream (a block's union, or all rows for the
// field initialisation) go through `nbuf` LDS buffers of S rows; up to nbuf-1
// rounds are in flight while one is consumed.  Round x's rows are
// base + (set bits [x*S, x*S+S) of mask), or base + x*S + i when mask == 0.
template <typename JT, int NS, int G>
struct RowPipe {
  typedef RowCopy<JT, NS> RC;
  char* ring;
  int S, nbuf, w, lane;
  const char* J;
  __device__ __forceinline__ int rows_in(int x, int total) const {
    const int r = total - x * S;
    return r < S ? r : S;
  }
  __device__ __forceinline__ void issue(int x, uint64_t mask, int base, int total) const {
    const int nr = rows_in(x, total);
    char* dst = ring + (x % nbuf) * S * RC::ROW_BYTES;
    if (mask) issue_rows<JT, NS, G>(dst, J, round_mask(mask, x, S), base, nr, w, lane);
    else issue_rows<JT, NS, G>(dst, J, 0, base + x * S, nr, w, lane);
  }
  __device__ __forceinline__ const char* buf(int x) const {
    return ring + (x % nbuf) * S * RC::ROW_BYTES;
  }
  // loads still allowed in flight when round j must be complete
  __device__ __forceinline__ int allowed(int j, int issued, int total) const {
    int n = 0;
    for (int x = j + 1; x < issued; ++x) n += round_loads<JT, NS, G>(rows_in(x, total), w);
    return n;
  }
};

template <typename FT, typename JT, int NS, int G>
__global__ __launch_bounds__(64 * G) void k_dense_shared(SharedArgs a) {
  typedef typename SharedThresh<FT>::type TT;
  typedef RowCopy<JT, NS> RC;
  constexpr int Npad = 64 * NS;
  constexpr int TILE_BYTES = 64 * 64 * (int)sizeof(JT);
  // the diagonal tile column of each lane lives in VGPRs when they allow it
  constexpr bool DREG = (G <= 8);
  extern __shared__ __align__(16) char smem[];
  char* sD = smem;                                   // 2 tiles
  char* ring = smem + 2 * TILE_BYTES;
  const int S = a.round_rows;
  const int nbuf = a.ring_rows / a.round_rows;
  uint64_t* sFlip = reinterpret_cast<uint64_t*>(ring + (int64_t)a.ring_rows * RC::ROW_BYTES);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x;
  const int r = b * G + w;
  const bool active = r < a.R;  // wave-uniform
  const int n = a.n;
  const char* J = reinterpret_cast<const char*>(a.J);
  RowPipe<JT, NS, G> pipe{ring, S, nbuf, w, lane, J};

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
    // F = h + sum_k s_k J[k][.] in k order; rows shared through the pipeline
    const FT* h = reinterpret_cast<const FT*>(a.h);
#pragma unroll
    for (int m = 0; m < NS; ++m) F.set(m, h[m * 64 + lane]);
    const int nrounds = (n + S - 1) / S;
    int issued = 0;
    for (; issued < nbuf - 1 && issued < nrounds; ++issued) pipe.issue(issued, 0, 0, n);
    for (int j = 0; j < nrounds; ++j) {
      wait_vmcnt(pipe.allowed(j, issued, n));
      lds_barrier();
      if (issued < nrounds) {
        pipe.issue(issued, 0, 0, n);
        ++issued;
      }
      const JT* buf = reinterpret_cast<const JT*>(pipe.buf(j));
      const int k0 = j * S;
      for (int jj = 0; jj < S && k0 + jj < n; ++jj) {
        const int k = k0 + jj;
        const uint64_t bal = ballot64((down >> (k >> 6)) & 1ull);
        const FT c = ((bal >> (k & 63)) & 1ull) ? (FT)(-1) : (FT)1;
        lds_row_axpy<FT, JT, NS>(F, buf + jj * Npad, lane, c);
      }
    }
    __syncthreads();
  }

  const uint32_t rep = (uint32_t)(a.rep_offset + (active ? r : 0));
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  constexpr int NQ = (NS + 3) / 4;
  const int nblocks = gridDim.x;
  uint32_t done = 0;  // 64-spin blocks completed by this workgroup
  int tb = 0;         // tile buffer holding the current block's diagonal tile

  if (a.T > 0) {
    issue_tile<JT, NS, G>(sD, J, 0, w, lane);
    wait_all_loads();
    __syncthreads();
  }
  uint64_t nflip = 0, nrow = 0;  // wave-uniform: this wave's flips, the workgroup's rows
  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    TT nb;
    if constexpr (sizeof(TT) == 8) nb = sweep_scalar(a.p0_d[t]);
    else nb = sweep_scalar(a.p0_f[t]);
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
        // ---- A: in-block decisions against the staged diagonal tile
        const JT* D = reinterpret_cast<const JT*>(sD + tb * TILE_BYTES);
        const int i = 64 * k + lane;
        uint64_t flips = 0, sb0 = 0;
        if (active) {
          FieldVec<FT, 64> Dc;  // Dc[i] = D[i][lane]
          if constexpr (DREG) {
#pragma unroll
            for (int q = 0; q < 64; ++q) Dc.set(q, (FT)D[q * 64 + lane]);
            __builtin_amdgcn_sched_barrier(0);
          }
          const TT th = (i < n) ? accept_log<TT>(wd[jj], a.rule) * nb : neg_inf<TT>();
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
            FT dv;
            if constexpr (DREG) dv = Dc.dyn(i0);
            else dv = (FT)D[i0 * 64 + lane];
            fk = fma_ft<FT>(c, dv, fk);
          }
          if ((flips >> lane) & 1ull) down ^= 1ull << k;
        }
        if (lane == 0) sFlip[w] = flips;
        lds_barrier();
        uint64_t U = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) U |= sFlip[g];
        nflip += (uint64_t)__builtin_popcountll(flips);
        nrow += (uint64_t)__builtin_popcountll(U);
        // soft XCD pacing (performance only, bounded)
        if (a.xcd_window > 0 && w == 0 && done > (uint32_t)a.xcd_window) {
          const uint32_t need = done - (uint32_t)a.xcd_window;
          for (int spin = 0; spin < a.xcd_spin; ++spin) {
            if (peer_min_progress(a.progress, b, nblocks, lane) >= need) break;
            __builtin_amdgcn_s_sleep(4);
          }
        }
        // next block's tile first (block 0 of the next sweep after the last)
        const int kn = (k + 1 < NS) ? k + 1 : 0;
        issue_tile<JT, NS, G>(sD + (tb ^ 1) * TILE_BYTES, J, kn, w, lane);
        // ---- B: union rows through the pipeline, in increasing spin order
        const int nU = __builtin_popcountll(U);
        const int nrounds = (nU + S - 1) / S;
        int issued = 0;
        for (; issued < nbuf - 1 && issued < nrounds; ++issued) pipe.issue(issued, U, 64 * k, nU);
        uint64_t Urest = U;
        for (int j = 0; j < nrounds; ++j) {
          wait_vmcnt(pipe.allowed(j, issued, nU));
          lds_barrier();
          if (issued < nrounds) {
            pipe.issue(issued, U, 64 * k, nU);
            ++issued;
          }
          const JT* buf = reinterpret_cast<const JT*>(pipe.buf(j));
          for (int jj2 = 0; jj2 < S && Urest; ++jj2) {
            const int u = __builtin_ctzll(Urest);
            Urest &= Urest - 1;
            if (active && ((flips >> u) & 1ull)) {
              const FT c = ((sb0 >> u) & 1ull) ? (FT)2 : (FT)(-2);
              lds_row_axpy<FT, JT, NS>(F, buf + jj2 * Npad, lane, c);
            }
          }
        }
        wait_all_loads();
        lds_barrier();
        tb ^= 1;
        ++done;
        if (a.xcd_window > 0 && tid == 0)
          __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }

  if (a.stats && lane == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    if (w == 0) atomicAdd(a.stats + 1, (unsigned long long)nrow);
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

bool all_resident(const void* kern, int blocks, int threads, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return false;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess)
    return false;
  return (int64_t)per_cu * cus >= blocks;
}

// LDS of k_dense_shared: two diagonal tiles, the row ring (ring_rows rows of
// 64*NS couplings of jbytes each) and the per-wave flip words.
size_t shared_lds_bytes(int NS, int G, int ring_rows, int jbytes) {
  return (size_t)jbytes * (2 * 64 * 64 + (size_t)ring_rows * 64 * NS) + 8 * (size_t)G + 16;
}

template <typename FT, typename JT, int NS, int G>
static hipError_t launch_shared_t(const SharedArgs& a, hipStream_t st) {
  const size_t lds = shared_lds_bytes(NS, G, a.ring_rows, (int)sizeof(JT));
  auto kern = k_dense_shared<FT, JT, NS, G>;
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
