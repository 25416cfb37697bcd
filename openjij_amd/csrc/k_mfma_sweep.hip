// k_mfma_sweep.hip -- dense classical sweep whose row updates run on the
// matrix cores.
//
// Semantics are the reference sweep's (include/openjij/updater/
// single_spin_flip.hpp:67-95): spins visited in index order, a spin's test
// reads the fields left by every earlier accepted flip, on accept the fields
// move by -2 s_i J_i.  Eight waves anneal MG = 8 replicas; per 64-spin block k:
//   0. the wave owning block k's columns publishes the block's fields of the
//      8 replicas in LDS;
//   A. wave w decides the block for replica w (rejection skipping against the
//      diagonal tile J[block k][block k], one column per lane in VGPRs);
//   B. every wave applies the block's flips to its own columns as one
//      contraction: dF[r][j] += sum_{i in block} c[r][i] J[i][j], with
//      c = -2 s_i(old) for a flipped spin and 0 otherwise, on the matrix cores.
//
// Matrix-core form.  A 16x16 MFMA computes D[m][n] += sum_k A[m][k] B[k][n]
// with K = 32 (bf16) or 64 (int8).  Eight replicas fill only half of the 16
// rows, so each instruction covers TWO 16-column tiles t, t' and half of K:
// the first half of the K lanes carries block rows for tile t, the second half
// the same block rows for tile t'; A rows 0..7 hold the replicas' coefficients
// in the first half of K (zero in the second), rows 8..15 the same
// coefficients in the second half.  Then D rows 0..7 are the updates of tile t
// and rows 8..15 those of tile t', so every accumulator lane holds a field:
// lane l, element j = replica 4 ((l / 16) % 2) + j, column 16 (t or t') +
// l % 16 (t for l < 32).  A block takes 64 / (K / 2) K steps.
// Coefficients and spins are exact in bf16 and int8, so
//   * FP32 fields (bf16 hi/lo couplings, v_mfma_f32_16x16x32_bf16): every
//     block update adds sum c (J_hi + J_lo) in FP32 accumulation;
//     |J - J_hi - J_lo| <= 2^-18 |J|;
//   * INT32 fields (int8 couplings, v_mfma_i32_16x16x64_i8): exact integer
//     updates, so trajectories equal the oracle replay bit for bit.
// The B operand is read from a packed copy of the couplings in which every
// (block, K step, hi / lo part, tile pair) operand is 1 KiB contiguous, so
// each wave load is one contiguous 1 KiB.  All 64 rows of a block are read
// whatever flipped: the cost per block is fixed (the block's 64 x Npad
// couplings per workgroup) and the VALU does no row arithmetic.
#include "ojmi_colown.h"

namespace ojmi {

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x4_t __attribute__((ext_vector_type(4)));

constexpr int MW = 8;  // waves per workgroup
constexpr int MG = 8;  // replicas per workgroup

template <typename FT>
struct MfmaKind;
template <>
struct MfmaKind<float> {
  typedef f32x4_t acc_t;
  typedef float JT;                 // diagonal tile: natural FP32 couplings
  static constexpr int KH = 16;     // block rows per K step (half of K = 32)
  static constexpr int PARTS = 2;   // J_hi, J_lo
};
template <>
struct MfmaKind<int32_t> {
  typedef i32x4_t acc_t;
  typedef int8_t JT;
  static constexpr int KH = 32;     // half of K = 64
  static constexpr int PARTS = 1;
};

// Coefficient operand of K step kq for this lane.  Lane l supplies A row
// m = l % 16 and K elements e0 .. e0 + EL - 1 with e0 = EL (l / 16); row m
// belongs to replica m % 8 in K half m / 8, so the lane is non-zero only when
// its K elements lie in that half.  fl, sg: that replica's flip and old-sign
// words of the block.
template <int EL>
__device__ __forceinline__ u32x4 coef_operand(uint64_t fl, uint64_t sg, int lane, int kq) {
  constexpr int KH = 2 * EL;  // block rows per K step
  const int half = lane >> 5, m = lane & 15;
  const int r0 = kq * KH + EL * ((lane >> 4) & 1);
  const bool use = (m >> 3) == half;
  const uint32_t f = use ? (uint32_t)(fl >> r0) & ((1u << EL) - 1u) : 0u;
  const uint32_t s = (uint32_t)(sg >> r0);
  u32x4 out;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    uint32_t wd = 0;
    if constexpr (EL == 8) {  // bf16: +2 = 0x4000 (spin was -1), -2 = 0xC000
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * q + h;
        const uint32_t v = ((f >> i) & 1u) ? (((s >> i) & 1u) ? 0x4000u : 0xC000u) : 0u;
        wd |= v << (16 * h);
      }
    } else {  // int8: +2 = 0x02, -2 = 0xFE
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        const int i = 4 * q + h;
        const uint32_t v = ((f >> i) & 1u) ? (((s >> i) & 1u) ? 0x02u : 0xFEu) : 0u;
        wd |= v << (8 * h);
      }
    }
    out[q] = wd;
  }
  return out;
}

template <typename FT>
__device__ __forceinline__ typename MfmaKind<FT>::acc_t mma(const u32x4& a, const u32x4& b,
                                                            typename MfmaKind<FT>::acc_t c) {
  if constexpr (std::is_floating_point<FT>::value) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                   __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(i32x4_t, a),
                                                 __builtin_bit_cast(i32x4_t, b), c, 0, 0, 0);
  }
}

}  // namespace

// Packed B operands.  Operand (block kb, K step kq, part p, tile pair P),
// 1 KiB at index ((kb * KQ + kq) * PARTS + p) * 2 NS + P with KQ = 64 / KH;
// lane l holds EL = KH / 2 consecutive rows of one column:
//   rows 64 kb + KH kq + EL ((l / 16) % 2) + e, e < EL,
//   column 16 (2 P + l / 32) + l % 16,
// of part p of J (bf16 J_hi / J_lo; int8 J).
__global__ __launch_bounds__(256) void k_pack_mfma_b16(const uint16_t* __restrict__ Jhi,
                                                      const uint16_t* __restrict__ Jlo, int NS,
                                                      uint16_t* __restrict__ out) {
  const int Npad = 64 * NS;
  const int64_t total = (int64_t)NS * 4 * 2 * (2 * NS) * 64;  // lanes of all operands
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e & 63);
    const int64_t op = e >> 6;
    const int P = (int)(op % (2 * NS));
    const int64_t q = op / (2 * NS);
    const int p = (int)(q & 1), kq = (int)((q >> 1) & 3), kb = (int)(q >> 3);
    const uint16_t* src = p ? Jlo : Jhi;
    const int col = 16 * (2 * P + (lane >> 5)) + (lane & 15);
    const int r0 = 64 * kb + 16 * kq + 8 * ((lane >> 4) & 1);
    uint16_t* dst = out + e * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) dst[i] = src[(int64_t)(r0 + i) * Npad + col];
  }
}

__global__ __launch_bounds__(256) void k_pack_mfma_i8(const int8_t* __restrict__ J8, int NS,
                                                     int8_t* __restrict__ out) {
  const int Npad = 64 * NS;
  const int64_t total = (int64_t)NS * 2 * (2 * NS) * 64;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int lane = (int)(e & 63);
    const int64_t op = e >> 6;
    const int P = (int)(op % (2 * NS));
    const int64_t q = op / (2 * NS);
    const int kq = (int)(q & 1), kb = (int)(q >> 1);
    const int col = 16 * (2 * P + (lane >> 5)) + (lane & 15);
    const int r0 = 64 * kb + 32 * kq + 16 * ((lane >> 4) & 1);
    int8_t* dst = out + e * 16;
#pragma unroll
    for (int i = 0; i < 16; ++i) dst[i] = J8[(int64_t)(r0 + i) * Npad + col];
  }
}

// a.J: natural couplings (FP32 or int8) for the diagonal tiles; Jpk: packed
// B operands; F_in: the fields of the current spins ([R][Npad]), required.
template <typename FT, int NS, int RULE>
__global__ __launch_bounds__(64 * MW) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_dense_mfma(SharedArgs a, const char* __restrict__ Jpk) {
  typedef MfmaKind<FT> K;
  typedef typename K::acc_t acc_t;
  typedef typename K::JT JT;
  typedef typename ColThresh<FT>::type TT;
  constexpr int Npad = 64 * NS;
  constexpr int NP = 2 * NS;          // tile pairs (32 columns each)
  constexpr int TP = NP / MW;         // tile pairs per wave
  constexpr int BPW = NS / MW;        // blocks per wave (2 pairs each)
  constexpr int KQ = 64 / K::KH;      // K steps per block
  constexpr int KP = KQ * K::PARTS;   // B operands per tile pair per block
  constexpr int EL = K::KH / 2;       // K elements per lane
  constexpr int NI = KP * TP;         // B operands per wave per block
  constexpr int GS = NI < 8 ? NI : 8;  // operands per load group
  constexpr int NG = NI / GS;
  constexpr int TL = 16 / MW;         // tile loads per wave
  constexpr int TILE_BYTES = 64 * 64 * (int)sizeof(JT);
  static_assert(NS % MW == 0 && NI % GS == 0 && TP >= 2, "shape");
  if (a.sel && *a.sel != a.sel_id) return;  // another kernel runs this segment

  extern __shared__ __align__(16) char smem[];
  char* sD = smem;                                             // diagonal tile
  FT* sF = reinterpret_cast<FT*>(smem + TILE_BYTES);           // [MG][64]
  uint64_t* sFlip = reinterpret_cast<uint64_t*>(sF + MG * 64);  // [MG]
  uint64_t* sSign = sFlip + MG;                                // [MG] old signs
  uint64_t* sDown = sSign + MG;                                // [MG][NS] spin words

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = blockIdx.x;
  const int n = a.n;
  const int rbase = b * MG;
  const JT* J = reinterpret_cast<const JT*>(a.J);

  // spins: word (g, kb) bit l = spin 64 kb + l of replica rbase + g is -1
  for (int p = w; p < MG * NS; p += MW) {
    const int g = p / NS, kb = p - g * NS;
    const int col = 64 * kb + lane;
    bool dn = false;
    if (rbase + g < a.R && col < n) dn = a.spins[(int64_t)(rbase + g) * n + col] < 0;
    const uint64_t m = ballot64(dn);
    if (lane == 0) sDown[p] = m;
  }
  // fields of this wave's tile pairs by buffer loads: element j of lane l is
  // replica 4 ((l / 16) % 2) + j, column 16 (2 P + l / 32) + l % 16 of pair
  // P = w TP + p; offsets of replicas past R fall outside the buffer (read 0,
  // writes dropped)
  const int fbytes_all = a.R * Npad * (int)sizeof(FT);
  int foff[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int g = 4 * ((lane >> 4) & 1) + j;
    foff[j] = (rbase + g < a.R)
                  ? ((rbase + g) * Npad + 32 * w * TP + 16 * (lane >> 5) + (lane & 15)) *
                        (int)sizeof(FT)
                  : fbytes_all;
  }
  acc_t acc[TP];
  {
    const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<void*>(a.F_in), (short)0, fbytes_all, 0x00020000);
#pragma unroll
    for (int p = 0; p < TP; ++p) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[p][j] = __builtin_bit_cast(
            FT, __builtin_amdgcn_raw_buffer_load_b32(frs, foff[j], 128 * p, 0));
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  const bool active = rbase + w < a.R;
  const uint32_t rep = (uint32_t)(a.rep_offset + (active ? rbase + w : 0));
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  const int nblocks = gridDim.x;
  uint32_t done = 0;
  uint64_t nflip = 0, nrow = 0;  // workgroup-uniform counters
  typename TileVec<JT>::type tile[TL];
  u32x4 buf[2][GS];
  // B operands by buffer loads: one VGPR offset per lane (its 16 bytes of
  // this wave's first tile pair), the operand offset in a scalar register
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char*>(Jpk), (short)0, (int)((size_t)NS * KP * NP * 1024), 0x00020000);
  const int voff = lane * 16 + w * TP * 1024;
  auto issue = [&](int kb, int grp, u32x4 (&bb)[GS]) {
    const int kbase = __builtin_amdgcn_readfirstlane(kb * KP * NP * 1024);
#pragma unroll
    for (int e = 0; e < GS; ++e) {
      const int q = grp * GS + e;
      const int kp = q / TP, p = q - kp * TP;
      bb[e] = __builtin_bit_cast(
          u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, kbase + (kp * NP + p) * 1024, 0));
    }
  };
  constexpr bool CROSS = (NG % 2) == 0;  // group 0 of the next block prefetched

  if (a.T > 0) {
    load_tile<JT, NS, MW>(tile, reinterpret_cast<const char*>(J), 0, w, lane);
    store_tile<JT, MW>(tile, sD, w, lane);
    if constexpr (CROSS) issue(0, 0, buf[0]);
  }
  __syncthreads();

  for (int t = 0; t < a.T; ++t) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
    TT nb;
    if constexpr (sizeof(TT) == 8) nb = a.p0_d[t];
    else nb = a.p0_f[t];
    uint32_t wd[4] = {0, 0, 0, 0};
#pragma unroll 1
    for (int k = 0; k < NS; ++k) {
      const int jj = k & 3;
      if (jj == 0) {
        const uint32_t c0 = (uint32_t)(16 * (k + (lane & 3)) + (lane >> 2));
        philox4x32_10(c0, tg, rep, 0u, slo, shi, wd[0], wd[1], wd[2], wd[3]);
        quad_transpose(wd[0], wd[1], wd[2], wd[3], lane);
      }
      const int ow = k / BPW, lb = k - ow * BPW;
      // ---- 0: the owner publishes block k's fields of the 8 replicas
      if (w == ow) {
#pragma unroll
        for (int c = 0; c < BPW; ++c) {
          if (c != lb) continue;
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 4; ++j)
              sF[(4 * ((lane >> 4) & 1) + j) * 64 + 16 * (2 * h + (lane >> 5)) + (lane & 15)] =
                  acc[2 * c + h][j];
        }
      }
      bar();
      // ---- A: wave w decides block k for its replica
      uint64_t flips = 0;
      const uint64_t sb = uniform64(sDown[w * NS + k]);
      if (active) {
        const JT* D = reinterpret_cast<const JT*>(sD);
        FieldVec<FT, 64> Dc;  // Dc[i] = D[i][lane]
#pragma unroll
        for (int q = 0; q < 64; ++q) Dc.set(q, (FT)D[q * 64 + lane]);
        __builtin_amdgcn_sched_barrier(0);
        const int i = 64 * k + lane;
        const TT th = (i < n) ? accept_log<TT>(pick4(wd, jj), RULE) * nb : neg_inf<TT>();
        uint64_t live = ballot64(i < n);
        // within the block every live lane keeps its sign: dE = s2 * F
        const FT s2 = ((sb >> lane) & 1ull) ? (FT)2 : (FT)(-2);
        FT fk = sF[w * 64 + lane];
        // the chain runs on the scalar unit: the next flip i0 and the lanes
        // <= i0 (done) from one subtraction, the coefficient from lane i0
        uint64_t cand = ballot64((TT)(s2 * fk) < th) & live;
        while (cand != 0ull) {
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
        sSign[w] = sb;
        sDown[w * NS + k] = sb ^ flips;
      }
      bar();
      // ---- B: the block's flips applied to this wave's columns
      uint64_t flm, sgm;
      {
        const int g = lane & 7;
        flm = sFlip[g];
        sgm = sSign[g];
      }
#pragma unroll
      for (int g = 0; g < MG; ++g)
        nflip += (uint64_t)__builtin_popcountll(uniform64(sFlip[g]));
      {
        const int left = n - 64 * k;  // rows of the block inside the problem
        nrow += (uint64_t)(left <= 0 ? 0 : (left < 64 ? left : 64));
      }
      u32x4 A[KQ];
#pragma unroll
      for (int kq = 0; kq < KQ; ++kq) A[kq] = coef_operand<EL>(flm, sgm, lane, kq);
      // soft XCD pacing (performance only, bounded)
      if (a.xcd_window > 0 && w == 0 && done > (uint32_t)a.xcd_window) {
        const uint32_t need = done - (uint32_t)a.xcd_window;
        for (int spin = 0; spin < a.xcd_spin; ++spin) {
          if (peer_min(a.progress, b, nblocks, lane) >= need) break;
          __builtin_amdgcn_s_sleep(4);
        }
      }
      const int kn = (k + 1 < NS) ? k + 1 : 0;
      if constexpr (!CROSS) issue(k, 0, buf[0]);
      load_tile<JT, NS, MW>(tile, reinterpret_cast<const char*>(J), kn, w, lane);
#pragma unroll
      for (int grp = 0; grp < NG; ++grp) {
        if (grp + 1 < NG) {
          issue(k, grp + 1, buf[(grp + 1) & 1]);
          __builtin_amdgcn_sched_barrier(0);
          if (grp == 0) wait_vm<GS + TL>();
          else wait_vm<GS>();
        } else if constexpr (CROSS) {
          issue(kn, 0, buf[0]);  // group 0 of the next block
          __builtin_amdgcn_sched_barrier(0);
          wait_vm<GS>();
        } else {
          wait_vm<0>();
        }
#pragma unroll
        for (int e = 0; e < GS; ++e) {
          const int q = grp * GS + e;
          const int kp = q / TP, p = q - kp * TP;
          acc[p] = mma<FT>(A[kp / K::PARTS], buf[grp & 1][e], acc[p]);
        }
      }
      store_tile<JT, MW>(tile, sD, w, lane);
      ++done;
      if (a.xcd_window > 0 && tid == 0)
        __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  wait_vm<0>();
  __syncthreads();

  if (a.stats && tid == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nrow);
  }
  for (int p = w; p < MG * NS; p += MW) {
    const int g = p / NS, kb = p - g * NS;
    const int col = 64 * kb + lane;
    if (rbase + g < a.R && col < n)
      a.spins[(int64_t)(rbase + g) * n + col] = ((sDown[p] >> lane) & 1ull) ? (int8_t)-1 : (int8_t)1;
  }
  if (a.F_out) {
    const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(
        a.F_out, (short)0, fbytes_all, 0x00020000);
#pragma unroll
    for (int p = 0; p < TP; ++p) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        // element copied to a scalar before the bit cast: __builtin_bit_cast
        // applied to a vector element expression compiled to element 0 for
        // every j (read from the ISA of this kernel)
        const FT x = acc[p][j];
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), frs, foff[j],
                                              128 * p, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

// Split-role form of the same sweep: waves 0..3 hold the fields (tile pairs
// [s NS/2, (s+1) NS/2) of wave s) and stream each block's update, waves 4..7
// decide (wave 4 + d for replicas d and d + 4, the two chains interleaved).
// Per block k, between two workgroup barriers B1 and B2:
//   before B1: the wave owning block k's columns applies block k-1's update
//              to those two tile pairs in temporaries (the same MFMAs in the
//              same order as its full update below, so the values are
//              identical) and publishes them;
//   B1 .. B2:  the deciders decide block k while the field waves apply block
//              k-1's update to all their pairs;
//   after B2:  block k's coefficient operands are built from its flip words.
// So each block costs about the larger of its decisions and its update
// stream, not their sum.  Same results as k_dense_mfma, bit for bit.
template <typename FT, int NS, int RULE>
__global__ __launch_bounds__(64 * MW) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_dense_mfma_split(SharedArgs a, const char* __restrict__ Jpk) {
  typedef MfmaKind<FT> K;
  typedef typename K::acc_t acc_t;
  typedef typename K::JT JT;
  typedef typename ColThresh<FT>::type TT;
  constexpr int SW = 4;               // field (streaming) waves; waves SW.. decide
  constexpr int Npad = 64 * NS;
  constexpr int NP = 2 * NS;          // tile pairs (32 columns each)
  constexpr int TP = NP / SW;         // tile pairs per field wave
  constexpr int BPW = NS / SW;        // blocks per field wave (2 pairs each)
  constexpr int KQ = 64 / K::KH;      // K steps per block
  constexpr int KP = KQ * K::PARTS;   // B operands per tile pair per block
  constexpr int EL = K::KH / 2;       // K elements per lane
  constexpr int NI = KP * TP;         // B operands per field wave per block
  constexpr int GS = 8;               // operands per load group
  constexpr int NG = NI / GS;
  constexpr int TL = 16 / (MW - SW);  // tile loads per deciding wave
  constexpr int TILE_BYTES = 64 * 64 * (int)sizeof(JT);
  static_assert(NS % SW == 0 && NI % GS == 0 && NG >= 2 && KP <= GS, "shape");
  if (a.sel && *a.sel != a.sel_id) return;  // another kernel runs this segment

  extern __shared__ __align__(16) char smem[];
  char* sD = smem;                                             // diagonal tile
  FT* sF = reinterpret_cast<FT*>(smem + TILE_BYTES);           // [MG][64]
  uint64_t* sFlip = reinterpret_cast<uint64_t*>(sF + MG * 64);  // [MG]
  uint64_t* sSign = sFlip + MG;                                // [MG] old signs
  uint64_t* sDown = sSign + MG;                                // [MG][NS] spin words

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const bool fw = w < SW;  // field wave
  const int b = blockIdx.x;
  const int n = a.n;
  const int rbase = b * MG;
  const JT* J = reinterpret_cast<const JT*>(a.J);

  for (int p = w; p < MG * NS; p += MW) {
    const int g = p / NS, kb = p - g * NS;
    const int col = 64 * kb + lane;
    bool dn = false;
    if (rbase + g < a.R && col < n) dn = a.spins[(int64_t)(rbase + g) * n + col] < 0;
    const uint64_t m = ballot64(dn);
    if (lane == 0) sDown[p] = m;
  }
  if (a.T > 0 && !fw) {
    typename TileVec<JT>::type tile[TL];
    load_tile<JT, NS, MW - SW>(tile, reinterpret_cast<const char*>(J), 0, w - SW, lane);
    store_tile<JT, MW - SW>(tile, sD, w - SW, lane);
  }
  __syncthreads();

  const int64_t total = (int64_t)a.T * NS;  // blocks of the run
  uint64_t nflip = 0, nrow = 0;             // workgroup-uniform counters (wave 0)
  if (fw) {
    // ================= field waves: tile pairs [w TP, (w+1) TP)
    const int fbytes_all = a.R * Npad * (int)sizeof(FT);
    int foff[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int g = 4 * ((lane >> 4) & 1) + j;
      foff[j] = (rbase + g < a.R)
                    ? ((rbase + g) * Npad + 32 * w * TP + 16 * (lane >> 5) + (lane & 15)) *
                          (int)sizeof(FT)
                    : fbytes_all;
    }
    acc_t acc[TP];
    {
      const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<void*>(a.F_in), (short)0, fbytes_all, 0x00020000);
#pragma unroll
      for (int p = 0; p < TP; ++p) {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[p][j] = __builtin_bit_cast(
              FT, __builtin_amdgcn_raw_buffer_load_b32(frs, foff[j], 128 * p, 0));
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    u32x4 buf[2][GS];
    u32x4 A[KQ];  // coefficient operands of the block whose update is pending
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(Jpk), (short)0, (int)((size_t)NS * KP * NP * 1024), 0x00020000);
    const int voff = lane * 16 + w * TP * 1024;
    // operand (block kb, K step / part kp, local pair p) of this wave
    auto opnd = [&](int kb, int kp, int p) {
      const int off =
          __builtin_amdgcn_readfirstlane(kb * KP * NP * 1024) + (kp * NP + p) * 1024;
      return __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, voff, off, 0));
    };
    auto issue = [&](int kb, int grp, u32x4 (&bb)[GS]) {
#pragma unroll
      for (int e = 0; e < GS; ++e) {
        const int q = grp * GS + e;
        bb[e] = opnd(kb, q / TP, q - (q / TP) * TP);
      }
    };
    // the pending update (block kb, operands A) on every pair of this wave,
    // with group 0 already in flight in buf[0]
    auto stream = [&](int kb) {
#pragma unroll
      for (int grp = 0; grp < NG; ++grp) {
        if (grp + 1 < NG) {
          issue(kb, grp + 1, buf[(grp + 1) & 1]);
          __builtin_amdgcn_sched_barrier(0);
          wait_vm<GS>();
        } else {
          wait_vm<0>();
        }
#pragma unroll
        for (int e = 0; e < GS; ++e) {
          const int q = grp * GS + e;
          const int kp = q / TP, p = q - kp * TP;
          acc[p] = mma<FT>(A[kp / K::PARTS], buf[grp & 1][e], acc[p]);
        }
      }
    };
    int pend = -1;  // block whose update is pending
#pragma unroll 1
    for (int64_t gb = 0; gb < total; ++gb) {
      const int k = (int)(gb % NS);
      const int ow = k / BPW, lb = k - ow * BPW;
      // ---- 1: the owner applies the pending update to block k's pairs
      // (operands in buf, loaded in the previous iteration) and publishes them
      if (w == ow) {
        acc_t tmp[2];
#pragma unroll
        for (int c = 0; c < BPW; ++c) {
          if (c != lb) continue;
          tmp[0] = acc[2 * c];
          tmp[1] = acc[2 * c + 1];
        }
        if (pend >= 0) {
          wait_vm<0>();
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int kp = 0; kp < KP; ++kp)
              tmp[h] = mma<FT>(A[kp / K::PARTS], buf[h][kp], tmp[h]);
          issue(pend, 0, buf[0]);  // group 0 of the pending update stream
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const FT x = tmp[h][j];
            sF[(4 * ((lane >> 4) & 1) + j) * 64 + 16 * (2 * h + (lane >> 5)) + (lane & 15)] = x;
          }
      }
      bar();  // B1
      // ---- 3: the pending update on all pairs while block k is decided
      if (pend >= 0) stream(pend);
      // operands of the next iteration: block k's update becomes pending
      const int nk = (k + 1 < NS) ? k + 1 : 0;
      if (gb + 1 < total && w == nk / BPW) {
        const int nlb = nk - (nk / BPW) * BPW;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int kp = 0; kp < KP; ++kp) buf[h][kp] = opnd(k, kp, 2 * nlb + h);
      } else {
        issue(k, 0, buf[0]);
      }
      bar();  // B2
      // ---- 5: block k's coefficient operands
      {
        const int g = lane & 7;
        const uint64_t flm = sFlip[g], sgm = sSign[g];
#pragma unroll
        for (int kq = 0; kq < KQ; ++kq) A[kq] = coef_operand<EL>(flm, sgm, lane, kq);
      }
      pend = k;
      if (w == 0) {
#pragma unroll
        for (int g = 0; g < MG; ++g)
          nflip += (uint64_t)__builtin_popcountll(uniform64(sFlip[g]));
        const int left = n - 64 * k;  // rows of the block inside the problem
        nrow += (uint64_t)(left <= 0 ? 0 : (left < 64 ? left : 64));
      }
    }
    if (pend >= 0) stream(pend);  // the last block's update
    wait_vm<0>();
    if (a.F_out) {
      const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(
          a.F_out, (short)0, fbytes_all, 0x00020000);
#pragma unroll
      for (int p = 0; p < TP; ++p) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          // element copied to a scalar before the bit cast (see k_dense_mfma)
          const FT x = acc[p][j];
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, x), frs, foff[j],
                                                128 * p, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  } else {
    // ================= deciding waves: replicas g0 = w - SW and g1 = g0 + 4
    const int dwv = w - SW;
    const int g0 = dwv, g1 = dwv + 4;
    const bool act0 = rbase + g0 < a.R, act1 = rbase + g1 < a.R;
    const uint32_t rep0 = (uint32_t)(a.rep_offset + (act0 ? rbase + g0 : 0));
    const uint32_t rep1 = (uint32_t)(a.rep_offset + (act1 ? rbase + g1 : 0));
    const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
    const int nblocks = gridDim.x;
    uint32_t done = 0;
    uint32_t wd0[4] = {0, 0, 0, 0}, wd1[4] = {0, 0, 0, 0};
    typename TileVec<JT>::type tile[TL];
#pragma unroll 1
    for (int64_t gb = 0; gb < total; ++gb) {
      const int t = (int)(gb / NS), k = (int)(gb - (int64_t)t * NS);
      const int jj = k & 3;
      if (jj == 0) {
        const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
        const uint32_t c0 = (uint32_t)(16 * (k + (lane & 3)) + (lane >> 2));
        philox4x32_10(c0, tg, rep0, 0u, slo, shi, wd0[0], wd0[1], wd0[2], wd0[3]);
        quad_transpose(wd0[0], wd0[1], wd0[2], wd0[3], lane);
        philox4x32_10(c0, tg, rep1, 0u, slo, shi, wd1[0], wd1[1], wd1[2], wd1[3]);
        quad_transpose(wd1[0], wd1[1], wd1[2], wd1[3], lane);
      }
      bar();  // B1
      // ---- 3: decide block k for replicas g0 and g1
      TT nb;
      if constexpr (sizeof(TT) == 8) nb = a.p0_d[t];
      else nb = a.p0_f[t];
      uint64_t flips0 = 0, flips1 = 0;
      const uint64_t sb0 = uniform64(sDown[g0 * NS + k]);
      const uint64_t sb1 = uniform64(sDown[g1 * NS + k]);
      if (act0 || act1) {
        const JT* D = reinterpret_cast<const JT*>(sD);
        FieldVec<FT, 64> Dc;  // Dc[i] = D[i][lane]
#pragma unroll
        for (int q = 0; q < 64; ++q) Dc.set(q, (FT)D[q * 64 + lane]);
        __builtin_amdgcn_sched_barrier(0);
        const int i = 64 * k + lane;
        const TT th0 =
            (act0 && i < n) ? accept_log<TT>(pick4(wd0, jj), RULE) * nb : neg_inf<TT>();
        const TT th1 =
            (act1 && i < n) ? accept_log<TT>(pick4(wd1, jj), RULE) * nb : neg_inf<TT>();
        uint64_t live0 = act0 ? ballot64(i < n) : 0ull;
        uint64_t live1 = act1 ? ballot64(i < n) : 0ull;
        // within the block every live lane keeps its sign: dE = s2 * F
        const FT s20 = ((sb0 >> lane) & 1ull) ? (FT)2 : (FT)(-2);
        const FT s21 = ((sb1 >> lane) & 1ull) ? (FT)2 : (FT)(-2);
        FT fk0 = sF[g0 * 64 + lane];
        FT fk1 = sF[g1 * 64 + lane];
        // two chains interleaved, each on the scalar unit as in the uniform
        // form (next flip and done lanes from one subtraction, coefficient
        // from the flipping lane)
        while (true) {
          const uint64_t cand0 = ballot64((TT)(s20 * fk0) < th0) & live0;
          const uint64_t cand1 = ballot64((TT)(s21 * fk1) < th1) & live1;
          if ((cand0 | cand1) == 0ull) break;
          if (cand0) {
            const uint64_t below = cand0 ^ (cand0 - 1ull);
            const int i0 = __builtin_ctzll(cand0);
            live0 &= ~below;
            flips0 |= cand0 & below;
            fk0 = fma_ft<FT>(lane_value(s20, i0), Dc.dyn(i0), fk0);
          }
          if (cand1) {
            const uint64_t below = cand1 ^ (cand1 - 1ull);
            const int i1 = __builtin_ctzll(cand1);
            live1 &= ~below;
            flips1 |= cand1 & below;
            fk1 = fma_ft<FT>(lane_value(s21, i1), Dc.dyn(i1), fk1);
          }
        }
      }
      if (lane == 0) {
        sFlip[g0] = flips0;
        sSign[g0] = sb0;
        sDown[g0 * NS + k] = sb0 ^ flips0;
        sFlip[g1] = flips1;
        sSign[g1] = sb1;
        sDown[g1 * NS + k] = sb1 ^ flips1;
      }
      // soft XCD pacing (performance only, bounded), by the first deciding
      // wave once its decisions are made (the field waves stream meanwhile)
      if (dwv == 0 && a.xcd_window > 0 && done > (uint32_t)a.xcd_window) {
        const uint32_t need = done - (uint32_t)a.xcd_window;
        for (int spin = 0; spin < a.xcd_spin; ++spin) {
          if (peer_min(a.progress, b, nblocks, lane) >= need) break;
          __builtin_amdgcn_s_sleep(4);
        }
      }
      // the next block's diagonal tile
      const int kn = (k + 1 < NS) ? k + 1 : 0;
      load_tile<JT, NS, MW - SW>(tile, reinterpret_cast<const char*>(J), kn, dwv, lane);
      bar();  // B2
      wait_vm<0>();
      store_tile<JT, MW - SW>(tile, sD, dwv, lane);
      if (dwv == 0) {
        ++done;
        if (a.xcd_window > 0 && lane == 0)
          __hip_atomic_store(a.progress + b, done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  __syncthreads();

  if (a.stats && tid == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nrow);
  }
  for (int p = w; p < MG * NS; p += MW) {
    const int g = p / NS, kb = p - g * NS;
    const int col = 64 * kb + lane;
    if (rbase + g < a.R && col < n)
      a.spins[(int64_t)(rbase + g) * n + col] = ((sDown[p] >> lane) & 1ull) ? (int8_t)-1 : (int8_t)1;
  }
}

__global__ void k_sweep_select(const unsigned long long* __restrict__ stats, int32_t* sel,
                               uint64_t attempts, int permille, int init) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  unsigned long long* snap = reinterpret_cast<unsigned long long*>(sel + 2);
  const unsigned long long now = stats[0];
  if (init) {
    sel[0] = 1;
    sel[1] = 0;
  } else {
    sel[1] += sel[0];  // the segment just run
    const unsigned long long f = now - *snap;
    sel[0] = (f * 1000ull >= (unsigned long long)permille * attempts) ? 1 : 0;
  }
  *snap = now;
}

hipError_t launch_sweep_select(const unsigned long long* stats, int32_t* sel, uint64_t attempts,
                               int permille, int init, hipStream_t st) {
  hipLaunchKernelGGL(k_sweep_select, dim3(1), dim3(64), 0, st, stats, sel, attempts, permille,
                     init);
  return hipGetLastError();
}

size_t mfma_sweep_lds_bytes(int NS, int fbytes, int jbytes) {
  return (size_t)64 * 64 * jbytes + (size_t)MG * 64 * fbytes + 2 * 8 * (size_t)MG +
         8 * (size_t)MG * NS;
}

bool mfma_sweep_supported(int NS) { return NS >= 16 && NS <= 64; }

size_t mfma_packed_bytes(int ftype, int NS) {
  // FP32: 4 K steps x 2 parts, int8: 2 K steps, per block and tile pair
  return (ftype == FT_F32 ? (size_t)8 : (size_t)2) * (size_t)NS * (2 * NS) * 1024;
}

hipError_t launch_pack_mfma(int ftype, const void* Jhi, const void* Jlo, int NS, void* out,
                            hipStream_t st) {
  const int64_t total = (int64_t)NS * (ftype == FT_F32 ? 8 : 2) * (2 * NS) * 64;
  const int grid = (int)std::min<int64_t>((total + 255) / 256, 8192);
  if (ftype == FT_F32)
    hipLaunchKernelGGL(k_pack_mfma_b16, dim3(grid), dim3(256), 0, st,
                       static_cast<const uint16_t*>(Jhi), static_cast<const uint16_t*>(Jlo), NS,
                       static_cast<uint16_t*>(out));
  else
    hipLaunchKernelGGL(k_pack_mfma_i8, dim3(grid), dim3(256), 0, st,
                       static_cast<const int8_t*>(Jhi), NS, static_cast<int8_t*>(out));
  return hipGetLastError();
}

namespace {
template <typename FT, int NS, int RULE>
hipError_t launch_mfma_t(const SharedArgs& a, const void* Jpk, int form, hipStream_t st) {
  const size_t lds = mfma_sweep_lds_bytes(NS, 4, (int)sizeof(typename MfmaKind<FT>::JT));
  auto kern = form ? k_dense_mfma_split<FT, NS, RULE> : k_dense_mfma<FT, NS, RULE>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  dim3 grid((a.R + MG - 1) / MG), block(64 * MW);
  SharedArgs b = a;
  if (b.xcd_window > 0 && !all_resident(reinterpret_cast<const void*>(kern), (int)grid.x,
                                         (int)block.x, lds))
    b.xcd_window = 0;  // pacing waits on peers that may not have started
  hipLaunchKernelGGL(kern, grid, block, lds, st, b, static_cast<const char*>(Jpk));
  return hipGetLastError();
}

template <typename FT, int RULE>
hipError_t launch_mfma_r(const SharedArgs& a, const void* Jpk, int NS, int form, hipStream_t st) {
  switch (NS) {
    case 16: return launch_mfma_t<FT, 16, RULE>(a, Jpk, form, st);
    case 32: return launch_mfma_t<FT, 32, RULE>(a, Jpk, form, st);
    case 64: return launch_mfma_t<FT, 64, RULE>(a, Jpk, form, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename FT>
hipError_t launch_mfma_f(const SharedArgs& a, const void* Jpk, int NS, int form, hipStream_t st) {
  if (a.rule == RULE_METROPOLIS) return launch_mfma_r<FT, RULE_METROPOLIS>(a, Jpk, NS, form, st);
  if (a.rule == RULE_HEAT_BATH) return launch_mfma_r<FT, RULE_HEAT_BATH>(a, Jpk, NS, form, st);
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t launch_dense_mfma(int ftype, const SharedArgs& a, const void* Jpk, int NS, int form,
                             hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (!mfma_sweep_supported(NS) || !a.F_in || !Jpk) return hipErrorInvalidValue;
  if (ftype == FT_F32) return launch_mfma_f<float>(a, Jpk, NS, form, st);
  if (ftype == FT_I32) return launch_mfma_f<int32_t>(a, Jpk, NS, form, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
