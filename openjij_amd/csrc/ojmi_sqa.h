// ojmi_sqa.h -- Suzuki-Trotter dense sweeps, one workgroup per replica: the
// kernel template, instantiated by k_sqa_*.hip (one unit per field type and
// slot range, so that they compile in parallel).
#pragma once
#include "ojmi_anneal.h"

namespace ojmi {

// ---------------------------------------------------------------------------
// Transverse (Suzuki-Trotter): one workgroup per replica, wave w owns the
// slice pairs p = w, w+W, ... (slices 2p and 2p+1), so every wave has one
// slice in the even phase and one in the odd phase.  Phase order follows the
// reference: even slices, odd slices, then slice M-1 when M is odd
// (include/openjij/updater/single_spin_flip.hpp:160-192).  The spin words of
// all slices live in LDS (one 64-bit ballot per slice and slot) so that each
// lane can read the up/down neighbours of its own spin.
//   dEcl = -2 s F,  dEtr = -2 s (s_up + s_dn),  dE = fma(a, dEcl, B*dEtr),
//   flip iff dE < -log(U)  (a = s*beta/M, B = 1/2 log tanh(beta*Gamma*(1-s)/M)).
template <typename FT, int NS, int PPW, int MAXT, int SL>
__global__ __launch_bounds__(MAXT) void k_sqa_wave(DenseArgs a) {
  typedef FT TT;
  extern __shared__ uint64_t s_down[];  // [M][NS], then (SL > 0) slots [SL][n][64] of every row
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = blockDim.x >> 6;
  const int r = blockIdx.x;
  const int n = a.n, M = a.M;
  const int P = (M + 1) / 2;
  const char* J = reinterpret_cast<const char*>(a.J);
  constexpr int64_t RB = row_bytes<FT, NS>();
  const FT* h = reinterpret_cast<const FT*>(a.h);

  // SL > 0: slots 0..SL-1 of every coupling row copied to LDS once
  // ([slot][row][64]); a flip reads those slots there and the others from L2
  FT* sJ = reinterpret_cast<FT*>(s_down + M * NS);
  if constexpr (SL > 0) {
    for (int e = threadIdx.x; e < SL * n * 64; e += blockDim.x) {
      const int m = e / (n * 64), rl = e - m * n * 64, row = rl >> 6, l = rl & 63;
      sJ[e] = reinterpret_cast<const FT*>(J + (int64_t)row * RB)[m * 64 + l];
    }
  }

  uint64_t down[2 * PPW];
  FieldVec<FT, NS> F[2 * PPW];
#pragma unroll
  for (int pp = 0; pp < PPW; ++pp) {
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      const int p = w + pp * W;
      const int t = 2 * p + par;
      down[2 * pp + par] = 0;
      if (p < P && t < M) {
        const int8_t* sp = a.spins + ((int64_t)r * M + t) * n;
        uint64_t d = 0;
#pragma unroll
        for (int m = 0; m < NS; ++m) {
          const int j = 64 * m + lane;
          if (j < n && sp[j] < 0) d |= 1ull << m;
        }
        down[2 * pp + par] = d;
#pragma unroll
        for (int m = 0; m < NS; ++m) {
          const uint64_t bal = ballot64((d >> m) & 1ull);
          if (lane == 0) s_down[t * NS + m] = bal;
        }
        init_fields_seq<FT, FT, NS>(F[2 * pp + par], J, h, d, n, lane);
      }
    }
  }
  __syncthreads();

  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  constexpr int NQ = (NS + 3) / 4;
  const int upper = (M & 1) ? M - 1 : M;

  uint64_t nflip = 0;  // wave-uniform: one row per flip (one wave per slice)
  for (int tt = 0; tt < a.T; ++tt) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)tt);
    const FT at = sweep_param<TT>(a, tt);
    const FT Bt = sweep_param1<TT>(a, tt);
    // phase 0: even slices < upper; phase 1: odd slices; phase 2: slice M-1 (M odd)
    for (int phase = 0; phase < 3; ++phase) {
      if (phase == 2 && !(M & 1)) break;
#pragma unroll
      for (int pp = 0; pp < PPW; ++pp) {
#pragma unroll
        for (int par = 0; par < 2; ++par) {
          const int p = w + pp * W;
          const int t = 2 * p + par;
          bool act;
          if (phase == 0) act = (par == 0) && (t < upper);
          else if (phase == 1) act = (par == 1) && (t < M);
          else act = (par == 0) && (t == M - 1);
          if (p >= P || !act) continue;
          const int tu = (t + 1) % M, td = (t + M - 1) % M;
          FieldVec<FT, NS>& Ft = F[2 * pp + par];
          uint64_t& dw = down[2 * pp + par];
          for (int Q = 0; Q < NQ; ++Q) {
            uint32_t wd[4];
            slot_words(slo, shi, tg, rep, (uint32_t)t, Q, lane, wd);
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
              const int k = 4 * Q + jj;
              if (k >= NS) break;
              const int i = 64 * k + lane;
              const TT th = (i < n) ? -log_uniform<TT>(wd[jj]) : neg_inf<TT>();
              uint64_t live = ballot64(i < n);
              uint64_t sb = ballot64((dw >> k) & 1ull);
              const uint64_t bu = s_down[tu * NS + k], bd = s_down[td * NS + k];
              // s_up + s_dn = 2 - 2(u + d)
              const int nb2 = 2 - 2 * (int)(((bu >> lane) & 1ull) + ((bd >> lane) & 1ull));
              uint64_t flips = 0;
              // (64 slots, or 8 slice pairs per wave: the first form moves
              // the fields to scratch, so those take the second, with per-flip
              // updates of sb, dw and the count.)
              if constexpr (NS <= 32 && PPW < 8) {
                // A lane leaves `live` when it flips (and so do the lanes
                // below a flip), so every live lane keeps its own spin, and
                // with it the Trotter term, for the whole slot: s2 and
                // xs = B dEtr are formed once per slot, dEcl = s2 F per
                // attempt.  The chain runs on the scalar unit, which the
                // waves of a SIMD share: the lowest candidate and the lanes
                // at or below it come from one subtraction (below = bits
                // 0..i0), the flip's coefficient is the flipping lane's own
                // s2 (its spin before the flip), and the slot's spin word and
                // flip count are formed once per slot.
                const bool dn = (dw >> k) & 1ull;
                const FT s2 = dn ? (FT)2 : (FT)(-2);
                const FT xs = Bt * (FT)(dn ? 2 * nb2 : -2 * nb2);
                auto candidates = [&]() -> uint64_t {
                  return ballot64(fma_ft<FT>(at, s2 * Ft.dyn(k), xs) < th) & live;
                };
                uint64_t cand = candidates();
                while (cand != 0) {
                  const uint64_t below = cand ^ (cand - 1);
                  const int i0 = __builtin_ctzll(cand);
                  live &= ~below;
                  flips |= cand & below;
                  const FT c = lane_value(s2, i0);
                  const int row = 64 * k + i0;
                  if constexpr (SL > 0) {
                    const FT* pg = reinterpret_cast<const FT*>(J + (int64_t)row * RB);
                    FT x[NS];
#pragma unroll
                    for (int m = 0; m < NS; ++m)
                      x[m] = (m < SL) ? sJ[(m * n + row) * 64 + lane] : pg[m * 64 + lane];
#pragma unroll
                    for (int m = 0; m < NS; ++m) Ft.set(m, fma_ft<FT>(c, x[m], Ft.get(m)));
                  } else {
                    row_axpy<FT, FT, NS>(Ft, J + (int64_t)row * RB, lane, c);
                  }
                  cand = candidates();
                }
              } else {
                while (true) {
                  const FT fk = Ft.dyn(k);
                  const bool dl = (sb >> lane) & 1ull;
                  const FT dEcl = dl ? (FT)2 * fk : (FT)(-2) * fk;
                  const FT dEtr = (FT)(dl ? 2 * nb2 : -2 * nb2);
                  const FT dE = fma_ft<FT>(at, dEcl, Bt * dEtr);
                  const uint64_t cand = ballot64(dE < th) & live;
                  if (cand == 0) break;
                  const int i0 = __builtin_ctzll(cand);
                  live &= lanes_above(i0);
                  const FT c = ((sb >> i0) & 1ull) ? (FT)2 : (FT)(-2);
                  sb ^= 1ull << i0;
                  if (lane == i0) dw ^= 1ull << k;
                  row_axpy<FT, FT, NS>(Ft, J + (int64_t)(64 * k + i0) * RB, lane, c);
                  ++nflip;
                }
              }
              sb ^= flips;
              nflip += (uint64_t)__builtin_popcountll(flips);
              dw ^= ((flips >> lane) & 1ull) << k;
              if (lane == 0) s_down[t * NS + k] = sb;
            }
          }
        }
      }
      __syncthreads();
    }
  }

  if (a.stats && lane == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nflip);
  }
#pragma unroll
  for (int pp = 0; pp < PPW; ++pp) {
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      const int p = w + pp * W;
      const int t = 2 * p + par;
      if (p < P && t < M) {
        int8_t* sp = a.spins + ((int64_t)r * M + t) * n;
#pragma unroll
        for (int m = 0; m < NS; ++m) {
          const int j = 64 * m + lane;
          if (j < n) sp[j] = ((down[2 * pp + par] >> m) & 1ull) ? (int8_t)-1 : (int8_t)1;
        }
      }
    }
  }
}

template <typename FT, int NS, int PPW, int MAXT, int SL>
inline hipError_t launch_sqa_one(const DenseArgs& a, int W, size_t lds, hipStream_t st) {
  auto kern = k_sqa_wave<FT, NS, PPW, MAXT, SL>;
  if constexpr (SL > 0) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, dim3(a.R), dim3(64 * W), lds, st, a);
  return hipGetLastError();
}

// W > 4 runs the 1024-thread variant (<= 128 VGPRs), W <= 4 the 256-thread one.
template <typename FT, int NS, int SL>
inline hipError_t launch_sqa_form(const DenseArgs& a, int PPW, int W, hipStream_t st) {
  const size_t lds = (size_t)a.M * NS * sizeof(uint64_t) + (size_t)SL * a.n * 64 * sizeof(FT);
  if (W > 4) {
    switch (PPW) {
      case 1: return launch_sqa_one<FT, NS, 1, 1024, SL>(a, W, lds, st);
      case 2: return launch_sqa_one<FT, NS, 2, 1024, SL>(a, W, lds, st);
      default: return hipErrorInvalidValue;
    }
  }
  switch (PPW) {
    case 1: return launch_sqa_one<FT, NS, 1, 256, SL>(a, W, lds, st);
    case 2: return launch_sqa_one<FT, NS, 2, 256, SL>(a, W, lds, st);
    case 4: return launch_sqa_one<FT, NS, 4, 256, SL>(a, W, lds, st);
    case 8: return launch_sqa_one<FT, NS, 8, 256, SL>(a, W, lds, st);
    default: return hipErrorInvalidValue;
  }
}

// Slots of every row that the LDS can hold for the largest system of NS
// slots (64 NS rows): built only where they fit 160 KiB.
template <typename FT, int NS, int SL>
constexpr bool sqa_slots_fit() {
  return SL <= NS && (size_t)SL * 64 * NS * 64 * sizeof(FT) <= 160 * 1024;
}

// a.cache_rows: slots of every coupling row held in LDS (0, 1, 2 or 4)
template <typename FT, int NS>
inline hipError_t launch_sqa_ns(const DenseArgs& a, int PPW, int W, hipStream_t st) {
  const int sl = a.cache_rows;
  if (sl == 0) return launch_sqa_form<FT, NS, 0>(a, PPW, W, st);
  if constexpr (sqa_slots_fit<FT, NS, 1>()) {
    if (sl == 1) return launch_sqa_form<FT, NS, 1>(a, PPW, W, st);
  }
  if constexpr (sqa_slots_fit<FT, NS, 2>()) {
    if (sl == 2) return launch_sqa_form<FT, NS, 2>(a, PPW, W, st);
  }
  if constexpr (sqa_slots_fit<FT, NS, 4>()) {
    if (sl == 4) return launch_sqa_form<FT, NS, 4>(a, PPW, W, st);
  }
  return hipErrorInvalidValue;
}

// launch_sqa_<ft>_<range>: NS in {1, 2, 4, 8} (small) or {16, 32} (large)
hipError_t launch_sqa_f32_small(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st);
hipError_t launch_sqa_f32_large(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st);
hipError_t launch_sqa_f64_small(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st);
hipError_t launch_sqa_f64_large(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st);
// NS = 64: FP64 with one slice pair per wave, FP32 with one or two
hipError_t launch_sqa_wide(int ftype, const DenseArgs& a, int PPW, int W, hipStream_t st);

}  // namespace ojmi
