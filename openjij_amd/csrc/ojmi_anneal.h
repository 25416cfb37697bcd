// ojmi_anneal.h -- dense single-spin-flip sweeps with one wave64 per replica
// (classical) or per Trotter slice (transverse), fields in VGPRs: the
// per-replica anneal shared by k_wave.hip, k_batch.hip and k_sqa.hip.
//
// Semantics follow the reference updater
// (include/openjij/updater/single_spin_flip.hpp:67-95 classical, :128-251 Trotter):
// spins are visited in index order 0..n-1 each sweep; a flip of spin i adds
// -2 s_i J[i][.] to every local field.  The only change is the random stream
// (Philox, drawn before dE is known) and the threshold form of the
// Metropolis test; see ojmi_device.h.
//
// Rejection skipping: inside a 64-spin slot all lanes evaluate their own test
// against the current fields; the lowest accepting lane is the next flip in
// sweep order (every lane below it was tested against up-to-date fields and
// rejected), so one ballot replaces a run of rejections.  After the flip only
// lanes above it are re-tested.  The trajectory equals the sequential one.
#pragma once
#include "ojmi_kernels.h"
#include "ojmi_wave.h"

namespace ojmi {

// Threshold type: fp64 for the fp64 and exact-integer paths, fp32 for the
// fp32 path.
template <typename FT>
struct ThreshT {
  typedef double type;
};
template <>
struct ThreshT<float> {
  typedef float type;
};

template <typename TT>
__device__ __forceinline__ TT sweep_param(const DenseArgs& a, int t);
template <>
__device__ __forceinline__ double sweep_param<double>(const DenseArgs& a, int t) {
  return a.p0_d[t];
}
template <>
__device__ __forceinline__ float sweep_param<float>(const DenseArgs& a, int t) {
  return a.p0_f[t];
}
template <typename TT>
__device__ __forceinline__ TT sweep_param1(const DenseArgs& a, int t);
template <>
__device__ __forceinline__ double sweep_param1<double>(const DenseArgs& a, int t) {
  return a.p1_d[t];
}
template <>
__device__ __forceinline__ float sweep_param1<float>(const DenseArgs& a, int t) {
  return a.p1_f[t];
}

// Four slots' worth of uniforms for this lane: w[j] is the word of spin
// 64*(4*Q+j) + lane.  Lane l evaluates counter 16*(4Q + (l&3)) + (l>>2) and a
// quad transpose hands each lane its own words.
__device__ __forceinline__ void slot_words(uint32_t seed_lo, uint32_t seed_hi, uint32_t tg,
                                           uint32_t rep, uint32_t slice, int Q, int lane,
                                           uint32_t w[4]) {
  const uint32_t c0 = (uint32_t)(16 * (4 * Q + (lane & 3)) + (lane >> 2));
  philox4x32_10(c0, tg, rep, slice, seed_lo, seed_hi, w[0], w[1], w[2], w[3]);
  quad_transpose(w[0], w[1], w[2], w[3], lane);
}

// One 64-spin slot k of a classical sweep: every lane tests its spin against
// its threshold th (-inf past n), the lowest accepting lane flips, rows are
// applied, lanes above it are tested again.  Returns the accepted flips.
template <typename FT, typename JT, int NS, typename TT>
__device__ __forceinline__ uint64_t decide_slot(FieldVec<FT, NS>& F, const char* J, int64_t RB,
                                                int k, int n, int lane, TT th, uint64_t& down,
                                                int ns) {
  const int i = 64 * k + lane;
  uint64_t live = ballot64(i < n);
  // a lane is never tested again once it has flipped, so every live
  // lane keeps its sign within the slot: dE = s2 * F
  const FT s2 = ((down >> k) & 1ull) ? (FT)2 : (FT)(-2);
  uint64_t flips = 0;
  // (64 slots: the form below moves the fields to scratch; INT32 fields with
  // int32 couplings and 8 or more slots: it takes up to 30 more registers;
  // those keep the second form)
  constexpr bool kReal = (FT)0.5 != (FT)0;
  constexpr bool kShortChain = NS <= 32 && (kReal || sizeof(JT) == 1 || NS < 8);
  if constexpr (kShortChain) {
    // The chain runs on the scalar unit, which the waves of a SIMD share:
    // the lowest candidate and the lanes at or below it come from one
    // subtraction (below = bits 0..i0).  The flip's coefficient is the
    // flipping lane's own s2, read from that lane for real fields; integer
    // fields take it as a select of +-2 from the sign word, which keeps the
    // row update a 32-bit multiply-add (a read-lane value is an arbitrary
    // integer to the compiler, and the update became 64-bit).
    uint64_t sb = 0;
    if constexpr (!kReal) sb = ballot64((down >> k) & 1ull);
    uint64_t cand = ballot64((TT)(s2 * F.dyn(k)) < th) & live;
    while (cand != 0) {
      const uint64_t below = cand ^ (cand - 1);
      const int i0 = __builtin_ctzll(cand);
      live &= ~below;
      flips |= cand & below;
      FT c;
      if constexpr (kReal) c = lane_value(s2, i0);
      else c = (sb & cand & below) ? (FT)2 : (FT)(-2);
      row_axpy<JT, FT, NS>(F, J + (int64_t)(64 * k + i0) * RB, lane, c, ns);
      cand = ballot64((TT)(s2 * F.dyn(k)) < th) & live;
    }
  } else {
    const uint64_t sb = ballot64((down >> k) & 1ull);
    while (true) {
      const FT fk = F.dyn(k);
      const FT dE = s2 * fk;
      const uint64_t cand = ballot64((TT)dE < th) & live;
      if (cand == 0) break;
      const int i0 = __builtin_ctzll(cand);
      live &= lanes_above(i0);
      const FT c = ((sb >> i0) & 1ull) ? (FT)2 : (FT)(-2);
      flips |= 1ull << i0;
      row_axpy<JT, FT, NS>(F, J + (int64_t)(64 * k + i0) * RB, lane, c, ns);
    }
  }
  down ^= ((flips >> lane) & 1ull) << k;
  return (uint64_t)__builtin_popcountll(flips);
}

// ---------------------------------------------------------------------------
// Classical: one wave anneals one replica through T sweeps.  `rep` and `slice`
// are the replica and fourth counter words of its random stream; p0[t] is the
// sweep's threshold factor -1/beta.
template <typename FT, typename JT, int NS>
__device__ __forceinline__ void anneal_replica(const char* J, const FT* h, int8_t* sp, int n,
                                               const FT* F_in, FT* F_out, uint32_t rep,
                                               uint32_t slice, uint64_t seed, uint64_t t0,
                                               int T, const typename ThreshT<FT>::type* p0,
                                               int rule, int lane,
                                               unsigned long long* stats, int ns = NS) {
  typedef typename ThreshT<FT>::type TT;
  // ns (wave-uniform, <= NS): the problem's slot count; NS is the capacity
  const int64_t RB = (ns == NS) ? row_bytes<JT, NS>() : row_bytes_rt<JT>(ns);
  uint64_t down = 0;
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    const int j = 64 * m + lane;
    if (j < n && sp[j] < 0) down |= 1ull << m;
  }
  FieldVec<FT, NS> F;
  if (F_in) {
#pragma unroll
    for (int m = 0; m < NS; ++m) F.set(m, F_in[64 * m + lane]);
  } else {
    init_fields_seq<JT, FT, NS>(F, J, h, down, n, lane, ns);
  }

  const uint32_t slo = (uint32_t)seed, shi = (uint32_t)(seed >> 32);
  const int NQ = (ns + 3) / 4;
  uint64_t nflip = 0;  // wave-uniform: one coupling row per accepted flip
  for (int t = 0; t < T; ++t) {
    const uint32_t tg = (uint32_t)(t0 + (uint64_t)t);
    const TT nb = sweep_scalar(p0[t]);
    for (int Q = 0; Q < NQ; ++Q) {
      uint32_t w[4];
      slot_words(slo, shi, tg, rep, slice, Q, lane, w);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int k = 4 * Q + jj;
        if (k >= ns) break;
        const int i = 64 * k + lane;
        const TT th = (i < n) ? accept_log<TT>(w[jj], rule) * nb : neg_inf<TT>();
        nflip += decide_slot<FT, JT, NS, TT>(F, J, RB, k, n, lane, th, down, ns);
      }
    }
  }
  if (stats && lane == 0) {
    atomicAdd(stats, (unsigned long long)nflip);
    atomicAdd(stats + 1, (unsigned long long)nflip);
  }

#pragma unroll
  for (int m = 0; m < NS; ++m) {
    const int j = 64 * m + lane;
    if (j < n) sp[j] = ((down >> m) & 1ull) ? (int8_t)-1 : (int8_t)1;
  }
  if (F_out) {
#pragma unroll
    for (int m = 0; m < NS; ++m) F_out[64 * m + lane] = F.get(m);
  }
}

template <typename TT>
__device__ __forceinline__ const TT* p0_ptr(const double* d, const float* f);
template <>
__device__ __forceinline__ const double* p0_ptr<double>(const double* d, const float*) {
  return d;
}
template <>
__device__ __forceinline__ const float* p0_ptr<float>(const double*, const float* f) {
  return f;
}

}  // namespace ojmi
