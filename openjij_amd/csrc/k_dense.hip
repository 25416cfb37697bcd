// k_dense.hip -- dense single-spin-flip sweeps, one wave64 per replica
// (classical) or per Trotter slice (transverse), fields in VGPRs.
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

// ---------------------------------------------------------------------------
// Classical: one wave anneals one replica through T sweeps.  `rep` and `slice`
// are the replica and fourth counter words of its random stream; p0[t] is the
// sweep's threshold factor -1/beta.
template <typename FT, typename JT, int NS>
__device__ __forceinline__ void anneal_replica(const char* J, const FT* h, int8_t* sp, int n,
                                               const FT* F_in, FT* F_out, uint32_t rep,
                                               uint32_t slice, uint64_t seed, uint64_t t0,
                                               int T, const typename ThreshT<FT>::type* p0,
                                               int rule, int lane) {
  typedef typename ThreshT<FT>::type TT;
  constexpr int64_t RB = row_bytes<JT, NS>();
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
    init_fields_seq<JT, FT, NS>(F, J, h, down, n, lane);
  }

  const uint32_t slo = (uint32_t)seed, shi = (uint32_t)(seed >> 32);
  constexpr int NQ = (NS + 3) / 4;
  for (int t = 0; t < T; ++t) {
    const uint32_t tg = (uint32_t)(t0 + (uint64_t)t);
    const TT nb = sweep_scalar(p0[t]);
    for (int Q = 0; Q < NQ; ++Q) {
      uint32_t w[4];
      slot_words(slo, shi, tg, rep, slice, Q, lane, w);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int k = 4 * Q + jj;
        if (k >= NS) break;
        const int i = 64 * k + lane;
        const TT th = (i < n) ? accept_log<TT>(w[jj], rule) * nb : neg_inf<TT>();
        uint64_t live = ballot64(i < n);
        // a lane is never tested again once it has flipped, so every live
        // lane keeps its sign within the slot: dE = s2 * F
        const uint64_t sb = ballot64((down >> k) & 1ull);
        const FT s2 = ((down >> k) & 1ull) ? (FT)2 : (FT)(-2);
        uint64_t flips = 0;
        while (true) {
          const FT fk = F.dyn(k);
          const FT dE = s2 * fk;
          const uint64_t cand = ballot64((TT)dE < th) & live;
          if (cand == 0) break;
          const int i0 = __builtin_ctzll(cand);
          live &= lanes_above(i0);
          const FT c = ((sb >> i0) & 1ull) ? (FT)2 : (FT)(-2);
          flips |= 1ull << i0;
          row_axpy<JT, FT, NS>(F, J + (int64_t)(64 * k + i0) * RB, lane, c);
        }
        down ^= ((flips >> lane) & 1ull) << k;
      }
    }
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

// One wave per replica of one problem.
template <typename FT, typename JT, int NS>
__global__ __launch_bounds__(256) void k_dense_wave(DenseArgs a) {
  typedef typename ThreshT<FT>::type TT;
  const int lane = threadIdx.x & 63;
  const int r = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (r >= a.R) return;
  const FT* F_in = a.F_in ? reinterpret_cast<const FT*>(a.F_in) + (int64_t)r * (64 * NS) : nullptr;
  FT* F_out = a.F_out ? reinterpret_cast<FT*>(a.F_out) + (int64_t)r * (64 * NS) : nullptr;
  anneal_replica<FT, JT, NS>(reinterpret_cast<const char*>(a.J), reinterpret_cast<const FT*>(a.h),
                             a.spins + (int64_t)r * a.n, a.n, F_in, F_out,
                             (uint32_t)(a.rep_offset + r), 0u, a.seed, a.t0, a.T,
                             p0_ptr<TT>(a.p0_d, a.p0_f), a.rule, lane);
}

// Batch of independent problems of different sizes (the same slot class NS):
// wave g anneals replica g % R of problem inst[g / R], whose couplings,
// fields, spins and per-sweep schedule start at its own offsets.  The problem
// index is the fourth word of the random counter.
// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD
// b % 8).  This bijection of [0, nb) keeps runs of `chunk` consecutive
// logical workgroups -- the replicas of one problem -- on one XCD, so they
// share its L2, while consecutive runs still rotate over the XCDs (problems of
// similar size spread evenly).  The last incomplete round keeps b.
__device__ __forceinline__ int xcd_chunked(int b, int nb, int chunk) {
  const int round = 8 * chunk;
  if (b >= (nb / round) * round) return b;
  const int xcd = b & 7, idx = b >> 3;
  return (idx / chunk) * round + xcd * chunk + (idx % chunk);
}

template <typename FT, typename JT, int NS>
__global__ __launch_bounds__(256) void k_batch_wave(BatchArgs b) {
  typedef typename ThreshT<FT>::type TT;
  const int lane = threadIdx.x & 63;
  const int wg = xcd_chunked(blockIdx.x, gridDim.x, (b.R + 3) >> 2);
  const int g = __builtin_amdgcn_readfirstlane(wg * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (g >= b.num_items * b.R) return;
  const int k = b.inst[g / b.R];
  const int r = g - (g / b.R) * b.R;
  const int n = b.n[k];
  const char* J = reinterpret_cast<const char*>(b.J) + b.j_off[k];
  const FT* h = reinterpret_cast<const FT*>(b.h) + b.h_off[k];
  int8_t* sp = b.spins + b.s_off[k] + (int64_t)r * n;
  const TT* p0 = p0_ptr<TT>(b.p0_d, b.p0_f) + (int64_t)k * b.T;
  anneal_replica<FT, JT, NS>(J, h, sp, n, nullptr, nullptr, (uint32_t)r, (uint32_t)k, b.seed,
                             b.t0, b.T, p0, b.rule, lane);
}

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
template <typename FT, int NS, int PPW, int MAXT>
__global__ __launch_bounds__(MAXT) void k_sqa_wave(DenseArgs a) {
  typedef FT TT;
  extern __shared__ uint64_t s_down[];  // [M][NS]
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = blockDim.x >> 6;
  const int r = blockIdx.x;
  const int n = a.n, M = a.M;
  const int P = (M + 1) / 2;
  const char* J = reinterpret_cast<const char*>(a.J);
  constexpr int64_t RB = row_bytes<FT, NS>();
  const FT* h = reinterpret_cast<const FT*>(a.h);

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
              while (true) {
                const FT fk = Ft.dyn(k);
                const bool dn = (sb >> lane) & 1ull;
                const FT dEcl = dn ? (FT)2 * fk : (FT)(-2) * fk;
                const FT dEtr = (FT)(dn ? 2 * nb2 : -2 * nb2);
                const FT x = Bt * dEtr;
                const FT dE = fma_ft<FT>(at, dEcl, x);
                const uint64_t cand = ballot64(dE < th) & live;
                if (cand == 0) break;
                const int i0 = __builtin_ctzll(cand);
                live &= lanes_above(i0);
                const FT c = ((sb >> i0) & 1ull) ? (FT)2 : (FT)(-2);
                sb ^= 1ull << i0;
                if (lane == i0) dw ^= 1ull << k;
                row_axpy<FT, FT, NS>(Ft, J + (int64_t)(64 * k + i0) * RB, lane, c);
              }
              if (lane == 0) s_down[t * NS + k] = sb;
            }
          }
        }
      }
      __syncthreads();
    }
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

// ---------------------------------------------------------------------------
// launchers

template <typename FT, typename JT>
static hipError_t launch_dense_ft(const DenseArgs& a, int NS, hipStream_t st) {
  const int wpb = 4;
  dim3 grid((a.R + wpb - 1) / wpb), block(64 * wpb);
  switch (NS) {
#define OJMI_CASE(ns)                                                          \
  case ns:                                                                     \
    hipLaunchKernelGGL((k_dense_wave<FT, JT, ns>), grid, block, 0, st, a);     \
    break;
    OJMI_CASE(1)
    OJMI_CASE(2)
    OJMI_CASE(4)
    OJMI_CASE(8)
    OJMI_CASE(16)
    OJMI_CASE(32)
    OJMI_CASE(64)
#undef OJMI_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename FT, typename JT>
static hipError_t launch_batch_ft(const BatchArgs& b, int NS, hipStream_t st) {
  const int wpb = 4;
  const int64_t waves = (int64_t)b.num_items * b.R;
  dim3 grid((unsigned)((waves + wpb - 1) / wpb)), block(64 * wpb);
  switch (NS) {
#define OJMI_CASE(ns)                                                          \
  case ns:                                                                     \
    hipLaunchKernelGGL((k_batch_wave<FT, JT, ns>), grid, block, 0, st, b);     \
    break;
    OJMI_CASE(1)
    OJMI_CASE(2)
    OJMI_CASE(4)
    OJMI_CASE(8)
    OJMI_CASE(16)
    OJMI_CASE(32)
    OJMI_CASE(64)
#undef OJMI_CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_batch_wave(int ftype, int jtype, const BatchArgs& b, int NS, hipStream_t st) {
  if (b.num_items <= 0 || b.R <= 0 || b.T <= 0) return hipSuccess;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_batch_ft<double, double>(b, NS, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_batch_ft<float, float>(b, NS, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_batch_ft<int32_t, int8_t>(b, NS, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_batch_ft<int32_t, int32_t>(b, NS, st);
  return hipErrorInvalidValue;
}

hipError_t launch_dense_wave(int ftype, int jtype, const DenseArgs& a, int NS, hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_dense_ft<double, double>(a, NS, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_dense_ft<float, float>(a, NS, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_dense_ft<int32_t, int8_t>(a, NS, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_dense_ft<int32_t, int32_t>(a, NS, st);
  return hipErrorInvalidValue;
}

template <typename FT, int NS>
static hipError_t launch_sqa_ns(const DenseArgs& a, int PPW, int W, hipStream_t st) {
  dim3 grid(a.R), block(64 * W);
  const size_t lds = (size_t)a.M * NS * sizeof(uint64_t);
  // W > 4 runs the 1024-thread variant (<= 128 VGPRs), W <= 4 the 256-thread one.
  if (W > 4) {
    switch (PPW) {
      case 1:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 1, 1024>), grid, block, lds, st, a);
        break;
      case 2:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 2, 1024>), grid, block, lds, st, a);
        break;
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (PPW) {
      case 1:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 1, 256>), grid, block, lds, st, a);
        break;
      case 2:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 2, 256>), grid, block, lds, st, a);
        break;
      case 4:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 4, 256>), grid, block, lds, st, a);
        break;
      case 8:
        hipLaunchKernelGGL((k_sqa_wave<FT, NS, 8, 256>), grid, block, lds, st, a);
        break;
      default:
        return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

template <typename FT>
static hipError_t launch_sqa_ft(const DenseArgs& a, int NS, int PPW, int W, hipStream_t st) {
  switch (NS) {
    case 1: return launch_sqa_ns<FT, 1>(a, PPW, W, st);
    case 2: return launch_sqa_ns<FT, 2>(a, PPW, W, st);
    case 4: return launch_sqa_ns<FT, 4>(a, PPW, W, st);
    case 8: return launch_sqa_ns<FT, 8>(a, PPW, W, st);
    case 16: return launch_sqa_ns<FT, 16>(a, PPW, W, st);
    case 32: return launch_sqa_ns<FT, 32>(a, PPW, W, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sqa_wave(int ftype, const DenseArgs& a, int NS, int PPW, int W,
                           hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (ftype == FT_F64) return launch_sqa_ft<double>(a, NS, PPW, W, st);
  if (ftype == FT_F32) return launch_sqa_ft<float>(a, NS, PPW, W, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
