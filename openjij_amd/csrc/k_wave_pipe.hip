// k_wave_pipe.hip -- the one-wave-per-replica anneal with its thresholds
// made ahead by helper waves, for few replicas of a small system, where each
// replica's sweep is a chain of dependent decisions and the GPU has spare
// waves.
//
// A workgroup anneals one replica: wave 0 decides (fields in its VGPRs, as in
// anneal_replica), waves 1..H (H <= 3) make thresholds.  Sweeps run in phases of
// PIPE_P; while wave 0 runs phase p with thresholds read from LDS, the helpers
// compute phase p+1's (the same Philox words and acceptance logarithm as
// anneal_replica, so the trajectory is the same bit for bit) into the other
// half of a two-phase ring.  One barrier per phase.  Coupling rows are read
// from LDS when they fit (copied once per workgroup), else from L2.
#include "ojmi_anneal.h"

namespace ojmi {

namespace {

constexpr int PIPE_P = 4;  // sweeps per phase

template <typename FT, typename JT, int NS, int H, bool ROWS_LDS>
__global__ __launch_bounds__(64 * (H + 1)) void k_dense_wave_pipe(DenseArgs a) {
  typedef typename ThreshT<FT>::type TT;
  constexpr int64_t RB = row_bytes<JT, NS>();
  constexpr int PH = PIPE_P * NS * 64;  // thresholds per phase
  extern __shared__ __align__(16) char smem[];
  TT* sTh = reinterpret_cast<TT*>(smem);          // [2][PIPE_P][NS][64]
  char* sJ = smem + 2 * PH * sizeof(TT);           // [n][RB] when ROWS_LDS
  const int n = a.n;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = blockIdx.x;
  if constexpr (ROWS_LDS) {
    const int64_t words = (int64_t)n * RB / 16;
    const uint4* src = reinterpret_cast<const uint4*>(a.J);
    uint4* dst = reinterpret_cast<uint4*>(sJ);
    for (int64_t e = threadIdx.x; e < words; e += blockDim.x) dst[e] = src[e];
  }
  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  const TT* p0 = p0_ptr<TT>(a.p0_d, a.p0_f);
  const int T = a.T;
  const int nph = (T + PIPE_P - 1) / PIPE_P;

  // helper w-1 fills the sweeps tl = w-1, w-1+H, ... of phase ph
  auto fill = [&](int ph) {
    TT* dst = sTh + (ph & 1) * PH;
    for (int tl = w - 1; tl < PIPE_P; tl += H) {
      const int t = ph * PIPE_P + tl;
      if (t >= T) break;
      const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)t);
      const TT nb = sweep_scalar(p0[t]);
#pragma unroll
      for (int Q = 0; Q < (NS + 3) / 4; ++Q) {
        uint32_t wd[4];
        slot_words(slo, shi, tg, rep, 0u, Q, lane, wd);
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int k = 4 * Q + jj;
          if (k >= NS) break;
          const int i = 64 * k + lane;
          dst[(tl * NS + k) * 64 + lane] =
              (i < n) ? accept_log<TT>(wd[jj], a.rule) * nb : neg_inf<TT>();
        }
      }
    }
  };

  int8_t* sp = a.spins + (int64_t)r * n;
  FieldVec<FT, NS> F;
  uint64_t down = 0, nflip = 0;
  if (w == 0) {
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      const int j = 64 * m + lane;
      if (j < n && sp[j] < 0) down |= 1ull << m;
    }
    if (a.F_in) {
      const FT* fi = reinterpret_cast<const FT*>(a.F_in) + (int64_t)r * (64 * NS);
#pragma unroll
      for (int m = 0; m < NS; ++m) F.set(m, fi[64 * m + lane]);
    } else {
      // from the global rows: the LDS copy may still be in progress
      init_fields_seq<JT, FT, NS>(F, reinterpret_cast<const char*>(a.J),
                                  reinterpret_cast<const FT*>(a.h), down, n, lane);
    }
  } else {
    fill(0);
  }
  __syncthreads();
  const char* J = ROWS_LDS ? static_cast<const char*>(sJ) : reinterpret_cast<const char*>(a.J);
  for (int ph = 0; ph < nph; ++ph) {
    if (w == 0) {
      const TT* src = sTh + (ph & 1) * PH;
      for (int tl = 0; tl < PIPE_P; ++tl) {
        if (ph * PIPE_P + tl >= T) break;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          const TT th = src[(tl * NS + k) * 64 + lane];
          nflip += decide_slot<FT, JT, NS, TT>(F, J, RB, k, n, lane, th, down, NS);
        }
      }
    } else if (ph + 1 < nph) {
      fill(ph + 1);
    }
    __syncthreads();
  }
  if (w != 0) return;
  if (a.stats && lane == 0) {
    atomicAdd(a.stats, (unsigned long long)nflip);
    atomicAdd(a.stats + 1, (unsigned long long)nflip);
  }
#pragma unroll
  for (int m = 0; m < NS; ++m) {
    const int j = 64 * m + lane;
    if (j < n) sp[j] = ((down >> m) & 1ull) ? (int8_t)-1 : (int8_t)1;
  }
  if (a.F_out) {
    FT* fo = reinterpret_cast<FT*>(a.F_out) + (int64_t)r * (64 * NS);
#pragma unroll
    for (int m = 0; m < NS; ++m) fo[64 * m + lane] = F.get(m);
  }
}

template <typename FT, typename JT, int NS, int H, bool ROWS_LDS>
hipError_t launch_pipe_t(const DenseArgs& a, hipStream_t st) {
  const size_t lds = wave_pipe_lds_bytes(a.n, NS, (int)sizeof(JT),
                                         (int)sizeof(typename ThreshT<FT>::type), ROWS_LDS);
  auto kern = k_dense_wave_pipe<FT, JT, NS, H, ROWS_LDS>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(a.R), dim3(64 * (H + 1)), lds, st, a);
  return hipGetLastError();
}

template <typename FT, typename JT, int NS>
hipError_t launch_pipe_ns(const DenseArgs& a, int H, bool rows_lds, hipStream_t st) {
  if (H == 1) return rows_lds ? launch_pipe_t<FT, JT, NS, 1, true>(a, st)
                              : launch_pipe_t<FT, JT, NS, 1, false>(a, st);
  if (H == 2) return rows_lds ? launch_pipe_t<FT, JT, NS, 2, true>(a, st)
                              : launch_pipe_t<FT, JT, NS, 2, false>(a, st);
  if (H == 3) return rows_lds ? launch_pipe_t<FT, JT, NS, 3, true>(a, st)
                              : launch_pipe_t<FT, JT, NS, 3, false>(a, st);
  return hipErrorInvalidValue;
}

template <typename FT, typename JT>
hipError_t launch_pipe_ft(const DenseArgs& a, int NS, int H, bool rows_lds, hipStream_t st) {
  switch (NS) {
    case 1: return launch_pipe_ns<FT, JT, 1>(a, H, rows_lds, st);
    case 2: return launch_pipe_ns<FT, JT, 2>(a, H, rows_lds, st);
    case 4: return launch_pipe_ns<FT, JT, 4>(a, H, rows_lds, st);
    case 8: return launch_pipe_ns<FT, JT, 8>(a, H, rows_lds, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

size_t wave_pipe_lds_bytes(int n, int NS, int jbytes, int tbytes, bool rows_lds) {
  return (size_t)2 * PIPE_P * NS * 64 * tbytes + (rows_lds ? wave_lds_bytes(n, NS, jbytes) : 0);
}

hipError_t launch_dense_wave_pipe(int ftype, int jtype, const DenseArgs& a, int NS, int H,
                                  bool rows_lds, hipStream_t st) {
  if (a.R <= 0 || a.T < 0) return hipSuccess;
  if (ftype == FT_F64 && jtype == JT_F64) return launch_pipe_ft<double, double>(a, NS, H, rows_lds, st);
  if (ftype == FT_F32 && jtype == JT_F32) return launch_pipe_ft<float, float>(a, NS, H, rows_lds, st);
  if (ftype == FT_I32 && jtype == JT_I8) return launch_pipe_ft<int32_t, int8_t>(a, NS, H, rows_lds, st);
  if (ftype == FT_I32 && jtype == JT_I32) return launch_pipe_ft<int32_t, int32_t>(a, NS, H, rows_lds, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
