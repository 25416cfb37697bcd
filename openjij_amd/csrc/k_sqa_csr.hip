// k_sqa_csr.hip -- Suzuki-Trotter sweeps of sparse (CSR) systems, one
// workgroup per replica, spins in level order.
//
// The reference's Trotter sweep (include/openjij/updater/
// single_spin_flip.hpp:128-251, over transverse_ising.hpp:426-596 for
// CSRSparse) visits the even slices, then the odd slices, then slice M-1 when
// M is odd; within a slice spins 0..n-1 in order, each reading its
// neighbours' current states in the same slice and its own spin in the two
// adjacent slices:
//   dEcl = -2 s_i F_i,  dEtr = -2 s_i (s_up + s_dn),
//   dE = fma(a, dEcl, B dEtr),  flip iff dE < -log(U)
// (a = s beta / M, B = 1/2 log tanh(beta Gamma (1 - s) / M)).  Slices of one
// phase are never adjacent, so they run in parallel, one wave per slice.
// Within a slice, the level order of the classical CSR kernel (k_csr.hip)
// reads exactly the states the sequential sweep reads: levels one after the
// other, the spins of a level 64 at a time.  Spins of every slice live in LDS
// as int8 ([M][n]); fields are recomputed per attempt in stored CSR order from
// the level-order ELL arrays; U comes from the Philox word of (spin, sweep,
// replica, slice), as in the dense Trotter kernel.
#include "ojmi_kernels.h"
#include "ojmi_csr.h"
#include "ojmi_wave.h"

namespace ojmi {

// LDS bytes of the slices' spins, rounded up to 16 for the word regions after them
__host__ __device__ inline size_t sqa_csr_spin_bytes(int n, int M) {
  return ((size_t)n * (size_t)M + 15) & ~(size_t)15;
}

// STAGE: each wave writes the Philox words of its slice's attempts to its own
// LDS region before the slice's passes (one Philox evaluation per 4 spins
// instead of one per spin); else every attempt evaluates its own.
// REC (neighbour lists of <= 16, n < 65536): a pass loads the per-position
// records (CsrRec in ojmi_csr.h, 16-byte chunks) instead of one word per
// column, value, index and field, one pass ahead also at 16 neighbours.
// V8 (with REC): the records hold int8 coupling quadruples and integer h.
template <typename FT, int DMAX, bool STAGE, bool REC, bool V8>
__global__ __launch_bounds__(1024) void k_sqa_csr(CsrArgs a) {
  extern __shared__ int8_t s_spin[];  // [M][n], then (STAGE) words [W][4 ng]
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = blockDim.x >> 6;
  const int r = blockIdx.x;
  const int n = a.n, M = a.M;
  const int ng = (n + 3) >> 2;
  uint32_t* s_w = reinterpret_cast<uint32_t*>(s_spin + sqa_csr_spin_bytes(n, M)) + (size_t)w * 4 * ng;
  int8_t* sp = a.spins + (int64_t)r * M * n;
  for (int e = threadIdx.x; e < M * n; e += blockDim.x) s_spin[e] = sp[e] < 0 ? (int8_t)-1 : (int8_t)1;
  __syncthreads();

  const uint32_t rep = (uint32_t)(a.rep_offset + r);
  const uint32_t slo = (uint32_t)a.seed, shi = (uint32_t)(a.seed >> 32);
  const FT* hv = reinterpret_cast<const FT*>(a.h);
  const FT* ev = reinterpret_cast<const FT*>(a.val);
  const int upper = (M & 1) ? M - 1 : M;
  uint64_t nflip = 0;  // wave-uniform

  for (int tt = 0; tt < a.T; ++tt) {
    const uint32_t tg = (uint32_t)(a.t0 + (uint64_t)tt);
    FT at, Bt;
    if constexpr (sizeof(FT) == 8) {
      at = sweep_scalar(a.p0_d[tt]);
      Bt = sweep_scalar(a.p1_d[tt]);
    } else {
      at = sweep_scalar(a.p0_f[tt]);
      Bt = sweep_scalar(a.p1_f[tt]);
    }
    // phase 0: even slices < upper; 1: odd slices; 2: slice M-1 (M odd)
    for (int phase = 0; phase < 3; ++phase) {
      if (phase == 2 && !(M & 1)) break;
      const int first = phase == 0 ? 0 : (phase == 1 ? 1 : M - 1);
      const int end = phase == 0 ? upper : (phase == 1 ? M : M);
      const int step = phase == 2 ? M : 2;
      for (int t = first + step * w; t < end; t += step * W) {
        int8_t* s = s_spin + (int64_t)t * n;
        const int8_t* su = s_spin + (int64_t)((t + 1) % M) * n;
        const int8_t* sd = s_spin + (int64_t)((t + M - 1) % M) * n;
        if constexpr (STAGE) {
          for (int g = lane; g < ng; g += 64) {
            uint32_t w0, w1, w2, w3;
            philox4x32_10((uint32_t)g, tg, rep, (uint32_t)t, slo, shi, w0, w1, w2, w3);
            s_w[4 * g + 0] = w0;
            s_w[4 * g + 1] = w1;
            s_w[4 * g + 2] = w2;
            s_w[4 * g + 3] = w3;
          }
        }
        // one attempt: field from the neighbour list (stored order), the
        // Trotter term from the adjacent slices, the Philox word's threshold
        // (with STAGE, the word and the spin states are read before the
        // field's gathers, so one LDS round trip serves them all)
        auto word = [&](int i) -> uint32_t {
          if constexpr (STAGE) {
            return s_w[i];
          } else {
            uint32_t u0, u1, u2, u3;
            philox4x32_10((uint32_t)(i >> 2), tg, rep, (uint32_t)t, slo, shi, u0, u1, u2, u3);
            return (i & 2) ? ((i & 1) ? u3 : u2) : ((i & 1) ? u1 : u0);
          }
        };
        auto decide = [&](int i, FT f, int si, int sud, uint32_t u) -> bool {
          const FT dEcl = (FT)(-2 * si) * f;
          const FT dEtr = (FT)(-2 * si * sud);
          const FT x = Bt * dEtr;
          const FT dE = fma_ft<FT>(at, dEcl, x);
          const bool fl = dE < -log_uniform<FT>(u);
          if (fl) s[i] = (int8_t)-si;
          return fl;
        };
        auto attempt = [&](int i, FT f) -> bool {
          return decide(i, f, s[i], (int)su[i] + (int)sd[i], word(i));
        };
        if constexpr (DMAX <= 8 || (REC && DMAX <= 16)) {
          // the next pass's neighbour lists in flight while this one runs
          int lv = 0, base = 0;
          const CsrBufs CB = REC ? CsrBufs{} : csr_bufs<FT>(a);
          CsrPass<FT, DMAX, V8, REC> cur;
          csr_load_pass<FT, DMAX, !REC, V8, false, REC>(a, CB, a.lvl_ptr, 0, 0, lane, cur);
          while (true) {
            const int b1 = __builtin_amdgcn_readfirstlane(a.lvl_ptr[lv + 1]);
            int nlv = lv, nbase = base + 64;
            if (nbase >= b1) {
              nlv = lv + 1;
              nbase = b1;
            }
            const bool last = nlv >= a.nlev;
            CsrPass<FT, DMAX, V8, REC> nxt;
            if (!last)
              csr_load_pass<FT, DMAX, !REC, V8, false, REC>(a, CB, a.lvl_ptr, nlv, nbase, lane,
                                                            nxt);
            bool flipped = false;
            if (cur.valid) {
              const int i = cur.i;
              const int si = s[i], sud = (int)su[i] + (int)sd[i];
              const uint32_t u = word(i);
              flipped = decide(i, csr_field(cur, s), si, sud, u);
            }
            nflip += (uint64_t)__builtin_popcountll(ballot64(flipped));
            if (last) break;
            cur = nxt;
            lv = nlv;
            base = nbase;
          }
        } else {
          for (int lv = 0; lv < a.nlev; ++lv) {
            const int b0 = __builtin_amdgcn_readfirstlane(a.lvl_ptr[lv]);
            const int b1 = __builtin_amdgcn_readfirstlane(a.lvl_ptr[lv + 1]);
            for (int base = b0; base < b1; base += 64) {
              const int pos = base + lane;
              bool flipped = false;
              if (pos < b1) {
                FT f = hv[pos];
                // eight neighbours in flight at a time (a full unroll of the
                // wider lists would not fit the registers of 16 waves per CU)
#pragma unroll 8
                for (int d = 0; d < DMAX; ++d) {
                  const int col = a.ell_col[(int64_t)d * n + pos];
                  const FT v = ev[(int64_t)d * n + pos];
                  f = (s[col] < 0) ? f - v : f + v;
                }
                flipped = attempt(a.order[pos], f);
              }
              nflip += (uint64_t)__builtin_popcountll(ballot64(flipped));
            }
          }
        }
      }
      __syncthreads();
    }
  }
  if (a.stats && lane == 0) atomicAdd(a.stats, (unsigned long long)nflip);
  for (int e = threadIdx.x; e < M * n; e += blockDim.x) sp[e] = s_spin[e];
}

size_t sqa_csr_lds_bytes(int n, int M) { return (size_t)n * (size_t)M; }

int sqa_csr_waves(int M) {
  const int P = (M + 1) / 2;
  return P < 16 ? P : 16;
}

size_t sqa_csr_staged_lds_bytes(int n, int M) {
  return sqa_csr_spin_bytes(n, M) + (size_t)sqa_csr_waves(M) * 16 * (size_t)((n + 3) / 4);
}

template <typename FT, int DMAX>
static hipError_t launch_sqa_csr_d(const CsrArgs& a, bool stage, dim3 grid, dim3 block,
                                   hipStream_t st) {
  const size_t lds = stage ? sqa_csr_staged_lds_bytes(a.n, a.M) : sqa_csr_lds_bytes(a.n, a.M);
  auto kern = stage ? k_sqa_csr<FT, DMAX, true, false, false>
                    : k_sqa_csr<FT, DMAX, false, false, false>;
  if constexpr (DMAX <= 16) {
    if (a.rec && a.pack8)
      kern = stage ? k_sqa_csr<FT, DMAX, true, true, true> : k_sqa_csr<FT, DMAX, false, true, true>;
    else if constexpr (DMAX <= 8 || sizeof(FT) == 4) {  // (FP64 values of 16: registers)
      if (a.rec)
        kern = stage ? k_sqa_csr<FT, DMAX, true, true, false>
                     : k_sqa_csr<FT, DMAX, false, true, false>;
    }
  }
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, grid, block, lds, st, a);
  return hipSuccess;
}

template <typename FT>
static hipError_t launch_sqa_csr_ft(const CsrArgs& a, bool stage, hipStream_t st) {
  const int W = sqa_csr_waves(a.M);
  dim3 grid(a.R), block(64 * W);
  hipError_t e = hipSuccess;
#define OJMI_CASE(dm)                                              \
  case dm:                                                         \
    e = launch_sqa_csr_d<FT, dm>(a, stage, grid, block, st);       \
    if (e != hipSuccess) return e;                                 \
    break;
  switch (a.D) {
    OJMI_CASE(8)
    OJMI_CASE(16)
    OJMI_CASE(32)
    OJMI_CASE(64)
    default:
      return hipErrorInvalidValue;
  }
#undef OJMI_CASE
  return hipGetLastError();
}

hipError_t launch_sqa_csr(int ftype, const CsrArgs& a, bool stage, hipStream_t st) {
  if (a.R <= 0 || a.T <= 0) return hipSuccess;
  if (a.M < 2) return hipErrorInvalidValue;
  if (ftype == FT_F64) return launch_sqa_csr_ft<double>(a, stage, st);
  if (ftype == FT_F32) return launch_sqa_csr_ft<float>(a, stage, st);
  return hipErrorInvalidValue;
}

}  // namespace ojmi
