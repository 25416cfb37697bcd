// k_batch.hip -- batches of independent dense problems, one wave64 per
// (problem, replica) (the anneal itself is in ojmi_anneal.h).
#include "ojmi_anneal.h"

namespace ojmi {

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

// NS is the slot capacity: a problem of n spins runs with its own slot count
// ns = ceil(n / 64), the layout of its rows, so one launch can hold every
// problem of a batch, largest first.
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
  const int ns = (n + 63) >> 6;
  anneal_replica<FT, JT, NS>(J, h, sp, n, nullptr, nullptr, (uint32_t)r, (uint32_t)k, b.seed,
                             b.t0, b.T, p0, b.rule, lane, b.stats, ns);
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

}  // namespace ojmi
