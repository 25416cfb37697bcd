// k_batch_persist.hip -- batches of independent dense problems in one
// persistent launch: every wave repeatedly claims one (problem, replica) item
// from the work queues (ojmi_persist.h) and anneals it with the per-replica
// sweep of ojmi_anneal.h at the problem's slot count, until every queue is
// drained.  Replaces the per-problem / per-read loops of the reference
// sampler (openjij/sampler/sampler.py:115-140) with one dispatch; a
// replica's trajectory is the one k_batch_wave computes (problem index as
// the fourth counter word), whatever wave runs it.
#include "ojmi_anneal.h"
#include "ojmi_persist.h"

namespace ojmi {

namespace {

__device__ __forceinline__ int slots_of(int n) {
  int ns = 1;
  while (64 * ns < n) ns <<= 1;
  return ns;
}

template <typename FT, typename JT, int NS>
__device__ __forceinline__ void run_item(const BatchArgs& b, int k, int r, int lane) {
  typedef typename ThreshT<FT>::type TT;
  const int n = b.n[k];
  const char* J = reinterpret_cast<const char*>(b.J) + b.j_off[k];
  const FT* h = reinterpret_cast<const FT*>(b.h) + b.h_off[k];
  int8_t* sp = b.spins + b.s_off[k] + (int64_t)r * n;
  const TT* p0 = p0_ptr<TT>(b.p0_d, b.p0_f) + (int64_t)k * b.T;
  anneal_replica<FT, JT, NS>(J, h, sp, n, nullptr, nullptr, (uint32_t)r, (uint32_t)k, b.seed,
                             b.t0, b.T, p0, b.rule, lane, b.stats, (n + 63) >> 6);
}

}  // namespace

template <typename FT, typename JT, int NSMAX>
__global__ __launch_bounds__(256) void k_batch_persist(BatchArgs b, BatchQueues q) {
  const int lane = threadIdx.x & 63;
  const int home = blockIdx.x & 7;  // workgroups are dispatched round-robin over the 8 XCDs
  for (int qi = 0; qi < 8;) {
    const int qq = (home + qi) & 7;
    uint32_t pos = 0;
    if (lane == 0) pos = atomicAdd(q.head + qq, 1u);
    pos = (uint32_t)__builtin_amdgcn_readfirstlane((int)pos);
    const int64_t idx = (int64_t)q.begin[qq] + pos;
    if (idx >= q.begin[qq + 1]) {  // this queue is drained: move on to the next
      ++qi;
      continue;
    }
    const int item = __builtin_amdgcn_readfirstlane(q.items[idx]);
    const int k = item / b.R, r = item - k * b.R;
    switch (slots_of(b.n[k])) {
      case 1: run_item<FT, JT, 1>(b, k, r, lane); break;
      case 2: run_item<FT, JT, 2>(b, k, r, lane); break;
      case 4: run_item<FT, JT, 4>(b, k, r, lane); break;
      case 8: run_item<FT, JT, 8>(b, k, r, lane); break;
      case 16: run_item<FT, JT, 16>(b, k, r, lane); break;
      case 32: if constexpr (NSMAX >= 32) run_item<FT, JT, 32>(b, k, r, lane); break;
      case 64: if constexpr (NSMAX >= 64) run_item<FT, JT, 64>(b, k, r, lane); break;
      default: break;
    }
  }
}

namespace {

template <typename FT, typename JT, int NSMAX>
const void* persist_kernel() {
  return reinterpret_cast<const void*>(k_batch_persist<FT, JT, NSMAX>);
}

const void* pick(int ftype, int jtype, int ns_max) {
  const bool small = ns_max <= PERSIST_NS_SMALL;
#define OJMI_PK(F, J)                                                                   \
  return small ? persist_kernel<F, J, PERSIST_NS_SMALL>() : persist_kernel<F, J, 64>();
  if (ftype == FT_F64 && jtype == JT_F64) OJMI_PK(double, double)
  if (ftype == FT_F32 && jtype == JT_F32) OJMI_PK(float, float)
  if (ftype == FT_I32 && jtype == JT_I8) OJMI_PK(int32_t, int8_t)
  if (ftype == FT_I32 && jtype == JT_I32) OJMI_PK(int32_t, int32_t)
#undef OJMI_PK
  return nullptr;
}

}  // namespace

int batch_persist_blocks(int ftype, int jtype, int ns_max, int waves) {
  const void* kern = pick(ftype, jtype, ns_max);
  if (!kern || waves < 1 || waves > 4) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * waves, 0) != hipSuccess)
    return 0;
  return per_cu * cus;
}

hipError_t launch_batch_persist(int ftype, int jtype, const BatchArgs& b, const BatchQueues& q,
                                int ns_max, int waves, int blocks, hipStream_t st) {
  if (b.R <= 0 || b.T <= 0 || blocks <= 0) return hipSuccess;
  const void* kern = pick(ftype, jtype, ns_max);
  if (!kern || waves < 1 || waves > 4) return hipErrorInvalidValue;
  void* args[] = {const_cast<BatchArgs*>(&b), const_cast<BatchQueues*>(&q)};
  return hipLaunchKernel(kern, dim3(blocks), dim3(64 * waves), args, 0, st);
}

}  // namespace ojmi
