// ojmi_persist.h -- the persistent batch kernel (k_batch_persist.hip): every
// (problem, replica) of a batch in one launch, pulled from per-XCD queues.
#pragma once
#include "ojmi_kernels.h"

namespace ojmi {

// Work queues: the batch's (problem, replica) items split into 8 queues, one
// per XCD.  Whole problems are assigned greedily (largest first, each to the
// queue with the least work so far: longest-processing-time order), their R
// replicas consecutive, so that a problem's replicas run on one XCD and share
// its L2; each queue lists problems in decreasing size.  A wave pulls items
// from its own XCD's queue and, once that is drained, from the others.
struct BatchQueues {
  const int32_t* items;  // [K * R] problem * R + replica
  const int32_t* begin;  // [9]: queue q holds items[begin[q] .. begin[q + 1])
  uint32_t* head;        // [8] next unclaimed position per queue, zero before a launch
};

// Largest slot class (64-spin slots, a power of two) a persistent launch
// handles: 16 (n <= 1024, fields of few registers) or 64 (n <= 4096).
constexpr int PERSIST_NS_SMALL = 16;

// Workgroups that can be resident at once (occupancy x CUs) for the
// persistent kernel of this type pair and slot bound, at `waves` waves per
// workgroup; 0 on error.
int batch_persist_blocks(int ftype, int jtype, int ns_max, int waves);
hipError_t launch_batch_persist(int ftype, int jtype, const BatchArgs& b, const BatchQueues& q,
                                int ns_max, int waves, int blocks, hipStream_t st);

}  // namespace ojmi
