// Deterministic FP64 tally reduction (host interface).
//
// The reference scores each history into a per-history map and adds the
// map to its thread's totals when the history ends, then merges the
// threads' totals serially (ScorableProxy::CommitHistory, Scorable.cpp:
// 101-106; FixedSource.cpp:30-33): S_b += s_{h,b} and S2_b += s_{h,b}^2, where
// s_{h,b} is history h's total score in bin b.
//
// Here every score is a record (key = bin << 32 | history, value).  The
// records of a batch arrive in the order of the GPU's schedule, so they are
// put in a fixed order first: two stable radix sorts order them by key and,
// inside one (bin, history), by the value's bit pattern.  s_{h,b} is then the
// sequential sum of one run, and each bin's S and S2 contributions are summed
// over its runs by one fixed reduction tree.  No floating-point atomics are
// involved, so the result depends on the scores only, not on the launch
// schedule, the bank layout or the thread timing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmc {

struct TallyWork {
  uint32_t cap = 0;            // records the buffers hold
  uint64_t* key_b = nullptr;   // sort buffers: keys and value bits, twice
  uint64_t* val_a = nullptr;
  uint64_t* val_b = nullptr;
  uint64_t* key_c = nullptr;
  double* run_s = nullptr;     // per record: its run's sum at the run head, else 0
  void* temp = nullptr;        // radix sort storage
  size_t temp_bytes = 0;
};

// Sizes the work buffers for `cap` records (no-op when they are large enough).
int TallyWorkEnsure(TallyWork& w, uint32_t cap);
void TallyWorkFree(TallyWork& w);

// Reduces n records (key, value bits in val; key and val are overwritten)
// over nb bins: sum[b] = sum over histories of s_{h,b}, sq[b] = sum of
// s_{h,b}^2 (sq may be null).  Bins without records get 0.  Asynchronous on s.
int TallyReduce(TallyWork& w, uint64_t* key, uint64_t* val, uint32_t n, int nb,
                double* sum, double* sq, hipStream_t s);

}  // namespace mmc
