// Process-wide state of libngsamd: device, stream and the small reduction
// workspace (per-workgroup partials, completion counters, device scalars).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace ngsamd {

constexpr int kMaxReduceBlocks = 4096;  // upper bound on workgroups of a reducing kernel
constexpr int kMaxPartialVals = 4;      // partial sums per workgroup
constexpr int kNumCounters = 16;        // completion counters (one per concurrent reducing kernel kind)
constexpr int kNumScalars = 64;

enum CounterSlot { kCounterDot = 0, kCounterCG1 = 1, kCounterCG2 = 2, kCounterMisc = 3 };

struct RuntimeState {
  std::mutex mutex;
  bool initialised = false;
  int device = -1;
  hipStream_t stream = nullptr;       // library stream
  hipStream_t user_stream = nullptr;  // stream set by ngsamd_set_stream
  double* partials = nullptr;         // kMaxReduceBlocks * kMaxPartialVals
  unsigned* counters = nullptr;       // kNumCounters, zero between launches
  double* dev_scalars = nullptr;      // kNumScalars
  double* host_scalars = nullptr;     // pinned mirror
};

RuntimeState& State();

}  // namespace ngsamd
