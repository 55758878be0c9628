// Process-wide state of libngsamd: device, stream and the small reduction
// workspace (per-workgroup partial sums, device scalars).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace ngsamd {

constexpr int kMaxReduceBlocks = 4096;  // upper bound on workgroups of a reducing kernel
// wide-grid kernels (the CSR operator's CG variant) write up to this many
// partial sums, folded to kFoldedPartials before CG's kernels read them
constexpr int kMaxWideBlocks = 16384;
constexpr int kFoldedPartials = 64;
constexpr int kNumScalars = 64;

struct RuntimeState {
  std::mutex mutex;
  bool initialised = false;
  int device = -1;
  hipStream_t stream = nullptr;       // library stream
  hipStream_t user_stream = nullptr;  // stream set by ngsamd_set_stream
  double* partials = nullptr;         // 2 x kMaxReduceBlocks partial sums
  double* wide_partials = nullptr;    // kMaxWideBlocks partial sums of wide-grid kernels
  double* dev_scalars = nullptr;      // kNumScalars
  double* host_scalars = nullptr;     // pinned mirror
};

RuntimeState& State();

}  // namespace ngsamd
