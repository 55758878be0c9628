// Process-wide state of libngsamd: device, stream and the small reduction
// workspace (per-workgroup partial sums, device scalars).
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

namespace ngsamd {

constexpr int kMaxReduceBlocks = 4096;  // upper bound on workgroups of a reducing kernel
// grid bound of the matrix-free gather kernel when it also writes partials of
// <x, A x> for CG (CG's update kernel reads all partials in every workgroup)
constexpr int kMaxDotBlocks = 1024;
constexpr int kNumScalars = 64;

struct RuntimeState {
  std::mutex mutex;
  bool initialised = false;
  int device = -1;
  hipStream_t stream = nullptr;       // library stream
  hipStream_t user_stream = nullptr;  // stream set by ngsamd_set_stream
  double* partials = nullptr;         // 2 x kMaxReduceBlocks partial sums
  double* dev_scalars = nullptr;      // kNumScalars
  double* host_scalars = nullptr;     // pinned mirror
};

RuntimeState& State();

}  // namespace ngsamd
