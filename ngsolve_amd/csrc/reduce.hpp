// Wavefront / workgroup reductions.  Grid-wide sums are two-stage: every
// workgroup stores its partial sum with a plain store, and a one-workgroup
// kernel launched next on the same stream sums the partials in index order.
// The kernel boundary makes the partials visible; no grid-wide fence is
// needed (on a multi-XCD device an agent-scope release writes back L2, which
// is far too expensive to do once per workgroup).  For a fixed launch shape
// the result is bit-reproducible.
#pragma once
#include <hip/hip_runtime.h>

namespace ngsamd {

__device__ __forceinline__ double WaveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a workgroup of BLOCK threads; result valid in thread 0.
// lds must hold BLOCK/64 doubles.
template <int BLOCK>
__device__ __forceinline__ double BlockSum(double v, double* lds) {
  v = WaveSum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) lds[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x < 64) {
    r = (threadIdx.x < BLOCK / 64) ? lds[threadIdx.x] : 0.0;
    r = WaveSum(r);
  }
  __syncthreads();
  return r;
}

// Workgroup sum of v stored to partials[blockIdx.x] (thread 0).
template <int BLOCK>
__device__ __forceinline__ void StorePartial(double v, double* partials, double* lds) {
  v = BlockSum<BLOCK>(v, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// Inside a one-workgroup kernel: deterministic sum of partials[0..count).
// Loads go 4 at a time (one round trip per 4 partials per thread); the
// additions keep the order of the plain strided loop.
template <int BLOCK>
__device__ __forceinline__ double SumPartials(const double* partials, int count, double* lds) {
  double v = 0.0;
  int i = threadIdx.x;
  for (; i + 3 * BLOCK < count; i += 4 * BLOCK) {
    const double a = partials[i], b = partials[i + BLOCK], c = partials[i + 2 * BLOCK], d = partials[i + 3 * BLOCK];
    v += a;
    v += b;
    v += c;
    v += d;
  }
  for (; i < count; i += BLOCK) v += partials[i];
  return BlockSum<BLOCK>(v, lds);
}

// The same sum, returned to every thread of the workgroup.  Every workgroup
// of a grid that calls it on the same partials gets the same value bit for
// bit (the order of the additions depends only on count), so a kernel can
// form a grid-wide scalar redundantly in each workgroup instead of in a
// one-workgroup kernel of its own.  lds must hold BLOCK/64 + 1 doubles.
template <int BLOCK>
__device__ __forceinline__ double SumPartialsAll(const double* partials, int count, double* lds) {
  const double v = SumPartials<BLOCK>(partials, count, lds);
  if (threadIdx.x == 0) lds[BLOCK / 64] = v;
  __syncthreads();
  return lds[BLOCK / 64];
}

}  // namespace ngsamd
