// Wavefront / workgroup reductions and the last-workgroup finalisation used by
// the dot products and the fused CG kernels.  Sums are formed in a fixed
// order for a fixed launch shape, so results are bit-reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>

namespace ngsamd {

__device__ __forceinline__ double WaveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ float WaveSumF(float v) {
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

// Publish this workgroup's partial sums (NV values) and return true in the
// workgroup that finishes last.  That workgroup then reads all partials
// (written by every workgroup before it) and resets the counter to zero, so
// the next launch (or graph replay) starts from a clean state.
template <int NV>
__device__ __forceinline__ bool PublishPartials(const double* vals, double* partials, unsigned* counter,
                                                bool* s_last) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) partials[(size_t)k * gridDim.x + blockIdx.x] = vals[k];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    *s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = *s_last;
  if (last) __threadfence();  // acquire: see every other workgroup's partials
  return last;
}

// In the last workgroup: deterministic sum of the NV partial arrays.
template <int BLOCK, int NV>
__device__ __forceinline__ void ReducePartials(const double* partials, double* out, double* lds) {
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double v = 0.0;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += BLOCK)
      v += partials[(size_t)k * gridDim.x + i];
    v = BlockSum<BLOCK>(v, lds);
    if (threadIdx.x == 0) out[k] = v;
  }
}

}  // namespace ngsamd
