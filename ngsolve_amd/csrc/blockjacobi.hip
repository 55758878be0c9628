// Block-Jacobi preconditioner on the device.
//
// BlockJacobiPrecond (linalg/blockjacobi.cpp): each block's dofs are sorted
// (:390), its diagonal block A[b][b] is taken from the sparse matrix, 0 where
// the pattern has no entry (:404), and inverted on the host (:498); MultAdd
// is y += s sum_b P_b^T inv_b P_b x (:596-632).  The device copy plays the
// part of DevBlockJacobiMatrix (ngscuda/dev_blockjacobi.cpp:92-131), which
// applies the inverses with atomic adds; here the apply is two atomic-free
// kernels: one thread per block row forms z[t] = (inv_b x_b)[r] (row r of
// block b), then one thread per dof sums the z of its block rows in row
// order (blocks may overlap), so every run gives the same result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "dense.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kBjBlock = 256;

// Host threads for the set-up: OMP_NUM_THREADS if set, else the hardware's,
// at most 64.
static int HostThreads() {
  const char* e = std::getenv("OMP_NUM_THREADS");
  int n = e ? std::atoi(e) : (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(n, 64));
}

// z[t] = sum_c inv_b[r][c] x[dof(b, c)] for block row t = row r of block b;
// 4 columns at a time from clamped indices so the loads are in flight
// together.  Inverses are stored column by column, so the rows of one block
// (neighbouring threads) read neighbouring words.
__global__ __launch_bounds__(kBjBlock) void BjRowsKernel(const int64_t* __restrict__ bptr,
                                                         const int64_t* __restrict__ vptr,
                                                         const int* __restrict__ bdofs,
                                                         const int* __restrict__ row_block,
                                                         const double* __restrict__ inv,
                                                         const double* __restrict__ x, double* __restrict__ z,
                                                         int64_t nrows, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t t = blockIdx.x * (int64_t)kBjBlock + threadIdx.x; t < nrows; t += (int64_t)gridDim.x * kBjBlock) {
    const int b = row_block[t];
    const int64_t b0 = bptr[b];
    const int sz = (int)(bptr[b + 1] - b0);
    const double* ir = inv + vptr[b] + (t - b0);
    const int* bd = bdofs + b0;
    double acc = 0.0;
    for (int c = 0; c < sz; c += 4) {
      double a[4], xv[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int cc = min(c + j, sz - 1);
        a[j] = ir[(int64_t)cc * sz];
        xv[j] = x[bd[cc]];
      }
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (c + j < sz) acc += a[j] * xv[j];
    }
    z[t] = acc;
  }
}

// y[d] = (accumulate ? y[d] : 0) + s * sum of z over the block rows of d
__global__ __launch_bounds__(kBjBlock) void BjGatherKernel(const int* __restrict__ gptr,
                                                           const int* __restrict__ grows,
                                                           const double* __restrict__ z, double* __restrict__ y,
                                                           int64_t n, double s, int accumulate,
                                                           const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t d = blockIdx.x * (int64_t)kBjBlock + threadIdx.x; d < n; d += (int64_t)gridDim.x * kBjBlock) {
    double acc = 0.0;
    for (int k = gptr[d]; k < gptr[d + 1]; k++) acc += z[grows[k]];
    y[d] = accumulate ? y[d] + s * acc : s * acc;
  }
}

BlockJacobiOp::BlockJacobiOp(int64_t n, const int64_t* rp, const int* cols, const double* vals, int64_t nb,
                             const int64_t* block_ptr, const int* block_dofs)
    : ngsamd_op_s(OpKind::BlockJacobi) {
  height = width = n;
  nblocks = nb;
  NGSAMD_REQUIRE(n >= 0 && n < (int64_t(1) << 31) && nb >= 0, "ngsamd_blockjacobi_create: bad sizes");
  NGSAMD_REQUIRE(rp[0] == 0, "ngsamd_blockjacobi_create: row pointer must start at 0");
  NGSAMD_REQUIRE(block_ptr[0] == 0, "ngsamd_blockjacobi_create: block pointer must start at 0");
  for (int64_t b = 0; b < nb; b++)
    NGSAMD_REQUIRE(block_ptr[b + 1] >= block_ptr[b], "ngsamd_blockjacobi_create: block pointer not monotone");
  nrows = block_ptr[nb];
  NGSAMD_REQUIRE(nrows < (int64_t(1) << 31), "ngsamd_blockjacobi_create: too many block rows");
  std::vector<int> bd(block_dofs, block_dofs + nrows);
  std::vector<int64_t> vp(nb + 1, 0);
  for (int64_t b = 0; b < nb; b++) {
    const int64_t sz = block_ptr[b + 1] - block_ptr[b];
    std::sort(bd.begin() + block_ptr[b], bd.begin() + block_ptr[b + 1]);
    for (int64_t k = block_ptr[b]; k < block_ptr[b + 1]; k++) {
      NGSAMD_REQUIRE(bd[k] >= 0 && bd[k] < n, "ngsamd_blockjacobi_create: dof number out of range");
      NGSAMD_REQUIRE(k == block_ptr[b] || bd[k] != bd[k - 1], "ngsamd_blockjacobi_create: dof twice in a block");
    }
    vp[b + 1] = vp[b] + sz * sz;
  }
  // diagonal blocks from the matrix, inverted; blocks are independent, so
  // host threads take contiguous block ranges (same result for any count).
  // The buffer is pinned host memory (the upload of the inverses, often
  // hundreds of MB, then runs at DMA rate) and left uninitialised: every
  // entry is written by the thread that owns its block.
  const size_t ninv = (size_t)std::max<int64_t>(vp[nb], 1);
  double* hinv_p = nullptr;
  NGSAMD_HIP(hipHostMalloc((void**)&hinv_p, ninv * sizeof(double), hipHostMallocDefault));
  std::unique_ptr<double, hipError_t (*)(void*)> hinv(hinv_p, hipHostFree);
  auto invert_range = [&](int64_t lo, int64_t hi) -> bool {
    bool ok = true;
    for (int64_t b = lo; b < hi; b++) {
      const int64_t b0 = block_ptr[b];
      const int sz = (int)(block_ptr[b + 1] - b0);
      double* m = hinv.get() + vp[b];
      for (int j = 0; j < sz; j++) {
        const int r = bd[b0 + j];
        const int* rc = cols + rp[r];
        const int* re = cols + rp[r + 1];
        for (int k = 0; k < sz; k++) {
          const int* pos = std::lower_bound(rc, re, bd[b0 + k]);
          m[(size_t)j * sz + k] = (pos != re && *pos == bd[b0 + k]) ? vals[rp[r] + (pos - rc)] : 0.0;
        }
      }
      if (sz && !InvertInPlace(m, sz)) ok = false;
      // stored column by column (the apply's access order)
      for (int j = 0; j < sz; j++)
        for (int k = j + 1; k < sz; k++) std::swap(m[(size_t)j * sz + k], m[(size_t)k * sz + j]);
    }
    return ok;
  };
  // ranges of equal work (sum of sz^2 entries; the large vertex-patch
  // blocks come first in the usual block lists)
  const int nthreads = (int)std::min<int64_t>(HostThreads(), std::max<int64_t>(1, vp[nb] / 100000));
  std::vector<int64_t> cut(nthreads + 1, nb);
  for (int t = 0; t < nthreads; t++)
    cut[t] = std::lower_bound(vp.begin(), vp.end() - 1, vp[nb] / nthreads * t) - vp.begin();
  std::vector<char> ok(nthreads, 1);
  {
    std::vector<std::thread> pool;
    int t = 1;
    try {
      for (; t < nthreads; t++) pool.emplace_back([&, t] { ok[t] = invert_range(cut[t], cut[t + 1]); });
    } catch (...) {
      // no more threads: the remaining ranges run here
      for (; t < nthreads; t++) ok[t] = invert_range(cut[t], cut[t + 1]);
    }
    ok[0] = invert_range(cut[0], cut[1]);
    for (auto& th : pool) th.join();
  }
  NGSAMD_REQUIRE(std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; }),
                 "ngsamd_blockjacobi_create: a diagonal block is singular");
  std::vector<int> rb(std::max<int64_t>(nrows, 1), 0);
  for (int64_t b = 0; b < nb; b++)
    for (int64_t k = block_ptr[b]; k < block_ptr[b + 1]; k++) rb[k] = (int)b;
  // per dof its block rows, in row order (counting sort)
  std::vector<int> gp(n + 1, 0), gr(std::max<int64_t>(nrows, 1), 0);
  for (int64_t t = 0; t < nrows; t++) gp[bd[t] + 1]++;
  for (int64_t d = 0; d < n; d++) gp[d + 1] += gp[d];
  {
    std::vector<int> fill(gp.begin(), gp.end() - 1);
    for (int64_t t = 0; t < nrows; t++) gr[fill[bd[t]]++] = (int)t;
  }
  bptr.alloc(nb + 1);
  bptr.upload(block_ptr, nb + 1);
  vptr.alloc(nb + 1);
  vptr.upload(vp.data(), nb + 1);
  bdofs.alloc(std::max<int64_t>(nrows, 1));
  if (nrows) bdofs.upload(bd.data(), nrows);
  row_block.alloc(rb.size());
  row_block.upload(rb.data(), rb.size());
  if (vp[nb] == 0) hinv.get()[0] = 0.0;
  inv.alloc(ninv);
  inv.upload(hinv.get(), ninv);
  gptr.alloc(n + 1);
  gptr.upload(gp.data(), n + 1);
  grows.alloc(gr.size());
  grows.upload(gr.data(), gr.size());
  z.alloc(std::max<int64_t>(nrows, 1));
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
}

void BlockJacobiOp::Apply(double s, const double* x, double* y, bool accumulate, const int* stop, hipStream_t st) {
  if (height == 0) return;
  if (nrows) {
    BjRowsKernel<<<GridFor(nrows, kBjBlock, 16384), kBjBlock, 0, st>>>(bptr.p, vptr.p, bdofs.p, row_block.p, inv.p,
                                                                       x, z.p, nrows, stop);
    NGSAMD_CHECK_LAUNCH();
  }
  BjGatherKernel<<<GridFor(height, kBjBlock, 16384), kBjBlock, 0, st>>>(gptr.p, grows.p, z.p, y, height, s,
                                                                        accumulate ? 1 : 0, stop);
  NGSAMD_CHECK_LAUNCH();
}

void BlockJacobiOp::MultAdd(double s, const double* x, double* y, hipStream_t st) { Apply(s, x, y, true, nullptr, st); }
void BlockJacobiOp::Mult(const double* x, double* y, hipStream_t st) { Apply(1.0, x, y, false, nullptr, st); }
void BlockJacobiOp::MultStop(const double* x, double* y, const int* stop, hipStream_t st) {
  Apply(1.0, x, y, false, stop, st);
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_blockjacobi_create(ngsamd_op_t* op, int64_t n, const int64_t* rowptr, const int* cols,
                                         const double* vals, int64_t nblocks, const int64_t* block_ptr,
                                         const int* block_dofs) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && rowptr && block_ptr, "ngsamd_blockjacobi_create: null argument");
    NGSAMD_REQUIRE(n >= 0 && nblocks >= 0, "ngsamd_blockjacobi_create: bad sizes");
    NGSAMD_REQUIRE(rowptr[n] == 0 || (cols && vals), "ngsamd_blockjacobi_create: null argument");
    NGSAMD_REQUIRE(block_ptr[nblocks] == 0 || block_dofs, "ngsamd_blockjacobi_create: null argument");
    *op = nullptr;
    *op = new BlockJacobiOp(n, rowptr, cols, vals, nblocks, block_ptr, block_dofs);
  });
}
