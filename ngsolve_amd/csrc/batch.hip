// Batched CG for many independent problems of uneven size (one block-diagonal
// system, one set of launches per iteration for all problems).
//
// Reference: CGSolver (python/krylovspace.py:227-291) applied to each problem
// separately, and the lockstep multi-right-hand-side CG
// CGSolver<double>::MultiMult (linalg/cg.cpp:132-250), generalised to
// independent operators: every problem keeps its own alpha, beta, residual
// history, stopping test and iteration count; a converged problem is frozen
// (its row blocks and vector chunks return at once) while the others go on.
//
// Layout: the problems' CSR matrices are stacked into one block-diagonal CSR
// (int32 row pointers / columns, FP64 values and optionally an FP32 copy of
// the values for the inner solves of iterative refinement).  Row blocks of
// the stream SpMV never straddle two problems, so each block's <s, A s>
// partial belongs to one problem; vectors are cut into chunks of at most
// kChunk dofs inside one problem for the update kernels.  Per-problem sums
// are formed by one workgroup per problem over its contiguous range of
// partials (deterministic).
//
// Tiles: the nonzeros per row block (1024/2048/4096) are picked when the
// operator is created by timing each candidate on the device.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "reduce.hpp"
#include "runs.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kBatchBlock = 256;
constexpr int kChunk = 4096;  // dofs per vector chunk

struct RowBlocks {
  std::vector<int> row;  // nblocks+1
  std::vector<int> seg;  // nblocks
  std::vector<int> seg_begin;  // nseg+1: first block of each problem
};

static RowBlocks MakeRowBlocks(const int64_t* rp, const std::vector<int64_t>& seg_rows, int block_nnz) {
  RowBlocks b;
  const int nseg = (int)seg_rows.size() - 1;
  for (int p = 0; p < nseg; p++) {
    b.seg_begin.push_back((int)b.seg.size());
    int64_t start = seg_rows[p];
    int64_t acc = 0;
    if (seg_rows[p + 1] == seg_rows[p]) continue;
    b.row.push_back((int)start);
    b.seg.push_back(p);
    for (int64_t r = seg_rows[p]; r < seg_rows[p + 1]; r++) {
      const int64_t len = rp[r + 1] - rp[r];
      if (r > start && acc + len > block_nnz) {
        b.row.push_back((int)r);
        b.seg.push_back(p);
        start = r;
        acc = 0;
      }
      acc += len;
    }
  }
  b.seg_begin.push_back((int)b.seg.size());
  b.row.push_back((int)seg_rows[nseg]);
  return b;
}

struct BatchOp : ngsamd_op_s {
  DevBuf<int> rowptr, cols;
  DevBuf<double> vals;
  DevBuf<float> vals32;
  // pattern-runs layout (see runs.hpp), used when it saves >= 25% of the
  // column numbers; blocks never straddle two problems
  bool use_runs = false;
  bool has_f32 = false;
  DevBuf<int> run_row, run_ptr, val_ptr, rcols, blk_run, blk_len;
  DevBuf<double> rvals;
  DevBuf<float> rvals32;
  int nseg = 0;
  std::vector<int64_t> seg_rows;
  int block_nnz = 2048;
  int nblocks = 0;
  DevBuf<int> blk_row, blk_seg, seg_blk;
  int nchunks = 0;
  DevBuf<int64_t> chunk_start;
  DevBuf<int> chunk_seg, seg_chunk;
  DevBuf<double> partials;
  // autotune: ms per apply of each candidate block size (1024, 2048, 4096),
  // empty when a single size was forced
  std::vector<float> tuned_ms;
  BatchOp(int64_t n, const int64_t* rp, const int* c, const double* v, int nseg, const int64_t* seg, bool with_f32);
  void SetBlocks(const RowBlocks& b);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
};

// ------------------------------------------------------------------ SpMV

__device__ __forceinline__ int XcdBlock(int b, int grid) {
  if ((grid & 7) != 0) return b;
  return (b & 7) * (grid >> 3) + (b >> 3);
}

// y = A x (ACCUM: y += s A x) over row blocks; with DOT the partial
// sum_r x[r] (A x)[r] of every row block goes to partials[block]; blocks of
// problems with done[seg] != 0 are skipped.
// Row block b (prod: NNZ doubles of LDS, lds: kBatchBlock / 64 doubles).
template <int NNZ, class V, class M, bool ACCUM, bool DOT>
__device__ __forceinline__ void BatchSpmvBlock(const int* __restrict__ rowptr, const int* __restrict__ cols,
                                               const M* __restrict__ vals, const V* __restrict__ x,
                                               V* __restrict__ y, double s, const int* __restrict__ blk_row,
                                               const int* __restrict__ blk_seg, const int* __restrict__ done,
                                               double* __restrict__ partials, int b, double* prod, double* lds) {
  constexpr int PER = NNZ / kBatchBlock;
  const int tid = threadIdx.x;
  {
    if (done && done[blk_seg[b]]) {
      if (DOT && tid == 0) partials[b] = 0.0;
      return;
    }
    const int r0 = blk_row[b], r1 = blk_row[b + 1];
    const int k0 = rowptr[r0], k1 = rowptr[r1];
    const int nz = k1 - k0;
    double dotacc = 0.0;
    if (nz > NNZ) {
      double a = 0.0;
      for (int k = k0 + tid; k < k1; k += kBatchBlock)
        a += (double)__builtin_nontemporal_load(&vals[k]) * (double)x[__builtin_nontemporal_load(&cols[k])];
      a = BlockSum<kBatchBlock>(a, lds);
      if (tid == 0) {
        if (ACCUM) y[r0] = (V)((double)y[r0] + s * a); else y[r0] = (V)(s * a);
        dotacc = (double)x[r0] * a;
        if (DOT) partials[b] = dotacc;
      }
      return;
    }
    // unconditional loads (lanes past the block read entry k0, in bounds
    // thanks to the padding) so that all of them are in flight at once
    int c[PER];
    double v[PER], xv[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kBatchBlock;
      const int kk = k0 + (k < nz ? k : 0);
      c[i] = __builtin_nontemporal_load(&cols[kk]);
      v[i] = (double)__builtin_nontemporal_load(&vals[kk]);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) xv[i] = (double)x[c[i]];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kBatchBlock;
      prod[k] = k < nz ? v[i] * xv[i] : 0.0;
    }
    __syncthreads();
    const int nr = r1 - r0;
    int tpr = 64;
    while (tpr > 1 && tpr * nr > kBatchBlock) tpr >>= 1;
    const int l = tid & (tpr - 1);
    for (int j0 = 0; j0 < nr; j0 += kBatchBlock / tpr) {
      const int j = j0 + tid / tpr;
      double sum = 0.0;
      if (j < nr) {
        const int s0 = rowptr[r0 + j] - k0, s1 = rowptr[r0 + j + 1] - k0;
        for (int k = s0 + l; k < s1; k += tpr) sum += prod[k];
      }
      for (int off = tpr / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (j < nr && l == 0) {
        if (ACCUM) y[r0 + j] = (V)((double)y[r0 + j] + s * sum); else y[r0 + j] = (V)(s * sum);
        if (DOT) dotacc += (double)x[r0 + j] * sum;
      }
    }
    if (DOT) {
      dotacc = BlockSum<kBatchBlock>(dotacc, lds);
      if (tid == 0) partials[b] = dotacc;
    }
    __syncthreads();
  }
}

template <int NNZ, class V, class M, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kBatchBlock) void BatchSpmvKernel(
    const int* __restrict__ rowptr, const int* __restrict__ cols, const M* __restrict__ vals, const V* __restrict__ x,
    V* __restrict__ y, double s, const int* __restrict__ blk_row, const int* __restrict__ blk_seg, int nblocks,
    int blocks_per_wg, const int* __restrict__ done, double* __restrict__ partials) {
  __shared__ double prod[NNZ];
  __shared__ double lds[kBatchBlock / 64];
  const int lb = XcdBlock(blockIdx.x, gridDim.x);
  const int b0 = lb * blocks_per_wg;
  const int b1 = min(b0 + blocks_per_wg, nblocks);
  for (int b = b0; b < b1; b++)
    BatchSpmvBlock<NNZ, V, M, ACCUM, DOT>(rowptr, cols, vals, x, y, s, blk_row, blk_seg, done, partials, b, prod,
                                          lds);
}

// Runs-layout form of BatchSpmvKernel (the runs kernel of spmv.hip with the
// frozen-problem skip and one partial per block).
// Block b of the runs layout (prod / lds as in BatchSpmvBlock).
template <int NNZ, class V, class M, bool ACCUM, bool DOT>
__device__ __forceinline__ void BatchRunsBlock(const int* __restrict__ run_row, const int* __restrict__ run_ptr,
                                               const int* __restrict__ val_ptr, const int* __restrict__ rcols,
                                               const M* __restrict__ rvals, const V* __restrict__ x,
                                               V* __restrict__ y, double s, const int* __restrict__ blk_run,
                                               const int* __restrict__ blk_len, const int* __restrict__ blk_seg,
                                               const int* __restrict__ done, double* __restrict__ partials, int b,
                                               double* prod, double* lds) {
  static_assert(NNZ <= 4096, "RunDiv is exact below 4096");
  constexpr int PER = NNZ / kBatchBlock;
  const int tid = threadIdx.x;
  {
    if (done && done[blk_seg[b]]) {
      if (DOT && tid == 0) partials[b] = 0.0;
      return;
    }
    const int u0 = blk_run[b], u1 = blk_run[b + 1];
    const int R = blk_len[b];
    const int inv = (65536 + R - 1) / R;
    const int c0 = run_ptr[u0];
    const int vb = val_ptr[u0];
    const int nv = val_ptr[u1] - vb;
    double dotacc = 0.0;
    if (nv > NNZ) {
      // one long row (R = 1): the whole workgroup strides over it
      double a = 0.0;
      for (int k = tid; k < nv; k += kBatchBlock)
        a += (double)__builtin_nontemporal_load(&rvals[vb + k]) * (double)x[__builtin_nontemporal_load(&rcols[c0 + k])];
      a = BlockSum<kBatchBlock>(a, lds);
      if (tid == 0) {
        const int r = run_row[u0];
        if (ACCUM) y[r] = (V)((double)y[r] + s * a); else y[r] = (V)(s * a);
        if (DOT) partials[b] = (double)x[r] * a;
      }
      return;
    }
    int c[PER];
    double v[PER], xv[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kBatchBlock;
      const int kk = k < nv ? k : 0;
      c[i] = __builtin_nontemporal_load(&rcols[c0 + RunDiv(kk, inv)]);
      v[i] = (double)__builtin_nontemporal_load(&rvals[vb + kk]);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) xv[i] = (double)x[c[i]];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kBatchBlock;
      prod[k] = k < nv ? v[i] * xv[i] : 0.0;
    }
    __syncthreads();
    const int row0 = run_row[u0];
    const int nr = (u1 - u0) * R;
    int tpr = 64;
    while (tpr > 1 && tpr * nr > kBatchBlock) tpr >>= 1;
    const int l = tid & (tpr - 1);
    for (int j0 = 0; j0 < nr; j0 += kBatchBlock / tpr) {
      const int j = j0 + tid / tpr;
      double sum = 0.0;
      if (j < nr) {
        const int run = RunDiv(j, inv), q = j - run * R;
        const int s0 = run_ptr[u0 + run] - c0, s1 = run_ptr[u0 + run + 1] - c0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = s0 + l;
        for (; k + 3 * tpr < s1; k += 4 * tpr) {
          a0 += prod[k * R + q];
          a1 += prod[(k + tpr) * R + q];
          a2 += prod[(k + 2 * tpr) * R + q];
          a3 += prod[(k + 3 * tpr) * R + q];
        }
        for (; k < s1; k += tpr) a0 += prod[k * R + q];
        sum = (a0 + a1) + (a2 + a3);
      }
      for (int off = tpr / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (j < nr && l == 0) {
        if (ACCUM) y[row0 + j] = (V)((double)y[row0 + j] + s * sum); else y[row0 + j] = (V)(s * sum);
        if (DOT) dotacc += (double)x[row0 + j] * sum;
      }
    }
    if (DOT) {
      dotacc = BlockSum<kBatchBlock>(dotacc, lds);
      if (tid == 0) partials[b] = dotacc;
    }
    __syncthreads();
  }
}

template <int NNZ, class V, class M, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kBatchBlock) void BatchRunsKernel(
    const int* __restrict__ run_row, const int* __restrict__ run_ptr, const int* __restrict__ val_ptr,
    const int* __restrict__ rcols, const M* __restrict__ rvals, const V* __restrict__ x, V* __restrict__ y, double s,
    const int* __restrict__ blk_run, const int* __restrict__ blk_len, const int* __restrict__ blk_seg, int nblocks,
    int blocks_per_wg, const int* __restrict__ done, double* __restrict__ partials) {
  __shared__ double prod[NNZ];
  __shared__ double lds[kBatchBlock / 64];
  const int lb = XcdBlock(blockIdx.x, gridDim.x);
  const int b0 = lb * blocks_per_wg;
  const int b1 = min(b0 + blocks_per_wg, nblocks);
  for (int b = b0; b < b1; b++)
    BatchRunsBlock<NNZ, V, M, ACCUM, DOT>(run_row, run_ptr, val_ptr, rcols, rvals, x, y, s, blk_run, blk_len,
                                          blk_seg, done, partials, b, prod, lds);
}

template <class V, class M, bool ACCUM, bool DOT>
static void LaunchBatchSpmv(const BatchOp& A, const M* vals, const V* x, V* y, double s, const int* done,
                            hipStream_t st, int block_nnz) {
  if (A.use_runs) {
    // vals selects the value copy: the FP64 CSR values stand for rvals and
    // the FP32 copy for rvals32
    const M* rv;
    if constexpr (sizeof(M) == 8) rv = A.rvals.p;
    else rv = A.rvals32.p;
    (void)vals;
    int64_t wg = A.nblocks < 16384 ? A.nblocks : 16384;
    if (wg > 8) wg &= ~int64_t(7);
    if (wg < 1) wg = 1;
    const int per = (int)((A.nblocks + wg - 1) / wg);
    auto rk = block_nnz == 1024   ? BatchRunsKernel<1024, V, M, ACCUM, DOT>
              : block_nnz == 4096 ? BatchRunsKernel<4096, V, M, ACCUM, DOT>
                                  : BatchRunsKernel<2048, V, M, ACCUM, DOT>;
    rk<<<(int)wg, kBatchBlock, 0, st>>>(A.run_row.p, A.run_ptr.p, A.val_ptr.p, A.rcols.p, rv, x, y, s, A.blk_run.p,
                                       A.blk_len.p, A.blk_seg.p, A.nblocks, per, done, A.partials.p);
    NGSAMD_CHECK_LAUNCH();
    return;
  }
  int64_t wg = A.nblocks < 8192 ? A.nblocks : 8192;
  if (wg > 8) wg &= ~int64_t(7);
  if (wg < 1) wg = 1;
  const int per = (int)((A.nblocks + wg - 1) / wg);
  auto kern = block_nnz == 1024   ? BatchSpmvKernel<1024, V, M, ACCUM, DOT>
              : block_nnz == 4096 ? BatchSpmvKernel<4096, V, M, ACCUM, DOT>
                                  : BatchSpmvKernel<2048, V, M, ACCUM, DOT>;
  kern<<<(int)wg, kBatchBlock, 0, st>>>(A.rowptr.p, A.cols.p, vals, x, y, s, A.blk_row.p, A.blk_seg.p, A.nblocks,
                                        per, done, A.partials.p);
  NGSAMD_CHECK_LAUNCH();
}

void BatchOp::SetBlocks(const RowBlocks& b) {
  nblocks = (int)b.seg.size();
  blk_row.alloc(b.row.size());
  blk_row.upload(b.row.data(), b.row.size());
  blk_seg.alloc(std::max<size_t>(b.seg.size(), 1));
  blk_seg.upload(b.seg.data(), b.seg.size());
  seg_blk.alloc(b.seg_begin.size());
  seg_blk.upload(b.seg_begin.data(), b.seg_begin.size());
}

BatchOp::BatchOp(int64_t n, const int64_t* rp, const int* c, const double* v, int ns, const int64_t* seg,
                 bool with_f32)
    : ngsamd_op_s(OpKind::Batch) {
  height = width = n;
  nseg = ns;
  NGSAMD_REQUIRE(n >= 0 && n < (int64_t(1) << 31), "ngsamd_batch_csr_create: size out of range");
  NGSAMD_REQUIRE(ns >= 1 && seg[0] == 0 && seg[ns] == n, "ngsamd_batch_csr_create: bad problem offsets");
  const int64_t nnz = rp[n];
  NGSAMD_REQUIRE(rp[0] == 0 && nnz < (int64_t(1) << 31), "ngsamd_batch_csr_create: bad row pointer");
  seg_rows.assign(seg, seg + ns + 1);
  for (int p = 0; p < ns; p++) {
    NGSAMD_REQUIRE(seg[p + 1] >= seg[p], "ngsamd_batch_csr_create: problem offsets not monotone");
    // block-diagonal: every column of a problem's rows lies inside the problem
    for (int64_t r = seg[p]; r < seg[p + 1]; r++)
      for (int64_t k = rp[r]; k < rp[r + 1]; k++)
        NGSAMD_REQUIRE(c[k] >= seg[p] && c[k] < seg[p + 1], "ngsamd_batch_csr_create: matrix is not block-diagonal");
  }
  has_f32 = with_f32;
  // 4 padding entries (column 0, value 0): lanes past the end of the last
  // row block read entry nnz
  static const int zc[4] = {0, 0, 0, 0};
  static const double zv[4] = {0, 0, 0, 0};
  static const float zf[4] = {0, 0, 0, 0};
  // pattern-runs layout when it saves >= 25% of the column numbers
  // (NGSAMD_BATCH_RUNS=0 keeps plain CSR)
  const char* renv = getenv("NGSAMD_BATCH_RUNS");
  if (!(renv && std::string(renv) == "0") && n > 0) {
    std::vector<int> rr, rpp, vp, rc, br, bl, bs;
    std::vector<double> rv;
    rc.reserve(nnz);
    rv.reserve(nnz);
    // runs of at most 1024 values (or one longer row), so that every block
    // size tried below holds whole runs
    const int64_t nsh = BuildRuns(n, rp, c, v, 1024, rr, rpp, vp, rc, rv, br, bl, seg, ns, &bs);
    if (4 * nsh <= 3 * nnz) {
      use_runs = true;
      block_nnz = 1024;
      nblocks = (int)br.size() - 1;
      run_row.alloc(rr.size());
      run_row.upload(rr.data(), rr.size());
      run_ptr.alloc(rpp.size());
      run_ptr.upload(rpp.data(), rpp.size());
      val_ptr.alloc(vp.size());
      val_ptr.upload(vp.data(), vp.size());
      rcols.alloc(nsh + 4);
      rcols.upload(rc.data(), nsh);
      rcols.upload_at(zc, nsh, 4);
      rvals.alloc(nnz + 4);
      rvals.upload(rv.data(), nnz);
      rvals.upload_at(zv, nnz, 4);
      if (with_f32) {
        std::vector<float> v32(rv.begin(), rv.end());
        rvals32.alloc(nnz + 4);
        rvals32.upload(v32.data(), nnz);
        rvals32.upload_at(zf, nnz, 4);
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
      }
      // problem of each run
      const int nruns = (int)rr.size() - 1;
      std::vector<int> rseg(nruns);
      {
        int q = 0;
        for (int u = 0; u < nruns; u++) {
          while (q < ns && rr[u] >= seg[q + 1]) q++;
          rseg[u] = q;
        }
      }
      auto set_blocks = [&](int bn) {
        GroupRunBlocks(rr, vp, rseg, bn, br, bl, &bs);
        nblocks = (int)br.size() - 1;
        blk_run.alloc(br.size());
        blk_run.upload(br.data(), br.size());
        blk_len.alloc(std::max<size_t>(bl.size(), 1));
        blk_len.upload(bl.data(), bl.size());
        blk_seg.alloc(std::max<size_t>(bs.size(), 1));
        blk_seg.upload(bs.data(), bs.size());
        // first block of each problem (blocks are in problem order)
        std::vector<int> sb(ns + 1);
        int b = 0;
        for (int q = 0; q <= ns; q++) {
          while (b < nblocks && bs[b] < q) b++;
          sb[q] = b;
        }
        seg_blk.alloc(sb.size());
        seg_blk.upload(sb.data(), sb.size());
        partials.alloc(std::max<size_t>({(size_t)nblocks, (size_t)ns, (size_t)1}));
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
      };
      // autotune the values per block of the runs kernel: each candidate
      // timed on the device (three applies), the fastest kept
      const char* benv = getenv("NGSAMD_SPMV_NNZ");
      std::vector<int> cand = benv ? std::vector<int>{atoi(benv)} : std::vector<int>{1024, 2048, 4096};
      int best = cand[0];
      float best_ms = 1e30f;
      if (cand.size() > 1) {
        hipStream_t st = Stream();
        DevBuf<double> tx(n), ty(n);
        LaunchFill(tx.p, n, 1.0, st);
        hipEvent_t e0, e1;
        NGSAMD_HIP(hipEventCreate(&e0));
        NGSAMD_HIP(hipEventCreate(&e1));
        for (int bn : cand) {
          set_blocks(bn);
          LaunchBatchSpmv<double, double, false, true>(*this, rvals.p, tx.p, ty.p, 1.0, nullptr, st, bn);
          NGSAMD_HIP(hipEventRecord(e0, st));
          for (int k = 0; k < 3; k++)
            LaunchBatchSpmv<double, double, false, true>(*this, rvals.p, tx.p, ty.p, 1.0, nullptr, st, bn);
          NGSAMD_HIP(hipEventRecord(e1, st));
          NGSAMD_HIP(hipEventSynchronize(e1));
          float ms = 0.f;
          NGSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
          tuned_ms.push_back(ms / 3);
          if (ms < best_ms) best_ms = ms, best = bn;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
      }
      NGSAMD_REQUIRE(best == 1024 || best == 2048 || best == 4096, "NGSAMD_SPMV_NNZ must be 1024, 2048 or 4096");
      block_nnz = best;
      set_blocks(best);
    }
  }
  if (!use_runs) {
    std::vector<int> rp32(n + 1);
    for (int64_t i = 0; i <= n; i++) rp32[i] = (int)rp[i];
    rowptr.alloc(n + 1);
    rowptr.upload(rp32.data(), n + 1);
    cols.alloc(nnz + 4);
    cols.upload(c, nnz);
    cols.upload_at(zc, nnz, 4);
    vals.alloc(nnz + 4);
    vals.upload(v, nnz);
    vals.upload_at(zv, nnz, 4);
    if (with_f32) {
      std::vector<float> v32(v, v + nnz);
      vals32.alloc(nnz + 4);
      vals32.upload(v32.data(), nnz);
      vals32.upload_at(zf, nnz, 4);
      NGSAMD_HIP(hipStreamSynchronize(Stream()));
    }
  }
  // vector chunks
  std::vector<int64_t> cs;
  std::vector<int> csg, sgc;
  for (int p = 0; p < ns; p++) {
    sgc.push_back((int)csg.size());
    for (int64_t a = seg[p]; a < seg[p + 1]; a += kChunk) {
      cs.push_back(a);
      csg.push_back(p);
    }
  }
  sgc.push_back((int)csg.size());
  cs.push_back(n);
  nchunks = (int)csg.size();
  chunk_start.alloc(cs.size());
  chunk_start.upload(cs.data(), cs.size());
  chunk_seg.alloc(std::max<size_t>(csg.size(), 1));
  chunk_seg.upload(csg.data(), csg.size());
  seg_chunk.alloc(sgc.size());
  seg_chunk.upload(sgc.data(), sgc.size());

  if (use_runs) {
    partials.alloc(std::max<size_t>({(size_t)nblocks, (size_t)nchunks, (size_t)1}));
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
    return;
  }
  tuned_ms.clear();
  // autotune the row-block size: time each candidate on the device
  hipStream_t st = Stream();
  const char* env = getenv("NGSAMD_SPMV_NNZ");
  std::vector<int> cand = env ? std::vector<int>{atoi(env)} : std::vector<int>{1024, 2048, 4096};
  int best = cand[0];
  float best_ms = 1e30f;
  RowBlocks best_blocks;
  size_t max_blocks = (size_t)nchunks;
  DevBuf<double> tx(std::max<int64_t>(n, 1)), ty(std::max<int64_t>(n, 1));
  LaunchFill(tx.p, n, 1.0, st);
  hipEvent_t e0, e1;
  NGSAMD_HIP(hipEventCreate(&e0));
  NGSAMD_HIP(hipEventCreate(&e1));
  for (int bn : cand) {
    NGSAMD_REQUIRE(bn == 1024 || bn == 2048 || bn == 4096, "NGSAMD_SPMV_NNZ must be 1024, 2048 or 4096");
    RowBlocks rb = MakeRowBlocks(rp, seg_rows, bn);
    SetBlocks(rb);
    partials.alloc(std::max<size_t>({rb.seg.size(), (size_t)nchunks, (size_t)1}));
    float ms = 0.f;
    if (cand.size() > 1 && n > 0) {
      LaunchBatchSpmv<double, double, false, true>(*this, vals.p, tx.p, ty.p, 1.0, nullptr, st, bn);
      NGSAMD_HIP(hipEventRecord(e0, st));
      for (int k = 0; k < 3; k++)
        LaunchBatchSpmv<double, double, false, true>(*this, vals.p, tx.p, ty.p, 1.0, nullptr, st, bn);
      NGSAMD_HIP(hipEventRecord(e1, st));
      NGSAMD_HIP(hipEventSynchronize(e1));
      NGSAMD_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    if (cand.size() > 1) tuned_ms.push_back(ms / 3);
    if (ms < best_ms) {
      best_ms = ms;
      best = bn;
      best_blocks = rb;
    }
    max_blocks = std::max(max_blocks, rb.seg.size());
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  block_nnz = best;
  SetBlocks(best_blocks);
  partials.alloc(std::max<size_t>(max_blocks, 1));
  NGSAMD_HIP(hipStreamSynchronize(st));
}

void BatchOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchBatchSpmv<double, double, true, false>(*this, vals.p, x, y, s, nullptr, st, block_nnz);
}

void BatchOp::Mult(const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchBatchSpmv<double, double, false, false>(*this, vals.p, x, y, 1.0, nullptr, st, block_nnz);
}

// ------------------------------------------------------- per-problem state

template <class T>
__global__ void FillKernelT(T* __restrict__ x, int64_t n, T a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = a;
}

template <class T>
static void Fill(T* x, int64_t n, T a, hipStream_t st) {
  if (n <= 0) return;
  FillKernelT<T><<<GridFor(n, 256, 8192), 256, 0, st>>>(x, n, a);
  NGSAMD_CHECK_LAUNCH();
}

struct BatchState {
  double* wdn;
  double* alpha;
  double* beta;
  double* final_res;
  double* tol;        // per problem relative tolerance
  double* last_res;
  long long* finish;  // device wall clock when the problem finished
  int* it;
  int* done;          // 0 running, 1 converged / maxiter, 2 breakdown
  int maxiter;
};

struct BatchWork {
  int nseg;
  DevBuf<double> wdn, alpha, beta, final_res, tol, last_res;
  DevBuf<long long> finish;
  DevBuf<int> it, done;
  explicit BatchWork(int n)
      : nseg(n), wdn(n), alpha(n), beta(n), final_res(n), tol(n), last_res(n), finish(n), it(n), done(n) {}
  BatchState state(int maxiter) {
    return BatchState{wdn.p, alpha.p, beta.p, final_res.p, tol.p, last_res.p, finish.p, it.p, done.p, maxiter};
  }
};

__device__ __forceinline__ long long DeviceClock() { return (long long)wall_clock64(); }

// reset the state of the problems with active[p] != 0 (all if active == null)
__global__ void BatchResetKernel(BatchState S, int nseg, const double* tol, double tol_all, const int* active) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nseg; p += gridDim.x * blockDim.x) {
    const bool act = active == nullptr || active[p] == 0;
    S.it[p] = 0;
    S.done[p] = act ? 0 : 1;
    S.tol[p] = tol ? tol[p] : tol_all;
    S.finish[p] = 0;
  }
}

template <class V>
__global__ __launch_bounds__(kBatchBlock) void BatchInitKernel(const V* __restrict__ inv, const V* __restrict__ d,
                                                               V* __restrict__ w, V* __restrict__ s,
                                                               const int64_t* __restrict__ cstart,
                                                               const int* __restrict__ cseg, BatchState S,
                                                               double* __restrict__ partials) {
  __shared__ double lds[kBatchBlock / 64];
  const int c = blockIdx.x;
  if (S.done[cseg[c]]) {
    if (threadIdx.x == 0) partials[c] = 0.0;
    return;
  }
  double acc = 0.0;
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += kBatchBlock) {
    const V di = d[i];
    const V wi = inv[i] * di;
    w[i] = wi;
    s[i] = wi;
    acc += (double)wi * (double)di;
  }
  StorePartial<kBatchBlock>(acc, partials, lds);
}

__device__ __forceinline__ void BatchCheck(BatchState& S, int p, double res) {
  const int it = ++S.it[p];
  S.last_res[p] = res;
  if (it == 1) S.final_res[p] = res * S.tol[p];
  if (it >= S.maxiter || res <= S.final_res[p]) {
    S.done[p] = 1;
    S.finish[p] = DeviceClock();
  }
}

// one workgroup per problem: sum of partials[first[p] .. first[p+1])
__device__ __forceinline__ double SegmentSum(const double* partials, const int* first, int p, double* lds) {
  double v = 0.0;
  for (int i = first[p] + threadIdx.x; i < first[p + 1]; i += kBatchBlock) v += partials[i];
  return BlockSum<kBatchBlock>(v, lds);
}

__global__ __launch_bounds__(kBatchBlock) void BatchInitCheckKernel(const double* __restrict__ partials,
                                                                    const int* __restrict__ seg_chunk, BatchState S) {
  __shared__ double lds[kBatchBlock / 64];
  const int p = blockIdx.x;
  if (S.done[p]) return;
  const double wdn = SegmentSum(partials, seg_chunk, p, lds);
  if (threadIdx.x == 0) {
    S.wdn[p] = wdn;
    BatchCheck(S, p, sqrt(fabs(wdn)));
  }
}

// alpha of problem p from its row blocks' <s, A s> partials (breakdown:
// done = 2); one workgroup
__device__ __forceinline__ void BatchAlphaProblem(const double* __restrict__ partials, const int* __restrict__ seg_blk,
                                                  BatchState& S, int p, double* lds) {
  if (S.done[p]) return;
  const double as_s = SegmentSum(partials, seg_blk, p, lds);
  if (threadIdx.x == 0) {
    const double wd = S.wdn[p];
    if (as_s == 0.0 || wd == 0.0) {
      S.done[p] = 2;
      S.finish[p] = DeviceClock();
    } else {
      S.alpha[p] = wd / as_s;
    }
  }
}

__global__ __launch_bounds__(kBatchBlock) void BatchAlphaKernel(const double* __restrict__ partials,
                                                                const int* __restrict__ seg_blk, BatchState S) {
  __shared__ double lds[kBatchBlock / 64];
  BatchAlphaProblem(partials, seg_blk, S, blockIdx.x, lds);
}

// u += al s, d -= al w, w = C d over dofs [i0, i1); returns this thread's
// part of <w, d>
template <class V>
__device__ __forceinline__ double BatchUpdateRange(const V* __restrict__ inv, V* __restrict__ u, V* __restrict__ d,
                                                   V* __restrict__ w, const V* __restrict__ s, int64_t i0,
                                                   int64_t i1, V al) {
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBatchBlock) {
    const V si = s[i], wi = w[i];
    u[i] += al * si;
    const V di = d[i] - al * wi;
    d[i] = di;
    const V zi = inv[i] * di;
    w[i] = zi;
    acc += (double)zi * (double)di;
  }
  return acc;
}

// s = beta s + w over dofs [i0, i1)
template <class V>
__device__ __forceinline__ void BatchDirectionRange(V* __restrict__ s, const V* __restrict__ w, int64_t i0,
                                                    int64_t i1, V beta) {
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBatchBlock) s[i] = beta * s[i] + w[i];
}

template <class V>
__global__ __launch_bounds__(kBatchBlock) void BatchUpdateKernel(const V* __restrict__ inv, V* __restrict__ u,
                                                                 V* __restrict__ d, V* __restrict__ w,
                                                                 const V* __restrict__ s,
                                                                 const int64_t* __restrict__ cstart,
                                                                 const int* __restrict__ cseg, BatchState S,
                                                                 double* __restrict__ partials) {
  __shared__ double lds[kBatchBlock / 64];
  const int c = blockIdx.x;
  const int p = cseg[c];
  if (S.done[p]) {
    if (threadIdx.x == 0) partials[c] = 0.0;
    return;
  }
  const double acc = BatchUpdateRange<V>(inv, u, d, w, s, cstart[c], cstart[c + 1], (V)S.alpha[p]);
  StorePartial<kBatchBlock>(acc, partials, lds);
}

// <w, d> of problem p from its chunks' partials, the stopping test and
// beta; one workgroup
__device__ __forceinline__ void BatchCheckProblem(const double* __restrict__ partials,
                                                  const int* __restrict__ seg_chunk, BatchState& S, int p,
                                                  double* lds) {
  if (S.done[p]) return;
  const double wdn = SegmentSum(partials, seg_chunk, p, lds);
  if (threadIdx.x == 0) {
    const double wd = S.wdn[p];
    S.wdn[p] = wdn;
    BatchCheck(S, p, sqrt(fabs(wdn)));
    S.beta[p] = wdn / wd;
  }
}

__global__ __launch_bounds__(kBatchBlock) void BatchCheckKernel(const double* __restrict__ partials,
                                                                const int* __restrict__ seg_chunk, BatchState S) {
  __shared__ double lds[kBatchBlock / 64];
  BatchCheckProblem(partials, seg_chunk, S, blockIdx.x, lds);
}

template <class V>
__global__ __launch_bounds__(kBatchBlock) void BatchDirectionKernel(V* __restrict__ s, const V* __restrict__ w,
                                                                    const int64_t* __restrict__ cstart,
                                                                    const int* __restrict__ cseg, BatchState S) {
  const int c = blockIdx.x;
  const int p = cseg[c];
  if (S.done[p]) return;
  BatchDirectionRange<V>(s, w, cstart[c], cstart[c + 1], (V)S.beta[p]);
}

// all problems done?  one workgroup
__global__ __launch_bounds__(kBatchBlock) void BatchAllDoneKernel(const int* __restrict__ done, int nseg,
                                                                  int* __restrict__ all) {
  __shared__ double lds[kBatchBlock / 64];
  double open = 0.0;
  for (int p = threadIdx.x; p < nseg; p += kBatchBlock) open += done[p] ? 0.0 : 1.0;
  open = BlockSum<kBatchBlock>(open, lds);
  if (threadIdx.x == 0) *all = open == 0.0 ? 1 : 0;
}

// ------------------------------------------------------ persistent block CG
//
// The whole FP64 block CG in one cooperative launch: every workgroup stays
// resident and walks the iterations, five phases each, separated by
// grid-wide barriers: (S) w = A s and <s, A s> per row block; (A) alpha per
// problem; (U) u += alpha s, d -= alpha w, w = C d and <w, d> per chunk; (B)
// the stopping test and beta per problem; (C) s = beta s + w per chunk.
// Row blocks and chunks are dealt to the workgroups round-robin (block b to
// workgroup b mod G), so every workgroup keeps a share of every problem's
// work while problems converge; problems go to workgroups p, p + G, ...
// Phases A and B run the graph form's per-problem bodies and S, U and C its
// per-block and per-chunk bodies, so both forms give the same iterates bit
// for bit.  done[] is written only in phases A and B, by the workgroup that
// owns the problem, so every workgroup reads the same done[] at the top of
// an iteration and all leave the loop together.

constexpr long kSpinLimit = 1L << 22;  // polls of a barrier wait (about a second)

// Grid-wide barrier: a monotone arrival counter; this is arrival round
// `round` of nwg workgroups.  Thread 0 publishes the workgroup's writes
// (agent-scope release: the XCD's L2 is written back), arrives, waits for
// the others and acquires (L2 invalidated).  The wait is bounded: after
// kSpinLimit polls, or when another workgroup has given up, *err is set and
// false returned, so a launch that lost co-residency ends instead of
// hanging.
__device__ __forceinline__ bool GridBarrier(unsigned long long* arrive, unsigned long long round, int nwg, int* err,
                                            int* lds_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long target = round * (unsigned long long)nwg;
    __threadfence();
    __hip_atomic_fetch_add(arrive, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int bad = 0;
    long spins = 0;
    while (__hip_atomic_load(arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++spins > kSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bad = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __threadfence();
    *lds_flag = bad;
  }
  __syncthreads();
  return *lds_flag == 0;
}

struct PersistentArgs {
  // plain CSR layout
  const int* rowptr;
  const int* cols;
  const double* vals;
  const int* blk_row;
  // runs layout
  const int* run_row;
  const int* run_ptr;
  const int* val_ptr;
  const int* rcols;
  const double* rvals;
  const int* blk_run;
  const int* blk_len;
  // blocks and chunks
  const int* blk_seg;
  const int* seg_blk;
  int nblocks;
  const int64_t* cstart;
  const int* cseg;
  const int* seg_chunk;
  int nchunks;
  int nseg;
  // vectors
  const double* inv;
  double* u;
  double* d;
  double* w;
  double* s;
  // partial sums
  double* bpart;  // nblocks: <s, A s> per row block
  double* cpart;  // nchunks: <w, d> per chunk
  BatchState S;
  unsigned long long* arrive;
  int* err;
};

template <int NNZ, bool RUNS>
__global__ __launch_bounds__(kBatchBlock) void BatchCgPersistentKernel(PersistentArgs P) {
  __shared__ double prod[NNZ];
  __shared__ double lds[kBatchBlock / 64 + 1];
  __shared__ int flag;
  const int G = gridDim.x;
  const int lb = XcdBlock(blockIdx.x, G);
  BatchState& S = P.S;
  unsigned long long round = 0;
  for (int k = 0; k < S.maxiter; k++) {
    // every problem done?  (done[] is written only in phases A and B)
    double open = 0.0;
    for (int p = threadIdx.x; p < P.nseg; p += kBatchBlock) open += S.done[p] ? 0.0 : 1.0;
    open = BlockSum<kBatchBlock>(open, lds);
    if (threadIdx.x == 0) lds[kBatchBlock / 64] = open;
    __syncthreads();
    if (lds[kBatchBlock / 64] == 0.0) break;
    // (S) w = A s, <s, A s> per row block
    for (int b = lb; b < P.nblocks; b += G) {
      if constexpr (RUNS)
        BatchRunsBlock<NNZ, double, double, false, true>(P.run_row, P.run_ptr, P.val_ptr, P.rcols, P.rvals, P.s, P.w,
                                                         1.0, P.blk_run, P.blk_len, P.blk_seg, S.done, P.bpart, b,
                                                         prod, lds);
      else
        BatchSpmvBlock<NNZ, double, double, false, true>(P.rowptr, P.cols, P.vals, P.s, P.w, 1.0, P.blk_row,
                                                         P.blk_seg, S.done, P.bpart, b, prod, lds);
    }
    if (!GridBarrier(P.arrive, ++round, G, P.err, &flag)) return;
    // (A) alpha per problem
    for (int p = lb; p < P.nseg; p += G) BatchAlphaProblem(P.bpart, P.seg_blk, S, p, lds);
    if (!GridBarrier(P.arrive, ++round, G, P.err, &flag)) return;
    // (U) updates and <w, d> per chunk
    for (int c = lb; c < P.nchunks; c += G) {
      const int p = P.cseg[c];
      if (S.done[p]) continue;
      const double acc = BlockSum<kBatchBlock>(
          BatchUpdateRange<double>(P.inv, P.u, P.d, P.w, P.s, P.cstart[c], P.cstart[c + 1], S.alpha[p]), lds);
      if (threadIdx.x == 0) P.cpart[c] = acc;
    }
    if (!GridBarrier(P.arrive, ++round, G, P.err, &flag)) return;
    // (B) stopping test and beta per problem
    for (int p = lb; p < P.nseg; p += G) BatchCheckProblem(P.cpart, P.seg_chunk, S, p, lds);
    if (!GridBarrier(P.arrive, ++round, G, P.err, &flag)) return;
    // (C) s = beta s + w per chunk
    for (int c = lb; c < P.nchunks; c += G) {
      const int p = P.cseg[c];
      if (!S.done[p]) BatchDirectionRange<double>(P.s, P.w, P.cstart[c], P.cstart[c + 1], S.beta[p]);
    }
    if (!GridBarrier(P.arrive, ++round, G, P.err, &flag)) return;
  }
}

template <class V, class M>
static void EnqueueBatchIteration(BatchOp& A, const M* vals, const V* inv, V* u, V* d, V* w, V* s, BatchState S,
                                  int* all_done, hipStream_t st) {
  LaunchBatchSpmv<V, M, false, true>(A, vals, s, w, 1.0, S.done, st, A.block_nnz);
  BatchAlphaKernel<<<A.nseg, kBatchBlock, 0, st>>>(A.partials.p, A.seg_blk.p, S);
  NGSAMD_CHECK_LAUNCH();
  BatchUpdateKernel<V><<<A.nchunks, kBatchBlock, 0, st>>>(inv, u, d, w, s, A.chunk_start.p, A.chunk_seg.p, S,
                                                           A.partials.p);
  NGSAMD_CHECK_LAUNCH();
  BatchCheckKernel<<<A.nseg, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, S);
  NGSAMD_CHECK_LAUNCH();
  BatchDirectionKernel<V><<<A.nchunks, kBatchBlock, 0, st>>>(s, w, A.chunk_start.p, A.chunk_seg.p, S);
  NGSAMD_CHECK_LAUNCH();
  BatchAllDoneKernel<<<1, kBatchBlock, 0, st>>>(S.done, A.nseg, all_done);
  NGSAMD_CHECK_LAUNCH();
}

// The iterations after the start check in one cooperative launch of
// BatchCgPersistentKernel (grid: as many workgroups as are resident at once,
// at most one per row block, chunk or problem).
template <int NNZ, bool RUNS>
static void LaunchPersistent(const PersistentArgs& P, hipStream_t st) {
  auto kern = BatchCgPersistentKernel<NNZ, RUNS>;
  int occ = 0;
  NGSAMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, kBatchBlock, 0));
  NGSAMD_REQUIRE(occ >= 1, "batch cg: persistent kernel does not fit a compute unit");
  // Two workgroups per compute unit by default (NGSAMD_BATCH_PERSISTENT_WG:
  // another count, at most the occupancy).  A grid at the full occupancy
  // has once timed out at a barrier (a workgroup not resident while the
  // others waited); well below it every workgroup finds a slot even when
  // the dispatcher does not spread them evenly, and 2 to 8 per CU measured
  // the same time.
  const char* wenv = getenv("NGSAMD_BATCH_PERSISTENT_WG");
  occ = std::max(1, std::min(occ, wenv ? atoi(wenv) : 2));
  int64_t g = std::min<int64_t>((int64_t)occ * DeviceCUs(), std::max({P.nblocks, P.nchunks, P.nseg}));
  if (g > 8) g &= ~int64_t(7);
  if (g < 1) g = 1;
  PersistentArgs args = P;
  void* kargs[] = {&args};
  NGSAMD_HIP(hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)g), dim3(kBatchBlock), kargs, 0, st));
}

static void RunPersistent(BatchOp& A, const double* inv, double* u, double* d, double* w, double* s, BatchState S,
                          hipStream_t st) {
  DevBuf<double> cpart(std::max(A.nchunks, 1));
  DevBuf<int> err(1);
  DevBuf<unsigned long long> arrive(1);
  NGSAMD_HIP(hipMemsetAsync(err.p, 0, sizeof(int), st));
  NGSAMD_HIP(hipMemsetAsync(arrive.p, 0, sizeof(unsigned long long), st));
  PersistentArgs P{};
  P.rowptr = A.rowptr.p;
  P.cols = A.cols.p;
  P.vals = A.vals.p;
  P.blk_row = A.blk_row.p;
  P.run_row = A.run_row.p;
  P.run_ptr = A.run_ptr.p;
  P.val_ptr = A.val_ptr.p;
  P.rcols = A.rcols.p;
  P.rvals = A.rvals.p;
  P.blk_run = A.blk_run.p;
  P.blk_len = A.blk_len.p;
  P.blk_seg = A.blk_seg.p;
  P.seg_blk = A.seg_blk.p;
  P.nblocks = A.nblocks;
  P.cstart = A.chunk_start.p;
  P.cseg = A.chunk_seg.p;
  P.seg_chunk = A.seg_chunk.p;
  P.nchunks = A.nchunks;
  P.nseg = A.nseg;
  P.inv = inv;
  P.u = u;
  P.d = d;
  P.w = w;
  P.s = s;
  P.bpart = A.partials.p;
  P.cpart = cpart.p;
  P.S = S;
  P.arrive = arrive.p;
  P.err = err.p;
  if (A.nblocks > 0 || A.nchunks > 0) {
    if (A.use_runs) {
      if (A.block_nnz == 1024) LaunchPersistent<1024, true>(P, st);
      else if (A.block_nnz == 4096) LaunchPersistent<4096, true>(P, st);
      else LaunchPersistent<2048, true>(P, st);
    } else {
      if (A.block_nnz == 1024) LaunchPersistent<1024, false>(P, st);
      else if (A.block_nnz == 4096) LaunchPersistent<4096, false>(P, st);
      else LaunchPersistent<2048, false>(P, st);
    }
  }
  int h_err = 0;
  NGSAMD_HIP(hipMemcpyAsync(&h_err, err.p, sizeof(int), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  if (h_err) throw Error(NGSAMD_E_HIP, "batch cg: a grid barrier of the persistent launch timed out");
}

// Runs the batched CG from d = rhs (u = 0) until every active problem is
// done; returns the number of launched iterations.
template <class V, class M>
static long long RunBatchCg(BatchOp& A, const M* vals, const V* inv, const V* rhs, V* u, BatchWork& W,
                            const double* d_tol, double tol, const int* d_frozen, int maxiter, int chunk,
                            hipStream_t st) {
  const int64_t n = A.height;
  DevBuf<V> d(std::max<int64_t>(n, 1)), w(std::max<int64_t>(n, 1)), s(std::max<int64_t>(n, 1));
  DevBuf<int> all(1);
  BatchState S = W.state(maxiter);
  BatchResetKernel<<<GridFor(A.nseg, 256, 64), 256, 0, st>>>(S, A.nseg, d_tol, tol, d_frozen);
  NGSAMD_CHECK_LAUNCH();
  if (n) {
    NGSAMD_HIP(hipMemcpyAsync(d.p, rhs, n * sizeof(V), hipMemcpyDeviceToDevice, st));
    Fill<V>(u, n, (V)0, st);
  }
  BatchInitKernel<V><<<A.nchunks, kBatchBlock, 0, st>>>(inv, d.p, w.p, s.p, A.chunk_start.p, A.chunk_seg.p, S,
                                                         A.partials.p);
  NGSAMD_CHECK_LAUNCH();
  BatchInitCheckKernel<<<A.nseg, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, S);
  NGSAMD_CHECK_LAUNCH();
  if constexpr (sizeof(V) == 8 && sizeof(M) == 8) {
    if (chunk < 0) {
      RunPersistent(A, inv, u, d.p, w.p, s.p, S, st);
      return 0;
    }
  } else {
    NGSAMD_REQUIRE(chunk >= 0, "batch cg: the persistent form is FP64 only");
  }
  BatchAllDoneKernel<<<1, kBatchBlock, 0, st>>>(S.done, A.nseg, all.p);
  NGSAMD_CHECK_LAUNCH();

  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = chunk > 0;
  if (use_graph) {
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      bool ok = true;
      try {
        for (int k = 0; k < chunk; k++) EnqueueBatchIteration<V, M>(A, vals, inv, u, d.p, w.p, s.p, S, all.p, st);
      } catch (...) {
        ok = false;
      }
      e = hipStreamEndCapture(st, &graph);
      if (!ok && e == hipSuccess) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (graph) (void)hipGraphDestroy(graph);
      graph = nullptr;
      exec = nullptr;
      use_graph = false;
    }
  }
  struct GraphGuard {
    hipGraph_t g;
    hipGraphExec_t x;
    ~GraphGuard() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard{graph, exec};
  int* pin = nullptr;
  NGSAMD_HIP(hipHostMalloc((void**)&pin, 2 * sizeof(int), hipHostMallocDefault));
  struct PinGuard {
    int* p;
    ~PinGuard() { (void)hipHostFree(p); }
  } pguard{pin};
  hipEvent_t ev[2];
  NGSAMD_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  NGSAMD_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() { (void)hipEventDestroy(e[0]); (void)hipEventDestroy(e[1]); }
  } eguard{ev};
  const int per = use_graph ? chunk : 1;
  long long launched = 0;
  auto enqueue = [&](int slot) {
    if (use_graph) NGSAMD_HIP(hipGraphLaunch(exec, st));
    else EnqueueBatchIteration<V, M>(A, vals, inv, u, d.p, w.p, s.p, S, all.p, st);
    launched += per;
    NGSAMD_HIP(hipMemcpyAsync(&pin[slot], all.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipEventRecord(ev[slot], st));
  };
  NGSAMD_HIP(hipMemcpyAsync(&pin[0], all.p, sizeof(int), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  if (!pin[0]) {
    enqueue(0);
    enqueue(1);
    int slot = 0;
    const long long max_launches = (long long)maxiter / per + 4;
    for (long long k = 0; k < max_launches; k++) {
      NGSAMD_HIP(hipEventSynchronize(ev[slot]));
      if (pin[slot]) break;
      enqueue(slot);
      slot ^= 1;
    }
  }
  NGSAMD_HIP(hipStreamSynchronize(st));
  return launched;
}

__global__ void ToF32ScaledSegKernel(const double* __restrict__ r, float* __restrict__ out,
                                     const int64_t* __restrict__ cstart, const int* __restrict__ cseg,
                                     const double* __restrict__ inv_norm, const int* __restrict__ frozen) {
  const int c = blockIdx.x;
  const int p = cseg[c];
  const double sc = frozen[p] ? 0.0 : inv_norm[p];
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += blockDim.x) out[i] = (float)(r[i] * sc);
}

__global__ void AddCorrectionSegKernel(double* __restrict__ u, const float* __restrict__ c32,
                                       const int64_t* __restrict__ cstart, const int* __restrict__ cseg,
                                       const double* __restrict__ norm, const int* __restrict__ frozen) {
  const int c = blockIdx.x;
  const int p = cseg[c];
  if (frozen[p]) return;
  const double sc = norm[p];
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += blockDim.x) u[i] += sc * (double)c32[i];
}

// per chunk: partial sum over free dofs of r^2
__global__ __launch_bounds__(kBatchBlock) void MaskedSqChunkKernel(const double* __restrict__ r,
                                                                   const double* __restrict__ inv,
                                                                   const int64_t* __restrict__ cstart,
                                                                   double* __restrict__ partials) {
  __shared__ double lds[kBatchBlock / 64];
  const int c = blockIdx.x;
  double acc = 0.0;
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += kBatchBlock)
    if (inv[i] != 0.0) acc += r[i] * r[i];
  StorePartial<kBatchBlock>(acc, partials, lds);
}

// per problem: out[p] = sqrt(sum of its chunk partials)
__global__ __launch_bounds__(kBatchBlock) void SegNormKernel(const double* __restrict__ partials,
                                                             const int* __restrict__ seg_chunk,
                                                             double* __restrict__ out) {
  __shared__ double lds[kBatchBlock / 64];
  const double v = SegmentSum(partials, seg_chunk, blockIdx.x, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(v);
}

// per problem: norm, convergence against tol*fnorm, inner tolerance of the next step
__global__ __launch_bounds__(kBatchBlock) void RefineCheckKernel(
    const double* __restrict__ partials, const int* __restrict__ seg_chunk, const double* __restrict__ fnorm,
    double tol, double inner_tol, int* __restrict__ frozen, double* __restrict__ rnorm, double* __restrict__ inv_rnorm,
    double* __restrict__ itol, double* __restrict__ relres, long long* __restrict__ finish, int* __restrict__ outer,
    int final_pass) {
  __shared__ double lds[kBatchBlock / 64];
  const int p = blockIdx.x;
  const double rn = sqrt(SegmentSum(partials, seg_chunk, p, lds));
  if (threadIdx.x == 0) {
    if (frozen[p]) return;
    const double fn = fnorm[p];
    relres[p] = fn > 0 ? rn / fn : 0.0;
    if (rn <= tol * fn || rn == 0.0 || final_pass) {
      frozen[p] = 1;
      finish[p] = DeviceClock();
      return;
    }
    outer[p] += 1;
    rnorm[p] = rn;
    inv_rnorm[p] = 1.0 / rn;
    double t = inner_tol;
    const double need = 0.3 * tol * fn / rn;
    if (need > t) t = need < 0.5 ? need : 0.5;
    itol[p] = t;
  }
}

// ------------------------------------------- refinement steps on the device
// Each problem runs its own FP64 refinement: when its inner FP32 CG has
// stopped (S.done[p] != 0) and it is not finished, the next outer block
// takes its step: u += ||r|| c, r = f - A u (FP64 SpMV over its rows only),
// the convergence test on ||r||_free, and either finishes it or restarts its
// inner CG on r / ||r|| with the tolerance still needed.  The outer block
// is part of the captured iteration graph, so problems leave the batch as
// soon as they converge.

// pending[p]: inner solve stopped, problem open; skip[p] = !pending[p] (the
// FP64 SpMV's skip flags); the inner iterations of the step are counted
__global__ void IrPendingKernel(const BatchState S, const int* __restrict__ frozen, int* __restrict__ pending,
                                int* __restrict__ skip, int* __restrict__ inner_sum, int nseg) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nseg; p += gridDim.x * blockDim.x) {
    const int pd = (!frozen[p] && S.done[p]) ? 1 : 0;
    pending[p] = pd;
    skip[p] = 1 - pd;
    if (pd) inner_sum[p] += S.it[p];
  }
}

// chunks of pending problems: u += ||r|| c; r = f
__global__ void IrCorrectKernel(double* __restrict__ u, const float* __restrict__ c32, double* __restrict__ r,
                                const double* __restrict__ f, const int64_t* __restrict__ cstart,
                                const int* __restrict__ cseg, const double* __restrict__ rnorm,
                                const int* __restrict__ pending) {
  const int c = blockIdx.x;
  const int p = cseg[c];
  if (!pending[p]) return;
  const double sc = rnorm[p];
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += blockDim.x) {
    u[i] += sc * (double)c32[i];
    r[i] = f[i];
  }
}

// chunks of pending problems: partial sums over free dofs of r^2
__global__ __launch_bounds__(kBatchBlock) void IrSqKernel(const double* __restrict__ r, const double* __restrict__ inv,
                                                          const int64_t* __restrict__ cstart,
                                                          const int* __restrict__ cseg,
                                                          const int* __restrict__ pending,
                                                          double* __restrict__ partials) {
  __shared__ double lds[kBatchBlock / 64];
  const int c = blockIdx.x;
  if (!pending[cseg[c]]) {
    if (threadIdx.x == 0) partials[c] = 0.0;
    return;
  }
  double acc = 0.0;
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += kBatchBlock)
    if (inv[i] != 0.0) acc += r[i] * r[i];
  StorePartial<kBatchBlock>(acc, partials, lds);
}

// per pending problem: the outer test (converged, or max_outer inner solves
// done: finished), else the next inner tolerance and restart[p] = 1
__global__ __launch_bounds__(kBatchBlock) void IrOuterCheckKernel(
    const double* __restrict__ partials, const int* __restrict__ seg_chunk, const double* __restrict__ fnorm,
    double tol, double inner_tol, int max_outer, const int* __restrict__ pending, int* __restrict__ frozen,
    int* __restrict__ restart, double* __restrict__ rnorm, double* __restrict__ inv_rnorm, double* __restrict__ itol,
    double* __restrict__ relres, long long* __restrict__ finish, int* __restrict__ outer) {
  __shared__ double lds[kBatchBlock / 64];
  const int p = blockIdx.x;
  if (!pending[p]) {
    if (threadIdx.x == 0) restart[p] = 0;
    return;
  }
  const double rn = sqrt(SegmentSum(partials, seg_chunk, p, lds));
  if (threadIdx.x == 0) {
    const double fn = fnorm[p];
    relres[p] = fn > 0 ? rn / fn : 0.0;
    if (rn <= tol * fn || rn == 0.0 || outer[p] >= max_outer) {
      frozen[p] = 1;
      finish[p] = DeviceClock();
      restart[p] = 0;
      return;
    }
    outer[p] += 1;
    rnorm[p] = rn;
    inv_rnorm[p] = 1.0 / rn;
    // the last step only needs to gain the factor still missing
    double t = inner_tol;
    const double need = 0.3 * tol * fn / rn;
    if (need > t) t = need < 0.5 ? need : 0.5;
    itol[p] = t;
    restart[p] = 1;
  }
}

// chunks of restarted problems: the inner CG from c = 0 on d = r / ||r||
// (FP32): w = C d, s = w, partials of <w, d>
__global__ __launch_bounds__(kBatchBlock) void IrRestartKernel(const double* __restrict__ r,
                                                               const double* __restrict__ inv_rnorm,
                                                               const float* __restrict__ inv32,
                                                               float* __restrict__ c32, float* __restrict__ d,
                                                               float* __restrict__ w, float* __restrict__ s,
                                                               const int64_t* __restrict__ cstart,
                                                               const int* __restrict__ cseg,
                                                               const int* __restrict__ restart,
                                                               double* __restrict__ partials) {
  __shared__ double lds[kBatchBlock / 64];
  const int c = blockIdx.x;
  const int p = cseg[c];
  if (!restart[p]) {
    if (threadIdx.x == 0) partials[c] = 0.0;
    return;
  }
  const double sc = inv_rnorm[p];
  double acc = 0.0;
  for (int64_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += kBatchBlock) {
    const float di = (float)(r[i] * sc);
    const float wi = inv32[i] * di;
    c32[i] = 0.0f;
    d[i] = di;
    w[i] = wi;
    s[i] = wi;
    acc += (double)wi * (double)di;
  }
  StorePartial<kBatchBlock>(acc, partials, lds);
}

// per restarted problem: the inner CG's state and its first check
__global__ __launch_bounds__(kBatchBlock) void IrRestartCheckKernel(const double* __restrict__ partials,
                                                                    const int* __restrict__ seg_chunk,
                                                                    const int* __restrict__ restart,
                                                                    const double* __restrict__ itol, BatchState S) {
  __shared__ double lds[kBatchBlock / 64];
  const int p = blockIdx.x;
  if (!restart[p]) return;
  const double wdn = SegmentSum(partials, seg_chunk, p, lds);
  if (threadIdx.x == 0) {
    S.it[p] = 0;
    S.done[p] = 0;
    S.tol[p] = itol[p];
    S.finish[p] = 0;
    S.wdn[p] = wdn;
    BatchCheck(S, p, sqrt(fabs(wdn)));
  }
}

// all problems finished?  one workgroup
__global__ __launch_bounds__(kBatchBlock) void IrAllFrozenKernel(const int* __restrict__ frozen, int nseg,
                                                                 int* __restrict__ all) {
  __shared__ double lds[kBatchBlock / 64];
  double open = 0.0;
  for (int p = threadIdx.x; p < nseg; p += kBatchBlock) open += frozen[p] ? 0.0 : 1.0;
  open = BlockSum<kBatchBlock>(open, lds);
  if (threadIdx.x == 0) *all = open == 0.0 ? 1 : 0;
}

struct IrDeviceWork {
  DevBuf<double> r, fnorm, rnorm, inv_rnorm, itol, relres;
  DevBuf<float> inv32, c32, d, w, s;
  // all: every problem finished; inner_all: the inner iterations' own
  // all-stopped flag (not used by the refinement)
  DevBuf<int> frozen, pending, skip, restart, outer, inner_sum, all, inner_all;
  DevBuf<long long> finish;
  IrDeviceWork(int64_t n, int ns)
      : r(n), fnorm(ns), rnorm(ns), inv_rnorm(ns), itol(ns), relres(ns), inv32(n), c32(n), d(n), w(n), s(n),
        frozen(ns), pending(ns), skip(ns), restart(ns), outer(ns), inner_sum(ns), all(1), inner_all(1), finish(ns) {}
};

// one outer block: the steps of every problem whose inner solve stopped
static void EnqueueOuterBlock(BatchOp& A, const double* f, double* u, const double* inv, IrDeviceWork& W,
                              BatchState S, double tol, double inner_tol, int max_outer, hipStream_t st) {
  const int ns = A.nseg;
  IrPendingKernel<<<GridFor(ns, 256, 64), 256, 0, st>>>(S, W.frozen.p, W.pending.p, W.skip.p, W.inner_sum.p, ns);
  NGSAMD_CHECK_LAUNCH();
  IrCorrectKernel<<<A.nchunks, 256, 0, st>>>(u, W.c32.p, W.r.p, f, A.chunk_start.p, A.chunk_seg.p, W.rnorm.p,
                                             W.pending.p);
  NGSAMD_CHECK_LAUNCH();
  LaunchBatchSpmv<double, double, true, false>(A, A.vals.p, u, W.r.p, -1.0, W.skip.p, st, A.block_nnz);
  IrSqKernel<<<A.nchunks, kBatchBlock, 0, st>>>(W.r.p, inv, A.chunk_start.p, A.chunk_seg.p, W.pending.p,
                                                A.partials.p);
  NGSAMD_CHECK_LAUNCH();
  IrOuterCheckKernel<<<ns, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, W.fnorm.p, tol, inner_tol, max_outer,
                                                 W.pending.p, W.frozen.p, W.restart.p, W.rnorm.p, W.inv_rnorm.p,
                                                 W.itol.p, W.relres.p, W.finish.p, W.outer.p);
  NGSAMD_CHECK_LAUNCH();
  IrRestartKernel<<<A.nchunks, kBatchBlock, 0, st>>>(W.r.p, W.inv_rnorm.p, W.inv32.p, W.c32.p, W.d.p, W.w.p, W.s.p,
                                                     A.chunk_start.p, A.chunk_seg.p, W.restart.p, A.partials.p);
  NGSAMD_CHECK_LAUNCH();
  IrRestartCheckKernel<<<ns, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, W.restart.p, W.itol.p, S);
  NGSAMD_CHECK_LAUNCH();
  IrAllFrozenKernel<<<1, kBatchBlock, 0, st>>>(W.frozen.p, ns, W.all.p);
  NGSAMD_CHECK_LAUNCH();
}

// Inner FP32 iterations and outer blocks of all problems, replayed from a
// graph of kIrGraphIters inner iterations with an outer block after every
// kIrOuterEvery; the host polls the all-finished flag once per replay.
constexpr int kIrOuterEvery = 8;
constexpr int kIrGraphIters = 16;

static void RunBatchIrDevice(BatchOp& A, const double* inv, const double* f, double* u, double tol,
                             double inner_tol, int max_outer, int max_inner, IrDeviceWork& W, BatchWork& B,
                             hipStream_t st) {
  const int64_t n = A.height;
  const int ns = A.nseg;
  BatchState S = B.state(max_inner);
  NGSAMD_REQUIRE(ngsamd_vec_f64_to_f32(n, inv, W.inv32.p) == 0, "ir: conversion failed");
  Fill<int>(W.frozen.p, ns, 0, st);
  Fill<int>(W.outer.p, ns, 0, st);
  Fill<int>(W.inner_sum.p, ns, 0, st);
  Fill<long long>(W.finish.p, ns, 0LL, st);
  Fill<double>(W.rnorm.p, ns, 0.0, st);
  Fill<float>(W.c32.p, n, 0.0f, st);
  Fill<double>(u, n, 0.0, st);
  // every inner solve "stopped" with zero iterations: the first outer block
  // forms r = f and starts the first inner solves
  BatchResetKernel<<<GridFor(ns, 256, 64), 256, 0, st>>>(S, ns, nullptr, 0.0, nullptr);
  NGSAMD_CHECK_LAUNCH();
  Fill<int>(S.done, ns, 1, st);
  // ||f||_free per problem
  MaskedSqChunkKernel<<<A.nchunks, kBatchBlock, 0, st>>>(f, inv, A.chunk_start.p, A.partials.p);
  NGSAMD_CHECK_LAUNCH();
  SegNormKernel<<<ns, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, W.fnorm.p);
  NGSAMD_CHECK_LAUNCH();
  EnqueueOuterBlock(A, f, u, inv, W, S, tol, inner_tol, max_outer, st);

  auto enqueue_iterations = [&](int count) {
    for (int k = 0; k < count; k++) {
      EnqueueBatchIteration<float, float>(A, A.vals32.p, W.inv32.p, W.c32.p, W.d.p, W.w.p, W.s.p, S,
                                          W.inner_all.p, st);
      if ((k + 1) % kIrOuterEvery == 0) EnqueueOuterBlock(A, f, u, inv, W, S, tol, inner_tol, max_outer, st);
    }
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  {
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      bool ok = true;
      try {
        enqueue_iterations(kIrGraphIters);
      } catch (...) {
        ok = false;
      }
      e = hipStreamEndCapture(st, &graph);
      if (!ok && e == hipSuccess) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (graph) (void)hipGraphDestroy(graph);
      graph = nullptr;
      exec = nullptr;
    }
  }
  struct GraphGuard {
    hipGraph_t g;
    hipGraphExec_t x;
    ~GraphGuard() {
      if (x) (void)hipGraphExecDestroy(x);
      if (g) (void)hipGraphDestroy(g);
    }
  } guard{graph, exec};
  // every problem finishes after at most max_outer inner solves of at most
  // max_inner iterations each; an outer block at least every kIrOuterEvery
  const long long max_replays =
      ((long long)(max_outer + 1) * (max_inner + kIrOuterEvery)) / kIrGraphIters + 4;
  int done = 0;
  for (long long k = 0; k < max_replays; k++) {
    if (exec) NGSAMD_HIP(hipGraphLaunch(exec, st));
    else enqueue_iterations(kIrGraphIters);
    NGSAMD_HIP(hipMemcpyAsync(&done, W.all.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipStreamSynchronize(st));
    if (done) break;
  }
  NGSAMD_REQUIRE(done, "batch_ir_cg: refinement did not terminate (internal error)");
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_batch_csr_create(ngsamd_op_t* op, int64_t n, const int64_t* rp, const int* cols,
                                       const double* vals, int nseg, const int64_t* seg, int with_f32) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(n >= 0 && nseg >= 0, "ngsamd_batch_csr_create: negative size");
    NGSAMD_REQUIRE(op && rp && seg && (rp[n] == 0 || (cols && vals)), "ngsamd_batch_csr_create: null argument");
    *op = nullptr;
    *op = new BatchOp(n, rp, cols, vals, nseg, seg, with_f32 != 0);
  });
}

extern "C" int ngsamd_batch_info(ngsamd_op_t op, int* nseg, int* block_nnz, int* nblocks) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && op->kind == OpKind::Batch, "ngsamd_batch_info: not a batched operator");
    auto* A = static_cast<BatchOp*>(op);
    if (nseg) *nseg = A->nseg;
    if (block_nnz) *block_nnz = A->block_nnz;
    if (nblocks) *nblocks = A->nblocks;
  });
}

extern "C" int ngsamd_batch_tuning(ngsamd_op_t op, int* h_count, double* h_ms) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && op->kind == OpKind::Batch && h_count, "ngsamd_batch_tuning: not a batched operator");
    auto* A = static_cast<BatchOp*>(op);
    *h_count = (int)A->tuned_ms.size();
    if (h_ms)
      for (size_t k = 0; k < A->tuned_ms.size(); k++) h_ms[k] = A->tuned_ms[k];
  });
}

static void ReadClock(const long long* d_finish, long long start, int nseg, double* h_ms, hipStream_t st) {
  std::vector<long long> fin(nseg);
  NGSAMD_HIP(hipMemcpyAsync(fin.data(), d_finish, nseg * sizeof(long long), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  int dev = 0, khz = 100000;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
  for (int p = 0; p < nseg; p++) h_ms[p] = fin[p] > 0 ? double(fin[p] - start) / double(khz) : -1.0;
}

__global__ void ClockKernel(long long* out) { *out = DeviceClock(); }

static long long StartClock(hipStream_t st) {
  DevBuf<long long> c(1);
  ClockKernel<<<1, 1, 0, st>>>(c.p);
  NGSAMD_CHECK_LAUNCH();
  long long h = 0;
  NGSAMD_HIP(hipMemcpyAsync(&h, c.p, sizeof(long long), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  return h;
}

extern "C" int ngsamd_batch_cg_solve(ngsamd_op_t op, ngsamd_op_t pre, const double* f, double* u, double tol,
                                     int maxiter, int chunk, int* h_iterations, double* h_final_res,
                                     double* h_finish_ms) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::Batch, "ngsamd_batch_cg_solve: not a batched operator");
    NGSAMD_REQUIRE(pre && pre->kind == OpKind::Jacobi && pre->height == op->height,
                   "ngsamd_batch_cg_solve: needs a Jacobi preconditioner of the same size");
    NGSAMD_REQUIRE(maxiter >= 1, "ngsamd_batch_cg_solve: maxiter must be >= 1");
    NGSAMD_REQUIRE(chunk >= 0 || chunk == NGSAMD_BATCH_PERSISTENT,
                   "ngsamd_batch_cg_solve: chunk must be >= 0 or NGSAMD_BATCH_PERSISTENT");
    auto& A = *static_cast<BatchOp*>(op);
    hipStream_t st = Stream();
    const double* inv = static_cast<JacobiOp*>(pre)->invdiag.p;
    BatchWork W(A.nseg);
    const long long t0 = StartClock(st);
    RunBatchCg<double, double>(A, A.vals.p, inv, f, u, W, nullptr, tol, nullptr, maxiter, chunk, st);
    if (h_iterations) {
      NGSAMD_HIP(hipMemcpyAsync(h_iterations, W.it.p, A.nseg * sizeof(int), hipMemcpyDeviceToHost, st));
    }
    if (h_final_res) {
      NGSAMD_HIP(hipMemcpyAsync(h_final_res, W.last_res.p, A.nseg * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    NGSAMD_HIP(hipStreamSynchronize(st));
    if (h_finish_ms) ReadClock(W.finish.p, t0, A.nseg, h_finish_ms, st);
  });
}

extern "C" int ngsamd_batch_ir_cg_solve(ngsamd_op_t op, ngsamd_op_t pre, const double* f, double* u, double tol,
                                        double inner_tol, int max_outer, int max_inner, int* h_outer,
                                        int* h_inner, double* h_relres, double* h_finish_ms) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::Batch, "ngsamd_batch_ir_cg_solve: not a batched operator");
    auto& A = *static_cast<BatchOp*>(op);
    NGSAMD_REQUIRE(A.has_f32, "ngsamd_batch_ir_cg_solve: operator created without FP32 values");
    NGSAMD_REQUIRE(pre && pre->kind == OpKind::Jacobi && pre->height == op->height,
                   "ngsamd_batch_ir_cg_solve: needs a Jacobi preconditioner of the same size");
    hipStream_t st = Stream();
    const int64_t n = A.height;
    const int ns = A.nseg;
    const double* inv = static_cast<JacobiOp*>(pre)->invdiag.p;
    NGSAMD_REQUIRE(max_outer >= 1 && max_inner >= 1, "ngsamd_batch_ir_cg_solve: iteration limits must be >= 1");
    const char* henv = getenv("NGSAMD_BATCH_IR_HOST");
    if (!(henv && std::string(henv) == "1")) {
      // every problem takes its refinement steps on the device
      IrDeviceWork W(std::max<int64_t>(n, 1), ns);
      BatchWork B(ns);
      const long long t0 = StartClock(st);
      RunBatchIrDevice(A, inv, f, u, tol, inner_tol, max_outer, max_inner, W, B, st);
      if (h_outer) NGSAMD_HIP(hipMemcpyAsync(h_outer, W.outer.p, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      if (h_inner) NGSAMD_HIP(hipMemcpyAsync(h_inner, W.inner_sum.p, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      if (h_relres) NGSAMD_HIP(hipMemcpyAsync(h_relres, W.relres.p, ns * sizeof(double), hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      if (h_finish_ms) ReadClock(W.finish.p, t0, ns, h_finish_ms, st);
      return;
    }
    // NGSAMD_BATCH_IR_HOST=1: the outer steps of all problems together, from
    // the host (every problem finishes at the same outer check)
    DevBuf<float> inv32(std::max<int64_t>(n, 1)), r32(std::max<int64_t>(n, 1)), c32(std::max<int64_t>(n, 1));
    DevBuf<double> r(std::max<int64_t>(n, 1));
    NGSAMD_REQUIRE(ngsamd_vec_f64_to_f32(n, inv, inv32.p) == 0, "ir: conversion failed");
    DevBuf<int> frozen(ns), outer(ns), inner_sum(ns);
    DevBuf<double> fnorm(ns), rnorm(ns), inv_rnorm(ns), itol(ns), relres(ns);
    DevBuf<long long> finish(ns);
    Fill<int>(frozen.p, ns, 0, st);
    Fill<int>(outer.p, ns, 0, st);
    Fill<long long>(finish.p, ns, 0LL, st);
    std::vector<int> hin(ns, 0), hit(ns);
    BatchWork W(ns);
    const long long t0 = StartClock(st);
    // ||f||_free per problem
    MaskedSqChunkKernel<<<A.nchunks, kBatchBlock, 0, st>>>(f, inv, A.chunk_start.p, A.partials.p);
    NGSAMD_CHECK_LAUNCH();
    SegNormKernel<<<ns, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, fnorm.p);
    NGSAMD_CHECK_LAUNCH();
    Fill<double>(u, n, 0.0, st);
    if (n) NGSAMD_HIP(hipMemcpyAsync(r.p, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    for (int k = 0; k <= max_outer; k++) {
      MaskedSqChunkKernel<<<A.nchunks, kBatchBlock, 0, st>>>(r.p, inv, A.chunk_start.p, A.partials.p);
      NGSAMD_CHECK_LAUNCH();
      RefineCheckKernel<<<ns, kBatchBlock, 0, st>>>(A.partials.p, A.seg_chunk.p, fnorm.p, tol, inner_tol, frozen.p,
                                                    rnorm.p, inv_rnorm.p, itol.p, relres.p, finish.p, outer.p,
                                                    k == max_outer ? 1 : 0);
      NGSAMD_CHECK_LAUNCH();
      std::vector<int> hf(ns);
      NGSAMD_HIP(hipMemcpyAsync(hf.data(), frozen.p, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      if (std::all_of(hf.begin(), hf.end(), [](int v) { return v != 0; })) break;
      // inner FP32 solve of A c = r/||r|| for the problems still open
      ToF32ScaledSegKernel<<<A.nchunks, 256, 0, st>>>(r.p, r32.p, A.chunk_start.p, A.chunk_seg.p, inv_rnorm.p,
                                                      frozen.p);
      NGSAMD_CHECK_LAUNCH();
      RunBatchCg<float, float>(A, A.vals32.p, inv32.p, r32.p, c32.p, W, itol.p, 0.0, frozen.p, max_inner, 16, st);
      NGSAMD_HIP(hipMemcpyAsync(hit.data(), W.it.p, ns * sizeof(int), hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      for (int p = 0; p < ns; p++)
        if (!hf[p]) hin[p] += hit[p];
      AddCorrectionSegKernel<<<A.nchunks, 256, 0, st>>>(u, c32.p, A.chunk_start.p, A.chunk_seg.p, rnorm.p, frozen.p);
      NGSAMD_CHECK_LAUNCH();
      // r = f - A u  (FP64)
      if (n) NGSAMD_HIP(hipMemcpyAsync(r.p, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      A.MultAdd(-1.0, u, r.p, st);
    }
    if (h_outer) NGSAMD_HIP(hipMemcpyAsync(h_outer, outer.p, ns * sizeof(int), hipMemcpyDeviceToHost, st));
    if (h_relres) NGSAMD_HIP(hipMemcpyAsync(h_relres, relres.p, ns * sizeof(double), hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipStreamSynchronize(st));
    if (h_inner)
      for (int p = 0; p < ns; p++) h_inner[p] = hin[p];
    if (h_finish_ms) ReadClock(finish.p, t0, ns, h_finish_ms, st);
  });
}
