// CSR sparse matrix-vector product on gfx950.
//
// Replaces SparseMatrix<double>::MultAdd (linalg/sparsematrix_impl.hpp:264-306,
// row kernel RowTimesVector linalg/sparsematrix.hpp:625-632) and the
// reference's cuSPARSE call (ngscuda/cuda_linalg.cpp:243-275).
//
// Three kernels:
//
// runs (default when it saves at least a quarter of the column numbers): up
// to 8 consecutive rows with the same column list (the dofs of one edge, face
// or cell of an H1 space) form a run that stores the list once and its values
// column by column.  Values still stream once (8 B per nonzero) but the column
// numbers drop to 4 B per run entry and each gathered x entry serves the
// whole run.  Blocks hold runs of one length and at most 2048 values; the
// block loop is the stream kernel's.
//
// stream: int32 row pointers, int32 column numbers, FP64 values (12 B per
// nonzero streamed once); the rows are cut on the host into blocks of consecutive
// rows holding at most 2048 nonzeros (NGSAMD_SPMV_NNZ=1024/4096 selects other
// sizes; a single longer row is a block of its own).  A 256-thread workgroup streams a block's column and value
// ranges with fully coalesced non-temporal loads (they are read once and
// must not evict x from L2), multiplies by the gathered x entries into LDS,
// and then reduces each row from LDS with a power-of-two group of lanes sized
// to the block's row count.  Every lane is busy whatever the row lengths.
//
// rowgroup: a group of G lanes works on one row, lane l of the group reading
// nonzeros l, l+G, ...; the partial sums are combined with cross-lane
// shuffles.  Kept for comparison (NGSAMD_SPMV=rowgroup).
//
// All three walk contiguous row ranges that are grouped per XCD (workgroup i runs
// on XCD i mod 8), so the x entries a row range gathers stay in one XCD's L2.
#include <algorithm>
#include <cstdlib>
#include <string>
#include <type_traits>

#include "common.hpp"
#include "csr.hpp"
#include "runs.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kSpmvBlock = 256;

// y[r] = (ACCUM ? y[r] : 0) + s * (A x)[r]; optionally the per-workgroup
// partial sums of x[r]*(A x)[r] (the <s, A s> of CG) into partials[].
template <int G, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kSpmvBlock) void CsrSpmvKernel(
    const int* __restrict__ rowptr, const int* __restrict__ cols, const double* __restrict__ vals,
    const double* __restrict__ x, double* __restrict__ y, int n, double s, int iters_per_wg,
    double* __restrict__ partials, const int* __restrict__ stop_flag) {
  if (stop_flag && *stop_flag) return;
  constexpr int RPB = kSpmvBlock / G;  // rows per workgroup pass
  const int lb = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int g = threadIdx.x / G;
  const int l = threadIdx.x % G;
  double dotacc = 0.0;
  const int row_begin = lb * iters_per_wg * RPB;
  for (int it = 0; it < iters_per_wg; it++) {
    const int r = row_begin + it * RPB + g;
    if (r < n) {
      const int k0 = rowptr[r], k1 = rowptr[r + 1];
      double a0 = 0.0, a1 = 0.0;
      int k = k0 + l;
      for (; k + G < k1; k += 2 * G) {
        const int c0 = cols[k], c1 = cols[k + G];
        const double v0 = vals[k], v1 = vals[k + G];
        a0 += v0 * x[c0];
        a1 += v1 * x[c1];
      }
      if (k < k1) a0 += vals[k] * x[cols[k]];
      double sum = a0 + a1;
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, G);
      if (l == 0) {
        const double val = s * sum;
        if (ACCUM) y[r] += val; else y[r] = val;
        if (DOT) dotacc += x[r] * sum;
      }
    }
  }
  if (DOT) {
    __shared__ double lds[kSpmvBlock / 64];
    StorePartial<kSpmvBlock>(dotacc, partials, lds);
  }
}


constexpr int kStreamBlock = 256;

// y[r] = (ACCUM ? y[r] : 0) + s * (A x)[r] over the row blocks
// [blk_row[b], blk_row[b+1]) of at most NNZ nonzeros; DOT as in CsrSpmvKernel.
template <int NNZ, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kStreamBlock) void CsrStreamKernel(
    const int* __restrict__ rowptr, const int* __restrict__ cols, const double* __restrict__ vals,
    const double* __restrict__ x, double* __restrict__ y, const int* __restrict__ blk_row, int nblocks,
    int blocks_per_wg, double s, double* __restrict__ partials, const int* __restrict__ stop_flag) {
  if (stop_flag && *stop_flag) return;
  __shared__ double prod[NNZ];
  __shared__ double lds[kStreamBlock / 64];
  const int tid = threadIdx.x;
  const int lb = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int b0 = lb * blocks_per_wg;
  const int b1 = min(b0 + blocks_per_wg, nblocks);
  double dotacc = 0.0;
  for (int b = b0; b < b1; b++) {
    const int r0 = blk_row[b], r1 = blk_row[b + 1];
    const int k0 = rowptr[r0], k1 = rowptr[r1];
    const int nz = k1 - k0;
    if (nz > NNZ) {
      // one long row: the whole workgroup strides over it
      double a = 0.0;
      for (int k = k0 + tid; k < k1; k += kStreamBlock)
        a += __builtin_nontemporal_load(&vals[k]) * x[__builtin_nontemporal_load(&cols[k])];
      a = BlockSum<kStreamBlock>(a, lds);
      if (tid == 0) {
        if (ACCUM) y[r0] += s * a; else y[r0] = s * a;
        if (DOT) dotacc += x[r0] * a;
      }
      continue;
    }
    // every load is unconditional (lanes past the block read entry k0,
    // which the padding keeps in bounds, and contribute 0) so all of them
    // are in flight at once instead of one round trip per branch
    constexpr int PER = NNZ / kStreamBlock;
    int c[PER];
    double v[PER], xv[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kStreamBlock;
      const int kk = k0 + (k < nz ? k : 0);
      c[i] = __builtin_nontemporal_load(&cols[kk]);
      v[i] = __builtin_nontemporal_load(&vals[kk]);
    }
#pragma unroll
    for (int i = 0; i < PER; i++) xv[i] = x[c[i]];
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kStreamBlock;
      prod[k] = k < nz ? v[i] * xv[i] : 0.0;
    }
    __syncthreads();
    const int nr = r1 - r0;
    int tpr = 64;
    while (tpr > 1 && tpr * nr > kStreamBlock) tpr >>= 1;
    const int l = tid & (tpr - 1);
    for (int j0 = 0; j0 < nr; j0 += kStreamBlock / tpr) {
      const int j = j0 + tid / tpr;
      double sum = 0.0;
      if (j < nr) {
        const int s0 = rowptr[r0 + j] - k0, s1 = rowptr[r0 + j + 1] - k0;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = s0 + l;
        for (; k + 3 * tpr < s1; k += 4 * tpr) {
          a0 += prod[k];
          a1 += prod[k + tpr];
          a2 += prod[k + 2 * tpr];
          a3 += prod[k + 3 * tpr];
        }
        for (; k < s1; k += tpr) a0 += prod[k];
        sum = (a0 + a1) + (a2 + a3);
      }
      for (int off = tpr / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (j < nr && l == 0) {
        if (ACCUM) y[r0 + j] += s * sum; else y[r0 + j] = s * sum;
        if (DOT) dotacc += x[r0 + j] * sum;
      }
    }
    __syncthreads();
  }
  if (DOT) StorePartial<kStreamBlock>(dotacc, partials, lds);
}


// y = (ACCUM ? y : 0) + s * A x over the run blocks [blk_run[b], blk_run[b+1])
// with blk_len[b] rows per run (at most NNZ values and NNZ rows per block);
// DOT as in CsrSpmvKernel.  Block b: lane k of its value range forms value k
// times x[column of shared slot k / R] into LDS (values stream coalesced, the
// R lanes of one slot read one column number and one x entry); row j = run R
// + q then sums the slots of its run at stride R.  A block holding more
// than NNZ values is a single row (R = 1) that the whole workgroup strides.
// NNZ <= 2048: at most 64 VGPRs, so 8 workgroups fit a CU (72 registers
// allowed 7; 394 vs 398 us at 32^3 p=4).
// Consecutive plain applies (Mult / MultAdd) alternate the direction of the
// block walk (reverse), so that each apply starts on the blocks the previous
// one read last, which the 256 MB Infinity Cache still holds; their values
// are read with plain loads for that (non-temporal ones are not kept):
// 374-375 against 383-384 us per apply at 32^3 p=4 (forward only with plain
// loads: 395).  Every row is summed by the same code in the same order
// either way.  Inside CG (DOT) the vector kernels between two applies take
// that cache, so the CG variant keeps the forward walk and non-temporal
// value loads (with the alternating walk and plain loads Jacobi-CG at 32^3
// p=4 took 0.142 against 0.140 s, and with the sparse Cholesky coarse solve
// 0.184 against 0.173 s).
template <int NNZ, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kStreamBlock, NNZ <= 2048 ? 8 : 4) void CsrRunsKernel(
    const int* __restrict__ run_row, const int* __restrict__ run_ptr, const int* __restrict__ val_ptr,
    const int* __restrict__ rcols, const double* __restrict__ rvals, const double* __restrict__ x,
    double* __restrict__ y, const int* __restrict__ blk_run, const int* __restrict__ blk_len, int nblocks,
    int blocks_per_wg, double s, double* __restrict__ partials, const int* __restrict__ stop_flag, int reverse) {
  static_assert(NNZ <= 4096, "RunDiv is exact below 4096");
  if (stop_flag && *stop_flag) return;
  constexpr int PER = NNZ / kStreamBlock;
  __shared__ double prod[NNZ];
  __shared__ double lds[kStreamBlock / 64];
  const int tid = threadIdx.x;
  // reverse: the first workgroups dispatched take the last blocks, each
  // workgroup walks its range backwards
  const int lx = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int lb = reverse ? (int)gridDim.x - 1 - lx : lx;
  const int b0 = lb * blocks_per_wg;
  const int b1 = min(b0 + blocks_per_wg, nblocks);
  auto blk = [&](int t) { return reverse ? b0 + b1 - 1 - t : t; };
  double dotacc = 0.0;

  // header and register-held loads of the block in flight
  int u0 = 0, u1 = 0, R = 1, inv = 65536, c0 = 0, vb = 0, nv = 0;
  int c[PER];
  double v[PER];
  auto fetch = [&](int b) {
    u0 = blk_run[b];
    u1 = blk_run[b + 1];
    R = blk_len[b];
    inv = (65536 + R - 1) / R;
    c0 = run_ptr[u0];
    vb = val_ptr[u0];
    nv = val_ptr[u1] - vb;
    if (nv > NNZ) return;  // long row: streamed when it is reached
    // unconditional loads (lanes past the block read entry 0 of it, kept in
    // bounds by the padding) so that every load is in flight at once
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kStreamBlock;
      const int kk = k < nv ? k : 0;
      c[i] = __builtin_nontemporal_load(&rcols[c0 + RunDiv(kk, inv)]);
      v[i] = DOT ? __builtin_nontemporal_load(&rvals[vb + kk]) : rvals[vb + kk];
    }
  };
  if (b0 < b1) fetch(blk(b0));
  for (int b = b0; b < b1; b++) {
    if (nv > NNZ) {
      // one long row: the whole workgroup strides over it
      double a = 0.0;
      for (int k = tid; k < nv; k += kStreamBlock)
        a += __builtin_nontemporal_load(&rvals[vb + k]) * x[__builtin_nontemporal_load(&rcols[c0 + k])];
      a = BlockSum<kStreamBlock>(a, lds);
      const int r = run_row[u0];
      if (tid == 0) {
        if (ACCUM) y[r] += s * a; else y[r] = s * a;
        if (DOT) dotacc += x[r] * a;
      }
      if (b + 1 < b1) fetch(blk(b + 1));
      continue;
    }
    {
      double xv[PER];
#pragma unroll
      for (int i = 0; i < PER; i++) xv[i] = x[c[i]];
#pragma unroll
      for (int i = 0; i < PER; i++) {
        const int k = tid + i * kStreamBlock;
        prod[k] = k < nv ? v[i] * xv[i] : 0.0;
      }
    }
    __syncthreads();
    // this block's header, then the next block's loads
    const int cu0 = u0, cR = R, cinv = inv, cc0 = c0;
    const int row0 = run_row[cu0];
    const int nr = (u1 - u0) * R;
    int tpr = 64;
    while (tpr > 1 && tpr * nr > kStreamBlock) tpr >>= 1;
    const int l = tid & (tpr - 1);
    for (int j0 = 0; j0 < nr; j0 += kStreamBlock / tpr) {
      const int j = j0 + tid / tpr;
      double sum = 0.0;
      if (j < nr) {
        const int run = RunDiv(j, cinv), q = j - run * cR;
        const int s0 = run_ptr[cu0 + run] - cc0, s1 = run_ptr[cu0 + run + 1] - cc0;
        // four independent partial sums so that the LDS reads overlap
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int k = s0 + l;
        for (; k + 3 * tpr < s1; k += 4 * tpr) {
          a0 += prod[k * cR + q];
          a1 += prod[(k + tpr) * cR + q];
          a2 += prod[(k + 2 * tpr) * cR + q];
          a3 += prod[(k + 3 * tpr) * cR + q];
        }
        for (; k < s1; k += tpr) a0 += prod[k * cR + q];
        sum = (a0 + a1) + (a2 + a3);
      }
      for (int off = tpr / 2; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
      if (j < nr && l == 0) {
        if (ACCUM) y[row0 + j] += s * sum; else y[row0 + j] = s * sum;
        if (DOT) dotacc += x[row0 + j] * sum;
      }
    }
    __syncthreads();
    if (b + 1 < b1) fetch(blk(b + 1));
  }
  if (DOT) StorePartial<kStreamBlock>(dotacc, partials, lds);
}

// Y = A X for K vectors stored interleaved (entry j of dof i at i K + j, the
// layout of a BaseVector with K components): each block's values and column
// numbers are loaded once and used for all K vectors; the x entries of G
// components of a value's column come in one pass (16-byte loads), then
// each component goes through the single kernel's products (16 KB of LDS)
// and row sums in the same order, so each column of Y is bit-identical to
// a single apply.
template <int NNZ, int G>
__global__ __launch_bounds__(kStreamBlock, NNZ <= 2048 && G == 1 ? 8 : 4) void CsrRunsMultiKernel(
    const int* __restrict__ run_row, const int* __restrict__ run_ptr, const int* __restrict__ val_ptr,
    const int* __restrict__ rcols, const double* __restrict__ rvals, const double* __restrict__ x,
    double* __restrict__ y, int K, const int* __restrict__ blk_run, const int* __restrict__ blk_len, int nblocks,
    int blocks_per_wg, const int* __restrict__ stop) {
  if (stop && *stop) return;
  constexpr int PER = NNZ / kStreamBlock;
  __shared__ double prod[NNZ];
  __shared__ double lds[kStreamBlock / 64];
  const int tid = threadIdx.x;
  const int lb = XcdContiguousBlock(blockIdx.x, gridDim.x);
  const int b0 = lb * blocks_per_wg;
  const int b1 = min(b0 + blocks_per_wg, nblocks);
  int u0 = 0, u1 = 0, R = 1, inv = 65536, c0 = 0, vb = 0, nv = 0;
  int c[PER];
  double v[PER];
  auto fetch = [&](int b) {
    u0 = blk_run[b];
    u1 = blk_run[b + 1];
    R = blk_len[b];
    inv = (65536 + R - 1) / R;
    c0 = run_ptr[u0];
    vb = val_ptr[u0];
    nv = val_ptr[u1] - vb;
    if (nv > NNZ) return;
#pragma unroll
    for (int i = 0; i < PER; i++) {
      const int k = tid + i * kStreamBlock;
      const int kk = k < nv ? k : 0;
      c[i] = __builtin_nontemporal_load(&rcols[c0 + RunDiv(kk, inv)]);
      v[i] = __builtin_nontemporal_load(&rvals[vb + kk]);
    }
  };
  if (b0 < b1) fetch(b0);
  for (int b = b0; b < b1; b++) {
    if (nv > NNZ) {
      const int r = run_row[u0];
      for (int j = 0; j < K; j++) {
        double a = 0.0;
        for (int k = tid; k < nv; k += kStreamBlock)
          a += rvals[vb + k] * x[(int64_t)rcols[c0 + k] * K + j];
        a = BlockSum<kStreamBlock>(a, lds);
        if (tid == 0) y[(int64_t)r * K + j] = a;
      }
      if (b + 1 < b1) fetch(b + 1);
      continue;
    }
    const int row0 = run_row[u0];
    const int nr = (u1 - u0) * R;
    int tpr = 64;
    while (tpr > 1 && tpr * nr > kStreamBlock) tpr >>= 1;
    const int l = tid & (tpr - 1);
    for (int j0 = 0; j0 < K; j0 += G) {
      // x of the G components of each value's column: loaded together
      double xg[PER][G];
#pragma unroll
      for (int i = 0; i < PER; i++) {
        const double* xp = x + (int64_t)c[i] * K + j0;
        if constexpr (G == 1) {
          xg[i][0] = xp[0];
        } else {
#pragma unroll
          for (int h = 0; h < G / 2; h++) {
            const double2 t = reinterpret_cast<const double2*>(xp)[h];
            xg[i][2 * h] = t.x;
            xg[i][2 * h + 1] = t.y;
          }
        }
      }
      // then, component by component, the single kernel's products and sums
#pragma unroll
      for (int g = 0; g < G; g++) {
#pragma unroll
        for (int i = 0; i < PER; i++) {
          const int k = tid + i * kStreamBlock;
          prod[k] = k < nv ? v[i] * xg[i][g] : 0.0;
        }
        __syncthreads();
        for (int jr = 0; jr < nr; jr += kStreamBlock / tpr) {
          const int j = jr + tid / tpr;
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
          if (j < nr && l == 0) y[(int64_t)(row0 + j) * K + j0 + g] = sum;
        }
        __syncthreads();
      }
    }
    if (b + 1 < b1) fetch(b + 1);
  }
}

// Component j of an interleaved K-vector to a plain vector and back (the
// other layouts apply the single-vector kernel per component).
__global__ void GatherComponentKernel(const double* __restrict__ x, int K, int j, double* __restrict__ out,
                                      int64_t n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = x[i * K + j];
}
__global__ void ScatterComponentKernel(const double* __restrict__ in, int K, int j, double* __restrict__ y,
                                       int64_t n, const int* __restrict__ stop) {
  if (stop && *stop) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i * K + j] = in[i];
}

template <bool ACCUM, bool DOT>
static int LaunchSpmv(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* partials,
                      const int* stop, int reverse = 0);

void CsrOp::MultMulti(int K, const double* x, double* y, hipStream_t st, const int* stop) {
  NGSAMD_REQUIRE(K >= 1 && K <= 64, "multi-vector apply: 1 <= K <= 64 vectors");
  if (height == 0) return;
  if (width == 0) {
    // no columns: Y = 0
    LaunchFill(y, height * K, 0.0, st);
    return;
  }
  if (algo == 2) {
    auto pick = [&](auto g) {
      constexpr int Gc = decltype(g)::value;
      return block_nnz == 1024 ? CsrRunsMultiKernel<1024, Gc>
             : block_nnz == 4096 ? CsrRunsMultiKernel<4096, Gc>
                                 : CsrRunsMultiKernel<2048, Gc>;
    };
    // x of four components per pass when K allows (two 16-byte loads per
    // value, more registers: 4 workgroups per CU); NGSAMD_MULTI_PASS=2 keeps
    // two per pass (for comparison)
    const char* pass_env = getenv("NGSAMD_MULTI_PASS");
    const bool four = K % 4 == 0 && !(pass_env && std::string(pass_env) == "2");
    auto kern = four ? pick(std::integral_constant<int, 4>{})
                : (K % 2 == 0) ? pick(std::integral_constant<int, 2>{})
                               : pick(std::integral_constant<int, 1>{});
    int wg = nblocks < max_grid ? nblocks : max_grid;
    if (wg > 8) wg &= ~7;
    if (wg < 1) wg = 1;
    const int per = (nblocks + wg - 1) / wg;
    kern<<<wg, kStreamBlock, 0, st>>>(run_row.p, run_ptr.p, val_ptr.p, rcols.p, rvals.p, x, y, K, blk_run.p,
                                      blk_len.p, nblocks, per, stop);
    NGSAMD_CHECK_LAUNCH();
    return;
  }
  if (xcomp.n < (size_t)width) xcomp.alloc(width);
  if (ycomp.n < (size_t)height) ycomp.alloc(height);
  for (int j = 0; j < K; j++) {
    GatherComponentKernel<<<GridFor(width, 256, 8192), 256, 0, st>>>(x, K, j, xcomp.p, width, stop);
    NGSAMD_CHECK_LAUNCH();
    LaunchSpmv<false, false>(*this, 1.0, xcomp.p, ycomp.p, st, nullptr, stop);
    ScatterComponentKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(ycomp.p, K, j, y, height, stop);
    NGSAMD_CHECK_LAUNCH();
  }
}

__global__ void RunsDiagKernel(const int* __restrict__ run_row, const int* __restrict__ run_ptr,
                               const int* __restrict__ val_ptr, const int* __restrict__ rcols,
                               const double* __restrict__ rvals, double* __restrict__ d, int nruns) {
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < nruns; u += gridDim.x * blockDim.x) {
    const int r0 = run_row[u], R = run_row[u + 1] - r0;
    const int c0 = run_ptr[u], L = run_ptr[u + 1] - c0;
    for (int q = 0; q < R; q++) {
      double val = 0.0;
      for (int j = 0; j < L; j++)
        if (rcols[c0 + j] == r0 + q) { val = rvals[val_ptr[u] + j * R + q]; break; }
      d[r0 + q] = val;
    }
  }
}

CsrOp::CsrOp(int64_t h, int64_t w, const int64_t* rp, const int* c, const double* v)
    : ngsamd_op_s(OpKind::CSR) {
  height = h;
  width = w;
  NGSAMD_REQUIRE(h >= 0 && w >= 0, "ngsamd_csr_create: negative size");
  NGSAMD_REQUIRE(h < (int64_t(1) << 31) && w < (int64_t(1) << 31), "ngsamd_csr_create: size exceeds int32");
  nnz = rp[h];
  NGSAMD_REQUIRE(rp[0] == 0 && nnz >= 0, "ngsamd_csr_create: row pointer must start at 0");
  NGSAMD_REQUIRE(nnz < (int64_t(1) << 31), "ngsamd_csr_create: nnz exceeds int32");
  std::vector<int> rp32(h + 1);
  for (int64_t i = 0; i <= h; i++) {
    if (i > 0) NGSAMD_REQUIRE(rp[i] >= rp[i - 1], "ngsamd_csr_create: row pointer not monotone");
    rp32[i] = (int)rp[i];
  }
  for (int64_t k = 0; k < nnz; k++)
    NGSAMD_REQUIRE(c[k] >= 0 && c[k] < w, "ngsamd_csr_create: column number out of range");
  for (int64_t i = 0; i < h; i++) max_row_len = std::max<int64_t>(max_row_len, rp[i + 1] - rp[i]);
  if (w == 0) {
    static const double z = 0.0;
    xpad.alloc(1);
    xpad.upload(&z, 1);
  }
  const char* nnz_env = getenv("NGSAMD_SPMV_NNZ");
  block_nnz = nnz_env ? atoi(nnz_env) : 0;
  NGSAMD_REQUIRE(block_nnz == 0 || block_nnz == 1024 || block_nnz == 2048 || block_nnz == 4096,
                 "NGSAMD_SPMV_NNZ must be 1024, 2048 or 4096");
  // grid of the stream/runs applies: 16384 measured best at 32^3 p=4 (runs
  // kernel 395 us; 8192 405, 4096 418, 65536 494)
  max_grid = 16384;
  const char* grid_env = getenv("NGSAMD_SPMV_GRID");
  if (grid_env) {
    max_grid = atoi(grid_env);
    NGSAMD_REQUIRE(max_grid >= 8 && max_grid <= 65536, "NGSAMD_SPMV_GRID must be 8..65536");
  }
  const char* algo_env = getenv("NGSAMD_SPMV");
  const std::string algo_name = algo_env ? algo_env : "";
  NGSAMD_REQUIRE(algo_name.empty() || algo_name == "rowgroup" || algo_name == "stream" || algo_name == "runs",
                 "NGSAMD_SPMV must be rowgroup, stream or runs");
  if (algo_name.empty() || algo_name == "runs") {
    // measured on the 32^3 p=4 matrix (us per apply, grid 16384 / 4096
    // workgroups): 2048 values per block 395 / 418, 4096 466 / 423
    const int saved = block_nnz;
    // below a million nonzeros (cache-resident, few blocks) 1024-value
    // blocks give twice the workgroups: 6.1 vs 7.1 us on the 64^2 p=3
    // triangle matrix (0.63M nonzeros)
    if (block_nnz == 0) block_nnz = nnz < (int64_t(1) << 20) ? 1024 : 2048;
    // the runs layout pays when it saves at least a quarter of the column
    // numbers (H1 matrices of order >= 2 share patterns between the dofs of
    // an edge, a face or a cell)
    std::vector<int> rr, rpp, vp, rc, br, bl;
    std::vector<double> rv;
    rc.reserve(nnz);
    rv.reserve(nnz);
    nshared = BuildRuns(h, rp, c, v, block_nnz, rr, rpp, vp, rc, rv, br, bl);
    if (algo_name == "runs" || 4 * nshared <= 3 * nnz) {
      algo = 2;
      nruns = (int64_t)rr.size() - 1;
      nblocks = (int)br.size() - 1;
      // Up to 4096 blocks: one block per workgroup (all resident at once;
      // 128^2 p=3, 1.2K blocks: 8.4 us against 12.1 us on 1024 workgroups).
      // Beyond: about three blocks per workgroup, so that each workgroup's
      // next block is in flight (16^3 p=4, 20M nonzeros: 57 us on 4096
      // workgroups against 78 us on one block per workgroup), at most 16384.
      if (!grid_env && nblocks > 4096) max_grid = std::max(4096, std::min(16384, nblocks / 3));
      run_row.alloc(rr.size());
      run_ptr.alloc(rpp.size());
      val_ptr.alloc(vp.size());
      // 4 padding entries (column 0, value 0): the kernel's lanes past the
      // end of the last block read entry nshared / nnz
      rcols.alloc(nshared + 4);
      rvals.alloc(nnz + 4);
      {
        const int zc[4] = {0, 0, 0, 0};
        const double zv[4] = {0, 0, 0, 0};
        rcols.upload_at(zc, nshared, 4);
        rvals.upload_at(zv, nnz, 4);
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
      }
      blk_run.alloc(br.size());
      blk_len.alloc(std::max<size_t>(bl.size(), 1));
      run_row.upload(rr.data(), rr.size());
      run_ptr.upload(rpp.data(), rpp.size());
      val_ptr.upload(vp.data(), vp.size());
      if (nshared) rcols.upload(rc.data(), nshared);
      if (nnz && v) rvals.upload(rv.data(), nnz);
      if (nnz && !v) NGSAMD_HIP(hipMemsetAsync(rvals.p, 0, sizeof(double) * nnz, Stream()));
      if (!v) {
        std::vector<int> rrun(std::max<int64_t>(h, 1), 0);
        for (int64_t u = 0; u < nruns; u++)
          for (int r = rr[u]; r < rr[u + 1]; r++) rrun[r] = (int)u;
        row_run.alloc(rrun.size());
        row_run.upload(rrun.data(), rrun.size());
        NGSAMD_HIP(hipStreamSynchronize(Stream()));
      }
      blk_run.upload(br.data(), br.size());
      if (!bl.empty()) blk_len.upload(bl.data(), bl.size());
      NGSAMD_HIP(hipStreamSynchronize(Stream()));
      return;
    }
    block_nnz = saved;
  }
  // stream kernel: 2048 values per block (443.8 us on the 32^3 p=4 matrix,
  // 4096 443.4 us, 1024 470.7 us)
  if (block_nnz == 0) block_nnz = 2048;
  rowptr.alloc(h + 1);
  // 4 padding entries (column 0, value 0) behind the last nonzero: lanes
  // past the last block read entry nnz
  cols.alloc(nnz + 4);
  vals.alloc(nnz + 4);
  rowptr.upload(rp32.data(), h + 1);
  cols.upload(c, nnz);
  if (v) vals.upload(v, nnz);
  else if (nnz) NGSAMD_HIP(hipMemsetAsync(vals.p, 0, sizeof(double) * nnz, Stream()));
  {
    const int zc[4] = {0, 0, 0, 0};
    const double zv[4] = {0, 0, 0, 0};
    NGSAMD_HIP(hipMemcpyAsync(cols.p + nnz, zc, sizeof(zc), hipMemcpyHostToDevice, Stream()));
    NGSAMD_HIP(hipMemcpyAsync(vals.p + nnz, zv, sizeof(zv), hipMemcpyHostToDevice, Stream()));
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  }
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  double mean = h ? double(nnz) / double(h) : 0.0;
  group = mean > 48 ? 32 : mean > 20 ? 16 : mean > 9 ? 8 : 4;
  // row blocks of at most block_nnz nonzeros (or one longer row)
  std::vector<int> br;
  br.push_back(0);
  int64_t acc = 0;
  for (int64_t r = 0; r < h; r++) {
    const int64_t len = rp[r + 1] - rp[r];
    if (r > br.back() && acc + len > block_nnz) {
      br.push_back((int)r);
      acc = 0;
    }
    acc += len;
  }
  if (br.back() != h || h == 0) br.push_back((int)h);
  nblocks = (int)br.size() - 1;
  // about two blocks per workgroup (P1 matrices: 1024^2 triangles, 3.6K
  // blocks, 24.7 us on 2048 workgroups against 39.1 us on one block each;
  // 96^3 tets, 6.6K blocks, 45 vs 64 us on 4096), 1024 to 16384 workgroups
  if (!grid_env && nblocks > 1024) max_grid = std::max(1024, std::min(16384, nblocks / 2));
  blk_row.alloc(br.size());
  blk_row.upload(br.data(), br.size());
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  algo = algo_name == "rowgroup" ? 0 : 1;
}

void CsrOp::Values(double* host_vals) const {
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  if (nnz == 0) return;
  if (algo != 2) {
    NGSAMD_HIP(hipMemcpy(host_vals, vals.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    return;
  }
  // runs layout: run u holds rows run_row[u].. with values [L][R]
  std::vector<int> rr(nruns + 1), vp(nruns + 1);
  std::vector<double> rv(nnz);
  NGSAMD_HIP(hipMemcpy(rr.data(), run_row.p, sizeof(int) * (nruns + 1), hipMemcpyDeviceToHost));
  NGSAMD_HIP(hipMemcpy(vp.data(), val_ptr.p, sizeof(int) * (nruns + 1), hipMemcpyDeviceToHost));
  NGSAMD_HIP(hipMemcpy(rv.data(), rvals.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  int64_t k = 0;  // CSR position: rows in order
  for (int64_t u = 0; u < nruns; u++) {
    const int R = rr[u + 1] - rr[u];
    const int64_t L = R ? (vp[u + 1] - vp[u]) / R : 0;
    for (int q = 0; q < R; q++)
      for (int64_t j = 0; j < L; j++) host_vals[k++] = rv[vp[u] + j * R + q];
  }
}

template <int G, bool ACCUM, bool DOT>
static int LaunchSpmvG(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* partials,
                       const int* stop) {
  constexpr int RPB = kSpmvBlock / G;
  const int64_t n = A.height;
  const int64_t passes = (n + RPB - 1) / RPB;
  // enough workgroups to fill the chip several times over; DOT variants are
  // limited by the size of the partials workspace
  const int64_t max_wg = DOT ? kMaxReduceBlocks : 65536;
  int64_t wg = passes < max_wg ? passes : max_wg;
  if (wg > 8) wg &= ~int64_t(7);
  if (wg < 1) wg = 1;
  const int iters = (int)((passes + wg - 1) / wg);
  CsrSpmvKernel<G, ACCUM, DOT><<<(int)wg, kSpmvBlock, 0, st>>>(A.rowptr.p, A.cols.p, A.vals.p, x, y, (int)n, s,
                                                               iters, partials, stop);
  NGSAMD_CHECK_LAUNCH();
  return (int)wg;
}

template <bool ACCUM, bool DOT>
static int LaunchStream(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* partials,
                        const int* stop) {
  auto kern = A.block_nnz == 1024   ? CsrStreamKernel<1024, ACCUM, DOT>
              : A.block_nnz == 4096 ? CsrStreamKernel<4096, ACCUM, DOT>
                                    : CsrStreamKernel<2048, ACCUM, DOT>;
  // a wide grid (many more workgroups than resident slots) lets the hardware
  // balance blocks of uneven cost: measured faster than one workgroup per
  // resident slot (412 vs 452 us for the runs kernel at 32^3 p=4); the
  // variant with partial sums stays within the partials workspace
  const int cap = DOT ? std::min(A.max_grid, kMaxWideBlocks) : A.max_grid;
  int wg = A.nblocks < cap ? A.nblocks : cap;
  if (wg > 8) wg &= ~7;
  if (wg < 1) wg = 1;
  const int per = (A.nblocks + wg - 1) / wg;
  kern<<<(int)wg, kStreamBlock, 0, st>>>(A.rowptr.p, A.cols.p, A.vals.p, x, y, A.blk_row.p, A.nblocks, per, s,
                                         partials, stop);
  NGSAMD_CHECK_LAUNCH();
  return (int)wg;
}

template <bool ACCUM, bool DOT>
static int LaunchRuns(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* partials,
                      const int* stop, int reverse) {
  auto kern = A.block_nnz == 1024   ? CsrRunsKernel<1024, ACCUM, DOT>
              : A.block_nnz == 4096 ? CsrRunsKernel<4096, ACCUM, DOT>
                                    : CsrRunsKernel<2048, ACCUM, DOT>;
  // a wide grid (many more workgroups than resident slots) lets the hardware
  // balance blocks of uneven cost: measured faster than one workgroup per
  // resident slot (412 vs 452 us for the runs kernel at 32^3 p=4); the
  // variant with partial sums stays within the partials workspace
  const int cap = DOT ? std::min(A.max_grid, kMaxWideBlocks) : A.max_grid;
  int wg = A.nblocks < cap ? A.nblocks : cap;
  if (wg > 8) wg &= ~7;
  if (wg < 1) wg = 1;
  const int per = (A.nblocks + wg - 1) / wg;
  kern<<<(int)wg, kStreamBlock, 0, st>>>(A.run_row.p, A.run_ptr.p, A.val_ptr.p, A.rcols.p, A.rvals.p, x, y,
                                         A.blk_run.p, A.blk_len.p, A.nblocks, per, s, partials, stop, reverse);
  NGSAMD_CHECK_LAUNCH();
  return (int)wg;
}

template <bool ACCUM, bool DOT>
static int LaunchSpmv(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* partials,
                      const int* stop, int reverse) {
  if (A.width == 0) x = A.xpad.p;
  if (A.algo == 2) return LaunchRuns<ACCUM, DOT>(A, s, x, y, st, partials, stop, reverse);
  if (A.algo == 1) return LaunchStream<ACCUM, DOT>(A, s, x, y, st, partials, stop);
  switch (A.group) {
    case 4: return LaunchSpmvG<4, ACCUM, DOT>(A, s, x, y, st, partials, stop);
    case 8: return LaunchSpmvG<8, ACCUM, DOT>(A, s, x, y, st, partials, stop);
    case 16: return LaunchSpmvG<16, ACCUM, DOT>(A, s, x, y, st, partials, stop);
    default: return LaunchSpmvG<32, ACCUM, DOT>(A, s, x, y, st, partials, stop);
  }
}

void CsrOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchSpmv<true, false>(*this, s, x, y, st, nullptr, nullptr, flip);
  flip ^= 1;
}

void CsrOp::Mult(const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchSpmv<false, false>(*this, 1.0, x, y, st, nullptr, nullptr, flip);
  flip ^= 1;
}

// Workgroups of the DOT variant's launch (one partial each), as LaunchRuns,
// LaunchStream and LaunchSpmvG choose them.
static int DotWorkgroups(const CsrOp& A) {
  if (A.algo == 0) return kMaxReduceBlocks;  // at most this many (LaunchSpmvG)
  const int cap = std::min(A.max_grid, kMaxWideBlocks);
  int wg = A.nblocks < cap ? A.nblocks : cap;
  if (wg > 8) wg &= ~7;
  return wg < 1 ? 1 : wg;
}

int CsrOp::MultDot(const double* x, double* y, double* partials, const int* d_stop, hipStream_t st,
                   double* wide) {
  // a small grid (at most 1,024 workgroups, e.g. 2D 64^2 p=3) writes its
  // partials straight to CG's, whose update kernel sums them: one launch
  // less per iteration
  if (DotWorkgroups(*this) <= 1024) return LaunchSpmv<false, true>(*this, 1.0, x, y, st, partials, d_stop);
  // the operator on its wide grid writes one partial per workgroup; a
  // 64-workgroup kernel folds them so that CG's kernels read 64 values
  if (!wide) wide = State().wide_partials;
  const int np = LaunchSpmv<false, true>(*this, 1.0, x, y, st, wide, d_stop);
  LaunchFoldPartials(wide, np, partials, d_stop, st);
  return kFoldedPartials;
}

__global__ void CsrDiagKernel(const int* __restrict__ rowptr, const int* __restrict__ cols,
                              const double* __restrict__ vals, double* __restrict__ d, int n) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int k = rowptr[r]; k < rowptr[r + 1]; k++)
      if (cols[k] == r) { v = vals[k]; break; }
    d[r] = v;
  }
}

void CsrOp::Diagonal(double* d, hipStream_t st) {
  if (height == 0) return;
  if (algo == 2) {
    RunsDiagKernel<<<GridFor(nruns, 256, 8192), 256, 0, st>>>(run_row.p, run_ptr.p, val_ptr.p, rcols.p, rvals.p, d,
                                                              (int)nruns);
    NGSAMD_CHECK_LAUNCH();
    return;
  }
  CsrDiagKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(rowptr.p, cols.p, vals.p, d, (int)height);
  NGSAMD_CHECK_LAUNCH();
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_csr_create(ngsamd_op_t* op, int64_t height, int64_t width, const int64_t* rp,
                                 const int* cols, const double* vals) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(height >= 0 && width >= 0, "ngsamd_csr_create: negative size");
    NGSAMD_REQUIRE(op && rp && (rp[height] == 0 || (cols && vals)), "ngsamd_csr_create: null argument");
    *op = nullptr;
    *op = new CsrOp(height, width, rp, cols, vals);
  });
}

extern "C" int ngsamd_csr_values(ngsamd_op_t op, double* h_vals) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::CSR, "ngsamd_csr_values: not a CSR operator");
    CsrOp* A = static_cast<CsrOp*>(op);
    NGSAMD_REQUIRE(h_vals || A->nnz == 0, "ngsamd_csr_values: null argument");
    A->Values(h_vals);
  });
}

extern "C" int ngsamd_csr_diagonal(ngsamd_op_t op, double* d) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::CSR, "ngsamd_csr_diagonal: not a CSR operator");
    static_cast<CsrOp*>(op)->Diagonal(d, Stream());
  });
}
