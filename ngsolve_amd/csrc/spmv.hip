// CSR sparse matrix-vector product on gfx950.
//
// Replaces SparseMatrix<double>::MultAdd (linalg/sparsematrix_impl.hpp:264-306,
// row kernel RowTimesVector linalg/sparsematrix.hpp:625-632) and the
// reference's cuSPARSE call (ngscuda/cuda_linalg.cpp:243-275).
//
// Layout: int32 row pointers, int32 column numbers, FP64 values (12 B per
// nonzero streamed once).  A group of G lanes of a 64-wide wavefront works on
// one row: lane l of the group reads nonzeros l, l+G, ... so every load
// instruction of a wavefront covers 64/G contiguous segments of the column
// and value streams; the G partial sums are combined with cross-lane
// shuffles.  G is picked from the mean row length.  The workgroups walk
// contiguous row ranges that are grouped per XCD (workgroup i runs on XCD
// i mod 8), so the x entries a row range gathers stay in one XCD's L2.
#include "common.hpp"
#include "csr.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kSpmvBlock = 256;

// Map the hardware workgroup id to a logical id so that the 8 XCDs each own a
// contiguous range of logical ids.
__device__ __forceinline__ int XcdContiguousBlock(int b, int grid) {
  if ((grid & 7) != 0) return b;
  return (b & 7) * (grid >> 3) + (b >> 3);
}

// y[r] = (ACCUM ? y[r] : 0) + s * (A x)[r]; optionally dot = sum_r x[r]*(A x)[r]
// (the <s, A s> of CG) published as per-workgroup partials.
template <int G, bool ACCUM, bool DOT>
__global__ __launch_bounds__(kSpmvBlock) void CsrSpmvKernel(
    const int* __restrict__ rowptr, const int* __restrict__ cols, const double* __restrict__ vals,
    const double* __restrict__ x, double* __restrict__ y, int n, double s, int rows_per_wg_iter,
    int iters_per_wg, double* __restrict__ partials, unsigned* __restrict__ counter,
    double* __restrict__ result, const int* __restrict__ stop_flag) {
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
  (void)rows_per_wg_iter;
  if (DOT) {
    __shared__ double lds[kSpmvBlock / 64];
    __shared__ bool s_last;
    double v = BlockSum<kSpmvBlock>(dotacc, lds);
    double vals1[1] = {v};
    if (PublishPartials<1>(vals1, partials, counter, &s_last)) {
      ReducePartials<kSpmvBlock, 1>(partials, result, lds);
      if (threadIdx.x == 0) *counter = 0u;
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
  rowptr.alloc(h + 1);
  cols.alloc(nnz);
  vals.alloc(nnz);
  rowptr.upload(rp32.data(), h + 1);
  cols.upload(c, nnz);
  vals.upload(v, nnz);
  NGSAMD_HIP(hipStreamSynchronize(Stream()));
  double mean = h ? double(nnz) / double(h) : 0.0;
  group = mean > 48 ? 32 : mean > 20 ? 16 : mean > 9 ? 8 : 4;
}

template <int G, bool ACCUM, bool DOT>
static void LaunchSpmvG(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* result,
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
  auto& S = State();
  CsrSpmvKernel<G, ACCUM, DOT><<<(int)wg, kSpmvBlock, 0, st>>>(
      A.rowptr.p, A.cols.p, A.vals.p, x, y, (int)n, s, RPB, iters, S.partials,
      S.counters + kCounterCG1, result, stop);
  NGSAMD_CHECK_LAUNCH();
}

template <bool ACCUM, bool DOT>
static void LaunchSpmv(const CsrOp& A, double s, const double* x, double* y, hipStream_t st, double* result,
                       const int* stop) {
  switch (A.group) {
    case 4: LaunchSpmvG<4, ACCUM, DOT>(A, s, x, y, st, result, stop); break;
    case 8: LaunchSpmvG<8, ACCUM, DOT>(A, s, x, y, st, result, stop); break;
    case 16: LaunchSpmvG<16, ACCUM, DOT>(A, s, x, y, st, result, stop); break;
    default: LaunchSpmvG<32, ACCUM, DOT>(A, s, x, y, st, result, stop); break;
  }
}

void CsrOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchSpmv<true, false>(*this, s, x, y, st, nullptr, nullptr);
}

void CsrOp::Mult(const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  LaunchSpmv<false, false>(*this, 1.0, x, y, st, nullptr, nullptr);
}

void CsrOp::MultDot(const double* x, double* y, double* d_result, const int* d_stop, hipStream_t st) {
  LaunchSpmv<false, true>(*this, 1.0, x, y, st, d_result, d_stop);
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
  CsrDiagKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(rowptr.p, cols.p, vals.p, d, (int)height);
  NGSAMD_CHECK_LAUNCH();
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_csr_create(ngsamd_op_t* op, int64_t height, int64_t width, const int64_t* rp,
                                 const int* cols, const double* vals) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && rp && (rp[height] == 0 || (cols && vals)), "ngsamd_csr_create: null argument");
    *op = nullptr;
    *op = new CsrOp(height, width, rp, cols, vals);
  });
}

extern "C" int ngsamd_csr_diagonal(ngsamd_op_t op, double* d) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && op->kind == OpKind::CSR, "ngsamd_csr_diagonal: not a CSR operator");
    static_cast<CsrOp*>(op)->Diagonal(d, Stream());
  });
}
