// Lockstep CG for K right-hand sides of one operator.
//
// CGSolver<double>::MultiMult (linalg/cg.cpp:132-250): vectors with K
// components per dof (entry j of dof i at i K + j); per component j its own
// alpha_j = wd_j / <s_j, A s_j> and beta_j = wdn_j / wd_j, one stopping test
// on the Euclidean norm of the vector (wdn_0 .. wdn_{K-1}) of w.d values:
// iterate while ||wdn|| > tol^2 ||wdn at start|| (tol^2 if absolute); if
// ||wdn|| is 0 at the start, wdn = 1; stop when ||kss|| = 0; the step count
// is the number of loop tests, as the reference's `steps`.  The one
// difference: a component with wd_j = 0 or kss_j = 0 gets alpha_j = 0 (and
// beta_j = 0) where the reference divides 0 by 0.
//
// The operator product reads the matrix once for all K components
// (CsrOp::MultMulti); the vector kernels handle all components of a dof in
// one thread, the update kernel forming the partials of the new w.d as it
// writes w and d; inner products are per-workgroup partials folded in a
// fixed order (deterministic).  The stopping test runs on the device: every
// kernel returns at once after it has stopped, and the host enqueues 16
// iterations at a time, reading the control word once per chunk.
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

constexpr int kMcBlock = 256;
constexpr int kMcGrid = 1024;  // workgroups of the dot kernels (fixed: same sums for every run)
constexpr int kMaxRhs = 8;

// partials[w K + j] = sum over the dofs of workgroup w of a[i K + j] b[i K + j]
__global__ __launch_bounds__(kMcBlock) void MultiDotPartialsKernel(const double* __restrict__ a,
                                                                   const double* __restrict__ b, int64_t n, int K,
                                                                   double* __restrict__ partials,
                                                                   const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double lds[kMcBlock / 64];
  double acc[kMaxRhs];
#pragma unroll
  for (int j = 0; j < kMaxRhs; j++) acc[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kMcBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kMcBlock) {
#pragma unroll
    for (int j = 0; j < kMaxRhs; j++)
      if (j < K) acc[j] += a[i * K + j] * b[i * K + j];
  }
  for (int j = 0; j < K; j++) {
    const double v = BlockSum<kMcBlock>(acc[j], lds);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * K + j] = v;
  }
}

// out[j] = sum over w < nwg of partials[w K + j], one workgroup, fixed order
__global__ __launch_bounds__(kMcBlock) void MultiDotFoldKernel(const double* __restrict__ partials, int nwg, int K,
                                                               double* __restrict__ out, const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double lds[kMcBlock / 64];
  for (int j = 0; j < K; j++) {
    double v = 0.0;
    for (int w = threadIdx.x; w < nwg; w += kMcBlock) v += partials[(int64_t)w * K + j];
    v = BlockSum<kMcBlock>(v, lds);
    if (threadIdx.x == 0) out[j] = v;
  }
}

__device__ __forceinline__ double SafeDiv(double a, double b) { return b != 0.0 ? a / b : 0.0; }

// w = C d (inv: Jacobi's 1/a_ii, 0 on fixed dofs; null: w = d)
__global__ __launch_bounds__(kMcBlock) void MultiPrecondKernel(const double* __restrict__ d,
                                                               const double* __restrict__ inv,
                                                               double* __restrict__ w, int64_t n, int K) {
  for (int64_t i = blockIdx.x * (int64_t)kMcBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kMcBlock) {
    const double c = inv ? inv[i] : 1.0;
    for (int j = 0; j < K; j++) w[i * K + j] = inv ? c * d[i * K + j] : d[i * K + j];
  }
}

// alpha_j = wd_j / kss_j;  u += alpha s;  d -= alpha w;  w = C d, and the
// workgroup partials of the new <w_j, d_j> (launched on the dot kernels'
// grid, so the sums are those MultiDotPartialsKernel would form)
__global__ __launch_bounds__(kMcBlock) void MultiCgUpdateKernel(double* __restrict__ u, double* __restrict__ d,
                                                                double* __restrict__ w,
                                                                const double* __restrict__ s,
                                                                const double* __restrict__ inv, int64_t n, int K,
                                                                const double* __restrict__ wd,
                                                                const double* __restrict__ kss,
                                                                double* __restrict__ partials,
                                                                const int* __restrict__ stop) {
  if (*stop) return;
  __shared__ double lds[kMcBlock / 64];
  double al[kMaxRhs], acc[kMaxRhs];
#pragma unroll
  for (int j = 0; j < kMaxRhs; j++) {
    al[j] = j < K ? SafeDiv(wd[j], kss[j]) : 0.0;
    acc[j] = 0.0;
  }
  for (int64_t i = blockIdx.x * (int64_t)kMcBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kMcBlock) {
    const double c = inv ? inv[i] : 1.0;
#pragma unroll
    for (int j = 0; j < kMaxRhs; j++) {
      if (j >= K) break;
      const int64_t k = i * K + j;
      u[k] += al[j] * s[k];
      const double dk = d[k] - al[j] * w[k];
      d[k] = dk;
      const double wk = inv ? c * dk : dk;
      w[k] = wk;
      acc[j] += wk * dk;
    }
  }
  for (int j = 0; j < K; j++) {
    const double v = BlockSum<kMcBlock>(acc[j], lds);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * K + j] = v;
  }
}

// beta_j = wdn_j / wd_j;  s = beta s + w
__global__ __launch_bounds__(kMcBlock) void MultiCgDirectionKernel(double* __restrict__ s,
                                                                   const double* __restrict__ w, int64_t n, int K,
                                                                   const double* __restrict__ wd,
                                                                   const double* __restrict__ wdn,
                                                                   const int* __restrict__ stop) {
  if (*stop) return;
  double be[kMaxRhs];
#pragma unroll
  for (int j = 0; j < kMaxRhs; j++) be[j] = j < K ? SafeDiv(wdn[j], wd[j]) : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kMcBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kMcBlock) {
#pragma unroll
    for (int j = 0; j < kMaxRhs; j++) {
      if (j >= K) break;
      const int64_t k = i * K + j;
      s[k] = s[k] * be[j] + w[k];
    }
  }
}

__device__ double NormK(const double* v, int K) {
  double s = 0.0;
  for (int j = 0; j < K; j++) s += v[j] * v[j];
  return sqrt(s);
}

// The loop test at the head of iteration it (1-based): go on while
// it - 1 < maxiter and ||wdn|| > err; else stop with steps = it.
// ctl[0] = stop flag, ctl[1] = steps.
__global__ void MultiCgHeadKernel(const double* __restrict__ wdn, int K, double err, int it, int maxiter,
                                  int* __restrict__ ctl) {
  if (threadIdx.x != 0 || ctl[0]) return;
  if (!(it - 1 < maxiter && NormK(wdn, K) > err)) {
    ctl[0] = 1;
    ctl[1] = it;
  }
}

// <s, A s> = 0 for every component: stop before the update (steps = it)
__global__ void MultiCgKssKernel(const double* __restrict__ kss, int K, int it, int* __restrict__ ctl) {
  if (threadIdx.x != 0 || ctl[0]) return;
  if (NormK(kss, K) == 0.0) {
    ctl[0] = 1;
    ctl[1] = it;
  }
}

// end of iteration it: hist[it] = sqrt(||wdn||)
__global__ void MultiCgTailKernel(const double* __restrict__ wdn, int K, int it, const int* __restrict__ ctl,
                                  double* __restrict__ hist) {
  if (threadIdx.x != 0 || ctl[0]) return;
  hist[it] = sqrt(NormK(wdn, K));
}

// d = f - d (initial residual with d = A u on entry)
__global__ __launch_bounds__(kMcBlock) void MultiResidualKernel(const double* __restrict__ f, double* __restrict__ d,
                                                                int64_t m) {
  for (int64_t k = blockIdx.x * (int64_t)kMcBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kMcBlock)
    d[k] = f[k] - d[k];
}

static double Norm(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_cg_solve_multi(ngsamd_op_t mat, ngsamd_op_t pre, int nrhs, const double* d_rhs, double* d_sol,
                                     int initialize, double tol, int absolute, int maxiter, int* h_steps,
                                     double* h_residuals) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(mat && d_rhs && d_sol && h_steps, "ngsamd_cg_solve_multi: null argument");
    NGSAMD_REQUIRE(mat->kind == OpKind::CSR, "ngsamd_cg_solve_multi: the operator must be a CSR matrix");
    NGSAMD_REQUIRE(!pre || pre->kind == OpKind::Jacobi, "ngsamd_cg_solve_multi: the preconditioner must be Jacobi");
    NGSAMD_REQUIRE(nrhs >= 1 && nrhs <= kMaxRhs, "ngsamd_cg_solve_multi: 1 <= nrhs <= 8");
    NGSAMD_REQUIRE(mat->height == mat->width && (!pre || pre->height == mat->height),
                   "ngsamd_cg_solve_multi: sizes differ");
    NGSAMD_REQUIRE(maxiter >= 0, "ngsamd_cg_solve_multi: maxiter < 0");
    CsrOp& A = *static_cast<CsrOp*>(mat);
    const double* inv = pre ? static_cast<JacobiOp*>(pre)->invdiag.p : nullptr;
    const int K = nrhs;
    const int64_t n = A.height, m = n * K;
    hipStream_t st = Stream();
    *h_steps = 0;
    if (n == 0) return;
    DevBuf<double> d(m), w(m), s(m), partials((size_t)kMcGrid * K), dots(3 * K);
    double* wd = dots.p;
    double* wdn = dots.p + K;
    double* kss = dots.p + 2 * K;
    const int grid = GridFor(n, kMcBlock, kMcGrid);
    DevBuf<int> ctl(2);
    DevBuf<double> hist((size_t)maxiter + 1);
    NGSAMD_HIP(hipMemsetAsync(ctl.p, 0, sizeof(int) * 2, st));
    auto dot = [&](const double* a, const double* b, double* out, const int* stop) {
      MultiDotPartialsKernel<<<grid, kMcBlock, 0, st>>>(a, b, n, K, partials.p, stop);
      NGSAMD_CHECK_LAUNCH();
      MultiDotFoldKernel<<<1, kMcBlock, 0, st>>>(partials.p, grid, K, out, stop);
      NGSAMD_CHECK_LAUNCH();
    };
    // d = f (u = 0) or f - A u
    if (initialize) {
      LaunchFill(d_sol, m, 0.0, st);
      NGSAMD_HIP(hipMemcpyAsync(d.p, d_rhs, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    } else {
      A.MultMulti(K, d_sol, d.p, st);
      MultiResidualKernel<<<GridFor(m, kMcBlock, 8192), kMcBlock, 0, st>>>(d_rhs, d.p, m);
      NGSAMD_CHECK_LAUNCH();
    }
    MultiPrecondKernel<<<GridFor(n, kMcBlock, 8192), kMcBlock, 0, st>>>(d.p, inv, w.p, n, K);
    NGSAMD_CHECK_LAUNCH();
    NGSAMD_HIP(hipMemcpyAsync(s.p, w.p, sizeof(double) * m, hipMemcpyDeviceToDevice, st));
    dot(w.p, d.p, wdn, nullptr);
    std::vector<double> h_wdn(K);
    NGSAMD_HIP(hipMemcpyAsync(h_wdn.data(), wdn, sizeof(double) * K, hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipStreamSynchronize(st));
    const double res0 = std::sqrt(Norm(h_wdn));  // recorded before the reset below, as the reference prints it
    if (Norm(h_wdn) == 0.0) {
      for (auto& v : h_wdn) v = 1.0;
      NGSAMD_HIP(hipMemcpyAsync(wdn, h_wdn.data(), sizeof(double) * K, hipMemcpyHostToDevice, st));
    }
    const double err = absolute ? tol * tol : tol * tol * Norm(h_wdn);
    // The iterations are enqueued a chunk at a time; the stopping test runs
    // on the device (every kernel of a stopped solve returns at once), so
    // the host reads the control word once per chunk, not per iteration.
    constexpr int kChunkIts = 16;
    int it = 0, h_ctl[2] = {0, 0};
    while (true) {
      for (int c = 0; c < kChunkIts && it <= maxiter; c++) {
        ++it;
        MultiCgHeadKernel<<<1, 64, 0, st>>>(wdn, K, err, it, maxiter, ctl.p);
        NGSAMD_CHECK_LAUNCH();
        A.MultMulti(K, s.p, w.p, st, ctl.p);  // w = A s
        std::swap(wd, wdn);                    // wd = wdn
        dot(s.p, w.p, kss, ctl.p);
        MultiCgKssKernel<<<1, 64, 0, st>>>(kss, K, it, ctl.p);
        NGSAMD_CHECK_LAUNCH();
        MultiCgUpdateKernel<<<grid, kMcBlock, 0, st>>>(d_sol, d.p, w.p, s.p, inv, n, K, wd, kss, partials.p,
                                                       ctl.p);
        NGSAMD_CHECK_LAUNCH();
        MultiDotFoldKernel<<<1, kMcBlock, 0, st>>>(partials.p, grid, K, wdn, ctl.p);
        NGSAMD_CHECK_LAUNCH();
        MultiCgDirectionKernel<<<GridFor(n, kMcBlock, 8192), kMcBlock, 0, st>>>(s.p, w.p, n, K, wd, wdn, ctl.p);
        NGSAMD_CHECK_LAUNCH();
        MultiCgTailKernel<<<1, 64, 0, st>>>(wdn, K, it, ctl.p, hist.p);
        NGSAMD_CHECK_LAUNCH();
      }
      NGSAMD_HIP(hipMemcpyAsync(h_ctl, ctl.p, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      // the head test of iteration maxiter + 1 always stops
      if (h_ctl[0]) break;
      NGSAMD_REQUIRE(it <= maxiter, "ngsamd_cg_solve_multi: internal error (no stop after maxiter)");
    }
    *h_steps = h_ctl[1];
    if (h_residuals) {
      h_residuals[0] = res0;
      if (h_ctl[1] > 1)
        NGSAMD_HIP(hipMemcpyAsync(h_residuals + 1, hist.p + 1, sizeof(double) * (h_ctl[1] - 1),
                                  hipMemcpyDeviceToHost, st));
    }
    NGSAMD_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int ngsamd_csr_mult_multi(ngsamd_op_t op, int nrhs, const double* d_x, double* d_y) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && d_x && d_y, "ngsamd_csr_mult_multi: null argument");
    NGSAMD_REQUIRE(op->kind == OpKind::CSR, "ngsamd_csr_mult_multi: not a CSR operator");
    static_cast<CsrOp*>(op)->MultMulti(nrhs, d_x, d_y, Stream());
  });
}
