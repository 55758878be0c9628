// Jacobi preconditioner, the device-resident preconditioned CG loop and the
// mixed-precision iterative refinement built on it.
//
// Jacobi: JacobiPrecond (linalg/jacobi.cpp:41-90): invdiag = 1/a_ii on free
// dofs and 0 elsewhere, y += s * invdiag .* x.
//
// CG: the recurrence and stopping rule of CGSolver._SolveImpl
// (python/krylovspace.py:263-291) and LinearSolver.CheckResidual (:129-150):
//   d = f - A u;  w = C d;  s = w;  wdn = <w,d>;  check(sqrt|wdn|)
//   loop: w = A s; wd = wdn; as = <s,w>; stop if as == 0 or wd == 0
//         alpha = wd/as; u += alpha s; d -= alpha w; w = C d; wdn = <w,d>
//         check(sqrt|wdn|); beta = wdn/wd; s = beta s + w
// check: iterations += 1; converged when iterations >= maxiter or
// residual <= final, final = tol*res0 (max'ed with atol when given).
//
// One iteration is three kernels after the operator: (1) w = A s with the
// partial sums of <s,w> (fused into the CSR kernel and into the matrix-free
// gather kernel), (2) the two updates, the Jacobi apply and the partials of
// <w,d>, each workgroup first forming alpha itself from the partials of (1),
// (3) s = beta s + w, each workgroup forming beta from the partials of (2);
// workgroup 0 of (3) also does the stopping test and keeps the iteration
// count and the residual history.  The operator kernels write one partial
// per workgroup of a wide grid; a small kernel folds them to 64 before (2).
// With a block-Jacobi preconditioner, (2) does only the two updates and the
// preconditioner's two kernels and a dot kernel for <w,d> follow it.  Every
// workgroup sums the same partials
// in the same order, so all of them hold the same alpha / beta bit for bit
// and every run gives the same iterates.  <w,d> is kept per iteration parity
// (wdn[it & 1]): kernel (3) of an even iteration reads wdn[0] and its
// workgroup 0 writes wdn[1], which the next iteration reads, so no kernel
// reads a scalar that it also writes.  Scalars stay on the device, so the
// loop needs no host round trip.  `chunk` (even) iterations are captured once
// into a HIP graph that is replayed until the device flag says done; every
// kernel returns at once when the flag is set, so the overshoot of the last
// chunk costs only launch latency.  (Reference: DevCGSolver,
// ngscuda/cuda_krylov.cpp:19-203, which uses a CUDA conditional WHILE node.)
//
// The vector type V is double for the FP64 solver and float for the inner
// solver of iterative refinement; inner products, alpha, beta and the
// stopping test are FP64 in both.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "csr.hpp"
#include "matfree.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

// ------------------------------------------------------------------ Jacobi

template <class V>
__global__ void InvDiagKernel(const double* __restrict__ diag, const uint8_t* __restrict__ free, V* __restrict__ inv,
                              int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    inv[i] = (free == nullptr || free[i]) ? (V)(1.0 / diag[i]) : (V)0;
}

JacobiOp::JacobiOp(int64_t n, const double* diag, bool on_device, const uint8_t* freedofs)
    : ngsamd_op_s(OpKind::Jacobi) {
  height = width = n;
  invdiag.alloc(n);
  if (n == 0) return;
  hipStream_t st = Stream();
  DevBuf<uint8_t> dfree;
  if (freedofs) {
    dfree.alloc(n);
    dfree.upload(freedofs, n);
  }
  DevBuf<double> ddiag;
  const double* src = diag;
  if (!on_device) {
    ddiag.alloc(n);
    ddiag.upload(diag, n);
    src = ddiag.p;
  }
  InvDiagKernel<double><<<GridFor(n, 256, 8192), 256, 0, st>>>(src, dfree.p, invdiag.p, n);
  NGSAMD_CHECK_LAUNCH();
  NGSAMD_HIP(hipStreamSynchronize(st));
}

__global__ void JacobiMultAddKernel(const double* __restrict__ inv, const double* __restrict__ x,
                                    double* __restrict__ y, int64_t n, double s, bool accum) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = s * (inv[i] * x[i]);
    y[i] = accum ? y[i] + v : v;
  }
}

void JacobiOp::MultAdd(double s, const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  JacobiMultAddKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(invdiag.p, x, y, height, s, true);
  NGSAMD_CHECK_LAUNCH();
}

void JacobiOp::Mult(const double* x, double* y, hipStream_t st) {
  if (height == 0) return;
  JacobiMultAddKernel<<<GridFor(height, 256, 8192), 256, 0, st>>>(invdiag.p, x, y, height, 1.0, false);
  NGSAMD_CHECK_LAUNCH();
}

// ---------------------------------------------------------- CG device state

struct CGScalars {
  double wdn[2];     // <w,d> of the current iteration, by iteration parity
  double alpha, beta;
  double final_res;  // stopping threshold
  double tol, atol;  // tol < 0: not given; atol < 0: not given
  int it;            // CheckResidual calls so far
  int done;          // 0 running, 1 converged / maxiter, 2 breakdown
  int maxiter;
  int pad;
};

constexpr int kCgBlock = 256;

__device__ __forceinline__ void CheckResidual(CGScalars* S, double* residuals, double res) {
  S->it += 1;
  if (S->it <= S->maxiter) residuals[S->it - 1] = res;
  if (S->it == 1) {
    if (S->tol < 0) S->final_res = S->atol;
    else {
      S->final_res = res * S->tol;
      if (S->atol >= 0 && S->atol > S->final_res) S->final_res = S->atol;
    }
  }
  if (S->it >= S->maxiter || res <= S->final_res) S->done = 1;
}

// w = C d (inv null: w already holds C d); s = w; partials of <w,d>
template <class V>
__global__ __launch_bounds__(kCgBlock) void CgInitKernel(const V* __restrict__ inv, const V* __restrict__ d,
                                                         V* __restrict__ w, V* __restrict__ s, int64_t n,
                                                         double* __restrict__ partials) {
  __shared__ double lds[kCgBlock / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock) {
    const V di = d[i];
    const V wi = inv ? inv[i] * di : w[i];
    w[i] = wi;
    s[i] = wi;
    acc += (double)wi * (double)di;
  }
  StorePartial<kCgBlock>(acc, partials, lds);
}

// one workgroup: wdn = sum of partials; first check
__global__ __launch_bounds__(kCgBlock) void CgInitCheckKernel(const double* __restrict__ partials, int count,
                                                              CGScalars* S, double* residuals) {
  __shared__ double lds[kCgBlock / 64];
  const double wdn = SumPartials<kCgBlock>(partials, count, lds);
  if (threadIdx.x == 0) {
    S->wdn[0] = wdn;
    CheckResidual(S, residuals, sqrt(fabs(wdn)));
  }
}

// as = <s, A s> from the partials pa of the operator kernel; breakdown when
// as == 0 or wd == 0, else alpha = wd / as and u += alpha s; d -= alpha w;
// w = C d; partials of <w,d> into pw.  P is the iteration parity.  inv null
// (a preconditioner other than Jacobi): only the two updates; C d and its
// partials follow in their own kernels.
template <class V>
__global__ __launch_bounds__(kCgBlock) void CgUpdateKernel(const V* __restrict__ inv, V* __restrict__ u,
                                                           V* __restrict__ d, V* __restrict__ w,
                                                           const V* __restrict__ s, int64_t n, CGScalars* S, int P,
                                                           const double* __restrict__ pa, int na,
                                                           double* __restrict__ pw) {
  if (S->done) return;
  __shared__ double lds[kCgBlock / 64 + 1];
  const double as_s = SumPartialsAll<kCgBlock>(pa, na, lds);
  const double wd = S->wdn[P];
  if (as_s == 0.0 || wd == 0.0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) S->done = 2;
    return;
  }
  const double alpha = wd / as_s;
  if (blockIdx.x == 0 && threadIdx.x == 0) S->alpha = alpha;
  const V alv = (V)alpha;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock) {
    const V si = s[i], wi = w[i];
    u[i] += alv * si;
    const V di = d[i] - alv * wi;
    d[i] = di;
    if (inv) {
      const V zi = inv[i] * di;
      w[i] = zi;
      acc += (double)zi * (double)di;
    }
  }
  if (inv) StorePartial<kCgBlock>(acc, pw, lds);
}

// wdn = <w,d> from the partials pw; beta = wdn / wd; s = beta s + w.
// Workgroup 0 stores wdn for the next iteration and does the stopping test
// (the other workgroups update s whatever its outcome: s is not used once
// the loop stops).
template <class V>
__global__ __launch_bounds__(kCgBlock) void CgDirectionKernel(V* __restrict__ s, const V* __restrict__ w, int64_t n,
                                                              CGScalars* S, int P, const double* __restrict__ pw,
                                                              int nw, double* residuals) {
  if (S->done) return;
  __shared__ double lds[kCgBlock / 64 + 1];
  const double wdn = SumPartialsAll<kCgBlock>(pw, nw, lds);
  const double beta = wdn / S->wdn[P];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->wdn[P ^ 1] = wdn;
    S->beta = beta;
    CheckResidual(S, residuals, sqrt(fabs(wdn)));
  }
  const V bv = (V)beta;
  for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock)
    s[i] = bv * s[i] + w[i];
}

// partials of <x,y> (FP64 accumulation), skipped when *done
template <class V>
__global__ __launch_bounds__(kCgBlock) void CgDotKernel(const V* __restrict__ x, const V* __restrict__ y, int64_t n,
                                                        const int* done, double* __restrict__ partials) {
  if (*done) return;
  __shared__ double lds[kCgBlock / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock)
    acc += (double)x[i] * (double)y[i];
  StorePartial<kCgBlock>(acc, partials, lds);
}

// partials of sum over free dofs (mask != 0) of x_i^2
__global__ __launch_bounds__(kCgBlock) void MaskedSqKernel(const double* __restrict__ x, const double* __restrict__ mask,
                                                           int64_t n, double* __restrict__ partials) {
  __shared__ double lds[kCgBlock / 64];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kCgBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCgBlock)
    if (mask[i] != 0.0) acc += x[i] * x[i];
  StorePartial<kCgBlock>(acc, partials, lds);
}

__global__ __launch_bounds__(kCgBlock) void SumToKernel(const double* __restrict__ partials, int count,
                                                        double* __restrict__ out) {
  __shared__ double lds[kCgBlock / 64];
  const double v = SumPartials<kCgBlock>(partials, count, lds);
  if (threadIdx.x == 0) *out = v;
}

__global__ void ScaledToF32Kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n,
                                  const double* __restrict__ scale) {
  const double s = *scale;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)(x[i] * s);
}

__global__ void AddScaledF32Kernel(double* __restrict__ u, const float* __restrict__ c, int64_t n,
                                   const double* __restrict__ scale) {
  const double s = *scale;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    u[i] += s * (double)c[i];
}

// Reducing grids stay at 1024 workgroups or below: the update and direction
// kernels read every partial in every workgroup.
constexpr int kCgMaxGrid = 1024;
static int CgGrid(int64_t n) { return GridFor(n, kCgBlock, kCgMaxGrid); }

// Partial sums of one CG instance: of the operator kernel (<s, A s>, a), of
// the update (<w, d>, w), each room for kMaxReduceBlocks, and of the
// operators' wide grids (kMaxWideBlocks, folded into a).
struct CgPartials {
  double* a;
  double* w;
  double* wide;
};
// the library's workspace (allocated by ngsamd_init)
static CgPartials LibraryPartials() {
  return {State().partials, State().partials + kMaxReduceBlocks, State().wide_partials};
}

// w = A s and the partials of <s, w> into B.a; returns their number
template <class V>
static int EnqueueMatDot(ngsamd_op_t A, const V* s, V* w, int64_t n, CGScalars* dS, const CgPartials& B,
                         hipStream_t st) {
  const int* done = &dS->done;
  if constexpr (sizeof(V) == 8) {
    if (A->kind == OpKind::CSR) return static_cast<CsrOp*>(A)->MultDot(s, w, B.a, done, st, B.wide);
    if (A->kind == OpKind::MatrixFree) {
      const int np = static_cast<MatFreeOp*>(A)->MultDot(s, w, B.a, done, st, B.wide);
      if (np > 0) return np;
    } else {
      A->Mult(s, w, st);
    }
  } else {
    NGSAMD_REQUIRE(A->kind == OpKind::MatrixFree, "cg: FP32 vectors need a matrix-free operator");
    const int np = static_cast<MatFreeOp*>(A)->MultDotF32(s, w, B.a, done, st, B.wide);
    if (np > 0) return np;
  }
  const int g = CgGrid(n);
  CgDotKernel<V><<<g, kCgBlock, 0, st>>>(s, w, n, done, B.a);
  NGSAMD_CHECK_LAUNCH();
  return g;
}

// w = C d for a preconditioner other than Jacobi (block Jacobi, two-level,
// sparse Cholesky, multigrid); its kernels return at once when *stop != 0
template <class V>
static void EnqueuePrecond(ngsamd_op_t pre, const V* d, V* w, const int* stop, hipStream_t st) {
  if constexpr (sizeof(V) == 8) {
    if (pre->kind == OpKind::BlockJacobi) static_cast<BlockJacobiOp*>(pre)->MultStop(d, w, stop, st);
    else if (pre->kind == OpKind::TwoLevel) static_cast<TwoLevelOp*>(pre)->MultStop(d, w, stop, st);
    else if (pre->kind == OpKind::SparseCholesky) static_cast<SparseCholeskyOp*>(pre)->Apply(1.0, d, w, false, stop, st);
    else if (pre->kind == OpKind::MultiGrid) MultiGridApply(pre, 1.0, d, w, false, stop, st);
    else throw Error(NGSAMD_E_ARG, "cg: unsupported preconditioner");
  } else {
    throw Error(NGSAMD_E_ARG, "cg: FP32 vectors need a Jacobi preconditioner");
  }
}

// One CG iteration of parity P on the stream (graph-capturable: kernels only).
// inv null: pre is applied between the update and the direction kernels.
template <class V>
static void EnqueueIteration(ngsamd_op_t A, const V* inv, ngsamd_op_t pre, V* u, V* d, V* w, V* s, int64_t n,
                             CGScalars* dS, double* residuals, int P, hipStream_t st,
                             const CgPartials& B = LibraryPartials()) {
  const int na = EnqueueMatDot<V>(A, s, w, n, dS, B, st);
  const int g = CgGrid(n);
  CgUpdateKernel<V><<<g, kCgBlock, 0, st>>>(inv, u, d, w, s, n, dS, P, B.a, na, B.w);
  NGSAMD_CHECK_LAUNCH();
  if (!inv) {
    EnqueuePrecond<V>(pre, d, w, &dS->done, st);
    CgDotKernel<V><<<g, kCgBlock, 0, st>>>(w, d, n, &dS->done, B.w);
    NGSAMD_CHECK_LAUNCH();
  }
  CgDirectionKernel<V><<<g, kCgBlock, 0, st>>>(s, w, n, dS, P, B.w, g, residuals);
  NGSAMD_CHECK_LAUNCH();
}

struct CgResult {
  int iterations = 0;
  int done = 0;
};

// Work vectors and device scalars of one CG instance.
template <class V>
struct CgWork {
  DevBuf<V> d, w, s;
  DevBuf<double> residuals;
  DevBuf<CGScalars> S;
  CgWork(int64_t n, int maxiter) : d(n), w(n), s(n), residuals(maxiter), S(1) {}
};

// A captured graph of `chunk` CG iterations.  The kernels read the
// tolerances and counters from the device scalars, so one instance serves
// every solve on the same operators and vectors (the refinement's inner
// solves capture it once per refinement, not once per step).
// It also keeps the pinned host copies of the device scalars and the two
// events the host polls, so that the inner solves of a refinement allocate
// them once.
struct CgGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool failed = false;  // capture not possible: launch the kernels directly
  CGScalars* pin = nullptr;  // 2 pinned host copies of the device scalars
  hipEvent_t ev[2] = {nullptr, nullptr};
  CgGraph() = default;
  CgGraph(const CgGraph&) = delete;
  ~CgGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (pin) (void)hipHostFree(pin);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void Prepare() {
    if (!pin) NGSAMD_HIP(hipHostMalloc((void**)&pin, 2 * sizeof(CGScalars), hipHostMallocDefault));
    for (hipEvent_t& e : ev)
      if (!e) NGSAMD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
};

// Iterates from an initialised state (d = f - A u already set, u set) until
// the device flag is raised.  Uses a graph of `chunk` iterations if possible,
// kept in *cache (captured on first use) when cache is not null.
template <class V>
static CgResult IterateCg(ngsamd_op_t A, const V* inv, ngsamd_op_t pre, V* u, CgWork<V>& W, int64_t n, double tol,
                          double atol, int maxiter, int chunk, hipStream_t st, CgGraph* cache = nullptr) {
  CGScalars hS{};
  hS.tol = tol;
  hS.atol = atol;
  hS.maxiter = maxiter;
  NGSAMD_HIP(hipMemcpyAsync(W.S.p, &hS, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  const int g = CgGrid(n);
  if (!inv) EnqueuePrecond<V>(pre, W.d.p, W.w.p, nullptr, st);
  CgInitKernel<V><<<g, kCgBlock, 0, st>>>(inv, W.d.p, W.w.p, W.s.p, n, LibraryPartials().w);
  NGSAMD_CHECK_LAUNCH();
  CgInitCheckKernel<<<1, kCgBlock, 0, st>>>(LibraryPartials().w, g, W.S.p, W.residuals.p);
  NGSAMD_CHECK_LAUNCH();

  // graphs hold an even number of iterations so that every replay starts
  // at parity 0
  if (chunk > 0) chunk += chunk & 1;
  CgGraph local;
  CgGraph& G = cache ? *cache : local;
  G.Prepare();
  CGScalars* pin = G.pin;
  hipEvent_t* ev = G.ev;
  if (chunk > 0 && !G.exec && !G.failed) {
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      bool ok = true;
      try {
        for (int k = 0; k < chunk; k++)
          EnqueueIteration<V>(A, inv, pre, u, W.d.p, W.w.p, W.s.p, n, W.S.p, W.residuals.p, k & 1, st);
      } catch (...) {
        ok = false;
      }
      e = hipStreamEndCapture(st, &G.graph);
      if (!ok && e == hipSuccess) e = hipErrorUnknown;
      if (e == hipSuccess) e = hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      if (G.graph) (void)hipGraphDestroy(G.graph);
      G.graph = nullptr;
      G.exec = nullptr;
      G.failed = true;
    }
  }
  const bool use_graph = chunk > 0 && G.exec != nullptr;
  hipGraphExec_t exec = G.exec;

  const int per_launch = use_graph ? chunk : 1;
  long long enqueued = 0;  // iterations enqueued without a graph (parity)
  auto enqueue = [&](int slot) {
    if (use_graph) NGSAMD_HIP(hipGraphLaunch(exec, st));
    else
      EnqueueIteration<V>(A, inv, pre, u, W.d.p, W.w.p, W.s.p, n, W.S.p, W.residuals.p, (int)(enqueued++ & 1), st);
    NGSAMD_HIP(hipMemcpyAsync(&pin[slot], W.S.p, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  };

  NGSAMD_HIP(hipMemcpyAsync(&pin[0], W.S.p, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  if (!pin[0].done) {
    // keep two launches in flight; poll the older one
    enqueue(0);
    NGSAMD_HIP(hipEventRecord(ev[0], st));
    enqueue(1);
    NGSAMD_HIP(hipEventRecord(ev[1], st));
    int slot = 0;
    const long long max_launches = (long long)maxiter / per_launch + 4;
    for (long long k = 0; k < max_launches; k++) {
      NGSAMD_HIP(hipEventSynchronize(ev[slot]));
      if (pin[slot].done) break;
      enqueue(slot);
      NGSAMD_HIP(hipEventRecord(ev[slot], st));
      slot ^= 1;
    }
  }
  NGSAMD_HIP(hipMemcpyAsync(&hS, W.S.p, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  NGSAMD_REQUIRE(hS.done != 0, "cg: iteration did not terminate within maxiter (internal error)");
  CgResult res;
  res.iterations = hS.it;
  res.done = hS.done;
  return res;
}

// ------------------------------------------------------ allocation-free CG
// Work space of ngsamd_cg_solve_noalloc: the device scalars, the residual
// history and the work vectors d, w, s, in one file-scope device array that
// each GPU holds (its own copy, at its own address) until the process ends.
// Host code reaches it only through the address hipGetSymbolAddress gives,
// kept per device by InitCgScratch (called by ngsamd_init with the device
// current); the array's name on the host is not a device address.
constexpr int64_t kNoAllocMaxDofs = int64_t(1) << 22;  // 4,194,304 dofs
constexpr int kNoAllocMaxIter = 1 << 16;
constexpr size_t kNoAllocResOffset = 256;  // CGScalars first
constexpr size_t kNoAllocVecOffset = kNoAllocResOffset + sizeof(double) * kNoAllocMaxIter;
constexpr size_t kNoAllocScratchBytes = kNoAllocVecOffset + 3 * sizeof(double) * kNoAllocMaxDofs;
static_assert(sizeof(CGScalars) <= kNoAllocResOffset, "CGScalars must fit the scratch header");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kScratchDevices = 64;
static void* g_scratch_ptr[kScratchDevices] = {};

void InitCgScratch(int device) {
  NGSAMD_REQUIRE(device >= 0 && device < kScratchDevices, "ngsamd_init: device number out of range");
  void* p = nullptr;
  NGSAMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = p;
}

// CG from an initialised d (= f - A u) to the stopping test, every kernel
// launched directly (no graph, no events, no pinned memory): `chunk`
// iterations are enqueued, then the device scalars are copied to the host
// and checked.  Same kernels in the same order as IterateCg: the same
// iterates.
static CgResult IterateCgDirect(ngsamd_op_t A, const double* inv, ngsamd_op_t pre, double* u, double* d, double* w,
                                double* s, double* residuals, CGScalars* dS, int64_t n, double tol, double atol,
                                int maxiter, hipStream_t st) {
  CGScalars hS{};
  hS.tol = tol;
  hS.atol = atol;
  hS.maxiter = maxiter;
  NGSAMD_HIP(hipMemcpyAsync(dS, &hS, sizeof(CGScalars), hipMemcpyHostToDevice, st));
  const int g = CgGrid(n);
  if (!inv) EnqueuePrecond<double>(pre, d, w, nullptr, st);
  CgInitKernel<double><<<g, kCgBlock, 0, st>>>(inv, d, w, s, n, LibraryPartials().w);
  NGSAMD_CHECK_LAUNCH();
  CgInitCheckKernel<<<1, kCgBlock, 0, st>>>(LibraryPartials().w, g, dS, residuals);
  NGSAMD_CHECK_LAUNCH();
  constexpr int kChunk = 16;  // even: every chunk starts at parity 0
  long long launched = 0;
  while (true) {
    NGSAMD_HIP(hipMemcpyAsync(&hS, dS, sizeof(CGScalars), hipMemcpyDeviceToHost, st));
    NGSAMD_HIP(hipStreamSynchronize(st));
    if (hS.done || launched > (long long)maxiter + kChunk) break;
    for (int k = 0; k < kChunk; k++) EnqueueIteration<double>(A, inv, pre, u, d, w, s, n, dS, residuals, k & 1, st);
    launched += kChunk;
  }
  NGSAMD_REQUIRE(hS.done != 0, "cg: iteration did not terminate within maxiter (internal error)");
  CgResult res;
  res.iterations = hS.it;
  res.done = hS.done;
  return res;
}

static void CheckCgOperators(ngsamd_op_t A, ngsamd_op_t pre, bool jacobi_only) {
  NGSAMD_REQUIRE(A && pre, "cg: null operator");
  const bool generic = pre->kind == OpKind::BlockJacobi || pre->kind == OpKind::TwoLevel ||
                       pre->kind == OpKind::SparseCholesky || pre->kind == OpKind::MultiGrid;
  NGSAMD_REQUIRE(pre->kind == OpKind::Jacobi || (!jacobi_only && generic),
                 jacobi_only ? "cg: this solver needs a Jacobi preconditioner"
                             : "cg: the device CG needs a Jacobi, block-Jacobi, two-level, sparse Cholesky or "
                               "multigrid preconditioner");
  NGSAMD_REQUIRE(A->height == A->width && pre->height == A->height, "cg: operator shapes do not match");
}

}  // namespace ngsamd

using namespace ngsamd;

extern "C" int ngsamd_jacobi_create(ngsamd_op_t* op, int64_t n, const double* diag, const uint8_t* freedofs) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && (n == 0 || diag), "ngsamd_jacobi_create: null argument");
    *op = nullptr;
    *op = new JacobiOp(n, diag, false, freedofs);
  });
}

extern "C" int ngsamd_jacobi_create_device(ngsamd_op_t* op, int64_t n, const double* d_diag,
                                           const uint8_t* freedofs) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op && (n == 0 || d_diag), "ngsamd_jacobi_create_device: null argument");
    *op = nullptr;
    *op = new JacobiOp(n, d_diag, true, freedofs);
  });
}

extern "C" int ngsamd_cg_solve(ngsamd_op_t A, ngsamd_op_t pre, const double* f, double* u, int initialize,
                               double tol, double atol, int maxiter, int chunk, int* h_iterations,
                               double* h_residuals) {
  return Guard([&] {
    RequireInit();
    CheckCgOperators(A, pre, false);
    NGSAMD_REQUIRE(maxiter >= 1, "cg: maxiter must be >= 1");
    const int64_t n = A->height;
    hipStream_t st = Stream();
    CgWork<double> W(n, maxiter);
    // d = f - A u
    if (initialize) LaunchFill(u, n, 0.0, st);
    if (n) NGSAMD_HIP(hipMemcpyAsync(W.d.p, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (!initialize) A->MultAdd(-1.0, u, W.d.p, st);
    const double* inv = pre->kind == OpKind::Jacobi ? static_cast<JacobiOp*>(pre)->invdiag.p : nullptr;
    CgResult r = IterateCg<double>(A, inv, pre, u, W, n, tol, atol, maxiter, chunk, st);
    if (h_iterations) *h_iterations = r.iterations;
    if (h_residuals && r.iterations > 0) {
      const int cnt = r.iterations < maxiter ? r.iterations : maxiter;
      NGSAMD_HIP(hipMemcpyAsync(h_residuals, W.residuals.p, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
    }
  });
}

extern "C" int ngsamd_cg_solve_noalloc(ngsamd_op_t A, ngsamd_op_t pre, const double* f, double* u, int initialize,
                                       double tol, double atol, int maxiter, int* h_iterations,
                                       double* h_residuals) {
  return Guard([&] {
    RequireInit();
    CheckCgOperators(A, pre, false);
    NGSAMD_REQUIRE(maxiter >= 1 && maxiter <= kNoAllocMaxIter, "cg_solve_noalloc: maxiter must be in 1..65536");
    const int64_t n = A->height;
    NGSAMD_REQUIRE(n <= kNoAllocMaxDofs, "cg_solve_noalloc: more than 4194304 dofs (the scratch array's limit)");
    const int dev = State().device;
    NGSAMD_REQUIRE(dev >= 0 && dev < kScratchDevices && g_scratch_ptr[dev],
                   "cg_solve_noalloc: no scratch address for this device (ngsamd_init)");
    unsigned char* base = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    CGScalars* dS = reinterpret_cast<CGScalars*>(base);
    double* residuals = reinterpret_cast<double*>(base + kNoAllocResOffset);
    double* d = reinterpret_cast<double*>(base + kNoAllocVecOffset);
    double* w = d + kNoAllocMaxDofs;
    double* s = w + kNoAllocMaxDofs;
    hipStream_t st = Stream();
    // d = f - A u
    if (initialize) LaunchFill(u, n, 0.0, st);
    if (n) NGSAMD_HIP(hipMemcpyAsync(d, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (!initialize) A->MultAdd(-1.0, u, d, st);
    const double* inv = pre->kind == OpKind::Jacobi ? static_cast<JacobiOp*>(pre)->invdiag.p : nullptr;
    CgResult r = IterateCgDirect(A, inv, pre, u, d, w, s, residuals, dS, n, tol, atol, maxiter, st);
    if (h_iterations) *h_iterations = r.iterations;
    if (h_residuals && r.iterations > 0) {
      const int cnt = r.iterations < maxiter ? r.iterations : maxiter;
      NGSAMD_HIP(hipMemcpyAsync(h_residuals, residuals, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    NGSAMD_HIP(hipStreamSynchronize(st));
  });
}

extern "C" int ngsamd_ir_cg_solve(ngsamd_op_t A, ngsamd_op_t inner, ngsamd_op_t pre, const double* f, double* u,
                                  double tol, double inner_tol, int max_outer, int max_inner, double* h_stats) {
  return Guard([&] {
    RequireInit();
    CheckCgOperators(A, pre, true);
    NGSAMD_REQUIRE(inner && inner->kind == OpKind::MatrixFree && inner->height == A->height,
                   "ir_cg: inner operator must be matrix-free of the same size");
    NGSAMD_REQUIRE(max_outer >= 1 && max_inner >= 1, "ir_cg: iteration limits must be >= 1");
    const int64_t n = A->height;
    hipStream_t st = Stream();
    auto& R = State();
    MatFreeOp* Ai = static_cast<MatFreeOp*>(inner);
    DevBuf<double> r(n);
    DevBuf<float> inv32(n), c32(n);
    DevBuf<double> scal(2);  // [0] = 1/||r||, [1] = ||r||
    // FP32 copy of the Jacobi inverse diagonal
    const double* inv64 = static_cast<JacobiOp*>(pre)->invdiag.p;
    ngsamd_vec_f64_to_f32(n, inv64, inv32.p);
    CgWork<float> W(n, max_inner);

    // residual norms over the free dofs (where the Jacobi inverse is nonzero)
    DevBuf<double> sq(1);
    auto free_norm = [&](const double* x) {
      const int g = CgGrid(n);
      MaskedSqKernel<<<g, kCgBlock, 0, st>>>(x, inv64, n, R.partials);
      NGSAMD_CHECK_LAUNCH();
      SumToKernel<<<1, kCgBlock, 0, st>>>(R.partials, g, sq.p);
      NGSAMD_CHECK_LAUNCH();
      double h = 0.0;
      NGSAMD_HIP(hipMemcpyAsync(&h, sq.p, sizeof(double), hipMemcpyDeviceToHost, st));
      NGSAMD_HIP(hipStreamSynchronize(st));
      return std::sqrt(h);
    };
    LaunchFill(u, n, 0.0, st);
    const double fn = n ? free_norm(f) : 0.0;
    int outer = 0;
    long long inner_total = 0;
    double relres = 1.0;
    if (n) NGSAMD_HIP(hipMemcpyAsync(r.p, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    // default inner tolerance by the inner operator's precision (the
    // fastest setting measured on the 32^3 p=4 lognormal-kappa problem)
    if (inner_tol <= 0.0)
      inner_tol = Ai->precision == NGSAMD_PREC_F32 || Ai->precision == NGSAMD_PREC_F16X3 ? 3e-4
                  : Ai->precision == NGSAMD_PREC_F16                                      ? 1e-3
                                                                                          : 3e-2;
    double hsc[2] = {1.0, 1.0};  // host source of the async copy below: lives until the next sync
    // inner iterations per captured graph; NGSAMD_CG_CHUNK=0 launches every
    // kernel directly (no graph), e.g. for a kernel trace
    int chunk = 16;
    if (const char* ce = getenv("NGSAMD_CG_CHUNK")) chunk = std::max(0, atoi(ce));
    // the inner solves of every step run on the same vectors: one graph
    CgGraph inner_graph;
    while (true) {
      const double rn = n ? free_norm(r.p) : 0.0;
      relres = fn > 0 ? rn / fn : 0.0;
      if (rn <= tol * fn || outer >= max_outer || rn == 0.0) break;
      // inner solve A c = r / ||r|| in FP32 vectors
      hsc[0] = 1.0 / rn;
      hsc[1] = rn;
      NGSAMD_HIP(hipMemcpyAsync(scal.p, hsc, sizeof(hsc), hipMemcpyHostToDevice, st));
      ScaledToF32Kernel<<<GridFor(n, 256, 8192), 256, 0, st>>>(r.p, W.d.p, n, scal.p);
      NGSAMD_CHECK_LAUNCH();
      LaunchFillF32(c32.p, n, 0.0f, st);
      // the last refinement step only needs to gain the factor still missing
      double itol = inner_tol;
      if (fn > 0) itol = std::max(itol, std::min(0.5, 0.3 * tol * fn / rn));
      CgResult ir = IterateCg<float>(Ai, inv32.p, pre, c32.p, W, n, itol, -1.0, max_inner, chunk, st, &inner_graph);
      inner_total += ir.iterations;
      // u += ||r|| c; r = f - A u (FP64)
      AddScaledF32Kernel<<<GridFor(n, 256, 8192), 256, 0, st>>>(u, c32.p, n, scal.p + 1);
      NGSAMD_CHECK_LAUNCH();
      NGSAMD_HIP(hipMemcpyAsync(r.p, f, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      A->MultAdd(-1.0, u, r.p, st);
      outer++;
    }
    (void)R;
    if (h_stats) {
      h_stats[0] = outer;
      h_stats[1] = (double)inner_total;
      h_stats[2] = relres;
    }
  });
}
