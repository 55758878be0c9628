// Runtime, memory and BaseVector-style vector operations of libngsamd.
#include <mutex>
#include <string>

#include "common.hpp"
#include "reduce.hpp"
#include "runtime_state.hpp"

namespace ngsamd {

static thread_local std::string g_last_error;
static RuntimeState g_state;

void SetLastError(const std::string& msg) { g_last_error = msg; }

RuntimeState& State() { return g_state; }

void RequireInit() {
  if (!g_state.initialised) throw Error(NGSAMD_E_STATE, "ngsamd_init has not been called");
}

hipStream_t Stream() { return g_state.user_stream ? g_state.user_stream : g_state.stream; }

int GridFor(int64_t n, int block, int max_blocks) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return (int)g;
}

__global__ __launch_bounds__(256) void FoldPartialsKernel(const double* __restrict__ in, int n,
                                                          double* __restrict__ out, const int* __restrict__ stop) {
  if (stop && *stop) return;
  __shared__ double lds[256 / 64];
  const int per = (n + kFoldedPartials - 1) / kFoldedPartials;
  const int b = blockIdx.x * per;
  const int cnt = max(0, min(per, n - b));
  const double v = SumPartials<256>(in + b, cnt, lds);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

void LaunchFoldPartials(const double* in, int n, double* out, const int* stop, hipStream_t st) {
  FoldPartialsKernel<<<kFoldedPartials, 256, 0, st>>>(in, n, out, stop);
  NGSAMD_CHECK_LAUNCH();
}

int DeviceCUs() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, val = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&val, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) val = 256;
    v = val;
  }
  return v;
}

int DeviceMaxLds() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, val = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&val, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) val = 65536;
    v = val;
  }
  return v;
}

// ------------------------------------------------------------ fill kernels
__global__ void FillKernel(double* __restrict__ x, int64_t n, double a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = a;
}
__global__ void FillKernelF32(float* __restrict__ x, int64_t n, float a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] = a;
}
void LaunchFill(double* x, int64_t n, double a, hipStream_t st) {
  if (n <= 0) return;
  FillKernel<<<GridFor(n, 256, 8192), 256, 0, st>>>(x, n, a);
  NGSAMD_CHECK_LAUNCH();
}
void LaunchFillF32(float* x, int64_t n, float a, hipStream_t st) {
  if (n <= 0) return;
  FillKernelF32<<<GridFor(n, 256, 8192), 256, 0, st>>>(x, n, a);
  NGSAMD_CHECK_LAUNCH();
}

// -------------------------------------------------------------- vector ops
__global__ void ScaleKernel(double* __restrict__ x, int64_t n, double a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    x[i] *= a;
}
__global__ void AxpyKernel(double* __restrict__ y, const double* __restrict__ x, int64_t n, double a) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
__global__ void AxpbyKernel(double* __restrict__ y, const double* __restrict__ x, int64_t n, double a, double b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}
__global__ void ToF32Kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
__global__ void ToF64Kernel(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (double)x[i];
}

// Dot product: fixed grid for a given n, per-workgroup partials, then a
// one-workgroup kernel sums them in index order -> deterministic result.
template <class T>
__global__ __launch_bounds__(256) void DotPartialKernel(const T* __restrict__ x, const T* __restrict__ y, int64_t n,
                                                        double* __restrict__ partials) {
  __shared__ double lds[4];
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    v += (double)x[i] * (double)y[i];
  StorePartial<256>(v, partials, lds);
}

__global__ __launch_bounds__(256) void SumPartialsKernel(const double* __restrict__ partials, int count,
                                                         double* __restrict__ result) {
  __shared__ double lds[4];
  double v = SumPartials<256>(partials, count, lds);
  if (threadIdx.x == 0) *result = v;
}

// Read-bandwidth calibration: every lane streams 16-byte non-temporal loads,
// four in flight, over a grid-stride loop; each workgroup stores one partial
// sum so that nothing is optimised away.
typedef double dbl2v __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void StreamReadKernel(const dbl2v* __restrict__ x, int64_t n2,
                                                        double* __restrict__ partials) {
  __shared__ double lds[4];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const dbl2v v0 = __builtin_nontemporal_load(x + i);
    const dbl2v v1 = __builtin_nontemporal_load(x + i + stride);
    const dbl2v v2 = __builtin_nontemporal_load(x + i + 2 * stride);
    const dbl2v v3 = __builtin_nontemporal_load(x + i + 3 * stride);
    a0 += v0.x + v0.y;
    a1 += v1.x + v1.y;
    a2 += v2.x + v2.y;
    a3 += v3.x + v3.y;
  }
  for (; i < n2; i += stride) {
    const dbl2v v = __builtin_nontemporal_load(x + i);
    a0 += v.x + v.y;
  }
  StorePartial<256>((a0 + a1) + (a2 + a3), partials, lds);
}

template <class T>
static double Dot(int64_t n, const T* x, const T* y) {
  RequireInit();
  auto& S = State();
  hipStream_t st = Stream();
  int grid = GridFor(n, 256, 2048);
  DotPartialKernel<T><<<grid, 256, 0, st>>>(x, y, n, S.partials);
  NGSAMD_CHECK_LAUNCH();
  SumPartialsKernel<<<1, 256, 0, st>>>(S.partials, grid, S.dev_scalars);
  NGSAMD_CHECK_LAUNCH();
  NGSAMD_HIP(hipMemcpyAsync(S.host_scalars, S.dev_scalars, sizeof(double), hipMemcpyDeviceToHost, st));
  NGSAMD_HIP(hipStreamSynchronize(st));
  return S.host_scalars[0];
}

}  // namespace ngsamd

using namespace ngsamd;

void ngsamd_op_s::Mult(const double* x, double* y, hipStream_t st) {
  LaunchFill(y, height, 0.0, st);
  MultAdd(1.0, x, y, st);
}

extern "C" {

int ngsamd_init(int device) {
  return Guard([&] {
    auto& S = State();
    std::lock_guard<std::mutex> lock(S.mutex);
    if (S.initialised) {
      if (device != S.device) throw Error(NGSAMD_E_STATE, "ngsamd_init: already initialised on another device");
      return;
    }
    NGSAMD_HIP(hipSetDevice(device));
    S.device = device;
    NGSAMD_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    NGSAMD_HIP(hipMalloc((void**)&S.partials, sizeof(double) * 2 * kMaxReduceBlocks));
    NGSAMD_HIP(hipMalloc((void**)&S.wide_partials, sizeof(double) * kMaxWideBlocks));
    NGSAMD_HIP(hipMalloc((void**)&S.dev_scalars, sizeof(double) * kNumScalars));
    NGSAMD_HIP(hipHostMalloc((void**)&S.host_scalars, sizeof(double) * kNumScalars, hipHostMallocDefault));
    InitCgScratch(device);
    NGSAMD_HIP(hipDeviceSynchronize());
    S.initialised = true;
  });
}

const char* ngsamd_last_error(void) { return g_last_error.c_str(); }

int ngsamd_set_stream(void* s) {
  return Guard([&] {
    RequireInit();
    State().user_stream = (hipStream_t)s;
  });
}

void* ngsamd_get_stream(void) { return (void*)Stream(); }

int ngsamd_synchronize(void) {
  return Guard([&] {
    RequireInit();
    NGSAMD_HIP(hipStreamSynchronize(Stream()));
  });
}

int ngsamd_malloc(void** p, size_t bytes) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(p != nullptr, "ngsamd_malloc: null output pointer");
    *p = nullptr;
    if (bytes == 0) return;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
      *p = nullptr;
      throw Error(NGSAMD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
  });
}

int ngsamd_free(void* p) {
  return Guard([&] {
    if (p) NGSAMD_HIP(hipFree(p));
  });
}

int ngsamd_memcpy_htod(void* d, const void* h, size_t bytes) {
  return Guard([&] {
    RequireInit();
    if (bytes) {
      NGSAMD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, Stream()));
      NGSAMD_HIP(hipStreamSynchronize(Stream()));
    }
  });
}

int ngsamd_memcpy_dtoh(void* h, const void* d, size_t bytes) {
  return Guard([&] {
    RequireInit();
    if (bytes) {
      NGSAMD_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, Stream()));
      NGSAMD_HIP(hipStreamSynchronize(Stream()));
    }
  });
}

int ngsamd_memcpy_dtod(void* d, const void* s, size_t bytes) {
  return Guard([&] {
    RequireInit();
    if (bytes) NGSAMD_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, Stream()));
  });
}

int ngsamd_event_create(void** ev) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(ev != nullptr, "ngsamd_event_create: null output");
    hipEvent_t e;
    NGSAMD_HIP(hipEventCreate(&e));
    *ev = (void*)e;
  });
}

int ngsamd_event_record(void* ev) {
  return Guard([&] {
    RequireInit();
    NGSAMD_HIP(hipEventRecord((hipEvent_t)ev, Stream()));
  });
}

int ngsamd_event_elapsed_ms(void* start, void* stop, float* ms) {
  return Guard([&] {
    NGSAMD_REQUIRE(ms != nullptr, "ngsamd_event_elapsed_ms: null output");
    NGSAMD_HIP(hipEventSynchronize((hipEvent_t)stop));
    NGSAMD_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  });
}

int ngsamd_event_destroy(void* ev) {
  return Guard([&] {
    if (ev) NGSAMD_HIP(hipEventDestroy((hipEvent_t)ev));
  });
}

int ngsamd_vec_dot(int64_t n, const double* x, const double* y, double* h_result) {
  return Guard([&] {
    NGSAMD_REQUIRE(h_result != nullptr, "ngsamd_vec_dot: null result");
    *h_result = Dot<double>(n, x, y);
  });
}

int ngsamd_vec_dot_f32(int64_t n, const float* x, const float* y, double* h_result) {
  return Guard([&] {
    NGSAMD_REQUIRE(h_result != nullptr, "ngsamd_vec_dot_f32: null result");
    *h_result = Dot<float>(n, x, y);
  });
}

int ngsamd_stream_read(int64_t n, const double* x, double* d_result) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(n >= 0 && n % 2 == 0, "ngsamd_stream_read: n must be even and >= 0");
    NGSAMD_REQUIRE(d_result && (x || n == 0), "ngsamd_stream_read: null argument");
    NGSAMD_REQUIRE(((uintptr_t)x & 15) == 0, "ngsamd_stream_read: x must be 16-byte aligned");
    auto& S = State();
    hipStream_t st = Stream();
    const int grid = GridFor(n / 2, 256 * 16, kMaxReduceBlocks);
    StreamReadKernel<<<grid, 256, 0, st>>>(reinterpret_cast<const dbl2v*>(x), n / 2, S.partials);
    NGSAMD_CHECK_LAUNCH();
    SumPartialsKernel<<<1, 256, 0, st>>>(S.partials, grid, d_result);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_vec_set(int64_t n, double a, double* x) {
  return Guard([&] {
    RequireInit();
    LaunchFill(x, n, a, Stream());
  });
}

int ngsamd_vec_scale(int64_t n, double a, double* x) {
  return Guard([&] {
    RequireInit();
    if (n <= 0) return;
    ScaleKernel<<<GridFor(n, 256, 8192), 256, 0, Stream()>>>(x, n, a);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_vec_axpy(int64_t n, double a, const double* x, double* y) {
  return Guard([&] {
    RequireInit();
    if (n <= 0) return;
    AxpyKernel<<<GridFor(n, 256, 8192), 256, 0, Stream()>>>(y, x, n, a);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_vec_axpby(int64_t n, double a, const double* x, double b, double* y) {
  return Guard([&] {
    RequireInit();
    if (n <= 0) return;
    AxpbyKernel<<<GridFor(n, 256, 8192), 256, 0, Stream()>>>(y, x, n, a, b);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_vec_f64_to_f32(int64_t n, const double* x, float* y) {
  return Guard([&] {
    RequireInit();
    if (n <= 0) return;
    ToF32Kernel<<<GridFor(n, 256, 8192), 256, 0, Stream()>>>(x, y, n);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_vec_f32_to_f64(int64_t n, const float* x, double* y) {
  return Guard([&] {
    RequireInit();
    if (n <= 0) return;
    ToF64Kernel<<<GridFor(n, 256, 8192), 256, 0, Stream()>>>(x, y, n);
    NGSAMD_CHECK_LAUNCH();
  });
}

int ngsamd_op_height(ngsamd_op_t op, int64_t* h) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && h, "ngsamd_op_height: null argument");
    *h = op->height;
  });
}

int ngsamd_op_width(ngsamd_op_t op, int64_t* w) {
  return Guard([&] {
    NGSAMD_REQUIRE(op && w, "ngsamd_op_width: null argument");
    *w = op->width;
  });
}

int ngsamd_op_mult_add(ngsamd_op_t op, double s, const double* x, double* y) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op != nullptr, "ngsamd_op_mult_add: null operator");
    op->MultAdd(s, x, y, Stream());
  });
}

int ngsamd_op_mult(ngsamd_op_t op, const double* x, double* y) {
  return Guard([&] {
    RequireInit();
    NGSAMD_REQUIRE(op != nullptr, "ngsamd_op_mult: null operator");
    op->Mult(x, y, Stream());
  });
}

int ngsamd_op_destroy(ngsamd_op_t op) {
  return Guard([&] {
    if (op) {
      // make sure no queued kernel still reads the operator's buffers
      if (State().initialised) NGSAMD_HIP(hipStreamSynchronize(Stream()));
      delete op;
    }
  });
}

}  // extern "C"
