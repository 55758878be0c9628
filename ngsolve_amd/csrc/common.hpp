// Internal runtime shared by the HIP translation units of libngsamd.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/ngsamd.h"

namespace ngsamd {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void SetLastError(const std::string& msg);
hipStream_t Stream();
void RequireInit();

#define NGSAMD_HIP(call)                                                                   \
  do {                                                                                     \
    hipError_t _e = (call);                                                                \
    if (_e != hipSuccess)                                                                  \
      throw ::ngsamd::Error(NGSAMD_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define NGSAMD_CHECK_LAUNCH() NGSAMD_HIP(hipGetLastError())

#define NGSAMD_REQUIRE(cond, msg)                                  \
  do {                                                             \
    if (!(cond)) throw ::ngsamd::Error(NGSAMD_E_ARG, (msg));       \
  } while (0)

// Wrap a C-ABI body: translate exceptions into status codes.
template <class F>
inline int Guard(F&& f) {
  try {
    f();
    return NGSAMD_OK;
  } catch (const Error& e) {
    SetLastError(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    SetLastError("host allocation failed");
    return NGSAMD_E_NOMEM;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return NGSAMD_E_ARG;
  }
}

// Device buffer owned by an operator.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) {
      hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
      if (e != hipSuccess) {
        p = nullptr;
        n = 0;
        throw Error(NGSAMD_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
      }
    }
  }
  void upload(const T* h, size_t count) {
    if (count > n) throw Error(NGSAMD_E_ARG, "DevBuf::upload: too many elements");
    if (count) NGSAMD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, Stream()));
  }
  void upload_at(const T* h, size_t offset, size_t count) {
    if (offset + count > n) throw Error(NGSAMD_E_ARG, "DevBuf::upload_at: out of range");
    if (count) NGSAMD_HIP(hipMemcpyAsync(p + offset, h, count * sizeof(T), hipMemcpyHostToDevice, Stream()));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

enum class OpKind { CSR, Jacobi, MatrixFree, Batch, BlockJacobi, SparseCholesky, TwoLevel, MultiGrid };

}  // namespace ngsamd

// Operator base: the C-ABI handle type.
struct ngsamd_op_s {
  ngsamd::OpKind kind;
  int64_t height = 0, width = 0;
  explicit ngsamd_op_s(ngsamd::OpKind k) : kind(k) {}
  virtual ~ngsamd_op_s() {}
  // y += s * A x on the given stream
  virtual void MultAdd(double s, const double* x, double* y, hipStream_t st) = 0;
  // y = A x (default: zero + MultAdd, linalg/basematrix.cpp:120-127)
  virtual void Mult(const double* x, double* y, hipStream_t st);
};

namespace ngsamd {
// launch helpers shared across units
void LaunchFill(double* x, int64_t n, double a, hipStream_t st);
void LaunchFillF32(float* x, int64_t n, float a, hipStream_t st);
int GridFor(int64_t n, int block, int max_blocks = 1 << 20);
// Sums consecutive groups of the n partial sums in[] into the 64 partials
// out[] (fixed order, deterministic); skipped when *stop != 0.
void LaunchFoldPartials(const double* in, int n, double* out, const int* stop, hipStream_t st);
// device properties of the current device (queried once)
int DeviceCUs();
int DeviceMaxLds();
// address of the allocation-free CG's scratch array on `device` (cg.hip),
// kept for ngsamd_cg_solve_noalloc; called by ngsamd_init with the device current
void InitCgScratch(int device);

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD
// b mod 8).  This maps the hardware id to a logical id such that each XCD owns
// a contiguous range of logical ids, so neighbouring work shares one L2.
__device__ __forceinline__ int XcdContiguousBlock(int b, int grid) {
  if ((grid & 7) != 0) return b;
  return (b & 7) * (grid >> 3) + (b >> 3);
}
}  // namespace ngsamd
