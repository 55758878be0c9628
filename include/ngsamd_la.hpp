// ngsamd_la.hpp -- C++ operator layer over the C ABI of include/ngsamd.h, with
// the virtual interface of NGSolve's ngla::BaseMatrix / BaseVector, so that a
// device backend module can hand these objects to NGSolve code that works on
// BaseMatrix references (Mult, MultAdd, CreateRowVector, ...).
//
// Header-only; link with libngsamd.so.  Reference interfaces mirrored:
//   BaseMatrix      VHeight / VWidth             linalg/basematrix.hpp:45-50
//                   IsSymmetric                  linalg/basematrix.hpp:64
//                   CreateRowVector / ColVector  linalg/basematrix.hpp:96-99
//                   Mult / MultAdd               linalg/basematrix.hpp:117-121
//                   MultTrans / MultTransAdd     linalg/basematrix.hpp:123-131
//                   defaults: Mult = SetZero + MultAdd, MultAdd = temporary + Mult,
//                   a safety check against recursion when neither is
//                   overridden                   linalg/basematrix.cpp:120-154
//   composites      SumMatrix, VScaleMatrix      linalg/basematrix.hpp:413,837
//   BaseVector      Size, SetScalar, Add(s, v), Scale, InnerProduct, FVDouble
//                   linalg/basevector.hpp:233-273; host/device mirror with
//                   dirty flags as UnifiedVector ngscuda/unifiedvector.cpp:288-330
//   DevSparseMatrix ngscuda/cuda_linalg.cpp:187-275 (SparseMatrix<double> storage)
//   JacobiPrecond   linalg/jacobi.cpp:41-90
//   BlockJacobiPrecond linalg/blockjacobi.cpp:390-632
//   CGSolver        linalg/cg.cpp:503-625 (GetSteps), python/krylovspace.py:227-291
// Errors: every failing ngsamd_* call throws ngsamd::Exception carrying
// ngsamd_last_error(), as NGSolve throws ngcore::Exception.
#ifndef NGSAMD_LA_HPP
#define NGSAMD_LA_HPP

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "ngsamd.h"

namespace ngsamd {

class Exception : public std::runtime_error {
 public:
  explicit Exception(const std::string& m) : std::runtime_error(m) {}
};

inline void Check(int status) {
  if (status != NGSAMD_OK) throw Exception(std::string("ngsamd: ") + ngsamd_last_error());
}

// ------------------------------------------------------------------ vectors

class BaseVector {
 public:
  virtual ~BaseVector() = default;
  virtual size_t Size() const = 0;
  // device memory, current (host changes are uploaded first)
  virtual double* DevData() const = 0;
  // host copy, current (device changes are downloaded first); after writing
  // through it, call HostModified()
  virtual double* FVDouble() const = 0;
  virtual void HostModified() = 0;
  virtual void DeviceModified() = 0;

  BaseVector& SetScalar(double a) {
    Check(ngsamd_vec_set((int64_t)Size(), a, DevData()));
    DeviceModified();
    return *this;
  }
  BaseVector& Scale(double a) {
    Check(ngsamd_vec_scale((int64_t)Size(), a, DevData()));
    DeviceModified();
    return *this;
  }
  // this += s * v
  BaseVector& Add(double s, const BaseVector& v) {
    if (v.Size() != Size()) throw Exception("BaseVector::Add: sizes differ");
    Check(ngsamd_vec_axpy((int64_t)Size(), s, v.DevData(), DevData()));
    DeviceModified();
    return *this;
  }
  BaseVector& Set(double s, const BaseVector& v) {
    if (v.Size() != Size()) throw Exception("BaseVector::Set: sizes differ");
    Check(ngsamd_vec_axpby((int64_t)Size(), s, v.DevData(), 0.0, DevData()));
    DeviceModified();
    return *this;
  }
  double InnerProduct(const BaseVector& v) const {
    if (v.Size() != Size()) throw Exception("BaseVector::InnerProduct: sizes differ");
    double r = 0.0;
    Check(ngsamd_vec_dot((int64_t)Size(), DevData(), v.DevData(), &r));
    return r;
  }
  double L2Norm() const {
    const double d = InnerProduct(*this);
    return d > 0 ? std::sqrt(d) : 0.0;
  }
};

inline double InnerProduct(const BaseVector& a, const BaseVector& b) { return a.InnerProduct(b); }

// A device buffer (ngsamd_malloc) with a host mirror; each side is copied to
// the other only when it is stale.
class DeviceVector : public BaseVector {
 public:
  explicit DeviceVector(size_t n) : n_(n), host_(n) {
    if (n_) Check(ngsamd_malloc((void**)&dev_, n_ * sizeof(double)));
  }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  ~DeviceVector() override {
    if (dev_) ngsamd_free(dev_);
  }
  size_t Size() const override { return n_; }
  double* DevData() const override {
    if (dev_stale_ && n_) {
      Check(ngsamd_memcpy_htod(dev_, host_.data(), n_ * sizeof(double)));
      dev_stale_ = false;
    }
    return dev_;
  }
  double* FVDouble() const override {
    if (host_stale_ && n_) {
      Check(ngsamd_memcpy_dtoh(host_.data(), dev_, n_ * sizeof(double)));
      host_stale_ = false;
    }
    return host_.data();
  }
  void HostModified() override {
    host_stale_ = false;
    dev_stale_ = true;
  }
  void DeviceModified() override {
    dev_stale_ = false;
    host_stale_ = true;
  }

 private:
  size_t n_;
  double* dev_ = nullptr;
  mutable std::vector<double> host_;
  mutable bool host_stale_ = false, dev_stale_ = true;
};

using AutoVector = std::unique_ptr<BaseVector>;

// ---------------------------------------------------------------- operators

class BaseMatrix {
 public:
  virtual ~BaseMatrix() = default;
  virtual int VHeight() const { throw Exception("BaseMatrix::VHeight not overloaded"); }
  virtual int VWidth() const { throw Exception("BaseMatrix::VWidth not overloaded"); }
  int Height() const { return VHeight(); }
  int Width() const { return VWidth(); }
  virtual bool IsSymmetric() const { return false; }
  virtual AutoVector CreateRowVector() const { return AutoVector(new DeviceVector(VWidth())); }
  virtual AutoVector CreateColVector() const { return AutoVector(new DeviceVector(VHeight())); }
  virtual std::string GetOperatorInfo() const { return "BaseMatrix"; }

  // y = A x; default: y = 0, then MultAdd(1, x, y)
  virtual void Mult(const BaseVector& x, BaseVector& y) const {
    Guard g(this, "Mult");
    y.SetScalar(0.0);
    MultAdd(1.0, x, y);
  }
  // y += s A x; default: a temporary t = A x, then y += s t
  virtual void MultAdd(double s, const BaseVector& x, BaseVector& y) const {
    Guard g(this, "MultAdd");
    AutoVector t = CreateColVector();
    Mult(x, *t);
    y.Add(s, *t);
  }
  // transposes: the operators here are symmetric, others must override
  virtual void MultTrans(const BaseVector& x, BaseVector& y) const {
    if (!IsSymmetric()) throw Exception("BaseMatrix::MultTrans not overloaded");
    Mult(x, y);
  }
  virtual void MultTransAdd(double s, const BaseVector& x, BaseVector& y) const {
    if (!IsSymmetric()) throw Exception("BaseMatrix::MultTransAdd not overloaded");
    MultAdd(s, x, y);
  }

 protected:
  // Mult and MultAdd call each other by default; an operator overriding
  // neither would recurse without end: the second entry throws instead.
  struct Guard {
    const BaseMatrix* m;
    Guard(const BaseMatrix* mm, const char* what) : m(mm) {
      if (m->in_default_ > 0)
        throw Exception(std::string("BaseMatrix::") + what + ": neither Mult nor MultAdd is overloaded");
      m->in_default_++;
    }
    ~Guard() { m->in_default_--; }
  };
  mutable int in_default_ = 0;
};

// An operator held by an ngsamd_op_t handle (owned: destroyed with it).
class DeviceOperator : public BaseMatrix {
 public:
  explicit DeviceOperator(ngsamd_op_t op, bool symmetric = true) : op_(op), symmetric_(symmetric) {
    int64_t h = 0, w = 0;
    Check(ngsamd_op_height(op_, &h));
    Check(ngsamd_op_width(op_, &w));
    h_ = (int)h;
    w_ = (int)w;
  }
  DeviceOperator(const DeviceOperator&) = delete;
  DeviceOperator& operator=(const DeviceOperator&) = delete;
  ~DeviceOperator() override {
    if (op_) ngsamd_op_destroy(op_);
  }
  int VHeight() const override { return h_; }
  int VWidth() const override { return w_; }
  bool IsSymmetric() const override { return symmetric_; }
  void Mult(const BaseVector& x, BaseVector& y) const override {
    CheckSizes(x, y);
    Check(ngsamd_op_mult(op_, x.DevData(), y.DevData()));
    y.DeviceModified();
  }
  void MultAdd(double s, const BaseVector& x, BaseVector& y) const override {
    CheckSizes(x, y);
    Check(ngsamd_op_mult_add(op_, s, x.DevData(), y.DevData()));
    y.DeviceModified();
  }
  ngsamd_op_t Handle() const { return op_; }

 protected:
  void CheckSizes(const BaseVector& x, const BaseVector& y) const {
    if ((int)x.Size() != w_ || (int)y.Size() != h_) throw Exception(GetOperatorInfo() + ": vector sizes do not match");
  }
  ngsamd_op_t op_ = nullptr;
  int h_ = 0, w_ = 0;
  bool symmetric_ = true;
};

// SparseMatrix<double> storage: firsti (size_t, height + 1), colnr (int,
// sorted per row), data (double).
class DevSparseMatrix : public DeviceOperator {
 public:
  DevSparseMatrix(int64_t height, int64_t width, const size_t* firsti, const int* colnr, const double* data,
                  bool symmetric = true)
      : DeviceOperator(Create(height, width, firsti, colnr, data), symmetric) {}
  std::string GetOperatorInfo() const override { return "DevSparseMatrix"; }

 private:
  static ngsamd_op_t Create(int64_t h, int64_t w, const size_t* firsti, const int* colnr, const double* data) {
    if (h < 0 || w < 0) throw Exception("DevSparseMatrix: negative size");
    std::vector<int64_t> rp(firsti, firsti + h + 1);
    ngsamd_op_t op = nullptr;
    Check(ngsamd_csr_create(&op, h, w, rp.data(), colnr, data));
    return op;
  }
};

// JacobiPrecond: y += s diag^-1 x on the free dofs (freedofs may be null).
class DevJacobi : public DeviceOperator {
 public:
  DevJacobi(int64_t n, const double* diag, const uint8_t* freedofs)
      : DeviceOperator(Create(n, diag, freedofs), true) {}
  std::string GetOperatorInfo() const override { return "DevJacobi"; }

 private:
  static ngsamd_op_t Create(int64_t n, const double* diag, const uint8_t* freedofs) {
    ngsamd_op_t op = nullptr;
    Check(ngsamd_jacobi_create(&op, n, diag, freedofs));
    return op;
  }
};

// BlockJacobiPrecond over the blocks (block_ptr offsets into block_dofs) of
// the CSR matrix given by (firsti, colnr, data).
class DevBlockJacobi : public DeviceOperator {
 public:
  DevBlockJacobi(int64_t n, const size_t* firsti, const int* colnr, const double* data, int64_t nblocks,
                 const int64_t* block_ptr, const int* block_dofs)
      : DeviceOperator(Create(n, firsti, colnr, data, nblocks, block_ptr, block_dofs), true) {}
  std::string GetOperatorInfo() const override { return "DevBlockJacobi"; }

 private:
  static ngsamd_op_t Create(int64_t n, const size_t* firsti, const int* colnr, const double* data, int64_t nb,
                            const int64_t* bp, const int* bd) {
    if (n < 0) throw Exception("DevBlockJacobi: negative size");
    std::vector<int64_t> rp(firsti, firsti + n + 1);
    ngsamd_op_t op = nullptr;
    Check(ngsamd_blockjacobi_create(&op, n, rp.data(), colnr, data, nb, bp, bd));
    return op;
  }
};

// The fused matrix-free H1 Laplace operator (arguments as ngsamd_mf_create).
class DevMatrixFreeH1 : public DeviceOperator {
 public:
  DevMatrixFreeH1(int dim, int nd, int nq, int64_t ndof, int nclass, const int64_t* class_offsets,
                  const double* bmat, const double* weights, const int* dofs, const double* metric,
                  const double* kappa, const int* elnr, int precision = NGSAMD_PREC_F64,
                  int scatter = NGSAMD_SCATTER_GATHER)
      : DeviceOperator(Create(dim, nd, nq, ndof, nclass, class_offsets, bmat, weights, dofs, metric, kappa, elnr,
                              precision, scatter),
                       true) {}
  std::string GetOperatorInfo() const override { return "DevMatrixFreeH1"; }

 private:
  static ngsamd_op_t Create(int dim, int nd, int nq, int64_t ndof, int nclass, const int64_t* off, const double* b,
                            const double* w, const int* dofs, const double* met, const double* kap, const int* el,
                            int prec, int scat) {
    ngsamd_op_t op = nullptr;
    Check(ngsamd_mf_create(&op, dim, nd, nq, ndof, nclass, off, b, w, dofs, met, kap, el, prec, scat));
    return op;
  }
};

// SumMatrix: (A + B) x
class SumMatrix : public BaseMatrix {
 public:
  SumMatrix(std::shared_ptr<BaseMatrix> a, std::shared_ptr<BaseMatrix> b) : a_(std::move(a)), b_(std::move(b)) {
    if (a_->Height() != b_->Height() || a_->Width() != b_->Width()) throw Exception("SumMatrix: shapes differ");
  }
  int VHeight() const override { return a_->Height(); }
  int VWidth() const override { return a_->Width(); }
  bool IsSymmetric() const override { return a_->IsSymmetric() && b_->IsSymmetric(); }
  void MultAdd(double s, const BaseVector& x, BaseVector& y) const override {
    a_->MultAdd(s, x, y);
    b_->MultAdd(s, x, y);
  }
  void Mult(const BaseVector& x, BaseVector& y) const override {
    a_->Mult(x, y);
    b_->MultAdd(1.0, x, y);
  }

 private:
  std::shared_ptr<BaseMatrix> a_, b_;
};

// VScaleMatrix: (c A) x
class ScaleMatrix : public BaseMatrix {
 public:
  ScaleMatrix(double c, std::shared_ptr<BaseMatrix> a) : c_(c), a_(std::move(a)) {}
  int VHeight() const override { return a_->Height(); }
  int VWidth() const override { return a_->Width(); }
  bool IsSymmetric() const override { return a_->IsSymmetric(); }
  void MultAdd(double s, const BaseVector& x, BaseVector& y) const override { a_->MultAdd(s * c_, x, y); }

 private:
  double c_;
  std::shared_ptr<BaseMatrix> a_;
};

// CGSolver: Mult(f, u) solves A u = f with the preconditioner (the device CG,
// CGSolver._SolveImpl's recurrence and stopping rule); GetSteps() is the
// number of residual checks, the initial one included.
class CGSolver : public BaseMatrix {
 public:
  CGSolver(std::shared_ptr<DeviceOperator> mat, std::shared_ptr<DeviceOperator> pre, double tol = 1e-8,
           int maxsteps = 200, double atol = -1.0)
      : mat_(std::move(mat)), pre_(std::move(pre)), tol_(tol), atol_(atol), maxsteps_(maxsteps) {}
  int VHeight() const override { return mat_->Width(); }
  int VWidth() const override { return mat_->Height(); }
  bool IsSymmetric() const override { return true; }
  void Mult(const BaseVector& f, BaseVector& u) const override {
    if ((int)f.Size() != mat_->Height() || (int)u.Size() != mat_->Width())
      throw Exception("CGSolver: vector sizes do not match");
    residuals_.assign(maxsteps_, 0.0);
    int its = 0;
    Check(ngsamd_cg_solve(mat_->Handle(), pre_->Handle(), f.DevData(), u.DevData(), 1, tol_, atol_, maxsteps_, 16,
                          &its, residuals_.data()));
    u.DeviceModified();
    steps_ = its;
    residuals_.resize(its < maxsteps_ ? its : maxsteps_);
  }
  int GetSteps() const { return steps_; }
  const std::vector<double>& Residuals() const { return residuals_; }

 private:
  std::shared_ptr<DeviceOperator> mat_, pre_;
  double tol_, atol_;
  int maxsteps_;
  mutable int steps_ = 0;
  mutable std::vector<double> residuals_;
};

}  // namespace ngsamd

#endif  // NGSAMD_LA_HPP
