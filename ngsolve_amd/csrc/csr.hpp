// Device CSR operator and Jacobi preconditioner types shared by the SpMV and
// CG translation units.
#pragma once
#include <vector>

#include "common.hpp"

namespace ngsamd {

struct CsrOp : ngsamd_op_s {
  DevBuf<int> rowptr, cols;
  DevBuf<double> vals;
  int64_t nnz = 0;
  int group = 32;  // lanes per row (rowgroup kernel)
  int algo = 1;    // 0: rowgroup, 1: stream
  DevBuf<int> blk_row;  // stream kernel row blocks
  int nblocks = 0;
  int block_nnz = 2048;
  int vec = 1;     // stream kernel: nonzeros per lane load (1 or 4)
  CsrOp(int64_t h, int64_t w, const int64_t* rp, const int* c, const double* v);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y = A x plus per-workgroup partial sums of <x, A x> in partials[];
  // returns the number of partials; skipped entirely when *d_stop != 0
  int MultDot(const double* x, double* y, double* partials, const int* d_stop, hipStream_t st);
  void Diagonal(double* d, hipStream_t st);
};

struct JacobiOp : ngsamd_op_s {
  DevBuf<double> invdiag;  // 1/a_ii on free dofs, 0 elsewhere
  JacobiOp(int64_t n, const double* diag, bool diag_on_device, const uint8_t* freedofs);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
};

}  // namespace ngsamd
