// Device CSR operator and Jacobi / block-Jacobi preconditioner types shared
// by the SpMV, preconditioner and CG translation units.
#pragma once
#include <vector>

#include "common.hpp"

namespace ngsamd {

// Pattern runs: up to kMaxRun consecutive rows with one column list share
// that list (the dofs of one mesh node have the same sparsity pattern); the
// values of a run are stored column by column, [L][R].
constexpr int kMaxRun = 8;

struct CsrOp : ngsamd_op_s {
  // rowgroup / stream layout (plain CSR)
  DevBuf<int> rowptr, cols;
  DevBuf<double> vals;
  // runs layout
  DevBuf<int> run_row;  // first row of each run, nruns + 1 entries
  DevBuf<int> run_ptr;  // first shared column of each run, nruns + 1
  DevBuf<int> val_ptr;  // first value of each run, nruns + 1
  DevBuf<int> rcols;    // shared column lists
  DevBuf<double> rvals; // values, run by run, [L][R]
  DevBuf<int> blk_run;  // runs kernel blocks: first run, nblocks + 1
  DevBuf<int> blk_len;  // runs kernel blocks: rows per run (all runs of a block alike)
  DevBuf<int> row_run;  // run of each row (structure-only operators, runs layout)
  int64_t nruns = 0, nshared = 0;
  int64_t max_row_len = 0;
  int64_t nnz = 0;
  // stands in for x when the width is 0 (padding lanes still read x[0])
  DevBuf<double> xpad;
  int group = 32;  // lanes per row (rowgroup kernel)
  int algo = 1;    // 0: rowgroup, 1: stream, 2: runs
  DevBuf<int> blk_row;  // stream kernel row blocks
  int nblocks = 0;
  int block_nnz = 2048;
  int max_grid = 16384;  // workgroups of the stream/runs kernels
  int flip = 0;          // direction of the next Mult / MultAdd block walk (alternates)
  // v null: the structure only, values zero until filled on the device
  CsrOp(int64_t h, int64_t w, const int64_t* rp, const int* c, const double* v);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y = A x plus per-workgroup partial sums of <x, A x> in partials[] (room
  // for kMaxReduceBlocks); returns the number of partials; skipped entirely
  // when *d_stop != 0.
  // wide: room for kMaxWideBlocks partials of the wide grid (null: the
  // library's)
  int MultDot(const double* x, double* y, double* partials, const int* d_stop, hipStream_t st,
              double* wide = nullptr);
  // Y = A X for K vectors stored interleaved ([n][K]); the matrix is read
  // once for all K in the runs layout
  // (all kernels return at once when *stop != 0, stop may be null)
  void MultMulti(int K, const double* x, double* y, hipStream_t st, const int* stop = nullptr);
  DevBuf<double> xcomp, ycomp;  // one component, for the other layouts
  void Diagonal(double* d, hipStream_t st);
  // the values in CSR order (row by row, columns as given at creation)
  void Values(double* host_vals) const;
};

struct JacobiOp : ngsamd_op_s {
  DevBuf<double> invdiag;  // 1/a_ii on free dofs, 0 elsewhere
  JacobiOp(int64_t n, const double* diag, bool diag_on_device, const uint8_t* freedofs);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
};

// Block Jacobi: y += s sum_b P_b^T inv(A_bb) P_b x.  Block rows are numbered
// block after block (each block's dofs sorted); the inverses are stored
// column by column.  The apply forms z[t] = (inv_b x_b)[r] for every block row t, then
// each dof sums the z of its rows in row order (blocks may overlap).
struct BlockJacobiOp : ngsamd_op_s {
  int64_t nblocks = 0, nrows = 0;
  DevBuf<int64_t> bptr;   // first block row of each block, nblocks + 1
  DevBuf<int64_t> vptr;   // first inverse entry of each block, nblocks + 1
  DevBuf<int> bdofs;      // dof of each block row
  DevBuf<int> row_block;  // block of each block row
  DevBuf<double> inv;     // inverses, column-major per block
  DevBuf<int> gptr;       // per dof: its block rows [gptr[d], gptr[d+1]) of grows
  DevBuf<int> grows;
  DevBuf<double> z;       // block-row results (preallocated: the apply allocates nothing)
  BlockJacobiOp(int64_t n, const int64_t* rp, const int* cols, const double* vals, int64_t nblocks,
                const int64_t* block_ptr, const int* block_dofs);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y = C x; every kernel returns at once when *stop != 0 (CG after convergence)
  void MultStop(const double* x, double* y, const int* stop, hipStream_t st);
  void Apply(double s, const double* x, double* y, bool accumulate, const int* stop, hipStream_t st);
};

// Sparse Cholesky inverse of the free block (sparsechol.hip): y = A_ff^-1 x
// on the free dofs, 0 elsewhere.  Factor data in dissection order: the
// supernode table, the inverses of the diagonal blocks and the
// off-diagonal blocks (both orientations), per forward level and per
// backward level the rows it solves.
// per supernode: first row, size m, boundary size b, offsets of L11^-1
// (m x m, row-major), of L21 (b x m, row-major), of L21^T (m x b,
// row-major) and of its boundary list
struct ScNode {
  int first, m, b, pad;
  int64_t inv, l21, l21t, bnd;
};

// Host result of the factorisation (the device copy's contents, before upload)
struct ScFactor {
  int nf = 0, nsuper = 0, nlevels = 0;
  int64_t nnz_factor = 0;
  std::vector<ScNode> tab;
  std::vector<double> lval;
  std::vector<int64_t> invt;
  std::vector<int> bnd, seg_ptr, seg_node, seg_row, rows_f, rows_b, fwd_level, bwd_level, sn_of_row, fdof, pos;
};
ScFactor FactorSparseCholesky(int64_t n, const int64_t* rp, const int* cols, const double* vals,
                              const uint8_t* freedofs);

struct SparseCholeskyOp : ngsamd_op_s {
  int nf = 0, nsuper = 0, nlevels = 0;
  int64_t nnz_factor = 0;
  std::vector<int> fwd_level, bwd_level;  // row ranges of each level (host)
  DevBuf<ScNode> d_nodes;
  DevBuf<double> d_lval;
  DevBuf<int64_t> d_invt;
  DevBuf<int> d_bnd, d_seg_ptr, d_seg_node, d_seg_row, d_rows_f, d_rows_b, d_sn_of_row, d_fdof, d_pos;
  DevBuf<double> d_b, d_t, d_y;  // work vectors (preallocated: the apply allocates nothing)
  SparseCholeskyOp(int64_t n, const int64_t* rp, const int* cols, const double* vals, const uint8_t* freedofs);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y = (accumulate ? y : 0) + s A_ff^-1 x; every kernel returns at once when *stop != 0
  void Apply(double s, const double* x, double* y, bool accumulate, const int* stop, hipStream_t st);
};

// Multigrid V-cycle for the P1 block of a structured mesh (mg.hip):
// y = (accumulate ? y : 0) + s M x on the first op->height dofs; every kernel
// returns at once when *stop != 0
void MultiGridApply(ngsamd_op_t op, double s, const double* x, double* y, bool accumulate, const int* stop,
                    hipStream_t st);

// Additive two-level preconditioner y = S x + E C E^T x (sparsechol.hip):
// S Jacobi or block Jacobi, C a sparse Cholesky inverse or a multigrid
// V-cycle of the first C->height dofs.  The operators are not owned.
struct TwoLevelOp : ngsamd_op_s {
  ngsamd_op_t smoother, coarse;
  TwoLevelOp(ngsamd_op_t smoother, ngsamd_op_t coarse);
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  void MultStop(const double* x, double* y, const int* stop, hipStream_t st);
  // y += s E C E^T x
  void CoarseApply(double s, const double* x, double* y, const int* stop, hipStream_t st);
};

}  // namespace ngsamd
