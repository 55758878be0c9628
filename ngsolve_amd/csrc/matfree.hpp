// Fused element-batched matrix-free H1 Laplace operator (device side).
#pragma once
#include <vector>

#include "common.hpp"
#include "csr.hpp"

namespace ngsamd {

struct MfClass {
  int64_t begin = 0, end = 0;  // element range in the (class-sorted) tables
};

struct MatFreeOp : ngsamd_op_s {
  int dim = 3, nd = 0, nq = 0, ncomp = 6;
  int ndp = 0;                // nd padded to a multiple of 4 (MFMA K)
  int nqp = 0;                // nq padded to a multiple of 16
  int precision = NGSAMD_PREC_F64;
  int scatter = NGSAMD_SCATTER_GATHER;
  bool use_mfma = false;
  bool ref_kernel = false;    // reference-matrix MFMA kernels (else per-quadrature-point)
  int64_t nel = 0;
  std::vector<MfClass> classes;
  // MFMA kernels' schedule: rounds of 16-element tiles (one per wavefront)
  DevBuf<int> round0;         // first round of each class, nclass + 1
  DevBuf<int64_t> cls_range;  // [begin, end) of each class
  int nrounds = 0;
  // tiles per round = wavefronts per workgroup of the MFMA kernels: 8, or 4
  // for split FP16 (whose 132 VGPRs allow one 8-wave workgroup per CU)
  int mf_waves = 8;
  mutable int occ[3][2] = {{1, 1}, {1, 1}, {1, 1}};  // resident workgroups per CU [mode][FP64 vectors]
  // per class reference gradients, MFMA layout [nqp][dim][ndp] (+1 pad per q),
  // in the contraction type (double / float / bf16 bits)
  DevBuf<double> bmat64;
  DevBuf<float> bmat32;
  DevBuf<uint16_t> bmat16;
  int64_t bstride = 0;        // elements per class block in bmat*
  int nqc = 0, k1p = 0, ks1 = 0, mt16 = 0;  // BF16 layout: point chunks of 32, padded dofs, k-steps, dof tiles
  // reference matrices S[class][c][i][j] (c over the packed metric
  // components), rows padded to ndp + 1 words (FP64/FP32) or ks1*32 + 8 halves
  DevBuf<double> smat64;
  DevBuf<double> smat_asm;    // FP64 S[class][k][i][j], for the CSR assembly
  DevBuf<float> smat32;
  DevBuf<uint16_t> smat16;
  int64_t sstride = 0;        // elements per class block in smat*
  // factored reference matrices S_(dd') = D_d D_d'^T (the default kernels):
  // rank r of the weighted gradient space, KG = ceil(r / 4) basis functions
  // per lane group, D in the kernels' LDS layouts per class
  bool fac_kernel = false;
  int fac_rank = 0, fac_kg = 0;
  DevBuf<double> fmat64;
  DevBuf<float> fmat32;
  DevBuf<uint16_t> fmat16;
  int64_t fstride = 0;        // elements per class block in fmat*
  DevBuf<double> weights;     // [nqp], 0 beyond nq
  DevBuf<double> bref;        // [nclass][nq][nd][dim] FP64 (scalar path, diagonal, gradient)
  DevBuf<int> dofs;           // [nel][nd]
  DevBuf<double> metric;      // [nel][ncomp]
  DevBuf<double> kappa;       // [ne_global] (owned copy)
  const double* kappa_ext = nullptr;  // set by ngsamd_mf_set_kappa
  DevBuf<int> elnr;           // [nel] global element numbers
  // tile-major copies for the factored kernels: tile T = 16 consecutive
  // elements of one class, dofs [T][nd][16], metric [T][ncomp][16], element
  // numbers [T][16] (lanes past the class end: the tile's first element);
  // tile_base[c] = first tile of class c; the 16-bit kernels' contribution
  // rows are stored [T][nd][16] as well
  DevBuf<int> t_dofs, t_elnr, tile_base;
  DevBuf<double> t_metric;
  int64_t ne_global = 0;
  // gather scatter: dof -> contribution slots (element row * nd + local),
  // dofs in the gather order g_order (by lowest element number)
  DevBuf<int> g_rowptr, g_slots, g_order;
  DevBuf<int> g_tslots;  // g_slots renumbered for the tile-major rows of the factored 16-bit kernels
  DevBuf<double> contrib64;
  DevBuf<float> contrib32;
  // colour-ordered direct scatter (NGSAMD_SCATTER_COLOR, set by
  // SetColoring): copies of the element tables sorted by colour (class order
  // kept inside a colour) and per colour its round schedule; one launch per
  // colour adds the element results straight into y
  int ncolors = 0;
  std::vector<int> c_nrounds;  // per colour
  DevBuf<int> c_round0;        // [ncolors][nclass + 1]
  DevBuf<int64_t> c_range;     // [ncolors][2 nclass]
  DevBuf<int> c_dofs, c_elnr;
  DevBuf<double> c_metric;
  // colour of every table row; refuses a colouring whose elements of one
  // colour share a dof
  void SetColoring(const int* colors);

  MatFreeOp(int dim, int nd, int nq, int64_t ndof, int nclass, const int64_t* class_offsets, const double* bmat,
            const double* w, const int* dofs, const double* metric, const double* kappa, const int* elnr,
            int precision, int scatter);
  const double* Kappa() const { return kappa_ext ? kappa_ext : kappa.p; }
  void MultAdd(double s, const double* x, double* y, hipStream_t st) override;
  void Mult(const double* x, double* y, hipStream_t st) override;
  // y = A x, all kernels return immediately when *d_stop != 0
  void MultStop(const double* x, double* y, const int* d_stop, hipStream_t st);
  void MultF32(const float* x, float* y, const int* d_stop, hipStream_t st);
  // y = A x plus per-workgroup partial sums of <x, A x> in partials[] (room
  // for kMaxReduceBlocks); returns their number, 0 when the scatter mode has
  // no gather kernel (then the caller forms the inner product itself).
  // wide: room for kMaxWideBlocks partials of the gather's wide grid (null:
  // the library's)
  int MultDot(const double* x, double* y, double* partials, const int* d_stop, hipStream_t st,
              double* wide = nullptr);
  int MultDotF32(const float* x, float* y, double* partials, const int* d_stop, hipStream_t st,
                 double* wide = nullptr);
  void Diagonal(double* d, hipStream_t st);
  void ElementGrad(const double* lam, const double* u, double* grad, hipStream_t st);
  // fills the values of A (same size, pattern holding every element pair)
  void AssembleCsr(CsrOp& A, hipStream_t st);
  void AssembleMassCsr(CsrOp& A, int nq_m, const double* h_w, const double* h_shape, const double* h_scale,
                       hipStream_t st);
  void AssembleCsrWith(CsrOp& A, const double* smat, const double* met, const double* kap, int nc,
                       hipStream_t st);
  // linear form of element quadrature values into d_out (overwritten)
  void LinearForm(int nq_lf, const double* h_w, const double* h_shape, const double* h_detj, const double* d_fvals,
                  double* d_out, hipStream_t st);
  // sparsity pattern of the form: h_cols null fills h_rowptr (height + 1);
  // then h_cols (h_rowptr[height] entries) receives the columns
  void Pattern(int64_t* h_rowptr, int* h_cols, hipStream_t st);
  int max_row_slots = 0;      // most element contributions of one dof
  DevBuf<int64_t> pat_rowptr; // row pointers of the last Pattern call

 private:
  template <class V>
  int Apply(double s, const V* x, V* y, bool accumulate, const int* d_stop, hipStream_t st,
            double* partials = nullptr, double* wide = nullptr);
};

}  // namespace ngsamd
