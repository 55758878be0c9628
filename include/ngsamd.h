/* ngsamd.h -- C ABI of the MI355X (gfx950) H1 operator / Krylov library.
 *
 * Every entry point returns 0 on success and a nonzero NGSAMD_E_* code on
 * failure; ngsamd_last_error() then returns a message (thread-local).
 * Pointers named d_* are device pointers (allocated with ngsamd_malloc or by
 * any HIP allocator on the current device); pointers named h_* are host.
 * All work is enqueued on the library stream (ngsamd_set_stream) in order;
 * calls that return host values synchronise that stream.
 *
 * Reference interfaces replaced (file:line in NGSolve):
 *   operators   ngla::BaseMatrix::Mult / MultAdd      linalg/basematrix.hpp:117-121
 *   CSR SpMV    SparseMatrix<double>::MultAdd          linalg/sparsematrix_impl.hpp:264-306
 *               DevSparseMatrix::MultAdd (cuSPARSE)    ngscuda/cuda_linalg.cpp:243-275
 *   matrix-free MatrixFreeBTDTB::MultAdd               comp/matrixfree.cpp:681-1087
 *               DevMatrixFreeBTDTB                     ngscuda/cuda_matrix_free_btdtb.cpp:16-427
 *   Jacobi      JacobiPrecond::MultAdd                 linalg/jacobi.cpp:41-90
 *   vectors     S_BaseVector<double>::InnerProduct     linalg/basevector.cpp:1127-1160
 *               UnifiedVector Add/Scale/SetScalar      ngscuda/unifiedvector.cpp:136-212,437-515
 *   CG          CGSolver._SolveImpl                    python/krylovspace.py:263-291
 *               DevCGSolver (graph-captured loop)      ngscuda/cuda_krylov.cpp:19-203
 *   block Jac.  BlockJacobiPrecond                     linalg/blockjacobi.cpp:390-632
 *   assembly    S_BilinearForm::DoAssemble             comp/bilinearform.cpp:1563-2215
 *               BilinearForm::GetGraph / MatrixGraph   comp/bilinearform.cpp:552-825
 *               SymbolicLinearFormIntegrator           fem/symbolicintegrator.cpp:1036-1110
 *   setup       H1HighOrderFESpace::Update, GetDofNrs  comp/h1hofespace.cpp:823-950,1389-1413
 */
#ifndef NGSAMD_H
#define NGSAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NGSAMD_OK 0
#define NGSAMD_E_HIP 1        /* HIP runtime error */
#define NGSAMD_E_ARG 2        /* invalid argument / shape mismatch */
#define NGSAMD_E_STATE 3      /* library not initialised, wrong handle kind */
#define NGSAMD_E_NOMEM 4      /* device allocation failed */

/* contraction precision of a matrix-free operator */
#define NGSAMD_PREC_F64 0
#define NGSAMD_PREC_F32 1
#define NGSAMD_PREC_BF16 2
#define NGSAMD_PREC_F16 3
/* split FP16: each operand as the sum of two FP16 values, three FP16 MFMA
 * products per contraction (about FP32 accuracy; matrix-free factored
 * kernels only, i.e. H1 triangles p <= 5 and tetrahedra p <= 4).  No
 * counterpart in the reference. */
#define NGSAMD_PREC_F16X3 4

/* scatter mode of a matrix-free operator */
#define NGSAMD_SCATTER_GATHER 0   /* deterministic two-pass owner-computes gather */
#define NGSAMD_SCATTER_ATOMIC 1   /* hardware FP64/FP32 atomics, as the reference kernel */
#define NGSAMD_SCATTER_COLOR 2    /* colour-ordered direct scatter (ngsamd_mf_set_coloring) */

typedef struct ngsamd_op_s* ngsamd_op_t;

/* ---------------------------------------------------------------- runtime */
int ngsamd_init(int device);
const char* ngsamd_last_error(void);
int ngsamd_set_stream(void* hip_stream);   /* NULL selects the library's own stream */
void* ngsamd_get_stream(void);
int ngsamd_synchronize(void);
int ngsamd_malloc(void** d_ptr, size_t bytes);
int ngsamd_free(void* d_ptr);
int ngsamd_memcpy_htod(void* d_dst, const void* h_src, size_t bytes);
int ngsamd_memcpy_dtoh(void* h_dst, const void* d_src, size_t bytes);
int ngsamd_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);

/* Timing utilities (no reference counterpart): HIP events recorded on the
 * library stream, so they bracket exactly the kernels this library enqueues. */
int ngsamd_event_create(void** ev);
int ngsamd_event_record(void* ev);
int ngsamd_event_elapsed_ms(void* start, void* stop, float* ms); /* synchronises `stop` */
int ngsamd_event_destroy(void* ev);
/* Read-bandwidth calibration kernel (no reference counterpart): reads the n
 * doubles of d_x (n even, 16-byte aligned) once with non-temporal 16-byte
 * loads and writes their sum to *d_result (device), asynchronously; timed
 * with the events above it gives the read ceiling that an HBM-bound,
 * read-dominated kernel such as the CSR SpMV is compared with. */
int ngsamd_stream_read(int64_t n, const double* d_x, double* d_result);

/* ------------------------------------------------- vectors (BaseVector ops) */
/* *h_result = sum_i x_i y_i, FP64 accumulation, fixed (deterministic) order */
int ngsamd_vec_dot(int64_t n, const double* d_x, const double* d_y, double* h_result);
int ngsamd_vec_dot_f32(int64_t n, const float* d_x, const float* d_y, double* h_result);
int ngsamd_vec_set(int64_t n, double a, double* d_x);                       /* x = a */
int ngsamd_vec_scale(int64_t n, double a, double* d_x);                     /* x *= a */
int ngsamd_vec_axpy(int64_t n, double a, const double* d_x, double* d_y);   /* y += a x */
int ngsamd_vec_axpby(int64_t n, double a, const double* d_x, double b, double* d_y); /* y = a x + b y */
int ngsamd_vec_f64_to_f32(int64_t n, const double* d_x, float* d_y);
int ngsamd_vec_f32_to_f64(int64_t n, const float* d_x, double* d_y);

/* ---------------------------------------------------- generic operators */
int ngsamd_op_height(ngsamd_op_t op, int64_t* h);
int ngsamd_op_width(ngsamd_op_t op, int64_t* w);
/* y += s * A x   (BaseMatrix::MultAdd, linalg/basematrix.hpp:121) */
int ngsamd_op_mult_add(ngsamd_op_t op, double s, const double* d_x, double* d_y);
/* y = A x        (BaseMatrix::Mult, linalg/basematrix.hpp:117) */
int ngsamd_op_mult(ngsamd_op_t op, const double* d_x, double* d_y);
int ngsamd_op_destroy(ngsamd_op_t op);

/* ------------------------------------------------------------- CSR matrix */
/* Copies a host CSR (int64 row pointers as NGSolve's size_t firsti, int32
 * column numbers sorted per row, FP64 values) to the device.
 * SparseMatrix<double> storage, linalg/sparsematrix.hpp. */
int ngsamd_csr_create(ngsamd_op_t* op, int64_t height, int64_t width,
                      const int64_t* h_rowptr, const int* h_cols, const double* h_vals);
/* diagonal of a CSR operator (0 where the row has no diagonal entry) */
int ngsamd_csr_diagonal(ngsamd_op_t op, double* d_diag);
/* Copies the values to the host in CSR order (row by row, columns as at
 * creation): the value vector of the matrix, BaseSparseMatrix::AsVector
 * (linalg/sparsematrix.hpp:312). */
int ngsamd_csr_values(ngsamd_op_t op, double* h_vals);

/* ---------------------------------------------------------------- Jacobi */
/* y += s * diag^{-1} x on free dofs, 0 on the others (linalg/jacobi.cpp:41-90).
 * h_freedofs may be NULL (all free). */
int ngsamd_jacobi_create(ngsamd_op_t* op, int64_t n, const double* h_diag, const uint8_t* h_freedofs);
int ngsamd_jacobi_create_device(ngsamd_op_t* op, int64_t n, const double* d_diag, const uint8_t* h_freedofs);
/* Block-Jacobi preconditioner y += s sum_b P_b^T inv(A_bb) P_b x over the
 * blocks b (h_block_ptr[nblocks + 1] offsets into h_block_dofs; each block's
 * dofs are sorted, blocks may overlap).  A_bb is taken from the host CSR
 * (int64 row pointers, sorted int32 columns, FP64 values; 0 where the
 * pattern has no entry) and inverted on the host; the apply runs on the
 * device without atomics.  BlockJacobiPrecond (linalg/blockjacobi.cpp:390,
 * 404, 498, 596-632) and its device copy DevBlockJacobiMatrix
 * (ngscuda/dev_blockjacobi.cpp:92-131).  Accepted by ngsamd_cg_solve. */
int ngsamd_blockjacobi_create(ngsamd_op_t* op, int64_t n, const int64_t* h_rowptr, const int* h_cols,
                              const double* h_vals, int64_t nblocks, const int64_t* h_block_ptr,
                              const int* h_block_dofs);

/* Sparse Cholesky inverse of the free block of a symmetric positive definite
 * matrix: y = A_ff^{-1} x on the free dofs (h_freedofs[i] != 0; NULL: all),
 * 0 on the others; Mult overwrites, MultAdd accumulates.  SparseCholesky
 * (linalg/sparsecholesky.cpp) behind a.mat.Inverse(freedofs,
 * inverse="sparsecholesky"), device solve as DevSparseCholesky
 * (ngscuda/dev_sparsecholesky.cpp:403-454).  Host: nested-dissection
 * ordering and supernodal (multifrontal) factorisation in FP64; device:
 * level-scheduled triangular solves (a few kernels per level of the
 * dissection tree, no waiting between workgroups, deterministic).  Accepted
 * by ngsamd_cg_solve as a preconditioner. */
int ngsamd_sparsecholesky_create(ngsamd_op_t* op, int64_t n, const int64_t* h_rowptr, const int* h_cols,
                                 const double* h_vals, const uint8_t* h_freedofs);
/* h_info[0] = free dofs, [1] = supernodes, [2] = levels, [3] = nonzeros of L */
int ngsamd_sparsecholesky_info(ngsamd_op_t op, int64_t* h_info);

/* Multigrid V-cycle for the P1 matrix of a structured mesh (the vertex
 * block of an H1 problem on MakeStructured2DMesh(dims[0], dims[1]) or
 * MakeStructured3DMesh(dims[0], dims[1], dims[2]); n = number of vertices,
 * vertices numbered as those builders number them): y ~= A_ff^-1 x on the
 * free vertices, 0 elsewhere.  Replaces MultigridPreconditioner::Mult / MGM
 * (multigrid/mgpre.cpp:205-318) with LinearProlongation
 * (multigrid/prolongation.cpp:77-144) over the hierarchy n, n/2, n/4, ...;
 * coarse matrices are the Galerkin products P^T A P, smoothing is
 * smoothing_steps damped-Jacobi sweeps (weight omega in (0, 2)) before and
 * after each coarse correction, the coarsest level (at most 1024 free
 * vertices, reached within max_levels levels) is solved with its dense
 * inverse.  Accepted by ngsamd_cg_solve and as the coarse operator of
 * ngsamd_twolevel_create. */
int ngsamd_multigrid_create(ngsamd_op_t* op, int dim, const int* dims, int64_t n, const int64_t* h_rowptr,
                            const int* h_cols, const double* h_vals, const uint8_t* h_freedofs, int smoothing_steps,
                            double omega, int max_levels);
/* *h_nlevels = number of levels; h_level_dofs (may be null, else room for
 * *h_nlevels entries, at most 32) = free dofs per level, finest first */
int ngsamd_multigrid_info(ngsamd_op_t op, int* h_nlevels, int64_t* h_level_dofs);

/* Additive two-level preconditioner y = S x + E C E^T x: S a Jacobi or
 * block-Jacobi operator of the whole space, C a sparse Cholesky inverse or a
 * multigrid operator acting on the first C->height dofs (E their embedding;
 * for a hierarchical H1 space the vertex dofs, i.e. the P1 space).  The
 * composite smoother + E @ a_lo.mat.Inverse(freedofs_lo) @ E.T of the
 * reference's sum and product matrices (linalg/basematrix.hpp:413,592), or
 * with the multigrid operator in place of the inverse.  The operators are
 * not owned: destroy the two-level operator first.  Accepted by
 * ngsamd_cg_solve. */
int ngsamd_twolevel_create(ngsamd_op_t* op, ngsamd_op_t smoother, ngsamd_op_t coarse);

/* ------------------------------------------------- matrix-free H1 Laplace */
/* Fused element-batched operator y += s * sum_e P_e^T B^T D_e B P_e x with
 * D_e = kappa_e w_q |det F_e| F_e^{-1} F_e^{-T} (comp/matrixfree.cpp:681-1087).
 *   dim        2 or 3
 *   nd         local dofs per element
 *   nq         quadrature points
 *   nclass     number of vertex-permutation classes
 *   h_class_offsets[nclass+1]  element ranges of the classes in the tables below
 *   h_bmat     [nclass][nq][nd][dim] reference gradients d phi_i / d xi at ip q
 *   h_weights  [nq]
 *   h_dofs     [nel][nd] global dof numbers (int32)
 *   h_metric   [nel][dim*(dim+1)/2] |det F| F^{-1}F^{-T}: (xx,yy,xy) in 2D,
 *              (xx,yy,zz,xy,xz,yz) in 3D
 *   h_kappa    [nel] element coefficient or NULL (=1)
 *   h_elnr     [nel] global element number of each table row or NULL (=row)
 *   precision  NGSAMD_PREC_*  (B and the contraction in that type, FP32
 *              accumulation for BF16 / FP16 / split FP16; input/output
 *              vectors stay FP64)
 *   scatter    NGSAMD_SCATTER_* */
int ngsamd_mf_create(ngsamd_op_t* op, int dim, int nd, int nq, int64_t ndof, int nclass,
                     const int64_t* h_class_offsets, const double* h_bmat, const double* h_weights,
                     const int* h_dofs, const double* h_metric, const double* h_kappa,
                     const int* h_elnr, int precision, int scatter);
/* number of global elements of a matrix-free operator: the largest element
 * number in h_elnr plus one (the row count of the tables if h_elnr was NULL);
 * element coefficients and element gradients are indexed by it */
int ngsamd_mf_num_elements(ngsamd_op_t op, int64_t* ne);
/* Colour-ordered atomic-free scatter, as the reference's coloured element
 * loop (IterateElements over FESpace::ElementColoring, comp/fespace.cpp:518-760,
 * 2137-2200; comp/bilinearform.cpp:4669-4750): h_colors[nel] gives the colour
 * of every table row (elements of one colour may share no dof: checked,
 * NGSAMD_E_ARG otherwise).  From then on the apply runs one launch per colour
 * that adds each element's result straight into y (no contribution buffer,
 * no gather).  Needs the factored kernels (H1 p <= 5 in 2D, p <= 4 in 3D). */
int ngsamd_mf_set_coloring(ngsamd_op_t op, const int* h_colors);
/* replace the element coefficient: d_kappa is a device array of n values
 * indexed by global element number, n must equal ngsamd_mf_num_elements; the
 * operator keeps the pointer (not a copy) until the next call, so the array
 * must stay allocated while the operator is applied */
int ngsamd_mf_set_kappa(ngsamd_op_t op, const double* d_kappa, int64_t n);
/* y = A x on FP32 vectors (inner solves of iterative refinement) */
int ngsamd_mf_mult_f32(ngsamd_op_t op, const float* d_x, float* d_y);
/* diagonal of the operator, FP64 */
int ngsamd_mf_diagonal(ngsamd_op_t op, double* d_diag);
/* g_e = -int_e grad(lam) . grad(u) = -(lam_e^T K_e u_e)/kappa_e, indexed by
 * global element number (adjoint gradient of an element coefficient); d_grad
 * holds ngsamd_mf_num_elements values. */
int ngsamd_mf_element_grad(ngsamd_op_t op, const double* d_lam, const double* d_u, double* d_grad);
/* Assembles the stiffness matrix of a matrix-free operator's form into a new
 * CSR operator on the device: A[r][c] = sum over the elements e holding dofs
 * r = dofs[e][i] and c = dofs[e][j] of K_e[i][j], with K_e = kappa_e times
 * the packed |det F| F^-1 F^-T contracted with the reference matrices, into
 * the sparsity pattern (h_rowptr, h_cols) (int64 row pointers, sorted int32
 * columns), which must hold every such (r, c).  Each value is summed in a
 * fixed order (deterministic).  Replaces BilinearForm::Assemble
 * (comp/bilinearform.cpp:832) -> S_BilinearForm::DoAssemble (:1563-2215) ->
 * SparseMatrixTM::AddElementMatrix (linalg/sparsematrix_impl.hpp:150-189)
 * for the grad-grad form. */
int ngsamd_mf_assemble_csr(ngsamd_op_t mf, const int64_t* h_rowptr, const int* h_cols, ngsamd_op_t* csr);
/* Assembles the mass form M[r][c] = sum_e h_scale[e] sum_q w_q phi_i(x_q)
 * phi_j(x_q) (r = dofs[e][i], c = dofs[e][j]) on the mesh and dof tables of
 * a matrix-free operator into a new CSR operator with the pattern (h_rowptr,
 * h_cols); h_scale per global element (rho_e |det F_e|, ne as at creation),
 * nq / h_weights the form's rule, h_shape phi per class [nclass][nq][nd] in
 * the creation's class order.  Each value is summed in a fixed order.
 * S_BilinearForm::DoAssemble (comp/bilinearform.cpp:1563-2215) for rho*u*v*dx
 * (SymbolicBilinearFormIntegrator, fem/symbolicintegrator.cpp:1549-1860). */
int ngsamd_mf_assemble_mass_csr(ngsamd_op_t mf, int nclass, int nq, int nd, int64_t ne, const double* h_weights,
                                const double* h_shape, const double* h_scale, const int64_t* h_rowptr,
                                const int* h_cols, ngsamd_op_t* csr);
/* Sparsity pattern of a matrix-free operator's form, built on the device:
 * row r is the sorted union of the dofs of the elements holding dof r, as
 * BilinearForm::GetGraph (comp/bilinearform.cpp:552-825) -> MatrixGraph
 * (linalg/sparsematrix.cpp:450-975) builds it.  Two calls: with h_cols NULL
 * it fills h_rowptr (height + 1 entries); then with h_cols of h_rowptr[height]
 * entries it fills the columns. */
int ngsamd_mf_pattern(ngsamd_op_t mf, int64_t* h_rowptr, int* h_cols);
/* Linear form f_i = sum_e |det F_e| sum_q w_q f(e, q) phi_i(x_q) on the
 * device, into d_out (height entries, overwritten), each dof's element
 * contributions summed in a fixed order through the operator's gather.
 * nq, h_weights: the form's own rule; h_shape: phi per class [nclass][nq][nd]
 * in the class order given at creation (nclass and nd as at creation);
 * h_detj: det F per global element (ne of them, as at creation); d_fvals: f
 * at the rule's points, [global element][nq], on the device.
 * SymbolicLinearFormIntegrator::T_CalcElementVector
 * (fem/symbolicintegrator.cpp:1036-1110) with LinearForm::Assemble's
 * scatter. */
int ngsamd_mf_linear_form(ngsamd_op_t mf, int nclass, int nq, int nd, int64_t ne, const double* h_weights,
                          const double* h_shape, const double* h_detj, const double* d_fvals, double* d_out);

/* ---------------------------------------------- mesh topology, H1 dof tables */
typedef struct ngsamd_h1_s* ngsamd_h1_t;
/* Builds on the device the edge and face numbering of a simplex mesh
 * (h_els: ne x (dim+1) vertex numbers) and the H1 space of the given order
 * on it: global edges / faces in ascending lexicographic order of their
 * sorted vertex tuples, per-element edge / face numbers, the edges (and in
 * 3D the face) of each boundary facet (h_bnd: nbnd x dim vertex numbers,
 * h_bnd_index: 1-based region; in 3D with nbnd = 0 the faces held by one
 * element become the boundary, region 1, in face order), the dof layout
 * (vertex dofs, p-1 per edge, (p-1)(p-2)/2 per face, cell dofs), the
 * element dof tables, the free-dof mask (dofs on facets of the regions
 * whose bit r-1 is set in dirichlet_mask are not free) and each element's
 * vertex-order class.  Replaces netgen's topology tables as NGSolve's
 * MeshAccess exposes them (comp/meshaccess.cpp), H1HighOrderFESpace::Update
 * (comp/h1hofespace.cpp:823-950), GetDofNrs (:1389-1413) and the free-dof
 * mask of FESpace::Update (comp/fespace.cpp:765-810). */
int ngsamd_h1_create(ngsamd_h1_t* h, int dim, int64_t nv, int64_t ne, const int* h_els, int64_t nbnd,
                     const int* h_bnd, const int* h_bnd_index, int order, uint64_t dirichlet_mask);
/* h_sizes[6] = {nedges, nfaces, nbnd, ndof, dofs per element, ne} */
int ngsamd_h1_sizes(ngsamd_h1_t h, int64_t* h_sizes);
/* copies the topology tables to the host (any pointer may be NULL):
 * edges [nedges][2], faces [nfaces][3], element edges [ne][3 or 6],
 * element faces [ne][4], boundary facets [nbnd][dim], their regions,
 * their edges [nbnd][1 or 3] and faces [nbnd] */
int ngsamd_h1_topology(ngsamd_h1_t h, int* h_edges, int* h_faces, int* h_el_edges, int* h_el_faces, int* h_bnd,
                       int* h_bnd_index, int* h_bnd_edges, int* h_bnd_faces);
/* copies the element dof tables [ne][nd], the free-dof mask [ndof] and the
 * element classes [ne] to the host (any may be NULL); *d_el_dofs (if not
 * NULL) receives the device copy of the dof tables, valid until destroy */
int ngsamd_h1_dofs(ngsamd_h1_t h, int* h_el_dofs, unsigned char* h_freedofs, int* h_classnr, const int** d_el_dofs);
int ngsamd_h1_destroy(ngsamd_h1_t h);

/* ----------------------------------------------------------------- CG */
/* Preconditioned CG with the recurrence and stopping rule of the reference's
 * CGSolver (python/krylovspace.py:227-291): stop when
 * sqrt|<w,d>| <= max(tol*sqrt|<w0,d0>|, atol) or after maxiter checks; the
 * iteration count includes the initial check.  pre must be a Jacobi
 * operator (applied inside the update kernel) or a block-Jacobi operator
 * (applied between the update and the direction kernels).  Scalars, dots and
 * convergence tests stay on the device; the iteration is replayed from a
 * captured HIP graph in chunks of `chunk` iterations (0 = no graph).
 * h_residuals (may be NULL) receives up to maxiter residuals. */
int ngsamd_cg_solve(ngsamd_op_t mat, ngsamd_op_t pre, const double* d_rhs, double* d_sol,
                    int initialize, double tol, double atol, int maxiter, int chunk,
                    int* h_iterations, double* h_residuals);

/* ngsamd_cg_solve without any allocation, for callers that cannot allocate
 * device or pinned memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Its
 * work vectors, device scalars and residual history live in a file-scope
 * device array of the library (each GPU holds its own copy, about 97 MiB,
 * until the process ends): at most 4,194,304 dofs and maxiter <= 65536, or
 * the call fails with NGSAMD_E_ARG.  Kernels are launched directly (no
 * graph); the host reads the device scalars every 16 iterations with a
 * plain copy.  Same kernels in the same order as ngsamd_cg_solve, so the
 * same iterates.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (several host threads or streams) race. */
int ngsamd_cg_solve_noalloc(ngsamd_op_t mat, ngsamd_op_t pre, const double* d_rhs, double* d_sol,
                            int initialize, double tol, double atol, int maxiter,
                            int* h_iterations, double* h_residuals);

/* Mixed-precision iterative refinement: outer FP64 residual r = f - A u with
 * `mat`, inner Jacobi-PCG on `inner` (FP32 vectors, FP64 dots and scalars)
 * to relative tolerance inner_tol, u += correction, until
 * ||r||_2 <= tol ||f||_2 (norms over the free dofs).  inner_tol <= 0 picks a
 * default by the inner operator's precision (FP32 3e-4, FP16 1e-3, BF16
 * 3e-2).  h_stats[0]=outer its, [1]=total inner its, [2]=final relative
 * residual.  No counterpart in the reference (it has no mixed-precision
 * solver); it sits beside CGSolver (python/krylovspace.py:227-291). */
int ngsamd_ir_cg_solve(ngsamd_op_t mat, ngsamd_op_t inner, ngsamd_op_t pre, const double* d_rhs,
                       double* d_sol, double tol, double inner_tol, int max_outer, int max_inner,
                       double* h_stats);

/* Lockstep CG for nrhs (1..8) right-hand sides of one CSR matrix, vectors
 * with nrhs components per dof (entry j of dof i at d_x[i * nrhs + j]), Jacobi
 * preconditioner or none (pre NULL): CGSolver<double>::MultiMult
 * (linalg/cg.cpp:132-250) -- per component alpha and beta, one stopping test
 * on the norm of the vector of w.d values (relative to its start, or tol^2
 * absolute), the matrix read once per iteration for all components.
 * *h_steps = the reference's `steps`; h_residuals (maxiter + 1 entries, may
 * be NULL) = sqrt(||wdn||) at the start and after each iteration.  A component
 * whose w.d or <s, A s> is 0 gets alpha = beta = 0 (the reference: 0/0). */
int ngsamd_cg_solve_multi(ngsamd_op_t mat, ngsamd_op_t pre, int nrhs, const double* d_rhs, double* d_sol,
                          int initialize, double tol, int absolute, int maxiter, int* h_steps, double* h_residuals);
/* Y = A X for nrhs vectors stored as above (CSR operators); the matrix is
 * read once for all of them.  Each component equals a single apply. */
int ngsamd_csr_mult_multi(ngsamd_op_t op, int nrhs, const double* d_x, double* d_y);

/* ------------------------------------------------ batched independent solves */
/* Many independent problems as one block-diagonal CSR system: problem p owns
 * rows/columns [h_seg[p], h_seg[p+1]).  with_f32 keeps an FP32 copy of the
 * values for the inner solves of ngsamd_batch_ir_cg_solve.  The nonzeros per
 * SpMV row block are autotuned when the operator is created.
 * Reference: CGSolver per problem (python/krylovspace.py:227-291) and the
 * lockstep multi-RHS CGSolver<double>::MultiMult (linalg/cg.cpp:132-250). */
int ngsamd_batch_csr_create(ngsamd_op_t* op, int64_t n, const int64_t* h_rowptr, const int* h_cols,
                            const double* h_vals, int nseg, const int64_t* h_seg, int with_f32);
int ngsamd_batch_info(ngsamd_op_t op, int* nseg, int* block_nnz, int* nblocks);
/* The block-size autotune of the batched operator: *h_count candidates
 * (0 when NGSAMD_SPMV_NNZ forced one), h_ms[k] (room for 3) the measured
 * milliseconds per apply of 1024, 2048 and 4096 values (nonzeros) per block;
 * ngsamd_batch_info reports the one kept. */
int ngsamd_batch_tuning(ngsamd_op_t op, int* h_count, double* h_ms);
/* Jacobi-PCG on every problem with the CGSolver recurrence and stopping rule
 * per problem (tol relative to each problem's initial residual).  Per problem:
 * h_iterations (checks incl. the initial one), h_final_res (last residual),
 * h_finish_ms (device time from the start of the call to its convergence).
 * chunk > 0: the iteration replayed from a HIP graph of chunk iterations;
 * 0: launched kernel by kernel; NGSAMD_BATCH_PERSISTENT (-1): all
 * iterations in one cooperative launch that stays resident (grid-wide
 * barriers between the phases of an iteration) -- the same iterates bit for
 * bit.  No counterpart in the reference's interface (its CG is one solve per
 * call); the persistent form is its task-queue idea (ngscuda/cuda_tasks.cpp:17-140)
 * applied to the lockstep MultiMult (linalg/cg.cpp:132-250). */
#define NGSAMD_BATCH_PERSISTENT (-1)
int ngsamd_batch_cg_solve(ngsamd_op_t op, ngsamd_op_t pre, const double* d_rhs, double* d_sol, double tol,
                          int maxiter, int chunk, int* h_iterations, double* h_final_res, double* h_finish_ms);
/* FP64 iterative refinement per problem around FP32 batched inner solves
 * (FP32 matrix values and vectors, FP64 inner products and scalars) until
 * ||f_p - A_p u_p||_free <= tol ||f_p||_free for every problem. */
int ngsamd_batch_ir_cg_solve(ngsamd_op_t op, ngsamd_op_t pre, const double* d_rhs, double* d_sol, double tol,
                             double inner_tol, int max_outer, int max_inner, int* h_outer, int* h_inner,
                             double* h_relres, double* h_finish_ms);

#ifdef __cplusplus
}
#endif
#endif /* NGSAMD_H */
