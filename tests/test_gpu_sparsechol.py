"""Sparse Cholesky inverse on the device (ngsamd_sparsecholesky_create) and
the additive two-level preconditioner (ngsamd_twolevel_create) against the
oracle's restatement (oracle/fe_oracle.py sparse_inverse, two_level; the
inverse of SparseCholesky, linalg/sparsecholesky.cpp, and the composite
smoother + E C E^T of linalg/basematrix.hpp:413,592): the apply within 1e-12
of the oracle's in the inf-norm for 2D and 3D H1 matrices (whole free block
and the vertex block) with lognormal kappa; Mult / MultAdd semantics, zero on
non-free dofs; device CG with the two-level preconditioner (Jacobi and
block-Jacobi smoothers, P1 coarse space) against the oracle's PCG with the
same preconditioner (iterations within 1, solution within 1e-10), with the
exact inverse as preconditioner (one step), through the Python API, and at
the 32^3 p=4 size with kappa = exp(N(0,1)) (seed 2)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


def _csr(P):
    return sp.csr_matrix((P["vals"], P["cols"], P["rowptr"]), shape=(P["ndof"], P["ndof"]))


def _vertex_block(P):
    nv = P["mesh"].nv
    A = _csr(P)[:nv][:, :nv].tocsr()
    A.sort_indices()
    return A, P["fes"].freedofs.astype(np.uint8)[:nv].copy()


def _create(A, free):
    return N.CreateSparseCholesky(A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                                  A.data.astype(np.float64), free)


@pytest.mark.parametrize("dim,n,p,block", [(2, 8, 1, "all"), (2, 12, 3, "all"), (3, 5, 2, "all"),
                                           (3, 4, 4, "all"), (3, 16, 4, "vertex"), (2, 64, 3, "vertex")])
def test_sparse_cholesky_apply_matches_oracle(device, dim, n, p, block):
    say(f"[sparsechol] dim {dim} n {n} p {p} block {block}")
    P = problem(dim, n, p, seed_kappa=2)
    if block == "all":
        A, free = _csr(P), P["fes"].freedofs.astype(np.uint8)
        A.sort_indices()
    else:
        A, free = _vertex_block(P)
    C = _create(A, free)
    info = N.SparseCholeskyInfo(C)
    assert info["nfree"] == int(free.sum())
    ref_apply = O.sparse_inverse(A.indptr, A.indices, A.data, free)
    x = np.random.default_rng(1).uniform(-1, 1, A.shape[0])
    ref = ref_apply(x)
    dx, dy = dvec(x), N.DeviceVector(A.shape[0], 8)
    dy.fill(3.0)
    C.mult(dx, dy)
    y = dy.download()
    err = relerr_inf(y, ref)
    say(f"[sparsechol] {info}: apply vs oracle {err:.2e}")
    assert err < 1e-12
    assert np.all(y[free == 0] == 0.0)
    # MultAdd accumulates; a second apply gives the same bits
    y0 = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    dy.upload(y0)
    C.mult_add(-0.5, dx, dy)
    assert relerr_inf(dy.download(), y0 - 0.5 * ref) < 1e-12
    dz = N.DeviceVector(A.shape[0], 8)
    C.mult(dx, dz)
    assert np.array_equal(dz.download(), y)


def _two_level_case(n, p, smoother, seed=2):
    P = problem(3, n, p, order=2 * p, seed_kappa=seed)
    free = P["fes"].freedofs.astype(np.uint8)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    Av, fv = _vertex_block(P)
    Cc = _create(Av, fv)
    coarse = O.sparse_inverse(Av.indptr, Av.indices, Av.data, fv)
    if smoother == "jacobi":
        S = N.CreateJacobi(diag, free)
        sm = lambda t: inv * t  # noqa: E731
    else:
        sp_, nv = P["fes"], P["mesh"].nv
        firsts = [sp_.first_edge_dof, sp_.first_face_dof, sp_.first_el_dof]
        starts = np.concatenate([np.arange(nv + 1)] + [np.asarray(f)[1:] for f in firsts])
        fb = free.astype(bool)
        blocks = [b[fb[b]] for b in np.split(np.arange(P["ndof"]), starts[1:-1])]
        blocks = [b for b in blocks if len(b)]
        bptr = np.zeros(len(blocks) + 1, dtype=np.int64)
        bptr[1:] = np.cumsum([len(b) for b in blocks])
        S = N.CreateBlockJacobi(P["ndof"], P["rowptr"], P["cols"], P["vals"], bptr,
                                np.concatenate(blocks).astype(np.int32))
        _, sm = O.block_jacobi(P["rowptr"], P["cols"], P["vals"], blocks)
    T = N.CreateTwoLevel(S, Cc)
    return P, A, T, O.two_level(sm, coarse, Av.shape[0]), free


@pytest.mark.parametrize("smoother", ["jacobi", "blockjacobi"])
def test_two_level_cg_matches_oracle(device, smoother):
    say(f"[twolevel] 6^3 p=3 {smoother}")
    P, A, T, pre, free = _two_level_case(6, 3, smoother)
    x = np.random.default_rng(4).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    T.mult(dx, dy)
    assert relerr_inf(dy.download(), pre(x)) < 1e-12
    f = np.where(free != 0, 1.0, 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, T, df, du, True, 1e-10, -1.0, 5000, 16)
    Am = _csr(P)
    u_ref, its_ref, _ = O.pcg(lambda t: Am @ t, pre, f, tol=1e-10, maxiter=5000)
    _, its_j, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"],
                                 O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)[1], f, tol=1e-10,
                                 maxiter=5000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[twolevel] {smoother}: its {its} oracle {its_ref} (Jacobi alone {its_j}), solution {err:.2e}")
    assert abs(its - its_ref) <= 1
    assert err <= 1e-10


def test_exact_inverse_as_preconditioner(device):
    """The sparse inverse of the whole free block as the preconditioner: the
    CG stops at its second check."""
    P = problem(3, 4, 3, seed_kappa=3)
    A, free = _csr(P), P["fes"].freedofs.astype(np.uint8)
    A.sort_indices()
    C = _create(A, free)
    Ad = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    f = np.where(free != 0, np.random.default_rng(5).uniform(0, 1, P["ndof"]), 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(Ad, C, df, du, True, 1e-10, -1.0, 100, 16)
    fd = np.flatnonzero(free)
    r = (A @ du.download())[fd] - f[fd]
    assert its == 2 and np.abs(r).max() <= 1e-10 * np.abs(f).max()


def test_python_api_inverse_and_two_level(device):
    import ngsolve_amd as ngs
    m = ngs.MakeStructured3DMesh(nx=5, ny=5, nz=5)
    fes = ngs.H1(m, order=3, dirichlet=".*")
    u, v = fes.TnT()
    a = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble()
    inv = a.mat.Inverse(fes.FreeDofs(), inverse="sparsecholesky")
    gfu = ngs.GridFunction(fes)
    gfu.vec.data = inv * f.vec
    A = sp.csr_matrix((a.mat.vals, a.mat.cols, a.mat.rowptr), shape=(fes.ndof, fes.ndof))
    fr = np.asarray(fes.FreeDofs().mask, dtype=bool)
    res = (f.vec.NumPy() - A @ gfu.vec.NumPy())[fr]
    assert np.abs(res).max() <= 1e-11 * np.abs(f.vec.NumPy()).max()
    nv = m.nv
    lo = ngs.LowOrderBlock(a.mat, nv)
    pre = ngs.TwoLevelPrecond(a.mat.CreateSmoother(fes.FreeDofs()), lo.Inverse(fr[:nv].astype(np.uint8)))
    sol = ngs.DevCGSolver(a.mat, pre, tol=1e-10, maxiter=2000)
    gfu2 = ngs.GridFunction(fes)
    gfu2.vec.data = sol * f.vec
    assert np.linalg.norm(gfu2.vec.NumPy() - gfu.vec.NumPy()) <= 1e-9 * np.linalg.norm(gfu.vec.NumPy())


@pytest.mark.slow
def test_two_level_full_size(device):
    """32^3 p=4, kappa = exp(N(0,1)) (seed 2), f = the assembled load vector:
    device CG with Jacobi + P1 coarse solve against the oracle's PCG with the
    same preconditioner."""
    say("[twolevel full] 32^3 p=4 set-up")
    P, A, T, pre, free = _two_level_case(32, 4, "jacobi")
    f = N.AssembleLinearForm(P["fes"], P["ir"], np.ones((P["mesh"].ne, len(P["ir"]))))
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, T, df, du, True, 1e-10, -1.0, 5000, 16)
    say(f"[twolevel full] device its {its}; oracle PCG with the same preconditioner")
    Am = _csr(P)
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(P["rowptr"], P["cols"], P["vals"], t), pre, f, tol=1e-10,
                              maxiter=5000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[twolevel full] its {its} oracle {its_ref}, solution {err:.2e}")
    assert abs(its - its_ref) <= 1
    assert err <= 1e-10
    del Am
