"""Multigrid V-cycle for the P1 block of structured meshes on the device
(ngsamd_multigrid_create) against the oracle's restatement (oracle/fe_oracle.py
mg_hierarchy, mg_vcycle, mg_apply: MultigridPreconditioner::MGM,
multigrid/mgpre.cpp:229-318, with LinearProlongation,
multigrid/prolongation.cpp:77-144, Galerkin coarse levels): the apply within
1e-12 of the oracle's in the inf-norm (2D and 3D, lognormal kappa, vertex
blocks of P1 and higher-order matrices, Mult / MultAdd, zero off the free
dofs); device CG preconditioned by it on P1 problems and by the two-level
preconditioner Jacobi + multigrid P1 coarse part on p=4 problems against the
oracle's PCG with the same preconditioner (iterations within 1, solution
within 1e-10), through the Python API, and at the 32^3 p=4 size with kappa =
exp(N(0,1)) (seed 2)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, relerr_inf, say

pytestmark = pytest.mark.gpu

NU, OMEGA = 2, 0.85


def _csr(P):
    return sp.csr_matrix((P["vals"], P["cols"], P["rowptr"]), shape=(P["ndof"], P["ndof"]))


def _vertex_block(P):
    nv = P["mesh"].nv
    A = _csr(P)[:nv][:, :nv].tocsr()
    A.sort_indices()
    return A, P["fes"].freedofs.astype(np.uint8)[:nv].copy()


def _create(dim, n, A, free, max_levels=32):
    return N.CreateMultiGrid(dim, [n] * dim, A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                             A.data.astype(np.float64), free, NU, OMEGA, max_levels)


@pytest.mark.parametrize("dim,n,p", [(3, 16, 1), (3, 16, 2), (2, 64, 1), (2, 64, 3), (3, 12, 1), (3, 8, 1)])
def test_multigrid_apply_matches_oracle(device, dim, n, p):
    say(f"[mg] apply {dim}D n={n} p={p}")
    P = problem(dim, n, p, seed_kappa=7)
    A, free = _vertex_block(P)
    M = _create(dim, n, A, free)
    levels = N.MultiGridInfo(M)
    H = O.mg_hierarchy(dim, [n] * dim, A.indptr, A.indices, A.data, free)
    assert levels == [a.shape[0] for a in H["A"]]
    x = np.random.default_rng(3).uniform(-1, 1, A.shape[0])
    ref = O.mg_apply(H, x, NU, OMEGA)
    dx, dy = dvec(x), N.DeviceVector(A.shape[0], 8)
    M.mult(dx, dy)
    y = dy.download()
    err = relerr_inf(y, ref)
    say(f"[mg] {dim}D n={n} p={p}: levels {levels}, apply vs oracle {err:.2e}")
    assert err < 1e-12
    assert np.all(y[free == 0] == 0.0)
    y0 = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    dy.upload(y0)
    M.mult_add(-0.5, dx, dy)
    assert relerr_inf(dy.download(), y0 - 0.5 * ref) < 1e-12
    dz = N.DeviceVector(A.shape[0], 8)
    M.mult(dx, dz)
    assert np.array_equal(dz.download(), y)


@pytest.mark.parametrize("nu", [1, 3])
def test_multigrid_smoothing_steps(device, nu):
    """Other numbers of smoothing sweeps (the fused first and last passes
    take other branches): the apply against the oracle's V-cycle."""
    P = problem(3, 16, 1, seed_kappa=8)
    A, free = _vertex_block(P)
    M = N.CreateMultiGrid(3, [16] * 3, A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                          A.data.astype(np.float64), free, nu, 0.7, 32)
    H = O.mg_hierarchy(3, [16] * 3, A.indptr, A.indices, A.data, free)
    x = np.random.default_rng(9).uniform(-1, 1, A.shape[0])
    dx, dy = dvec(x), N.DeviceVector(A.shape[0], 8)
    M.mult(dx, dy)
    err = relerr_inf(dy.download(), O.mg_apply(H, x, nu, 0.7))
    say(f"[mg] nu={nu}: apply vs oracle {err:.2e}")
    assert err < 1e-12


def test_multigrid_rejects_bad_input(device):
    P = problem(3, 6, 1)
    A, free = _vertex_block(P)
    with pytest.raises(RuntimeError, match="vertex count"):
        N.CreateMultiGrid(3, [6, 6, 5], A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                          A.data, free, NU, OMEGA, 32)
    with pytest.raises(RuntimeError, match="omega"):
        N.CreateMultiGrid(3, [6, 6, 6], A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int32),
                          A.data, free, NU, 2.0, 32)
    # 26^3: 15,625 free vertices, then 13^3 with 1,728, which cannot be halved
    P = problem(3, 26, 1)
    A, free = _vertex_block(P)
    with pytest.raises(RuntimeError, match="coarsest"):
        _create(3, 26, A, free)


@pytest.mark.parametrize("dim,n", [(3, 32), (2, 128)])
def test_multigrid_cg_p1_matches_oracle(device, dim, n):
    """P1 Poisson with lognormal kappa: CG with the V-cycle needs a few
    iterations independent of n where Jacobi-CG needs hundreds."""
    say(f"[mg] P1 CG {dim}D n={n}")
    P = problem(dim, n, 1, seed_kappa=2)
    A, free = _vertex_block(P)
    M = _create(dim, n, A, free)
    H = O.mg_hierarchy(dim, [n] * dim, A.indptr, A.indices, A.data, free)
    Ad = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    f = np.where(free != 0, np.random.default_rng(4).uniform(0, 1, P["ndof"]), 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(Ad, M, df, du, True, 1e-10, -1.0, 2000, 16)
    Am = _csr(P)
    u_ref, its_ref, _ = O.pcg(lambda t: Am @ t, lambda t: O.mg_apply(H, t, NU, OMEGA), f, tol=1e-10, maxiter=2000)
    _, its_j, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"],
                                 O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)[1], f, tol=1e-10,
                                 maxiter=5000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[mg] P1 {dim}D n={n}: its {its} oracle {its_ref} (Jacobi {its_j}), solution {err:.2e}")
    assert abs(its - its_ref) <= 1
    assert err <= 1e-10
    assert its < its_j / 4


def _two_level_mg_case(n, p, seed=2):
    P = problem(3, n, p, order=2 * p, seed_kappa=seed)
    free = P["fes"].freedofs.astype(np.uint8)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    Av, fv = _vertex_block(P)
    M = _create(3, n, Av, fv)
    H = O.mg_hierarchy(3, [n] * 3, Av.indptr, Av.indices, Av.data, fv)
    S = N.CreateJacobi(diag, free)
    T = N.CreateTwoLevel(S, M)
    pre = O.two_level(lambda t: inv * t, lambda t: O.mg_apply(H, t, NU, OMEGA), Av.shape[0])
    return P, A, T, pre, free


def test_two_level_multigrid_cg_matches_oracle(device):
    say("[mg] two-level Jacobi + multigrid, 16^3 p=4")
    P, A, T, pre, free = _two_level_mg_case(16, 4)
    x = np.random.default_rng(4).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    T.mult(dx, dy)
    assert relerr_inf(dy.download(), pre(x)) < 1e-12
    f = np.where(free != 0, 1.0, 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, T, df, du, True, 1e-10, -1.0, 5000, 16)
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(P["rowptr"], P["cols"], P["vals"], t), pre, f, tol=1e-10,
                              maxiter=5000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[mg] two-level 16^3 p=4: its {its} oracle {its_ref}, solution {err:.2e}")
    assert abs(its - its_ref) <= 1
    assert err <= 1e-10


def test_python_api_multigrid(device):
    import ngsolve_amd as ngs
    m = ngs.MakeStructured3DMesh(nx=16, ny=16, nz=16)
    fes = ngs.H1(m, order=1, dirichlet=".*")
    u, v = fes.TnT()
    a = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble()
    mg = ngs.MultiGrid(a.mat, (16, 16, 16), fes.FreeDofs())
    assert mg.levels == [3375, 343]
    sol = ngs.DevCGSolver(a.mat, mg, tol=1e-10, maxiter=200)
    gfu = ngs.GridFunction(fes)
    gfu.vec.data = sol * f.vec
    A = sp.csr_matrix((a.mat.vals, a.mat.cols, a.mat.rowptr), shape=(fes.ndof, fes.ndof))
    fr = np.asarray(fes.FreeDofs().mask, dtype=bool)
    res = (f.vec.NumPy() - A @ gfu.vec.NumPy())[fr]
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f.vec.NumPy()[fr])
    assert sol.iterations < 30
    # p=3: two-level with the multigrid P1 part
    fes3 = ngs.H1(m, order=3, dirichlet=".*")
    u3, v3 = fes3.TnT()
    a3 = ngs.BilinearForm(ngs.grad(u3) * ngs.grad(v3) * ngs.dx).Assemble()
    f3 = ngs.LinearForm(1 * v3 * ngs.dx).Assemble()
    nv = m.nv
    fr3 = np.asarray(fes3.FreeDofs().mask, dtype=np.uint8)
    pre = ngs.TwoLevelPrecond(a3.mat.CreateSmoother(fes3.FreeDofs()),
                              ngs.MultiGrid(ngs.LowOrderBlock(a3.mat, nv), (16, 16, 16), fr3[:nv]))
    sol3 = ngs.DevCGSolver(a3.mat, pre, tol=1e-10, maxiter=2000)
    g3 = ngs.GridFunction(fes3)
    g3.vec.data = sol3 * f3.vec
    A3 = sp.csr_matrix((a3.mat.vals, a3.mat.cols, a3.mat.rowptr), shape=(fes3.ndof, fes3.ndof))
    fr3b = fr3.astype(bool)
    r3 = (f3.vec.NumPy() - A3 @ g3.vec.NumPy())[fr3b]
    assert np.linalg.norm(r3) <= 1e-9 * np.linalg.norm(f3.vec.NumPy()[fr3b])


@pytest.mark.slow
def test_two_level_multigrid_full_size(device):
    """32^3 p=4, kappa = exp(N(0,1)) (seed 2), f = the assembled load vector:
    device CG with Jacobi + multigrid P1 part against the oracle's PCG with the
    same preconditioner."""
    say("[mg full] 32^3 p=4 set-up")
    P, A, T, pre, free = _two_level_mg_case(32, 4)
    f = N.AssembleLinearForm(P["fes"], P["ir"], np.ones((P["mesh"].ne, len(P["ir"]))))
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, T, df, du, True, 1e-10, -1.0, 5000, 16)
    say(f"[mg full] device its {its}; oracle PCG with the same preconditioner")
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(P["rowptr"], P["cols"], P["vals"], t), pre, f, tol=1e-10,
                              maxiter=5000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[mg full] its {its} oracle {its_ref}, solution {err:.2e}")
    assert abs(its - its_ref) <= 1
    assert err <= 1e-10
