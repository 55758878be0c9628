"""Block-Jacobi preconditioner on the device (ngsamd_blockjacobi_create)
against the oracle's restatement of BlockJacobiPrecond
(oracle/fe_oracle.py block_jacobi): the apply within 1e-12 of the largest
entry for per-node blocks, random overlapping blocks and singleton blocks
(= Jacobi); Mult / MultAdd semantics; the device CG with the block
preconditioner (per-node blocks and the reference's H1 smoothing block types)
against the oracle's PCG with the same preconditioner (iterations within 1,
solution within 1e-10); the Python CGSolver with it; singular blocks
refused."""
import numpy as np
import pytest

import ngsolve_amd as ngs
import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


def _create(P, blocks):
    bptr = np.zeros(len(blocks) + 1, dtype=np.int64)
    bptr[1:] = np.cumsum([len(b) for b in blocks])
    bd = np.concatenate([np.asarray(b, dtype=np.int32) for b in blocks]) if blocks else np.zeros(0, np.int32)
    return N.CreateBlockJacobi(P["ndof"], P["rowptr"], P["cols"], P["vals"], bptr, bd)


def _node_blocks(P):
    sp = P["fes"]
    nv = P["mesh"].nv
    firsts = [sp.first_edge_dof, sp.first_el_dof] if P["dim"] == 2 else [sp.first_edge_dof, sp.first_face_dof,
                                                                       sp.first_el_dof]
    starts = np.concatenate([np.arange(nv + 1)] + [np.asarray(f)[1:] for f in firsts])
    free = sp.freedofs.astype(bool)
    return [b[free[b]] for b in np.split(np.arange(P["ndof"]), starts[1:-1])]


def _overlapping_blocks(P, rng):
    """Blocks of 1..12 dofs drawn at random: most dofs in several blocks."""
    free = np.flatnonzero(P["fes"].freedofs)
    return [np.sort(rng.choice(free, size=rng.integers(1, 13), replace=False)) for _ in range(len(free))]


CASES = [(2, 6, 3), (2, 5, 5), (3, 3, 2), (3, 3, 4)]


@pytest.mark.parametrize("dim,n,p", CASES)
@pytest.mark.parametrize("kind", ["node", "overlap", "single"])
def test_blockjacobi_apply_matches_oracle(device, dim, n, p, kind):
    say(f"[blockjacobi {kind}] dim={dim} n={n} p={p}")
    P = problem(dim, n, p, seed_kappa=3)
    rng = np.random.default_rng(7)
    if kind == "node":
        blocks = _node_blocks(P)
    elif kind == "overlap":
        blocks = _overlapping_blocks(P, rng)
    else:
        blocks = [[d] for d in np.flatnonzero(P["fes"].freedofs)]
    C = _create(P, blocks)
    _, apply = O.block_jacobi(P["rowptr"], P["cols"], P["vals"], blocks)
    x = rng.uniform(-1, 1, P["ndof"])
    ref = apply(x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    C.mult(dx, dy)
    assert relerr_inf(dy.download(), ref) < 1e-12
    y0 = rng.uniform(-1, 1, P["ndof"])
    dy.upload(y0)
    C.mult_add(-0.5, dx, dy)
    assert relerr_inf(dy.download(), y0 - 0.5 * ref) < 1e-12
    if kind == "single":
        free = P["fes"].freedofs.astype(np.uint8)
        _, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
        assert relerr_inf(ref, inv * x) < 1e-14


@pytest.mark.parametrize("dim,n,p", [(2, 16, 3), (3, 4, 4)])
@pytest.mark.parametrize("chunk", [0, 16])
def test_devcg_blockjacobi_matches_oracle(device, dim, n, p, chunk):
    say(f"[blockjacobi cg] dim={dim} n={n} p={p} chunk={chunk}")
    P = problem(dim, n, p, seed_kappa=5)
    blocks = _node_blocks(P)
    C = _create(P, blocks)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    free = P["fes"].freedofs.astype(bool)
    f = np.where(free, np.random.default_rng(3).uniform(0, 1, P["ndof"]), 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, res = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 2000, chunk)
    _, apply = O.block_jacobi(P["rowptr"], P["cols"], P["vals"], blocks)
    mv = lambda v: oracle.spmv(P["rowptr"], P["cols"], P["vals"], v)
    u_ref, its_ref, res_ref = O.pcg(mv, apply, f, tol=1e-10, maxiter=2000)
    assert abs(its - its_ref) <= 1
    u = du.download()
    assert np.abs(u - u_ref).max() <= 1e-10 * np.abs(u_ref).max()


@pytest.mark.parametrize("dim,p,blocktype", [(2, 4, 1), (2, 4, 2), (2, 3, 4), (3, 3, 3), (3, 4, 4)])
def test_devcg_smoothing_blocks_match_oracle(device, dim, p, blocktype):
    """The reference's H1 smoothing blocks (type 4, its default, overlaps:
    vertex-edge patches) in the device CG against the oracle's PCG."""
    say(f"[blockjacobi cg] smoothing blocks type {blocktype} dim={dim} p={p}")
    m = ngs.MakeStructured2DMesh(nx=12, ny=12) if dim == 2 else ngs.MakeStructured3DMesh(nx=4)
    fes = ngs.H1(m, order=p, dirichlet=".*")
    u, v = fes.TnT()
    kappa = ngs.ElementCoefficient(np.exp(np.random.default_rng(4).standard_normal(m.ne)))
    a = ngs.BilinearForm(kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    blocks = fes.CreateSmoothingBlocks(blocktype)
    pre = a.mat.CreateBlockSmoother(blocks)
    free = fes.space.freedofs.astype(bool)
    f = np.where(free, np.random.default_rng(3).uniform(0, 1, fes.ndof), 0.0)
    df, du = dvec(f), N.DeviceVector(fes.ndof, 8)
    its, _ = N.CGSolve(a.mat.op, pre.op, df, du, True, 1e-10, -1.0, 3000, 16)
    rp, c, vals = a.mat.rowptr, a.mat.cols, a.mat.vals
    _, apply = O.block_jacobi(rp, c, vals, blocks)
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(rp, c, vals, t), apply, f, tol=1e-10, maxiter=3000)
    assert abs(its - its_ref) <= 1
    assert np.abs(du.download() - u_ref).max() <= 1e-10 * np.abs(u_ref).max()


def test_python_cgsolver_with_block_smoother(device):
    say("[blockjacobi] CGSolver + CreateBlockSmoother")
    m = ngs.MakeStructured3DMesh(nx=3)
    fes = ngs.H1(m, order=3, dirichlet=".*")
    u, v = fes.TnT()
    a = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble()
    pre = a.mat.CreateBlockSmoother(fes.CreateSmoothingBlocks())
    gfu1 = ngs.GridFunction(fes)
    gfu2 = ngs.GridFunction(fes)
    inv1 = ngs.CGSolver(a.mat, pre, tol=1e-12, maxiter=500)
    inv2 = ngs.DevCGSolver(a.mat, pre, tol=1e-12, maxiter=500)
    gfu1.vec.data = inv1 * f.vec
    gfu2.vec.data = inv2 * f.vec
    d = gfu1.vec.NumPy() - gfu2.vec.NumPy()
    assert np.abs(d).max() <= 1e-10 * np.abs(gfu1.vec.NumPy()).max()
    assert abs(inv1.iterations - inv2.iterations) <= 1


def test_singular_block_refused(device):
    say("[blockjacobi] singular block")
    P = problem(2, 3, 1)
    vals = P["vals"].copy()
    # zero row and column 0 inside the matrix: the block {0} is singular
    rows = np.repeat(np.arange(P["ndof"]), np.diff(P["rowptr"]))
    vals[(rows == 0) | (P["cols"] == 0)] = 0.0
    Q = dict(P, vals=vals)
    with pytest.raises(RuntimeError, match="singular"):
        _create(Q, [[0], [1, 2]])
