"""The reference's Python operator surface on the device path:
BaseMatrix subclassing (tests/pytest/test_basematrix.py:1-66), vector
expressions, BilinearForm/LinearForm/CGSolver, and the p=4 exactness
known answer (tests/pytest/test_solvers.py:58-77) solved on the GPU."""
import numpy as np
import pytest

import ngsolve_amd as ngs
from helpers import say

pytestmark = pytest.mark.gpu


class MyMatrix(ngs.BaseMatrix):
    def __init__(self, n):
        super().__init__()
        self.n = n

    def IsComplex(self):
        return False

    def Mult(self, x, y):
        for i in range(len(x)):
            y[i] = sum(x[:i])

    def MultTransAdd(self, s, x, y):
        for i in range(len(x)):
            y[i] = s * sum(x[i:])

    def CreateColVector(self):
        return ngs.CreateVVector(self.n)

    def Height(self):
        return self.n

    def Width(self):
        return self.n


def test_derive_basematrix(device):
    say("[api] derived BaseMatrix")
    m = MyMatrix(5)
    x = ngs.CreateVVector(5)
    y = ngs.CreateVVector(5)
    for i in range(len(x)):
        x[i] = 1 + i
    x *= 10
    for i in range(len(x)):
        assert x[i] == 10 + 10 * i
    y.data = 2.0 * m * x
    assert list(y) == [0.0, 20.0, 60.0, 120.0, 200.0]
    y.data = 2.0 * m.T * x
    assert list(y) == [300.0, 280.0, 240.0, 180.0, 100.0]


def test_derive_basematrix_lifetime(device):
    m = MyMatrix(5) @ MyMatrix(5)
    m1 = MyMatrix(5)
    x, y, y1 = ngs.CreateVVector(5), ngs.CreateVVector(5), ngs.CreateVVector(5)
    for i in range(len(x)):
        x[i] = 1 + i
    y.data = m * x
    y1.data = (m1 @ m1) * x
    assert list(y) == list(y1)


def test_vector_expressions(device):
    say("[api] vector expressions")
    a = ngs.BaseVector.from_numpy(np.arange(6.0))
    b = ngs.BaseVector.from_numpy(np.ones(6))
    c = a.CreateVector()
    c.data = 2 * a - b
    assert np.array_equal(c.NumPy(), 2 * np.arange(6.0) - 1)
    c.data += 3 * b
    assert np.array_equal(c.NumPy(), 2 * np.arange(6.0) + 2)
    c.data -= a
    assert np.array_equal(c.NumPy(), np.arange(6.0) + 2)
    assert a.InnerProduct(b) == 15.0
    c *= 0.5
    assert np.array_equal(c.NumPy(), (np.arange(6.0) + 2) / 2)
    c[:] = 0
    assert not c.NumPy().any()


def test_matrix_expression_with_alias(device):
    """`d.data = rhs - mat * d` must not overwrite d while it is read."""
    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 7))
    rp = np.arange(0, 50, 7, dtype=np.int64)
    mat = ngs.DevSparseMatrix(7, 7, rp, np.tile(np.arange(7, dtype=np.int32), 7), A.ravel(), symmetric=False)
    d0, r0 = rng.standard_normal(7), rng.standard_normal(7)
    d, rhs = ngs.BaseVector.from_numpy(d0), ngs.BaseVector.from_numpy(r0)
    d.data = rhs - mat * d
    assert np.allclose(d.NumPy(), r0 - A @ d0, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("solver", ["CGSolver", "DevCGSolver"])
@pytest.mark.parametrize("mesh", ["structured", "square.vol.gz"])
def test_p4_exact_on_gpu(device, solver, mesh):
    """Order-4 Poisson with f = 32(y(1-y)+x(1-x)) is exact (error < 1e-12)."""
    say(f"[api] p4 exact solver={solver} mesh={mesh}")
    import os
    from conftest import GOLDEN
    m = ngs.MakeStructured2DMesh(nx=5, ny=5) if mesh == "structured" else ngs.Mesh(os.path.join(GOLDEN, mesh))
    fes = ngs.H1(m, order=4, dirichlet=".*")
    u, v = fes.TnT()
    f = ngs.LinearForm(32 * (ngs.y * (1 - ngs.y) + ngs.x * (1 - ngs.x)) * v * ngs.dx).Assemble()
    a = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    pre = a.mat.CreateSmoother(fes.FreeDofs)
    inv = getattr(ngs, solver)(mat=a.mat, pre=pre, tol=1e-15, maxiter=2000)
    gfu = ngs.GridFunction(fes)
    gfu.vec.data = inv * f.vec
    exact = 16 * ngs.x * (1 - ngs.x) * ngs.y * (1 - ngs.y)
    err = gfu.L2Error(exact)
    say(f"[api] iterations={inv.iterations} error={err:.3e}")
    assert err < 1e-12


def test_cgsolver_and_devcgsolver_agree(device):
    """Host-driven CGSolver and the graph-captured DevCGSolver: same iterations."""
    say("[api] CGSolver vs DevCGSolver")
    m = ngs.MakeStructured2DMesh(nx=16, ny=16)
    fes = ngs.H1(m, order=3, dirichlet=".*")
    u, v = fes.TnT()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble()
    a = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx(bonus_intorder=2)).Assemble()
    pre = a.mat.CreateSmoother(fes.FreeDofs)
    s1 = ngs.CGSolver(mat=a.mat, pre=pre, tol=1e-10, maxiter=1000)
    s2 = ngs.DevCGSolver(a.mat, pre, tol=1e-10, maxiter=1000)
    u1, u2 = f.vec.CreateVector(), f.vec.CreateVector()
    u1.data = s1 * f.vec
    u2.data = s2 * f.vec
    assert abs(s1.iterations - s2.iterations) <= 1
    assert np.allclose(s1.residuals[:5], s2.residuals[:5], rtol=1e-12)
    assert np.linalg.norm(u1.NumPy() - u2.NumPy()) <= 1e-10 * np.linalg.norm(u2.NumPy())


def test_matrix_free_bilinearform(device):
    say("[api] matrix-free BilinearForm")
    m = ngs.MakeStructured3DMesh(nx=3)
    fes = ngs.H1(m, order=4, dirichlet=".*")
    u, v = fes.TnT()
    kappa = ngs.ElementCoefficient(np.exp(np.random.default_rng(2).standard_normal(m.ne)))
    a = ngs.BilinearForm(kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx(bonus_intorder=2)).Assemble()
    amf = ngs.BilinearForm(kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx(bonus_intorder=2), mf=ngs.MFOpts()).Assemble()
    x = ngs.BaseVector.from_numpy(np.random.default_rng(1).uniform(-1, 1, fes.ndof))
    y1, y2 = x.CreateVector(), x.CreateVector()
    y1.data = a.mat * x
    y2.data = amf.mat * x
    d = y1.NumPy() - y2.NumPy()
    assert np.abs(d).max() <= 1e-13 * np.abs(y1.NumPy()).max()


@pytest.mark.parametrize("precision", ["f32", "f16x3", "bf16"])
def test_ircgsolver_with_matrix_free_inner(device, precision):
    """IRCGSolver: FP64 matrix-free residuals around a low-precision
    matrix-free inner CG (MFOpts(precision=...)) reaches the FP64 CG's
    solution."""
    say(f"[api] IRCGSolver inner {precision}")
    m = ngs.MakeStructured3DMesh(nx=4)
    fes = ngs.H1(m, order=4, dirichlet=".*")
    u, v = fes.TnT()
    kappa = ngs.ElementCoefficient(np.exp(np.random.default_rng(2).standard_normal(m.ne)))
    form = kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx(bonus_intorder=2)
    a64 = ngs.BilinearForm(form, mf=ngs.MFOpts()).Assemble()
    alo = ngs.BilinearForm(form, mf=ngs.MFOpts(precision=precision)).Assemble()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble()
    pre = a64.mat.CreateSmoother(fes.FreeDofs)
    ir = ngs.IRCGSolver(a64.mat, alo.mat, pre, tol=1e-10)
    cg = ngs.DevCGSolver(a64.mat, pre, tol=1e-12, maxiter=5000)
    u1, u2 = f.vec.CreateVector(), f.vec.CreateVector()
    u1.data = ir * f.vec
    u2.data = cg * f.vec
    say(f"[api] outer {ir.outer_iterations} inner {ir.inner_iterations} residual {ir.relative_residual:.2e}")
    assert ir.relative_residual <= 1e-10
    assert np.linalg.norm(u1.NumPy() - u2.NumPy()) <= 1e-8 * np.linalg.norm(u2.NumPy())
