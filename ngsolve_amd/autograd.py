"""Differentiating through the solve with torch.autograd.

SolvePoisson.apply(kappa, problem) returns u(kappa) = A(kappa)^{-1} f for the
H1 Laplace operator A(kappa) = sum_e kappa_e K_e restricted to the free dofs
(Dirichlet values 0), with kappa a float64 CUDA tensor of one value per
element.  The forward pass is the device matrix-free operator with the new
coefficient plus a Jacobi-PCG solve; the backward pass is the adjoint solve
A lam = dJ/du (A is symmetric) followed by the element kernel
dJ/dkappa_e = -lam_e^T K_e u_e / kappa_e = -int_e grad(lam).grad(u)
(ngsamd_mf_element_grad).

ApplySymmetric.apply(x, op) is y = M x for a symmetric device operator M
(e.g. a mass matrix), with the backward pass M g.

Torch tensors are used for device memory only; every product and solve runs
in libngsamd.  The library runs on its own stream, so the wrappers
synchronise torch's stream before and the library's stream after each call.
"""
import numpy as np
import torch

from ._native import N, ensure_device
from .comp import DevMatrixFree


def _view(t):
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return N.DeviceVector.wrap(t.data_ptr(), t.numel(), 8)


class PoissonProblem:
    """Everything the solve needs except kappa: the FP64 matrix-free
    operator, the free dofs, the right-hand side and the CG settings."""

    def __init__(self, fes, f, order=None, tol=1e-12, maxiter=200000):
        ensure_device()
        self.fes = fes
        order = 2 * fes.order - 2 if order is None else order
        self.mf = DevMatrixFree(fes, order, None, "f64")
        self.ndof = fes.ndof
        self.ne = fes.mesh.ne
        self.free = np.ascontiguousarray(fes.space.freedofs)
        self.f = torch.as_tensor(np.asarray(f, dtype=np.float64), device="cuda").contiguous()
        self.tol, self.maxiter = tol, maxiter
        self._diag = N.DeviceVector(self.ndof, 8)
        self.iterations = []

    def bind(self, kappa):
        """Point the operator at kappa and rebuild the Jacobi preconditioner."""
        torch.cuda.synchronize()
        self._kappa = kappa
        self._kview = _view(kappa)
        self.mf.op.set_kappa(self._kview)
        self.mf.op.diagonal(self._diag)
        self.pre = N.CreateJacobiDevice(self._diag, self.free)

    def solve(self, rhs):
        torch.cuda.synchronize()
        u = torch.zeros_like(rhs)
        its, _ = N.CGSolve(self.mf.op, self.pre, _view(rhs), _view(u), True, self.tol, -1.0, self.maxiter, 16)
        self.iterations.append(its)
        return u

    def element_grad(self, lam, u):
        torch.cuda.synchronize()
        g = torch.empty(self.ne, dtype=torch.float64, device="cuda")
        self.mf.op.element_grad(_view(lam), _view(u), _view(g))
        N.synchronize()
        return g


class SolvePoisson(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, problem):
        kappa = kappa.detach().contiguous()
        problem.bind(kappa)
        u = problem.solve(problem.f)
        ctx.problem = problem
        ctx.save_for_backward(kappa, u)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        kappa, u = ctx.saved_tensors
        P = ctx.problem
        P.bind(kappa)
        lam = P.solve(grad_u.contiguous())
        return P.element_grad(lam, u), None


class ApplySymmetric(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op):
        x = x.contiguous()
        torch.cuda.synchronize()
        y = torch.empty_like(x)
        op.mult(_view(x), _view(y))
        N.synchronize()
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, g):
        g = g.contiguous()
        torch.cuda.synchronize()
        out = torch.empty_like(g)
        ctx.op.mult(_view(g), _view(out))
        N.synchronize()
        return out, None
