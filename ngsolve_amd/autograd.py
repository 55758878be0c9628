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
kappa may be a CPU or a CUDA tensor; u comes back on the same device.  CUDA
tensors need one HIP runtime in the process: torch ships its own copy of
libamdhip64 under the system soname, so import torch before ngsolve_amd
(then libngsamd binds to torch's copy).
"""
import numpy as np
import torch

from ._native import N, ensure_device
from .comp import DevMatrixFree


def _hip_runtimes():
    with open("/proc/self/maps") as f:
        return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})


def _check_single_runtime():
    libs = _hip_runtimes()
    if len(libs) > 1:
        raise RuntimeError("two HIP runtimes are loaded (" + ", ".join(libs) + "); import torch before "
                           "ngsolve_amd to use CUDA tensors, or pass CPU tensors")


def _view(t):
    assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
    return N.DeviceVector.wrap(t.data_ptr(), t.numel(), 8)


def _to_device_vector(t):
    """Owned device copy of a CPU tensor."""
    v = N.DeviceVector(t.numel(), 8)
    v.upload(t.detach().to(torch.float64).contiguous().numpy())
    return v


def _in(t):
    """Owned device vector holding the tensor's values (a device-to-device
    copy for CUDA tensors, an upload for CPU tensors)."""
    if t.is_cuda:
        _check_single_runtime()
        tc = t.detach().to(torch.float64).contiguous()
        torch.cuda.synchronize()
        v = N.DeviceVector(tc.numel(), 8)
        v.copy_from(_view(tc))
        N.synchronize()
        return v
    return _to_device_vector(t)


def _out(v, like):
    """Tensor on like's device with the values of device vector v."""
    if like.is_cuda:
        t = torch.empty(v.size, dtype=torch.float64, device=like.device)
        torch.cuda.synchronize()
        _view(t).copy_from(v)
        N.synchronize()
        return t
    return torch.from_numpy(np.asarray(v.download(), dtype=np.float64).copy())


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
        self.f = N.DeviceVector(self.ndof, 8)
        self.f.upload(np.asarray(f, dtype=np.float64))
        self.tol, self.maxiter = tol, maxiter
        self._diag = N.DeviceVector(self.ndof, 8)
        self.iterations = []

    def bind(self, kappa):
        """Point the operator at kappa (a tensor, one value per element) and
        rebuild the Jacobi preconditioner."""
        if kappa.numel() != self.ne:
            raise ValueError(f"kappa has {kappa.numel()} values, the mesh {self.ne} elements")
        self._kappa_ref = kappa  # keeps a CUDA tensor alive while bound
        self._kdev = _in(kappa)
        self.mf.op.set_kappa(self._kdev)
        self.mf.op.diagonal(self._diag)
        self.pre = N.CreateJacobiDevice(self._diag, self.free)

    def solve(self, rhs):
        """Device vector u with A u = rhs on the free dofs (u = 0 elsewhere)."""
        u = N.DeviceVector(self.ndof, 8)
        its, _ = N.CGSolve(self.mf.op, self.pre, rhs, u, True, self.tol, -1.0, self.maxiter, 16)
        self.iterations.append(its)
        return u

    def element_grad(self, lam, u):
        g = N.DeviceVector(self.ne, 8)
        self.mf.op.element_grad(lam, u, g)
        return g


class SolvePoisson(torch.autograd.Function):
    @staticmethod
    def forward(ctx, kappa, problem):
        problem.bind(kappa)
        u_dev = problem.solve(problem.f)
        u = _out(u_dev, kappa)
        ctx.problem = problem
        ctx.u_dev = u_dev
        ctx.save_for_backward(kappa)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        (kappa,) = ctx.saved_tensors
        P = ctx.problem
        P.bind(kappa)
        lam = P.solve(_in(grad_u))
        return _out(P.element_grad(lam, ctx.u_dev), kappa), None


class ApplySymmetric(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, op):
        y = N.DeviceVector(x.numel(), 8)
        op.mult(_in(x), y)
        ctx.op = op
        return _out(y, x)

    @staticmethod
    def backward(ctx, g):
        out = N.DeviceVector(g.numel(), 8)
        ctx.op.mult(_in(g), out)
        return _out(out, g), None
