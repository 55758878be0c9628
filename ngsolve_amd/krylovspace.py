"""Krylov solvers with the reference's Python semantics.

CGSolver      python/krylovspace.py:227-291 (recurrence, CheckResidual
              :129-150: `iterations` counts the initial check, stopping on the
              preconditioned residual sqrt|<w,d>|).  Works on any BaseMatrix;
              every inner product returns to the host, as in the reference.
DevCGSolver   the device-resident loop (ngscuda/python_ngscuda.cpp:243-266,
              ngscuda/cuda_krylov.cpp:19-203): same recurrence and stopping
              rule, scalars on the device, iterations replayed from a HIP
              graph.  Needs device operators and a Jacobi preconditioner.
IRCGSolver    mixed-precision iterative refinement: FP64 residual f - A u
              with `mat`, inner Jacobi-PCG on a low-precision matrix-free
              operator (FP32 vectors, FP64 dots/scalars).
"""
from math import log, sqrt

from .la import (BaseMatrix, BlockJacobiPrecond, DeviceMatrix, JacobiPrecond, MultiGrid, Projector, SparseCholesky,
                 TwoLevelPrecond)


class LinearSolver(BaseMatrix):
    name = "LinearSolver"

    def __init__(self, mat, pre=None, freedofs=None, tol=None, maxiter=100, atol=None, callback=None,
                 callback_sol=None, printrates=False, plotrates=False):
        super().__init__()
        if atol is None and tol is None:
            tol = 1e-12
        self.mat = mat
        assert (freedofs is None) != (pre is None)  # either pre or freedofs must be given
        self.pre = pre if pre else Projector(freedofs, True)
        self.tol = tol
        self.atol = atol
        self.maxiter = maxiter
        self.callback = callback
        self.callback_sol = callback_sol
        self.printrates = printrates
        self.residuals = []
        self.iterations = 0

    def Solve(self, rhs, sol=None, initialize=True):
        self.iterations = 0
        self.residuals = []
        if sol is None:
            sol = rhs.CreateVector()
            initialize = True
        if initialize:
            sol[:] = 0
        self.sol = sol
        self._SolveImpl(rhs=rhs, sol=sol)
        return sol

    def Height(self):
        return self.mat.width

    def Width(self):
        return self.mat.height

    def IsSymmetric(self):
        return True

    def Mult(self, x, y):
        self.Solve(rhs=x, sol=y, initialize=True)

    def CheckResidual(self, residual):
        self.iterations += 1
        self.residuals.append(residual)
        if len(self.residuals) == 1:
            if self.tol is None:
                self._final_residual = self.atol
            else:
                self._final_residual = residual * self.tol
                if self.atol is not None:
                    self._final_residual = max(self._final_residual, self.atol)
        else:
            if self.callback is not None:
                self.callback(self.iterations, residual)
            if self.callback_sol is not None:
                self.callback_sol(self.sol)
        if self.printrates:
            print("{} iteration {}, residual = {}".format(self.name, self.iterations, residual))
            if self.iterations == self.maxiter and residual > self._final_residual:
                print("WARNING: {} did not converge to TOL".format(self.name))
        return self.iterations >= self.maxiter or residual <= self._final_residual


class CGSolver(LinearSolver):
    name = "CG"

    def __init__(self, *args, conjugate=False, abstol=None, maxsteps=None, printing=False, **kwargs):
        if printing:
            kwargs["printrates"] = printing
        if abstol is not None:
            kwargs["atol"] = abstol
        if maxsteps is not None:
            kwargs["maxiter"] = maxsteps
        super().__init__(*args, **kwargs)
        self.conjugate = conjugate

    @property
    def errors(self):
        return self.residuals

    def _SolveImpl(self, rhs, sol):
        d, w, s = [sol.CreateVector() for i in range(3)]
        conjugate = self.conjugate
        d.data = rhs - self.mat * sol
        w.data = self.pre * d
        s.data = w
        wdn = w.InnerProduct(d, conjugate=conjugate)
        if self.CheckResidual(sqrt(abs(wdn))):
            return

        while True:
            w.data = self.mat * s
            wd = wdn
            as_s = s.InnerProduct(w, conjugate=conjugate)
            if as_s == 0 or wd == 0:
                break
            alpha = wd / as_s
            sol.data += alpha * s
            d.data += (-alpha) * w

            w.data = self.pre * d

            wdn = w.InnerProduct(d, conjugate=conjugate)
            if self.CheckResidual(sqrt(abs(wdn))):
                return

            beta = wdn / wd
            s *= beta
            s.data += w


def CG(mat, rhs, pre=None, sol=None, tol=1e-12, maxsteps=100, printrates=True, initialize=True, **kwargs):
    solver = CGSolver(mat=mat, pre=pre, tol=tol, maxiter=maxsteps, printrates=printrates)
    return solver.Solve(rhs=rhs, sol=sol, initialize=initialize)


class DevCGSolver(LinearSolver):
    """Device-resident CG (graph-captured iteration, device scalars)."""

    name = "DevCG"

    def __init__(self, mat, pre, tol=1e-12, maxiter=400, atol=None, chunk=16, precision=None, maxsteps=None,
                 printrates=False):
        if precision is not None:
            tol = precision
        if maxsteps is not None:
            maxiter = maxsteps
        super().__init__(mat, pre=pre, tol=tol, maxiter=maxiter, atol=atol, printrates=printrates)
        if not isinstance(mat, DeviceMatrix) or not isinstance(pre, (JacobiPrecond, BlockJacobiPrecond,
                                                                     TwoLevelPrecond, SparseCholesky, MultiGrid)):
            raise TypeError("DevCGSolver needs a device matrix and a JacobiPrecond, BlockJacobiPrecond, "
                            "TwoLevelPrecond, SparseCholesky or MultiGrid preconditioner")
        self.chunk = chunk

    def _SolveImpl(self, rhs, sol):
        from ._native import N
        its, res = N.CGSolve(self.mat.op, self.pre.op, rhs._device(), sol._write_device(), False,
                             -1.0 if self.tol is None else float(self.tol),
                             -1.0 if self.atol is None else float(self.atol), int(self.maxiter), int(self.chunk))
        self.iterations = its
        self.residuals = [float(r) for r in res]
        if self.printrates and self.residuals:
            print(f"{self.name} converged in {its} iterations to residual {self.residuals[-1]}")


class IRCGSolver(LinearSolver):
    """FP64 iterative refinement around a low-precision matrix-free inner CG.

    `mat` is the FP64 operator for the residual, `inner` the FP32 / FP16 /
    BF16 / split-FP16 ("f16x3") matrix-free operator; stops when
    ||f - A u||_free <= tol ||f||_free.  inner_tol=None uses the solver's
    default for the inner precision (FP32 and split FP16 3e-4, FP16 1e-3,
    BF16 3e-2)."""

    name = "IRCG"

    def __init__(self, mat, inner, pre, tol=1e-10, inner_tol=None, maxouter=50, maxinner=20000):
        super().__init__(mat, pre=pre, tol=tol, maxiter=maxouter)
        self.inner = inner
        self.inner_tol = inner_tol
        self.maxinner = maxinner
        self.outer_iterations = 0
        self.inner_iterations = 0
        self.relative_residual = None

    def _SolveImpl(self, rhs, sol):
        from ._native import N
        outer, inner, rel = N.IRCGSolve(self.mat.op, self.inner.op, self.pre.op, rhs._device(), sol._write_device(),
                                        float(self.tol), -1.0 if self.inner_tol is None else float(self.inner_tol),
                                        int(self.maxiter),
                                        int(self.maxinner))
        self.outer_iterations, self.inner_iterations, self.relative_residual = outer, inner, rel
        self.iterations = outer
