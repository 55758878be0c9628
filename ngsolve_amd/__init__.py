"""ngsolve_amd: MI355X-native H1 operator application and Krylov solves
behind NGSolve's BaseMatrix / BaseVector / CGSolver API.

The hot path (CSR SpMV, fused matrix-free apply, CG vector kernels) runs in
libngsamd.so (hand-written HIP for gfx950); the host setup (meshes, H1 dofs,
element classes, assembly) is C++ in _ngsamd.  Importing this package
without the built libraries raises ImportError: there is no CPU fallback.
"""
from ._native import N as _N  # noqa: F401  (fails loudly when unbuilt)
from ._native import ensure_device
from .comp import (H1, BilinearForm, CoefficientFunction, DevMatrixFree, ElementCoefficient, GridFunction,
                   Integrate, LinearForm, Mesh, MFOpts, dx, grad, x, y, z)
from .krylovspace import CG, CGSolver, DevCGSolver, IRCGSolver, LinearSolver
from .la import (BaseMatrix, BaseVector, BitArray, BlockJacobiPrecond, CreateVVector, DeviceMatrix, DevSparseMatrix,
                 InnerProduct, JacobiPrecond, LowOrderBlock, MultiGrid, Norm, Projector, SparseCholesky,
                 TwoLevelPrecond)
from .meshes import MakeStructured2DMesh, MakeStructured3DMesh

__all__ = [
    "H1", "BilinearForm", "CoefficientFunction", "DevMatrixFree", "ElementCoefficient", "GridFunction",
    "Integrate", "LinearForm", "Mesh", "MFOpts", "dx", "grad", "x", "y", "z", "CG", "CGSolver", "DevCGSolver", "IRCGSolver",
    "LinearSolver", "BaseMatrix", "BaseVector", "BitArray", "CreateVVector", "DeviceMatrix", "DevSparseMatrix",
    "InnerProduct", "JacobiPrecond", "BlockJacobiPrecond", "LowOrderBlock", "MultiGrid", "Norm", "Projector",
    "SparseCholesky",
    "TwoLevelPrecond", "MakeStructured2DMesh", "MakeStructured3DMesh",
    "ensure_device",
]
