"""Discretisation layer with the reference's Python names.

Mesh          comp/meshaccess.cpp (structured meshes python/meshes.py, .vol files)
H1            H1HighOrderFESpace (comp/h1hofespace.cpp:823-950,1389-1413),
              dirichlet regex over boundary names (comp/fespace.cpp:765-810)
BilinearForm  grad(u)*grad(v)*dx with an element-wise coefficient and/or
              u*v*dx; Assemble() builds the CSR matrix (comp/bilinearform.cpp:
              552-825, 1563-2215) or, with matrix-free options, the fused
              operator (comp/matrixfree.cpp:1090-1464); `.mat` is a BaseMatrix
              whose products run on the GPU.
LinearForm    f*v*dx (fem/symbolicintegrator.cpp:1036-1110)
MFOpts        MatFreeOptions (comp/bilinearform.hpp:24-52): fused, atomic, fp32,
              plus `precision` ("f64"/"f32"/"bf16"/"f16"/"f16x3" = split FP16), which
              the reference lacks
"""
import gzip
import re

import numpy as np

from ._native import N, PREC, SCATTER, ensure_device
from .la import BaseVector, BitArray, BlockTable, DeviceMatrix, DevSparseMatrix, JacobiPrecond


# ------------------------------------------------------------------- mesh

class Mesh:
    def __init__(self, ngmesh):
        if isinstance(ngmesh, str):
            opener = gzip.open if ngmesh.endswith(".gz") else open
            with opener(ngmesh, "rt") as f:
                ngmesh = N.ReadVolMesh(f.read())
        self.ngmesh = ngmesh

    @property
    def dim(self):
        return self.ngmesh.dim

    @property
    def ne(self):
        return self.ngmesh.ne

    @property
    def nv(self):
        return self.ngmesh.nv

    def GetNE(self, vb=None):
        return self.ne

    def GetBoundaries(self):
        names = self.ngmesh.boundary_names
        idx = self.ngmesh.boundary_index
        return [names[i - 1] if 0 < i <= len(names) else "bc%d" % i for i in idx]

    def __repr__(self):
        return f"Mesh(dim={self.dim}, nv={self.nv}, ne={self.ne})"


# ------------------------------------------------- coefficient functions

class CoefficientFunction:
    """Pointwise function of the physical coordinates, evaluated at the
    quadrature points of every element; or an element-wise constant."""

    def __init__(self, fn, elementwise=None):
        self.fn = fn                      # fn(X) with X (E, Q, dim) -> (E, Q)
        self.elementwise = elementwise    # array (E,) for element-wise constants

    def Evaluate(self, X):
        return np.broadcast_to(self.fn(X), X.shape[:2])

    def _bin(self, o, op):
        o = _cf(o)
        return CoefficientFunction(lambda X: op(self.Evaluate(X), o.Evaluate(X)))

    def __add__(self, o): return self._bin(o, np.add)
    def __radd__(self, o): return _cf(o)._bin(self, np.add)
    def __sub__(self, o): return self._bin(o, np.subtract)
    def __rsub__(self, o): return _cf(o)._bin(self, np.subtract)
    def __neg__(self): return CoefficientFunction(lambda X: -self.Evaluate(X))

    def __mul__(self, o):
        if isinstance(o, (_Proxy, _FormTerm)):
            return _FormTerm(self, []) * o
        if isinstance(o, _GfEnergy):
            return o * self
        return self._bin(o, np.multiply)

    def __rmul__(self, o): return _cf(o)._bin(self, np.multiply)
    def __truediv__(self, o): return self._bin(o, np.divide)
    def __pow__(self, k): return CoefficientFunction(lambda X: self.Evaluate(X) ** k)


def _cf(o):
    if isinstance(o, CoefficientFunction):
        return o
    c = float(o)
    return CoefficientFunction(lambda X: np.full(X.shape[:2], c))


x = CoefficientFunction(lambda X: X[..., 0])
y = CoefficientFunction(lambda X: X[..., 1])
z = CoefficientFunction(lambda X: X[..., 2])


def ElementCoefficient(values):
    """Element-wise constant coefficient (one value per element)."""
    v = np.ascontiguousarray(values, dtype=np.float64)
    return CoefficientFunction(lambda X: np.broadcast_to(v[:, None], X.shape[:2]), elementwise=v)


# ----------------------------------------------------------- symbolic forms

class _Proxy:
    """Trial or test function (or its gradient) in a symbolic form."""

    def __init__(self, fes, trial, deriv=False):
        self.fes, self.trial, self.deriv = fes, trial, deriv

    def __mul__(self, o):
        return _FormTerm(None, [self]) * o

    def __rmul__(self, o):
        return _FormTerm(None, [self]) * o


def grad(p):
    if isinstance(p, GridFunction):
        return _GfGrad(p)
    if not isinstance(p, _Proxy):
        raise TypeError("grad() of a trial or test function or of a GridFunction")
    return _Proxy(p.fes, p.trial, True)


class _GfGrad:
    """grad(gfu) inside an integrand of Integrate."""

    def __init__(self, gf):
        self.gf = gf

    def __mul__(self, o):
        if isinstance(o, _GfGrad):
            return _GfEnergy(self.gf, o.gf, None)
        return NotImplemented


class _GfEnergy:
    """c * grad(a) . grad(b) for GridFunctions a, b of one space and an
    element-wise constant (or scalar) c."""

    def __init__(self, a, b, coef):
        if a.space is not b.space:
            raise ValueError("grad(a)*grad(b): a and b must live in one space")
        self.a, self.b, self.coef = a, b, coef

    def __rmul__(self, c):
        if isinstance(c, (int, float)):
            val = float(c)
        else:
            c = _cf(c)
            if c.elementwise is None:
                raise ValueError("grad(a)*grad(b) takes element-wise constant coefficients only")
            val = np.asarray(c.elementwise, dtype=np.float64)
        return _GfEnergy(self.a, self.b, val if self.coef is None else self.coef * val)

    __mul__ = __rmul__


def Integrate(cf, mesh, element_wise=False, order=5):
    """Integral of cf over the mesh, or per element (element_wise=True), as
    Integrate (comp/python_comp.cpp:3697-3810).  c * grad(a) * grad(b) of
    GridFunctions is integrated on the device, element by element with the
    matrix-free operator's element matrices (ngsamd_mf_element_grad: exact
    for the space's polynomials); any other coefficient function is
    evaluated at the points of a rule of the given order and summed with the
    weights |det F| w_q."""
    if isinstance(cf, _GfGrad):
        raise TypeError("Integrate: grad(gfu) alone is a vector; integrate grad(a)*grad(b)")
    if isinstance(cf, _GfEnergy):
        fes = cf.a.space
        if cf.b.space.mesh.ngmesh is not fes.mesh.ngmesh or mesh.ngmesh is not fes.mesh.ngmesh:
            raise ValueError("Integrate: the grid functions must live on the mesh that is integrated over")
        mf = DevMatrixFree(fes, max(2 * fes.order - 2, 1), None, "f64")
        g = N.DeviceVector(mf.op.ne, 8)
        mf.op.element_grad(cf.a.vec._device(), cf.b.vec._device(), g)
        vals = -g.download()  # element_grad gives -(a_e^T K_e b_e) / kappa_e with kappa = 1
        if cf.coef is not None:
            vals = vals * (np.asarray(cf.coef, dtype=np.float64) * np.ones(mesh.ne))
    else:
        ir = N.TrigRule(order) if mesh.dim == 2 else N.TetRule(order)
        X, W = N.QuadraturePoints(mesh.ngmesh, ir)
        fv = np.broadcast_to(np.asarray(_cf(cf).Evaluate(X), dtype=np.float64), W.shape)
        vals = (np.abs(W) * fv).sum(axis=1)
    return vals if element_wise else float(vals.sum())


class _FormTerm:
    """Product of a coefficient and trial/test proxies."""

    def __init__(self, coef, proxies):
        self.coef, self.proxies = coef, proxies

    def __mul__(self, o):
        if isinstance(o, _Measure):
            return _Integral([(self, o)])
        if isinstance(o, _Proxy):
            return _FormTerm(self.coef, self.proxies + [o])
        if isinstance(o, _FormTerm):
            c = o.coef if self.coef is None else (self.coef if o.coef is None else self.coef * o.coef)
            return _FormTerm(c, self.proxies + o.proxies)
        c = _cf(o)
        return _FormTerm(c if self.coef is None else self.coef * c, self.proxies)

    __rmul__ = __mul__

    @property
    def bilinear(self):
        return len(self.proxies) == 2

    @property
    def kind(self):
        if len(self.proxies) == 1:
            if self.proxies[0].trial or self.proxies[0].deriv:
                raise ValueError("linear forms support f*v only")
            return "source"
        if len(self.proxies) != 2 or self.proxies[0].trial == self.proxies[1].trial:
            raise ValueError("need one trial and one test function")
        if self.proxies[0].deriv != self.proxies[1].deriv:
            raise ValueError("only grad(u)*grad(v) and u*v are supported")
        return "laplace" if self.proxies[0].deriv else "mass"

    @property
    def fes(self):
        return self.proxies[0].fes


class _Measure:
    def __init__(self, bonus_intorder=0, intorder=None):
        self.bonus_intorder, self.intorder = bonus_intorder, intorder

    def __call__(self, bonus_intorder=0, intorder=None):
        return _Measure(bonus_intorder, intorder)

    def __rmul__(self, term):
        if isinstance(term, _FormTerm):
            return _Integral([(term, self)])
        return NotImplemented


dx = _Measure()


class _Integral:
    def __init__(self, parts):
        self.parts = parts

    def __add__(self, o):
        return _Integral(self.parts + o.parts)


# ------------------------------------------------------------------- space

def smoothing_blocks(m, sp, blocktype=4):
    """Smoothing blocks of the free dofs of the H1 space sp on the mesh m
    (native objects) for BlockJacobiPrecond, as
    H1HighOrderFESpace::CreateSmoothingBlocks (comp/h1hofespace.cpp:1478-1620)
    builds them (non-free dofs filtered out, so some blocks are empty), as a
    BlockTable (flat pointer and dof arrays, iterable as a list of blocks):
      1  2D: V + E + I   one block per vertex, edge and element interior
      2  2D: VE + I      per vertex its dof and the dofs of its edges; interiors
      3  V + E + F + I   one block per node
      4  VE + F + I      (the default) vertex-edge patches, faces, cells
    """
    dim = m.dim
    if blocktype not in (1, 2, 3, 4) or (dim == 3 and blocktype in (1, 2)):
        raise ValueError("block types 1 and 2 (2D), 3 and 4 are provided")
    nv, ned = m.nv, m.nedges
    fe = np.asarray(sp.first_edge_dof, dtype=np.int64)
    fi = np.asarray(sp.first_el_dof, dtype=np.int64)
    ff = np.asarray(sp.first_face_dof, dtype=np.int64) if dim == 3 else np.zeros(1, dtype=np.int64)
    nfa = len(ff) - 1
    ni = len(fi) - 1

    def ranges(first, base):
        """(block, dof) pairs for blocks base + i holding [first[i], first[i+1])."""
        cnt = np.diff(first)
        blk = np.repeat(np.arange(len(cnt), dtype=np.int64) + base, cnt)
        return blk, np.arange(first[0], first[-1], dtype=np.int64)

    parts = []
    vb, vd = np.arange(nv, dtype=np.int64), np.arange(nv, dtype=np.int64)
    if blocktype in (2, 4):
        ev = np.asarray(m.edges, dtype=np.int64).reshape(-1, 2)
        eb, ed = ranges(fe, 0)  # edge number, dof
        parts += [(vb, vd), (ev[eb, 0], ed), (ev[eb, 1], ed)]
        nblk_ve = nv if blocktype == 4 else nv + ned
    else:
        eb, ed = ranges(fe, nv)
        parts += [(vb, vd), (eb, ed)]
        nblk_ve = nv + ned
    if dim == 3:
        fb, fd = ranges(ff, nblk_ve)
        parts.append((fb, fd))
    ib, idof = ranges(fi, nblk_ve + (nfa if dim == 3 else 0))
    parts.append((ib, idof))
    nblocks = nblk_ve + (nfa if dim == 3 else 0) + ni
    blk = np.concatenate([p[0] for p in parts])
    dof = np.concatenate([p[1] for p in parts])
    keep = sp.freedofs.astype(bool)[dof]
    blk, dof = blk[keep], dof[keep]
    order = np.argsort(blk, kind="stable")
    blk, dof = blk[order], dof[order].astype(np.int32)
    ptr = np.searchsorted(blk, np.arange(nblocks + 1))
    return BlockTable(ptr, dof)


class _FreeDofs(BitArray):
    """A BitArray that can also be called (fes.FreeDofs(coupling=False))."""

    def __call__(self, coupling=False):
        return self


class H1:
    def __init__(self, mesh, order=1, dirichlet=None, **kwargs):
        if not isinstance(mesh, Mesh):
            mesh = Mesh(mesh)
        self.mesh = mesh
        self.order = int(order)
        mask = 0
        if dirichlet:
            names = mesh.ngmesh.boundary_names
            idx = sorted(set(mesh.ngmesh.boundary_index.tolist()))
            for r in idx:
                name = names[r - 1] if 0 < r <= len(names) else "bc%d" % r
                if re.fullmatch(dirichlet, name):
                    mask |= 1 << (r - 1)
        self.dirichlet_mask = mask
        self.space = N.H1Space(mesh.ngmesh, self.order, mask)

    @property
    def ndof(self):
        return self.space.ndof

    @property
    def FreeDofs(self):
        """The free dofs; fes.FreeDofs() as in NGSolve (FESpace::GetFreeDofs,
        comp/fespace.cpp:765-810) and fes.FreeDofs alike."""
        return _FreeDofs(self.space.freedofs.astype(bool))

    def GetDofNrs(self, el):
        return tuple(int(d) for d in self.space.el_dofs[el])

    def CreateSmoothingBlocks(self, blocktype=4):
        """Blocks of free dofs for BlockJacobiPrecond (smoothing_blocks)."""
        return smoothing_blocks(self.mesh.ngmesh, self.space, blocktype)

    def TrialFunction(self):
        return _Proxy(self, True)

    def TestFunction(self):
        return _Proxy(self, False)

    def TnT(self):
        return self.TrialFunction(), self.TestFunction()

    def integration_rule(self, order):
        return N.TrigRule(order) if self.mesh.dim == 2 else N.TetRule(order)


# ------------------------------------------------------------- operators

class MFOpts:
    """MatFreeOptions (comp/bilinearform.hpp:24-52)."""

    def __init__(self, fused=True, atomic=False, fp32=False, precision=None, coloring=False, **kwargs):
        """atomic=True: hardware atomics (the reference's default);
        coloring=True: the colour-ordered plain += of the reference's
        non-atomic path (one launch per element colour); neither: the
        deterministic two-pass gather (the fastest here)."""
        self.fused = fused
        self.atomic = atomic
        self.coloring = coloring
        self.precision = precision or ("f32" if fp32 else "f64")


class DevMatrixFree(DeviceMatrix):
    """Fused element-batched operator on the device (MatrixFreeBTDTB)."""

    def __init__(self, fes, order, kappa=None, precision="f64", atomic=False, coloring=False):
        ensure_device()
        mesh = fes.mesh.ngmesh
        ir = fes.integration_rule(order)
        classes = N.BuildClasses(fes.space, ir)
        detj, ginv = N.ComputeGeometry(mesh)
        dim = mesh.dim
        el = np.concatenate([c["elements"] for c in classes]).astype(np.int32)
        offsets = np.cumsum([0] + [len(c["elements"]) for c in classes]).astype(np.int64)
        bmat = np.concatenate([c["bmat"].ravel() for c in classes])
        dofs = fes.space.el_dofs[el]
        G = detj[el, None, None] * ginv[el]
        if dim == 2:
            metric = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 0, 1]], axis=1)
        else:
            metric = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 2, 2], G[:, 0, 1], G[:, 0, 2], G[:, 1, 2]], axis=1)
        kap = None if kappa is None else np.ascontiguousarray(np.asarray(kappa, dtype=np.float64)[el])
        op = N.CreateMatrixFree(dim, fes.space.nd_el, len(ir), fes.ndof, offsets, np.ascontiguousarray(bmat),
                                np.ascontiguousarray(ir.weights), np.ascontiguousarray(dofs, dtype=np.int32),
                                np.ascontiguousarray(metric), kap, el, PREC[precision],
                                SCATTER["atomic" if atomic else "gather"])
        if coloring:
            # FESpace element colouring (comp/fespace.cpp:518-760), in table order
            op.set_coloring(np.ascontiguousarray(N.ElementColoring(mesh)[el], dtype=np.int32))
        super().__init__(op, True)
        self.fes, self.nq, self.nclass, self.precision = fes, len(ir), len(classes), precision

    def Diagonal(self):
        d = BaseVector(self.Height())
        self.op.diagonal(d._write_device())
        return d

    def CreateSmoother(self, freedofs=None):
        return JacobiPrecond(self.Diagonal(), freedofs)


class BilinearForm:
    def __init__(self, space, integrals=None, mf=None, matrix_free_bdb=False, device_assembly=False, **kwargs):
        """device_assembly=True builds the sparsity pattern and the CSR
        values on the device (ngsamd_mf_pattern, ngsamd_mf_assemble_csr /
        ngsamd_mf_assemble_mass_csr; one grad(u)*grad(v) or u*v term)."""
        if isinstance(space, _Integral):  # BilinearForm(grad(u)*grad(v)*dx)
            integrals, space = space, None
        self.integrals = []
        self.space = space
        self.device_assembly = device_assembly
        self.mf = mf if mf is not None else (MFOpts() if matrix_free_bdb else None)
        self.mat = None
        if integrals is not None:
            self += integrals

    def __iadd__(self, integral):
        for term, meas in integral.parts:
            if not term.bilinear:
                raise ValueError("linear term in a bilinear form")
            self.integrals.append((term, meas))
            if self.space is None:
                self.space = _space_of(term)
        return self

    def _order(self, term, meas):
        p = self.space.order
        if meas.intorder is not None:
            return meas.intorder
        base = 2 * p - 2 if term.kind == "laplace" else 2 * p
        return base + meas.bonus_intorder

    def _coef(self, term):
        if term.coef is None:
            return None
        if term.coef.elementwise is None:
            raise ValueError("bilinear-form coefficients must be element-wise constants")
        return term.coef.elementwise

    def Assemble(self):
        fes = self.space
        if self.mf is not None:
            if len(self.integrals) != 1 or self.integrals[0][0].kind != "laplace":
                raise ValueError("the matrix-free operator supports one grad(u)*grad(v) term")
            term, meas = self.integrals[0]
            self.mat = DevMatrixFree(fes, self._order(term, meas), self._coef(term), self.mf.precision,
                                     self.mf.atomic, getattr(self.mf, "coloring", False))
            return self
        if self.device_assembly:
            if len(self.integrals) != 1:
                raise ValueError("device assembly supports one grad(u)*grad(v) or u*v term")
            term, meas = self.integrals[0]
            order = self._order(term, meas)
            if term.kind == "laplace":
                mf = DevMatrixFree(fes, order, self._coef(term), "f64")
                rowptr, cols = N.PatternDevice(mf.op)
                op = N.AssembleCSRDevice(mf.op, rowptr, cols)
            else:
                # the mesh and dof tables of a matrix-free operator, the mass
                # form's own rule and coefficient
                mf = DevMatrixFree(fes, max(2 * fes.order - 2, 1), None, "f64")
                rowptr, cols = N.PatternDevice(mf.op)
                ir = fes.integration_rule(order)
                shape = np.ascontiguousarray(np.stack([c["shape"] for c in N.BuildClasses(fes.space, ir)]))
                detj, _ = N.ComputeGeometry(fes.mesh.ngmesh)
                rho = self._coef(term)
                scale = np.abs(detj) if rho is None else np.asarray(rho, dtype=np.float64) * np.abs(detj)
                op = N.AssembleMassCSRDevice(mf.op, np.ascontiguousarray(ir.weights), shape,
                                             np.ascontiguousarray(scale), rowptr, cols)
            self.mat = DevSparseMatrix(fes.ndof, fes.ndof, rowptr, cols, None, symmetric=True, op=op)
            return self
        rowptr, cols = N.BuildPattern(fes.space)
        vals = np.zeros(len(cols))
        for term, meas in self.integrals:
            ir = fes.integration_rule(self._order(term, meas))
            lap = term.kind == "laplace"
            c = self._coef(term)
            vals += N.AssembleCSR(fes.space, rowptr, cols, ir, lap, c if lap else None, not lap,
                                  None if lap else c)
        self.mat = DevSparseMatrix(fes.ndof, fes.ndof, rowptr, cols, vals, symmetric=True)
        return self


def _space_of(term):
    return term.fes


class LinearForm:
    def __init__(self, space, integrals=None):
        if isinstance(space, _Integral):
            integrals, space = space, None
        self.space = space
        self.integrals = []
        self.vec = None
        if integrals is not None:
            self += integrals

    def __iadd__(self, integral):
        for term, meas in integral.parts:
            if term.bilinear:
                raise ValueError("bilinear term in a linear form")
            self.integrals.append((term, meas))
            if self.space is None:
                self.space = _space_of(term)
        return self

    def Assemble(self):
        fes = self.space
        f = np.zeros(fes.ndof)
        for term, meas in self.integrals:
            order = meas.intorder if meas.intorder is not None else 2 * fes.order + meas.bonus_intorder
            ir = fes.integration_rule(order)
            X, _ = N.QuadraturePoints(fes.mesh.ngmesh, ir)
            coef = term.coef if term.coef is not None else _cf(1.0)
            fv = np.ascontiguousarray(coef.Evaluate(X), dtype=np.float64)
            f += N.AssembleLinearForm(fes.space, ir, fv)
        self.vec = BaseVector.from_numpy(f)
        return self


class GridFunction:
    def __init__(self, space):
        self.space = space
        self.vec = BaseVector(space.ndof)

    def L2Error(self, exact, order=None):
        fes = self.space
        ir = fes.integration_rule(order or 2 * fes.order + 4)
        X, W = N.QuadraturePoints(fes.mesh.ngmesh, ir)
        uh = N.EvaluateAtPoints(fes.space, ir, np.ascontiguousarray(self.vec.NumPy()))
        e = uh - _cf(exact).Evaluate(X)
        return float(np.sqrt(np.sum(W * e * e)))
