"""Linear-algebra layer with the reference's Python surface.

BaseVector   ngla::BaseVector / UnifiedVector (linalg/basevector.hpp,
             ngscuda/unifiedvector.cpp:288-330): device storage with a lazily
             synchronised host mirror.
BaseMatrix   ngla::BaseMatrix (linalg/basematrix.hpp:25-225,
             linalg/basematrix.cpp:120-154): Mult = SetZero + MultAdd,
             MultAdd = temporary + Mult; composites Sum/Product/Transpose/Scale
             (linalg/basematrix.hpp:413-860); `mat * vec` is a lazy expression
             evaluated on `vec.data = ...` (linalg/python_linalg.cpp:1165-1167).
"""
import numbers

import numpy as np

from ._native import ensure_device


# ---------------------------------------------------------------- vectors

class BaseVector:
    """FP64 vector in device memory with a host mirror.

    Device operations run on the device copy; element access and NumPy views
    go through the host copy.  Whichever side was written last is copied to
    the other side on demand.
    """

    def __init__(self, size, _dev=None):
        N = ensure_device()
        self._size = int(size)
        self._dev = _dev if _dev is not None else N.DeviceVector(self._size, 8)
        self._host = None
        self._host_valid = False
        self._dev_valid = _dev is not None

    # -- synchronisation
    def _device(self):
        if not self._dev_valid:
            if self._host_valid:
                self._dev.upload(self._host)
            else:
                self._dev.fill(0.0)
            self._dev_valid = True
        return self._dev

    def _write_device(self):
        d = self._device()
        self._host_valid = False
        return d

    def _hostarr(self):
        if not self._host_valid:
            if self._dev_valid:
                self._host = np.asarray(self._dev.download(), dtype=np.float64)
            else:
                self._host = np.zeros(self._size)
            self._host_valid = True
        return self._host

    def _write_host(self):
        h = self._hostarr()
        self._dev_valid = False
        return h

    # -- reference API
    def __len__(self):
        return self._size

    @property
    def size(self):
        return self._size

    def CreateVector(self, copy=False):
        v = BaseVector(self._size)
        if copy:
            v.data = self
        return v

    def Copy(self):
        return self.CreateVector(copy=True)

    def FV(self):
        return _FlatVector(self)

    def NumPy(self):
        return self._write_host()

    def __getitem__(self, i):
        return self._hostarr()[i]

    def __setitem__(self, i, v):
        if isinstance(i, slice) and i == slice(None) and isinstance(v, numbers.Number):
            self._write_device().fill(float(v))
            return
        h = self._write_host()
        h[i] = v

    def __iter__(self):
        return iter(self._hostarr().copy())

    def SetScalar(self, v):
        self._write_device().fill(float(v))

    def InnerProduct(self, other, conjugate=False):
        return self._device().dot(other._device())

    def Norm(self):
        return float(np.sqrt(max(self.InnerProduct(self), 0.0)))

    def __imul__(self, s):
        self._write_device().scale(float(s))
        return self

    def __itruediv__(self, s):
        self._write_device().scale(1.0 / float(s))
        return self

    def __iadd__(self, other):
        _assign(self, _as_expr(other), accumulate=True)
        return self

    def __isub__(self, other):
        _assign(self, _scale_expr(-1.0, _as_expr(other)), accumulate=True)
        return self

    @property
    def data(self):
        return _DataProxy(self)

    @data.setter
    def data(self, expr):
        if isinstance(expr, _DataProxy):
            return  # result of `v.data += ...`, already applied
        _assign(self, _as_expr(expr), accumulate=False)

    # -- expression algebra
    def __add__(self, o):
        return _as_expr(self) + o

    def __radd__(self, o):
        return _as_expr(o) + self

    def __sub__(self, o):
        return _as_expr(self) - o

    def __rsub__(self, o):
        return _as_expr(o) - self

    def __neg__(self):
        return _scale_expr(-1.0, _as_expr(self))

    def __mul__(self, s):
        if isinstance(s, numbers.Number):
            return _scale_expr(float(s), _as_expr(self))
        return NotImplemented

    __rmul__ = __mul__

    # -- interop
    @property
    def device_ptr(self):
        return self._device().ptr

    @staticmethod
    def from_numpy(a):
        v = BaseVector(len(a))
        v._host = np.array(a, dtype=np.float64, copy=True)
        v._host_valid = True
        v._dev_valid = False
        return v

    def __repr__(self):
        return f"BaseVector(size={self._size})"


class _FlatVector:
    def __init__(self, v):
        self._v = v

    def NumPy(self):
        return self._v.NumPy()

    def __len__(self):
        return len(self._v)

    def __getitem__(self, i):
        return self._v[i]

    def __setitem__(self, i, x):
        self._v[i] = x


class _DataProxy:
    """Return value of `vec.data`: supports `vec.data += expr`, `-=`."""

    def __init__(self, v):
        self.v = v

    def __iadd__(self, expr):
        _assign(self.v, _as_expr(expr), accumulate=True)
        return self

    def __isub__(self, expr):
        _assign(self.v, _scale_expr(-1.0, _as_expr(expr)), accumulate=True)
        return self


def CreateVVector(size):
    return BaseVector(size)


def InnerProduct(x, y, conjugate=False):
    return x.InnerProduct(y, conjugate=conjugate)


def Norm(x):
    return x.Norm()


# ------------------------------------------------------------ expressions

class _Expr:
    """Linear combination sum_k c_k * term_k, term = vector or (matrix, vector)."""

    def __init__(self, terms):
        self.terms = terms  # list of (coef, BaseVector) or (coef, (BaseMatrix, BaseVector))

    def __add__(self, o):
        return _Expr(self.terms + _as_expr(o).terms)

    __radd__ = __add__

    def __sub__(self, o):
        return _Expr(self.terms + _scale_expr(-1.0, _as_expr(o)).terms)

    def __rsub__(self, o):
        return _as_expr(o) - self

    def __neg__(self):
        return _scale_expr(-1.0, self)

    def __mul__(self, s):
        if isinstance(s, numbers.Number):
            return _scale_expr(float(s), self)
        return NotImplemented

    __rmul__ = __mul__


def _as_expr(x):
    if isinstance(x, _Expr):
        return x
    if isinstance(x, BaseVector):
        return _Expr([(1.0, x)])
    if isinstance(x, _DataProxy):
        return _Expr([(1.0, x.v)])
    raise TypeError(f"cannot use {type(x).__name__} in a vector expression")


def _scale_expr(s, e):
    return _Expr([(s * c, t) for c, t in e.terms])


def _assign(y, expr, accumulate):
    terms = expr.terms
    # a matrix term reading y must not overwrite it while computing
    aliased = any(isinstance(t, tuple) and t[1] is y for _, t in terms) and not accumulate
    if not accumulate and len(terms) > 1:
        aliased = aliased or any(isinstance(t, tuple) and t[1] is y for _, t in terms[1:])
    target = y.CreateVector() if aliased else y
    first = not accumulate
    for c, t in terms:
        if isinstance(t, tuple):
            M, x = t
            if first:
                M.Mult(x, target)
                if c != 1.0:
                    target *= c
            else:
                M.MultAdd(c, x, target)
        else:
            if t is target and first:
                if c != 1.0:
                    target *= c
            elif first:
                target._write_device().copy_from(t._device())
                if c != 1.0:
                    target *= c
            else:
                target._write_device().axpy(c, t._device())
        first = False
    if target is not y:
        y._write_device().copy_from(target._device())


# --------------------------------------------------------------- matrices

class BaseMatrix:
    """Operator base class with the reference's default semantics.

    Subclasses override Height/Width and Mult and/or MultAdd (and the
    transposed variants); everything else derives from those.
    """

    def __init__(self):
        pass

    # -- shape
    def Height(self):
        raise NotImplementedError

    def Width(self):
        raise NotImplementedError

    @property
    def height(self):
        return self.Height()

    @property
    def width(self):
        return self.Width()

    @property
    def shape(self):
        return (self.height, self.width)

    def IsComplex(self):
        return False

    def IsSymmetric(self):
        return False

    def CreateRowVector(self):
        return BaseVector(self.Width())

    def CreateColVector(self):
        return BaseVector(self.Height())

    def CreateVector(self, colvector=False):
        return self.CreateColVector() if colvector else self.CreateRowVector()

    # -- products (linalg/basematrix.cpp:120-154)
    def Mult(self, x, y):
        if type(self).MultAdd is BaseMatrix.MultAdd:
            raise NotImplementedError(f"{type(self).__name__} overrides neither Mult nor MultAdd")
        y[:] = 0.0
        self.MultAdd(1.0, x, y)

    def MultAdd(self, s, x, y):
        if type(self).Mult is BaseMatrix.Mult:
            raise NotImplementedError(f"{type(self).__name__} overrides neither Mult nor MultAdd")
        tmp = self.CreateColVector()
        self.Mult(x, tmp)
        y.data += s * tmp

    def MultTrans(self, x, y):
        if self.IsSymmetric():
            return self.Mult(x, y)
        if type(self).MultTransAdd is BaseMatrix.MultTransAdd:
            raise NotImplementedError(f"{type(self).__name__} has no transposed product")
        y[:] = 0.0
        self.MultTransAdd(1.0, x, y)

    def MultTransAdd(self, s, x, y):
        if self.IsSymmetric():
            return self.MultAdd(s, x, y)
        if type(self).MultTrans is BaseMatrix.MultTrans:
            raise NotImplementedError(f"{type(self).__name__} has no transposed product")
        tmp = self.CreateRowVector()
        self.MultTrans(x, tmp)
        y.data += s * tmp

    # -- algebra
    @property
    def T(self):
        return TransposeMatrix(self)

    def __mul__(self, o):
        if isinstance(o, BaseVector):
            return _Expr([(1.0, (self, o))])
        if isinstance(o, _Expr):
            # matrix times expression: materialise the expression first
            tmp = self.CreateRowVector()
            tmp.data = o
            return _Expr([(1.0, (self, tmp))])
        if isinstance(o, numbers.Number):
            return ScaleMatrix(float(o), self)
        return NotImplemented

    def __rmul__(self, s):
        if isinstance(s, numbers.Number):
            return ScaleMatrix(float(s), self)
        return NotImplemented

    def __matmul__(self, o):
        if isinstance(o, BaseMatrix):
            return ProductMatrix(self, o)
        if isinstance(o, BaseVector):
            return self * o
        return NotImplemented

    def __add__(self, o):
        return SumMatrix(self, o)

    def __sub__(self, o):
        return SumMatrix(self, ScaleMatrix(-1.0, o))

    def __neg__(self):
        return ScaleMatrix(-1.0, self)

    def CreateDeviceMatrix(self):
        return self


class TransposeMatrix(BaseMatrix):
    def __init__(self, m):
        super().__init__()
        self.m = m

    def Height(self):
        return self.m.Width()

    def Width(self):
        return self.m.Height()

    def Mult(self, x, y):
        self.m.MultTrans(x, y)

    def MultAdd(self, s, x, y):
        self.m.MultTransAdd(s, x, y)

    def MultTrans(self, x, y):
        self.m.Mult(x, y)

    def MultTransAdd(self, s, x, y):
        self.m.MultAdd(s, x, y)


class ScaleMatrix(BaseMatrix):
    def __init__(self, s, m):
        super().__init__()
        self.s, self.m = s, m

    def Height(self):
        return self.m.Height()

    def Width(self):
        return self.m.Width()

    def Mult(self, x, y):
        self.m.Mult(x, y)
        y *= self.s

    def MultAdd(self, s, x, y):
        self.m.MultAdd(s * self.s, x, y)

    def MultTrans(self, x, y):
        self.m.MultTrans(x, y)
        y *= self.s

    def MultTransAdd(self, s, x, y):
        self.m.MultTransAdd(s * self.s, x, y)


class SumMatrix(BaseMatrix):
    def __init__(self, a, b):
        super().__init__()
        self.a, self.b = a, b

    def Height(self):
        return self.a.Height()

    def Width(self):
        return self.a.Width()

    def Mult(self, x, y):
        self.a.Mult(x, y)
        self.b.MultAdd(1.0, x, y)

    def MultAdd(self, s, x, y):
        self.a.MultAdd(s, x, y)
        self.b.MultAdd(s, x, y)

    def MultTrans(self, x, y):
        self.a.MultTrans(x, y)
        self.b.MultTransAdd(1.0, x, y)

    def MultTransAdd(self, s, x, y):
        self.a.MultTransAdd(s, x, y)
        self.b.MultTransAdd(s, x, y)


class ProductMatrix(BaseMatrix):
    def __init__(self, a, b):
        super().__init__()
        self.a, self.b = a, b

    def Height(self):
        return self.a.Height()

    def Width(self):
        return self.b.Width()

    def Mult(self, x, y):
        tmp = self.b.CreateColVector()
        self.b.Mult(x, tmp)
        self.a.Mult(tmp, y)

    def MultAdd(self, s, x, y):
        tmp = self.b.CreateColVector()
        self.b.Mult(x, tmp)
        self.a.MultAdd(s, tmp, y)

    def MultTrans(self, x, y):
        tmp = self.a.CreateRowVector()
        self.a.MultTrans(x, tmp)
        self.b.MultTrans(tmp, y)

    def MultTransAdd(self, s, x, y):
        tmp = self.a.CreateRowVector()
        self.a.MultTrans(x, tmp)
        self.b.MultTransAdd(s, tmp, y)


class DeviceMatrix(BaseMatrix):
    """A BaseMatrix backed by a libngsamd operator handle."""

    def __init__(self, op, symmetric=True):
        super().__init__()
        self.op = op
        self._symmetric = symmetric

    def Height(self):
        return self.op.height

    def Width(self):
        return self.op.width

    def IsSymmetric(self):
        return self._symmetric

    def Mult(self, x, y):
        self.op.mult(x._device(), y._write_device())

    def MultAdd(self, s, x, y):
        self.op.mult_add(float(s), x._device(), y._write_device())

    def MultTrans(self, x, y):
        if not self._symmetric:
            raise NotImplementedError("transposed product of a non-symmetric device operator")
        self.Mult(x, y)

    def MultTransAdd(self, s, x, y):
        if not self._symmetric:
            raise NotImplementedError("transposed product of a non-symmetric device operator")
        self.MultAdd(s, x, y)


class DevSparseMatrix(DeviceMatrix):
    """CSR matrix on the device (SparseMatrix<double> + DevSparseMatrix)."""

    def __init__(self, height, width, rowptr, cols, vals, symmetric=True, op=None):
        """From host CSR arrays, or (vals None, op given) around a CSR operator
        whose values were assembled on the device."""
        N = ensure_device()
        if op is None:
            op = N.CreateCSR(int(height), int(width), np.ascontiguousarray(rowptr, dtype=np.int64),
                             np.ascontiguousarray(cols, dtype=np.int32), np.ascontiguousarray(vals, dtype=np.float64))
        super().__init__(op, symmetric)
        self.rowptr, self.cols, self._vals = rowptr, cols, vals

    @property
    def vals(self):
        """The values in CSR order (copied from the device when they were
        assembled there)."""
        if self._vals is None:
            self._vals = self.op.csr_values()
        return self._vals

    @property
    def nze(self):
        return int(self.rowptr[-1])

    def __getitem__(self, ij):
        i, j = ij
        s, e = int(self.rowptr[i]), int(self.rowptr[i + 1])
        k = np.searchsorted(self.cols[s:e], j)
        if k < e - s and self.cols[s + k] == j:
            return float(self.vals[s + k])
        return 0.0

    def Diagonal(self):
        d = BaseVector(self.Height())
        self.op.diagonal(d._write_device())
        return d

    def CreateSmoother(self, freedofs=None):
        return JacobiPrecond(self.Diagonal(), freedofs)

    def CreateBlockSmoother(self, blocks):
        """BaseSparseMatrix::CreateBlockJacobiPrecond (Python CreateBlockSmoother)."""
        return BlockJacobiPrecond(self, blocks)

    def Inverse(self, freedofs=None, inverse="sparsecholesky"):
        """The inverse of the free block (a.mat.Inverse(freedofs,
        inverse="sparsecholesky"), linalg/sparsecholesky.cpp): factorised on
        the host, solved on the device."""
        if inverse not in ("sparsecholesky", None, ""):
            raise ValueError(f"inverse type {inverse!r}: only 'sparsecholesky' is available on the device")
        return SparseCholesky(self, freedofs)


class JacobiPrecond(DeviceMatrix):
    """JacobiPrecond (linalg/jacobi.cpp:41-90): y += s * diag^{-1} x on free dofs."""

    def __init__(self, diag, freedofs=None):
        N = ensure_device()
        fr = None if freedofs is None else np.ascontiguousarray(_bits(freedofs), dtype=np.uint8)
        if isinstance(diag, BaseVector):
            op = N.CreateJacobiDevice(diag._device(), fr)
        else:
            op = N.CreateJacobi(np.ascontiguousarray(diag, dtype=np.float64), fr)
        super().__init__(op, True)


class BlockTable:
    """Blocks of dof numbers as two flat arrays (block b holds
    dofs[ptr[b]:ptr[b+1]]), the form NGSolve's Table<int> has; indexing and
    iteration give the blocks as arrays."""

    def __init__(self, ptr, dofs):
        self.ptr = np.ascontiguousarray(ptr, dtype=np.int64)
        self.dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        if self.ptr.ndim != 1 or len(self.ptr) < 1 or self.ptr[0] != 0 or self.ptr[-1] != len(self.dofs) \
                or np.any(np.diff(self.ptr) < 0):
            raise ValueError("BlockTable: ptr must rise from 0 to len(dofs)")

    def __len__(self):
        return len(self.ptr) - 1

    def __getitem__(self, b):
        if b < 0:
            b += len(self)
        if not 0 <= b < len(self):
            raise IndexError("block index out of range")
        return self.dofs[self.ptr[b]:self.ptr[b + 1]]

    def __iter__(self):
        return (self.dofs[self.ptr[b]:self.ptr[b + 1]] for b in range(len(self)))


class BlockJacobiPrecond(DeviceMatrix):
    """BlockJacobiPrecond (linalg/blockjacobi.cpp) on the device:
    y += s * sum_b P_b^T inv(A_bb) P_b x for the blocks (lists of dof
    numbers; sorted, may overlap).  The blocks are inverted on the host, the
    apply runs on the device (DevBlockJacobiMatrix,
    ngscuda/dev_blockjacobi.cpp)."""

    def __init__(self, mat, blocks):
        N = ensure_device()
        if not isinstance(blocks, BlockTable):
            blocks = [np.asarray(b, dtype=np.int32).ravel() for b in blocks]
            bptr = np.zeros(len(blocks) + 1, dtype=np.int64)
            bptr[1:] = np.cumsum([len(b) for b in blocks])
            blocks = BlockTable(bptr, np.concatenate(blocks) if len(blocks) else np.zeros(0, dtype=np.int32))
        bptr, bd = blocks.ptr, blocks.dofs
        op = N.CreateBlockJacobi(int(mat.Height()), np.ascontiguousarray(mat.rowptr, dtype=np.int64),
                                 np.ascontiguousarray(mat.cols, dtype=np.int32),
                                 np.ascontiguousarray(mat.vals, dtype=np.float64), bptr, bd)
        super().__init__(op, True)
        self.blocks = blocks


class SparseCholesky(DeviceMatrix):
    """Sparse Cholesky inverse of the free block of a symmetric positive
    definite device matrix (SparseCholesky, linalg/sparsecholesky.cpp; device
    solve DevSparseCholesky, ngscuda/dev_sparsecholesky.cpp:403-454):
    y = A_ff^-1 x on the free dofs, 0 elsewhere."""

    def __init__(self, mat, freedofs=None):
        N = ensure_device()
        fr = None if freedofs is None else np.ascontiguousarray(_bits(freedofs), dtype=np.uint8)
        op = N.CreateSparseCholesky(int(mat.Height()), np.ascontiguousarray(mat.rowptr, dtype=np.int64),
                                    np.ascontiguousarray(mat.cols, dtype=np.int32),
                                    np.ascontiguousarray(mat.vals, dtype=np.float64), fr)
        super().__init__(op, True)
        self.info = N.SparseCholeskyInfo(op)


class MultiGrid(DeviceMatrix):
    """Multigrid V-cycle preconditioner for the P1 matrix of a structured
    mesh (MultigridPreconditioner, multigrid/mgpre.cpp:205-318, with
    LinearProlongation, multigrid/prolongation.cpp:77-144): y ~= A_ff^-1 x.
    mat: the P1 matrix over the mesh's vertices (for a hierarchical H1 space
    LowOrderBlock(a.mat, mesh.nv)); dims: the cells per direction the mesh
    was built with (MakeStructured2DMesh(nx, ny) -> (nx, ny),
    MakeStructured3DMesh(nx, ny, nz) -> (nx, ny, nz)).  The hierarchy halves
    the grid while every size is even and the level has more than 1024 free
    vertices; coarse levels are Galerkin products; smoothing_steps damped
    Jacobi sweeps (weight omega) before and after each coarse correction.
    DevCGSolver accepts it, and TwoLevelPrecond as its coarse part."""

    def __init__(self, mat, dims, freedofs=None, smoothing_steps=2, omega=0.85, max_levels=32):
        N = ensure_device()
        fr = None if freedofs is None else np.ascontiguousarray(_bits(freedofs), dtype=np.uint8)
        dims = [int(d) for d in dims]
        op = N.CreateMultiGrid(len(dims), dims, int(mat.Height()), np.ascontiguousarray(mat.rowptr, dtype=np.int64),
                               np.ascontiguousarray(mat.cols, dtype=np.int32),
                               np.ascontiguousarray(mat.vals, dtype=np.float64), fr, int(smoothing_steps),
                               float(omega), int(max_levels))
        super().__init__(op, True)
        self.levels = N.MultiGridInfo(op)


class TwoLevelPrecond(DeviceMatrix):
    """Additive two-level preconditioner y = S x + E C E^T x: S a Jacobi or
    block-Jacobi preconditioner, C a SparseCholesky or a MultiGrid acting on
    the first C.Height() dofs (E their embedding: the vertex dofs of a
    hierarchical H1 space, i.e. its P1 part).  The device form of the
    composite smoother + E @ a_lo.mat.Inverse(freedofs_lo) @ E.T (SumMatrix /
    ProductMatrix, linalg/basematrix.hpp:413,592), or with a multigrid cycle
    for the inverse; DevCGSolver accepts it."""

    def __init__(self, smoother, coarse):
        N = ensure_device()
        if not isinstance(smoother, (JacobiPrecond, BlockJacobiPrecond)) or \
                not isinstance(coarse, (SparseCholesky, MultiGrid)):
            raise TypeError("TwoLevelPrecond needs a Jacobi or block-Jacobi smoother and a SparseCholesky or "
                            "MultiGrid")
        super().__init__(N.CreateTwoLevel(smoother.op, coarse.op), True)
        self.smoother, self.coarse = smoother, coarse


def LowOrderBlock(mat, nlow):
    """The leading nlow x nlow block of a device CSR matrix (for a
    hierarchical H1 space with nlow = number of vertices: the P1 matrix of the
    same form), as a DevSparseMatrix."""
    rp, cols, vals = np.asarray(mat.rowptr), np.asarray(mat.cols), np.asarray(mat.vals)
    keep = cols[:rp[nlow]] < nlow
    rows = np.repeat(np.arange(nlow), np.diff(rp[:nlow + 1]))[keep]
    brp = np.zeros(nlow + 1, dtype=np.int64)
    np.add.at(brp, rows + 1, 1)
    return DevSparseMatrix(nlow, nlow, np.cumsum(brp), cols[:rp[nlow]][keep], vals[:rp[nlow]][keep])


class Projector(BaseMatrix):
    """Projector onto the set (range=True) or complement of a bit mask."""

    def __init__(self, mask, range=True):
        super().__init__()
        m = _bits(mask).astype(bool)
        self.mask = m if range else ~m

    def Height(self):
        return len(self.mask)

    def Width(self):
        return len(self.mask)

    def IsSymmetric(self):
        return True

    def Mult(self, x, y):
        y.NumPy()[:] = np.where(self.mask, x.NumPy(), 0.0)

    def MultAdd(self, s, x, y):
        y.NumPy()[:] += s * np.where(self.mask, x.NumPy(), 0.0)


def _bits(mask):
    if hasattr(mask, "mask"):
        return np.asarray(mask.mask, dtype=np.uint8)
    return np.asarray(mask, dtype=np.uint8)


class BitArray:
    def __init__(self, mask):
        self.mask = np.asarray(mask, dtype=bool)

    def __len__(self):
        return len(self.mask)

    def __getitem__(self, i):
        return bool(self.mask[i])

    def NumSet(self):
        return int(self.mask.sum())

    def __invert__(self):
        return BitArray(~self.mask)
