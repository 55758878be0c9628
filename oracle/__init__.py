"""CPU oracle of the H1 Laplace hot path -- test infrastructure only.

fe_oracle.py : numpy restatement of the FE setup (meshes, dofs, shapes,
               quadrature, element matrices, CSR pattern/assembly).
oracle.c     : C kernels (CSR SpMV, matrix-free BTDB apply, Jacobi-PCG) used
               as the checker and as bench.py's CPU baseline.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        I64 = ctypes.c_int64
        D = ctypes.c_double
        I = ctypes.c_int
        L.oracle_csr_spmv.argtypes = [I64, P, P, P, P, P, D, I]
        L.oracle_mf_apply.argtypes = [I, I, I, I64, P, P, P, P, P, P, P, P, P, D]
        L.oracle_pcg_csr.argtypes = [I64, P, P, P, P, P, P, D, D, I, P]
        L.oracle_pcg_csr.restype = I
        L.oracle_num_threads.restype = I
        L.oracle_set_threads.argtypes = [I]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def num_threads():
    return lib().oracle_num_threads()


def set_threads(n):
    lib().oracle_set_threads(int(n))


def spmv(rowptr, cols, vals, x, y=None, s=1.0, accumulate=False):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(rowptr) - 1
    if y is None:
        y = np.zeros(n)
    lib().oracle_csr_spmv(n, _p(rowptr), _p(cols), _p(vals), _p(x), _p(y), float(s), int(accumulate))
    return y


def mf_apply(dim, bref, el_class, weights, dofs, finv, detf, kappa, x, y=None, s=1.0):
    """bref (nclass, nq, nd, dim); el_class (nel,); dofs (nel, nd); finv (nel, dim, dim)."""
    bref = np.ascontiguousarray(bref, dtype=np.float64)
    nclass, nq, nd, _ = bref.shape
    el_class = np.ascontiguousarray(el_class, dtype=np.int32)
    weights = np.ascontiguousarray(weights, dtype=np.float64)
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    finv = np.ascontiguousarray(finv, dtype=np.float64)
    detf = np.ascontiguousarray(np.abs(detf), dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    kap = None if kappa is None else np.ascontiguousarray(kappa, dtype=np.float64)
    if y is None:
        y = np.zeros(len(x))
    lib().oracle_mf_apply(dim, nd, nq, len(dofs), _p(bref), _p(el_class), _p(weights), _p(dofs), _p(finv),
                          _p(detf), None if kap is None else _p(kap), _p(x), _p(y), float(s))
    return y


def pcg_csr(rowptr, cols, vals, invdiag, f, tol=1e-12, atol=None, maxiter=100):
    """Jacobi-PCG with the reference CGSolver recurrence; returns (u, iterations, residuals)."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    invdiag = np.ascontiguousarray(invdiag, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    u = np.zeros(len(f))
    res = np.zeros(max(maxiter, 1))
    its = lib().oracle_pcg_csr(len(f), _p(rowptr), _p(cols), _p(vals), _p(invdiag), _p(f), _p(u),
                               -1.0 if tol is None else float(tol), -1.0 if atol is None else float(atol),
                               int(maxiter), _p(res))
    return u, its, res[:min(its, maxiter)].copy()
