"""Shared builders for the GPU parity tests."""
import sys

import numpy as np

from ngsolve_amd import _ngsamd as N
from oracle import fe_oracle as O

ALL = 0xFFFFFFFFFFFFFFFF


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(dim, n, p, order=None, seed_kappa=None):
    """Product host setup + oracle on the same structured mesh."""
    m = N.MakeStructured2DMesh(n, n) if dim == 2 else N.MakeStructured3DMesh(n, n, n)
    om = O.structured_2d(n) if dim == 2 else O.structured_3d(n)
    order = 2 * p - 2 if order is None else order
    fes = N.H1Space(m, p, ALL)
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    kappa = None
    if seed_kappa is not None:
        kappa = np.exp(np.random.default_rng(seed_kappa).standard_normal(m.ne))
    rp, cols = N.BuildPattern(fes)
    vals = N.AssembleCSR(fes, rp, cols, ir, True, kappa, False, None)
    return dict(mesh=m, omesh=om, fes=fes, ir=ir, order=order, kappa=kappa, rowptr=rp, cols=cols, vals=vals,
                ndof=fes.ndof, dim=dim, p=p)


def mf_operator(P, precision=0, scatter=0):
    """Matrix-free operator built from the product's host setup."""
    fes, ir, dim = P["fes"], P["ir"], P["dim"]
    classes = N.BuildClasses(fes, ir)
    detj, ginv = N.ComputeGeometry(P["mesh"])
    el = np.concatenate([c["elements"] for c in classes]).astype(np.int32)
    offsets = np.cumsum([0] + [len(c["elements"]) for c in classes]).astype(np.int64)
    bmat = np.ascontiguousarray(np.concatenate([c["bmat"].ravel() for c in classes]))
    G = detj[el, None, None] * ginv[el]
    if dim == 2:
        metric = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 0, 1]], axis=1)
    else:
        metric = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 2, 2], G[:, 0, 1], G[:, 0, 2], G[:, 1, 2]], axis=1)
    kap = None if P["kappa"] is None else np.ascontiguousarray(P["kappa"][el])
    return N.CreateMatrixFree(dim, fes.nd_el, len(ir), fes.ndof, offsets, bmat, np.ascontiguousarray(ir.weights),
                              np.ascontiguousarray(fes.el_dofs[el], dtype=np.int32), np.ascontiguousarray(metric),
                              kap, el, precision, scatter)


def dvec(a, itemsize=8):
    v = N.DeviceVector(len(a), itemsize)
    v.upload(np.ascontiguousarray(a, dtype=np.float64 if itemsize == 8 else np.float32))
    return v


def relerr_inf(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))
