"""Product host FE setup (C++ in _ngsamd) against the oracle: dof numbering,
free dofs, classes, rules and CSR pattern bit-exact; values to rounding.
CPU only (no device calls)."""
import gzip
import os

import numpy as np
import pytest

from ngsolve_amd import _ngsamd as N
from oracle import fe_oracle as O
from conftest import GOLDEN

ALL = 0xFFFFFFFFFFFFFFFF


def _meshes(dim, n):
    if dim == 2:
        return N.MakeStructured2DMesh(n, n), O.structured_2d(n)
    return N.MakeStructured3DMesh(n, n, n), O.structured_3d(n)


@pytest.mark.parametrize("dim,n,p", [(2, 1, 1), (2, 4, 2), (2, 4, 3), (2, 6, 5), (2, 3, 7),
                                     (3, 1, 1), (3, 2, 2), (3, 2, 4), (3, 3, 3), (3, 2, 5)])
def test_dofs_pattern_values(dim, n, p):
    m, om = _meshes(dim, n)
    assert np.array_equal(m.elements, om["els"]) and np.array_equal(m.coords, om["coords"])
    fes = N.H1Space(m, p, ALL)
    order = 2 * p
    pr = O.Problem(om, p, intorder=order)
    assert fes.ndof == pr.ndof
    assert np.array_equal(fes.el_dofs, pr.dofs)
    assert np.array_equal(fes.freedofs, pr.free)
    assert np.array_equal(fes.classnr, O.class_number(dim, om["els"]))
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    rp, cols = N.BuildPattern(fes)
    orp, ocols = pr.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    kappa = np.exp(np.random.default_rng(5).standard_normal(m.ne))
    vals = N.AssembleCSR(fes, rp, cols, ir, True, kappa, True, None)
    _, _, ov = pr.assemble(kappa=kappa, mass=True)
    assert np.abs(vals - ov).max() <= 1e-13 * np.abs(ov).max()


@pytest.mark.parametrize("dim,order", [(2, o) for o in range(0, 13)] + [(3, o) for o in range(0, 11)])
def test_rules_match_oracle(dim, order):
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    pts, w = O.rule(dim, order)
    assert ir.points.shape == pts.shape
    assert np.abs(ir.points - pts).max() < 1e-14 and np.abs(ir.weights - w).max() < 1e-15


@pytest.mark.parametrize("dim,p,order", [(2, 3, 6), (2, 3, 4), (3, 4, 8), (3, 4, 6), (2, 5, 8)])
def test_reference_gradients(dim, p, order):
    """Per-class B tensors = the oracle's per-element reference gradients."""
    m, om = _meshes(dim, 3 if dim == 2 else 2)
    fes = N.H1Space(m, p, ALL)
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    classes = N.BuildClasses(fes, ir)
    pr = O.Problem(om, p, intorder=order)
    val, grad = pr.shapes()
    for c in classes:
        for e in c["elements"][:5]:
            assert np.abs(c["bmat"] - grad[e]).max() < 1e-13
            assert np.abs(c["shape"] - val[e]).max() < 1e-14


def test_golden_reference_gradients():
    """Committed B tensors (trig p3 Q=12, tet p4 Q=125) of class 0."""
    g = np.load(os.path.join(GOLDEN, "bmat.npz"))
    for key, (dim, p, order) in {"trig_p3_q12": (2, 3, 6), "tet_p4_q125": (3, 4, 8)}.items():
        m, _ = _meshes(dim, 2)
        fes = N.H1Space(m, p, ALL)
        ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
        c0 = [c for c in N.BuildClasses(fes, ir) if c["classnr"] == 0][0]
        assert np.abs(c0["bmat"] - g[key]).max() < 1e-13


def test_golden_trig64_p3_pattern_and_apply():
    """64x64 trigs, p=3: pattern digest, values sample and y = A x."""
    g = np.load(os.path.join(GOLDEN, "trig64_p3.npz"))
    m = N.MakeStructured2DMesh(64, 64)
    fes = N.H1Space(m, 3, ALL)
    assert fes.ndof == 37249 and int(fes.freedofs.sum()) == 36481
    rp, cols = N.BuildPattern(fes)
    assert rp[-1] == 628609
    assert np.array_equal(rp, g["rowptr"]) and np.array_equal(cols, g["cols"])
    vals = N.AssembleCSR(fes, rp, cols, N.TrigRule(6), True, None, False, None)
    x = np.random.default_rng(0).uniform(-1, 1, fes.ndof)
    import scipy.sparse as sp
    y = sp.csr_matrix((vals, cols, rp), shape=(fes.ndof, fes.ndof)) @ x
    assert np.abs(y - g["y"]).max() <= 1e-13 * np.abs(g["y"]).max()


def test_full_size_counts():
    """Sizes of the 3D configuration (32^3 Kuhn tets, p=4) and its classes."""
    m = N.MakeStructured3DMesh(32, 32, 32)
    fes = N.H1Space(m, 4, ALL)
    assert fes.ndof == 2146689
    assert int((fes.freedofs == 0).sum()) == 98306
    vals, cnt = np.unique(fes.classnr, return_counts=True)
    assert vals.tolist() == [0, 2, 26] and cnt.tolist() == [98304, 65536, 32768]


@pytest.mark.parametrize("name", ["square.vol.gz", "cube.vol.gz"])
def test_vol_reader_matches_oracle(name):
    with gzip.open(os.path.join(GOLDEN, name), "rt") as f:
        text = f.read()
    m = N.ReadVolMesh(text)
    om = O.read_vol(text)
    assert np.array_equal(m.elements, om["els"]) and np.allclose(m.coords, om["coords"])
    fes = N.H1Space(m, 3, ALL)
    pr = O.Problem(om, 3)
    assert np.array_equal(fes.el_dofs, pr.dofs) and np.array_equal(fes.freedofs, pr.free)


@pytest.mark.parametrize("dim,n", [(2, 1), (2, 7), (3, 1), (3, 5)])
def test_element_coloring_matches_oracle(dim, n):
    """FESpace element colouring (comp/fespace.cpp:518-760): the host C++
    colouring equals the oracle's, bit for bit, and is valid (elements of one
    colour share no vertex, hence no H1 dof)."""
    m, om = _meshes(dim, n)
    col = np.asarray(N.ElementColoring(m))
    ocol = O.element_coloring(om["els"], int(om["els"].max()) + 1)
    assert np.array_equal(col, ocol)
    els = np.asarray(om["els"])
    for c in np.unique(col):
        v = els[col == c].ravel()
        assert len(np.unique(v)) == len(v)
