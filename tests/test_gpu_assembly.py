"""Sparsity pattern on the device (ngsamd_mf_pattern): bit-exact against
the oracle's pattern (numpy.unique of all element dof pairs) and the
product's host pattern, 2D and 3D, p = 1..5, structured and unstructured.
CSR assembly on the device (ngsamd_mf_assemble_csr) against the oracle's
assembly (oracle/fe_oracle.py Problem.assemble: per-element matrices from
the per-point F^-T F^-1 form, scipy COO -> CSR): values within 1e-13 of the
largest entry, in every CSR layout the operator may choose (runs, stream,
rowgroup).  Also against the product's host assembly on the same pattern,
the assembled operator's apply against the oracle SpMV, determinism (two
assemblies are bit-identical) and the refusal of a pattern that lacks a
pair of dofs of one element.  Mass assembly on the device
(ngsamd_mf_assemble_mass_csr) against the oracle's mass matrix (pinned by
the reference's known mass entry), with an element coefficient."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, mf_operator, problem, relerr_inf, say

pytestmark = pytest.mark.gpu

CASES = [(2, 6, 1, None), (2, 8, 3, 6), (2, 5, 5, None), (3, 3, 1, None), (3, 3, 2, None), (3, 3, 4, 8),
         (3, 2, 5, None)]
ALGOS = ["runs", "stream", "rowgroup"]


@pytest.mark.parametrize("dim,n,p,order", CASES)
@pytest.mark.parametrize("algo", ALGOS)
def test_assembly_matches_oracle(device, monkeypatch, dim, n, p, order, algo):
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    say(f"[assemble {algo}] dim={dim} n={n} p={p} order={order}")
    P = problem(dim, n, p, order=order, seed_kappa=3)
    A = N.AssembleCSRDevice(mf_operator(P), P["rowptr"], P["cols"])
    vals = A.csr_values()
    pr = O.Problem(P["omesh"], p, intorder=P["order"])
    orp, ocols, ovals = pr.assemble(kappa=P["kappa"])
    assert np.array_equal(orp, P["rowptr"]) and np.array_equal(ocols, P["cols"])
    scale = np.abs(ovals).max()
    assert np.abs(vals - ovals).max() / scale < 1e-13
    assert np.abs(vals - P["vals"]).max() / scale < 1e-13
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(orp, ocols, ovals, x)) < 1e-13


PATTERN_CASES = [(2, 7, 1), (2, 6, 3), (2, 4, 5), (3, 3, 1), (3, 3, 2), (3, 3, 4), (3, 2, 5)]


@pytest.mark.parametrize("dim,n,p", PATTERN_CASES)
def test_pattern_matches_oracle(device, dim, n, p):
    say(f"[pattern] dim={dim} n={n} p={p}")
    P = problem(dim, n, p)
    rp, cols = N.PatternDevice(mf_operator(P))
    orp, ocols = O.Problem(P["omesh"], p).pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    assert np.array_equal(rp, P["rowptr"]) and np.array_equal(cols, P["cols"])


@pytest.mark.parametrize("name", ["square.vol.gz", "cube.vol.gz"])
@pytest.mark.parametrize("p", [1, 3, 4])
def test_pattern_on_vol_mesh(device, name, p):
    import gzip
    import os
    say(f"[pattern] {name} p={p}")
    text = gzip.open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name), "rt").read()
    m = N.ReadVolMesh(text)
    fes = N.H1Space(m, p, 0)
    ir = N.TrigRule(2 * p) if m.dim == 2 else N.TetRule(2 * p)
    P = dict(mesh=m, fes=fes, ir=ir, dim=m.dim, kappa=None)
    rp, cols = N.PatternDevice(mf_operator(P))
    hrp, hcols = N.BuildPattern(fes)
    orp, ocols = O.Problem(O.read_vol(text), p).pattern()
    assert np.array_equal(rp, hrp) and np.array_equal(cols, hcols)
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)


def test_assembly_deterministic_and_kappa_one(device):
    say("[assemble] determinism, kappa = 1")
    P = problem(3, 4, 3)
    M = mf_operator(P)
    v1 = N.AssembleCSRDevice(M, P["rowptr"], P["cols"]).csr_values()
    v2 = N.AssembleCSRDevice(M, P["rowptr"], P["cols"]).csr_values()
    assert np.array_equal(v1, v2)
    # A 1 = 0 without boundary conditions: the constant 1 is the sum of the
    # vertex functions (dofs [0, nv)), every other dof coefficient 0
    one = np.zeros(P["ndof"])
    one[:P["mesh"].nv] = 1.0
    rows = np.repeat(np.arange(P["ndof"]), np.diff(P["rowptr"]))
    a1 = np.bincount(rows, weights=v1 * one[P["cols"]], minlength=P["ndof"])
    assert np.abs(a1).max() < 1e-13 * np.abs(v1).max()


def test_assembly_rejects_short_pattern(device):
    say("[assemble] pattern without an element pair")
    P = problem(2, 3, 2)
    rp, cols = P["rowptr"].copy(), P["cols"].copy()
    # drop the last column of row 0 (an element neighbour of dof 0)
    keep = np.ones(len(cols), dtype=bool)
    keep[rp[1] - 1] = False
    cols2 = cols[keep]
    rp2 = rp.copy()
    rp2[1:] -= 1
    with pytest.raises(RuntimeError, match="lacks a pair"):
        N.AssembleCSRDevice(mf_operator(P), rp2, cols2)


def test_bilinearform_device_assembly(device):
    """BilinearForm(..., device_assembly=True) gives the host assembly's matrix."""
    import ngsolve_amd as ngs
    say("[assemble] BilinearForm device_assembly")
    m = ngs.MakeStructured3DMesh(nx=3)
    fes = ngs.H1(m, order=3, dirichlet=".*")
    u, v = fes.TnT()
    kappa = ngs.ElementCoefficient(np.exp(np.random.default_rng(2).standard_normal(m.ne)))
    form = kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx(bonus_intorder=2)
    a_host = ngs.BilinearForm(form).Assemble()
    a_dev = ngs.BilinearForm(form, device_assembly=True).Assemble()
    vh, vd = np.asarray(a_host.mat.vals), np.asarray(a_dev.mat.vals)
    assert vh.shape == vd.shape
    assert np.abs(vh - vd).max() < 1e-13 * np.abs(vh).max()
    assert abs(a_dev.mat[0, 0] - a_host.mat[0, 0]) < 1e-13 * abs(a_host.mat[0, 0])


@pytest.mark.parametrize("dim,n,p", [(2, 6, 3), (2, 4, 5), (3, 3, 2), (3, 3, 4)])
def test_linear_form_matches_oracle(device, dim, n, p):
    """Device linear form (ngsamd_mf_linear_form) with its own rule (order
    2p) against the oracle's linear form and the host one, f varying per
    element and point."""
    say(f"[linear form] dim={dim} n={n} p={p}")
    P = problem(dim, n, p)
    M = mf_operator(P)
    order = 2 * p
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    classes = N.BuildClasses(P["fes"], ir)
    shape = np.ascontiguousarray(np.stack([c["shape"] for c in classes]))
    detj, _ = N.ComputeGeometry(P["mesh"])
    ne, nq = P["mesh"].ne, len(ir)
    fvals = np.random.default_rng(9).uniform(-1, 2, (ne, nq))
    dfv, dout = dvec(fvals.ravel()), N.DeviceVector(P["ndof"], 8)
    N.LinearFormDevice(M, np.ascontiguousarray(ir.weights), shape, np.ascontiguousarray(detj), dfv, dout)
    got = dout.download()
    pr = O.Problem(P["omesh"], p, intorder=order)
    ref = pr.linear_form(fvals)
    assert relerr_inf(got, ref) < 1e-13
    assert relerr_inf(got, N.AssembleLinearForm(P["fes"], ir, fvals)) < 1e-13


MASS_CASES = [(2, 6, 1), (2, 5, 3), (2, 4, 5), (3, 3, 1), (3, 3, 2), (3, 2, 4)]


@pytest.mark.parametrize("dim,n,p", MASS_CASES)
@pytest.mark.parametrize("algo", ["runs", "stream"])
def test_mass_assembly_matches_oracle(device, monkeypatch, dim, n, p, algo):
    """rho * u * v * dx with its own rule (order 2p) and a lognormal element
    coefficient, against the oracle's mass matrix (1e-13 of the largest
    entry) and its SpMV."""
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    say(f"[mass {algo}] dim={dim} n={n} p={p}")
    P = problem(dim, n, p)
    M = mf_operator(P)
    order = 2 * p
    ir = N.TrigRule(order) if dim == 2 else N.TetRule(order)
    shape = np.ascontiguousarray(np.stack([c["shape"] for c in N.BuildClasses(P["fes"], ir)]))
    detj, _ = N.ComputeGeometry(P["mesh"])
    rho = np.exp(np.random.default_rng(11).standard_normal(P["mesh"].ne))
    A = N.AssembleMassCSRDevice(M, np.ascontiguousarray(ir.weights), shape, rho * np.abs(detj), P["rowptr"],
                                P["cols"])
    pr = O.Problem(P["omesh"], p, intorder=order)
    orp, ocols, ovals = pr.assemble(stiffness=False, mass=True, rho=rho)
    assert np.array_equal(orp, P["rowptr"]) and np.array_equal(ocols, P["cols"])
    vals = A.csr_values()
    assert np.abs(vals - ovals).max() < 1e-13 * np.abs(ovals).max()
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(orp, ocols, ovals, x)) < 1e-13


def test_mass_bilinear_form_on_device(device):
    """BilinearForm(rho*u*v*dx, device_assembly=True) against the host
    assembly; c^T M c = integral of rho for the coefficients c of the
    constant 1."""
    import ngsolve_amd as ngs
    say("[mass] BilinearForm(device_assembly=True)")
    m = ngs.MakeStructured3DMesh(nx=3)
    fes = ngs.H1(m, order=3)
    u, v = fes.TnT()
    rho = np.exp(np.random.default_rng(2).standard_normal(m.ne))
    c = ngs.ElementCoefficient(rho)
    a_dev = ngs.BilinearForm(c * u * v * ngs.dx, device_assembly=True).Assemble()
    a_host = ngs.BilinearForm(c * u * v * ngs.dx).Assemble()
    vh, vd = a_host.mat.vals, a_dev.mat.vals
    assert np.abs(vh - vd).max() < 1e-13 * np.abs(vh).max()
    one = np.zeros(fes.ndof)
    one[:m.ngmesh.nv] = 1.0  # the constant 1 is the sum of the vertex functions (the first nv dofs)
    detj, _ = N.ComputeGeometry(m.ngmesh)
    assert abs(one @ oracle.spmv(a_dev.mat.rowptr, a_dev.mat.cols, vd, one) - np.sum(rho * np.abs(detj)) / 6) < 1e-12
