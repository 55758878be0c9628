"""The device path on the reference's own unstructured meshes
(tests/pytest/square.vol.gz, cube.vol.gz, copied to tests/golden): CSR SpMV
in every layout, the matrix-free apply (both kernel forms, FP64 and FP32),
the device assembly and the Jacobi-PCG, each against the oracle assembled
from its own reading of the same .vol file (1e-13 relative for FP64,
iterations within 1 and 1e-10 for the CG)."""
import gzip
import os

import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import ALL, N, dvec, mf_operator, relerr_inf, say

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vol_problem(name, p, seed_kappa=7):
    text = gzip.open(os.path.join(GOLDEN, name), "rt").read()
    m = N.ReadVolMesh(text)
    om = O.read_vol(text)
    order = 2 * p
    fes = N.H1Space(m, p, ALL)
    ir = N.TrigRule(order) if m.dim == 2 else N.TetRule(order)
    kappa = np.exp(np.random.default_rng(seed_kappa).standard_normal(m.ne))
    pr = O.Problem(om, p, intorder=order, dirichlet_regions=None)
    rp, cols, vals = pr.assemble(kappa=kappa)
    prp, pcols = N.BuildPattern(fes)
    assert np.array_equal(prp, rp) and np.array_equal(pcols, cols)
    return dict(mesh=m, omesh=om, fes=fes, ir=ir, order=order, kappa=kappa, rowptr=rp, cols=cols, vals=vals,
                ndof=fes.ndof, dim=m.dim, p=p, pr=pr)


MESHES = ["square.vol.gz", "cube.vol.gz"]


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("p", [1, 2, 4])
@pytest.mark.parametrize("algo", ["runs", "stream", "rowgroup"])
def test_spmv_on_vol_mesh(device, monkeypatch, name, p, algo):
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    say(f"[vol spmv {algo}] {name} p={p}")
    P = vol_problem(name, p)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)) < 1e-13


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("p", [1, 3, 4])
@pytest.mark.parametrize("kernel", ["ref", "quad"])
def test_matfree_on_vol_mesh(device, monkeypatch, name, p, kernel):
    monkeypatch.setenv("NGSAMD_MF_KERNEL", kernel)
    say(f"[vol mf {kernel}] {name} p={p}")
    P = vol_problem(name, p)
    x = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
    ref_y = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    for prec, tol in ((0, 1e-13), (1, 1e-5)):
        A = mf_operator(P, precision=prec)
        dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
        A.mult(dx, dy)
        assert relerr_inf(dy.download(), ref_y) < tol


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("p", [2, 4])
def test_assembly_on_vol_mesh(device, name, p):
    say(f"[vol assemble] {name} p={p}")
    P = vol_problem(name, p)
    A = N.AssembleCSRDevice(mf_operator(P), P["rowptr"], P["cols"])
    vals = A.csr_values()
    assert np.abs(vals - P["vals"]).max() < 1e-13 * np.abs(P["vals"]).max()


@pytest.mark.parametrize("name", MESHES)
def test_cg_on_vol_mesh(device, name):
    say(f"[vol cg] {name} p=3")
    P = vol_problem(name, 3)
    # Dirichlet on every boundary region, in the oracle and the product alike
    free = np.asarray(P["pr"].free, dtype=np.uint8)
    assert np.array_equal(free, P["fes"].freedofs.astype(np.uint8)) and 0 < free.sum() < P["ndof"]
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, np.random.default_rng(3).uniform(0, 1, P["ndof"]), 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 5000, 16)
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(P["rowptr"], P["cols"], P["vals"], t), inv, f, tol=1e-10,
                              maxiter=5000)
    assert abs(its - its_ref) <= 1
    assert np.abs(du.download() - u_ref).max() <= 1e-10 * np.abs(u_ref).max()
