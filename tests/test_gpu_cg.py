"""Device CG against the oracle's CG (reference CGSolver recurrence):
iteration count within +-1, solution within 1e-10 relative."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


def _jacobi(P):
    _, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], P["fes"].freedofs)
    diag = 1.0 / np.where(inv == 0, 1, inv)
    return inv, diag


@pytest.mark.parametrize("chunk", [16, 0])
def test_cg_trig64_p3(device, chunk):
    """64x64 trigs, p=3, kappa=1, f=1, Jacobi CG to 1e-10."""
    say(f"[cg] 64x64 trigs p=3 chunk={chunk}")
    P = problem(2, 64, 3, order=6)
    f = N.AssembleLinearForm(P["fes"], P["ir"], np.ones((P["mesh"].ne, len(P["ir"]))))
    inv, diag = _jacobi(P)
    u_ref, its_ref, res_ref = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-10, maxiter=5000)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, P["fes"].freedofs)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, res = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 5000, chunk)
    u = du.download()
    say(f"[cg] device its={its} oracle its={its_ref}")
    assert abs(its - its_ref) <= 1
    assert np.linalg.norm(u - u_ref) <= 1e-10 * np.linalg.norm(u_ref)
    assert len(res) == its and abs(res[0] - res_ref[0]) <= 1e-12 * res_ref[0]
    # true residual on the free dofs
    r = f - oracle.spmv(P["rowptr"], P["cols"], P["vals"], u)
    fr = P["fes"].freedofs.astype(bool)
    assert np.linalg.norm(r[fr]) < 1e-8 * np.linalg.norm(f[fr])


@pytest.mark.parametrize("algo", ["runs", "stream", "rowgroup"])
def test_cg_kappa_3d(device, monkeypatch, algo):
    """Small 3D p=4 problem with lognormal kappa (the precision configuration's
    shape), with each SpMV kernel's fused <s, As> variant."""
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    say(f"[cg] 3D p=4 lognormal kappa, SpMV kernel {algo}")
    P = problem(3, 4, 4, order=8, seed_kappa=2)
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, P["ndof"])
    inv, diag = _jacobi(P)
    u_ref, its_ref, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-10, maxiter=20000)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, P["fes"].freedofs)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 20000, 16)
    say(f"[cg] device its={its} oracle its={its_ref}")
    assert abs(its - its_ref) <= 1
    assert np.linalg.norm(du.download() - u_ref) <= 1e-10 * np.linalg.norm(u_ref)


def test_cg_maxiter_and_zero_rhs(device):
    say("[cg] maxiter / zero rhs")
    P = problem(2, 8, 2)
    inv, diag = _jacobi(P)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, P["fes"].freedofs)
    f = np.random.default_rng(3).uniform(-1, 1, P["ndof"])
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, res = N.CGSolve(A, C, df, du, True, 1e-30, -1.0, 5, 16)
    _, its_ref, res_ref = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-30, maxiter=5)
    assert its == its_ref == 5 and np.allclose(res, res_ref, rtol=1e-12)
    dz = dvec(np.zeros(P["ndof"]))
    its, res = N.CGSolve(A, C, dz, du, True, 1e-10, -1.0, 100, 16)
    assert its == 1 and not du.download().any()


def test_cg_atol(device):
    say("[cg] atol")
    P = problem(2, 16, 3)
    inv, diag = _jacobi(P)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, P["fes"].freedofs)
    f = np.random.default_rng(8).uniform(-1, 1, P["ndof"])
    _, its_ref, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=None, atol=1e-6, maxiter=1000)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, res = N.CGSolve(A, C, df, du, True, -1.0, 1e-6, 1000, 16)
    assert abs(its - its_ref) <= 1 and res[-1] <= 1e-6
