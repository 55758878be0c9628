"""Mixed-precision iterative refinement: FP64 residual around an FP32, BF16,
FP16 or split-FP16 matrix-free inner CG reaches 1e-10 relative residual and
agrees with the FP64 solution."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, mf_operator, problem, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [1, 2, 3, 4])
@pytest.mark.parametrize("dim,n,p,order", [(3, 4, 4, 8), (2, 32, 3, 6)])
def test_ir_reaches_1e10(device, precision, dim, n, p, order):
    say(f"[ir prec={precision}] dim={dim} n={n} p={p}")
    P = problem(dim, n, p, order=order, seed_kappa=2)
    A64 = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    A32 = mf_operator(P, precision=precision)
    _, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], P["fes"].freedofs)
    diag = 1.0 / np.where(inv == 0, 1, inv)
    C = N.CreateJacobi(diag, P["fes"].freedofs)
    f = np.ones(P["ndof"])
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    outer, inner, rel = N.IRCGSolve(A64, A32, C, df, du, 1e-10, 1e-4, 50, 20000)
    say(f"[ir] outer={outer} inner={inner} rel={rel:.2e}")
    assert rel <= 1e-10
    u = du.download()
    fr = P["fes"].freedofs.astype(bool)
    r = f - oracle.spmv(P["rowptr"], P["cols"], P["vals"], u)
    assert np.linalg.norm(r[fr]) <= 1.0001e-10 * np.linalg.norm(f[fr])
    u_ref, _, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-12, maxiter=50000)
    assert np.linalg.norm(u - u_ref) <= 1e-8 * np.linalg.norm(u_ref)
