"""Multi-vector CSR apply (ngsamd_csr_mult_multi) and the lockstep multi-RHS
CG (ngsamd_cg_solve_multi, CGSolver<double>::MultiMult): each component of
A X bit-identical to a single apply in the runs layout (the same summation
order) and within 1e-15 in the other layouts; the CG against the oracle's
cg_multi (pinned against the single CG and closed forms) -- steps within 1,
solutions within 1e-10 of the largest entry -- with Jacobi and without a
preconditioner, from zero and from a given start, with a zero column."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, say

pytestmark = pytest.mark.gpu


def _eliminated(P):
    """The matrix with the rows and columns of the fixed dofs replaced by the
    identity (symmetric positive definite, as CG without a preconditioner
    needs)."""
    free = P["fes"].freedofs.astype(bool)
    rows = np.repeat(np.arange(P["ndof"]), np.diff(P["rowptr"]))
    vals = P["vals"].copy()
    fixed = ~free[rows] | ~free[P["cols"]]
    vals[fixed] = 0.0
    vals[fixed & (rows == P["cols"])] = 1.0
    return dict(P, vals=vals)


def _multi_spmv(P, X):
    return np.stack([oracle.spmv(P["rowptr"], P["cols"], P["vals"], np.ascontiguousarray(X[:, j]))
                     for j in range(X.shape[1])], 1)


@pytest.mark.parametrize("dim,n,p", [(2, 10, 3), (3, 3, 4)])
@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 6, 8, 12])
def test_mult_multi_matches_single(device, dim, n, p, K):
    say(f"[csr multi] dim={dim} n={n} p={p} K={K}")
    P = problem(dim, n, p)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    X = np.random.default_rng(K).uniform(-1, 1, (P["ndof"], K))
    dX, dY = dvec(X.ravel()), N.DeviceVector(P["ndof"] * K, 8)
    N.CSRMultMulti(A, dX, dY, K)
    Y = dY.download().reshape(P["ndof"], K)
    dx, dy = N.DeviceVector(P["ndof"], 8), N.DeviceVector(P["ndof"], 8)
    for j in range(K):
        dx.upload(np.ascontiguousarray(X[:, j]))
        A.mult(dx, dy)
        assert np.array_equal(Y[:, j], dy.download())
    ref = _multi_spmv(P, X)
    assert np.abs(Y - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("algo", ["stream", "rowgroup"])
def test_mult_multi_other_layouts(device, monkeypatch, algo):
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    say(f"[csr multi {algo}]")
    P = problem(2, 8, 3)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    X = np.random.default_rng(3).uniform(-1, 1, (P["ndof"], 3))
    dX, dY = dvec(X.ravel()), N.DeviceVector(P["ndof"] * 3, 8)
    N.CSRMultMulti(A, dX, dY, 3)
    ref = _multi_spmv(P, X)
    assert np.abs(dY.download().reshape(-1, 3) - ref).max() <= 1e-15 * np.abs(ref).max() * 10


@pytest.mark.parametrize("dim,n,p", [(2, 12, 3), (3, 3, 3)])
@pytest.mark.parametrize("K", [1, 3, 4, 8])
@pytest.mark.parametrize("jacobi", [True, False])
def test_cg_multi_matches_oracle(device, dim, n, p, K, jacobi):
    if not jacobi:
        # unpreconditioned CG on the hierarchical basis is ill-conditioned
        # (1000+ steps at p = 3, where rounding alone moves the count by ~1%):
        # P1 cases
        n, p = (6, 1) if dim == 2 else (3, 1)
    say(f"[cg multi] dim={dim} n={n} p={p} K={K} jacobi={jacobi}")
    P = _eliminated(problem(dim, n, p, seed_kappa=2))
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    F = np.random.default_rng(K).uniform(-1, 1, (P["ndof"], K)) * free[:, None]
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, free) if jacobi else None
    dF, dU = dvec(F.ravel()), N.DeviceVector(P["ndof"] * K, 8)
    steps, hist = N.CGSolveMulti(A, C, dF, dU, K, True, 1e-10, False, 5000)
    U = dU.download().reshape(-1, K)
    Uo, so, ho = O.cg_multi(lambda V: _multi_spmv(P, V), inv if jacobi else None, F, tol=1e-10, maxiter=5000)
    assert abs(steps - so) <= 1
    assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max()
    if jacobi:  # unpreconditioned, the histories part after a few steps (rounding); the results agree
        m = min(len(hist), len(ho), 10)
        assert np.allclose(hist[:m], ho[:m], rtol=1e-7)


def test_cg_multi_start_vector_and_zero_column(device):
    say("[cg multi] given start, zero column, maxiter")
    P = _eliminated(problem(2, 10, 2, seed_kappa=4))
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    rng = np.random.default_rng(6)
    F = rng.uniform(-1, 1, (P["ndof"], 3)) * free[:, None]
    F[:, 1] = 0.0
    U0 = rng.uniform(-1, 1, (P["ndof"], 3)) * free[:, None]
    U0[:, 1] = 0.0
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, free)
    dF, dU = dvec(F.ravel()), dvec(U0.ravel())
    steps, _ = N.CGSolveMulti(A, C, dF, dU, 3, False, 1e-10, False, 5000)
    U = dU.download().reshape(-1, 3)
    Uo, so, _ = O.cg_multi(lambda V: _multi_spmv(P, V), inv, F, tol=1e-10, maxiter=5000, U0=U0)
    assert abs(steps - so) <= 1
    assert np.all(np.isfinite(U)) and np.all(U[:, 1] == 0.0)
    assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max()
    # maxiter bounds the steps (steps = maxiter + 1 loop tests)
    dU2 = N.DeviceVector(P["ndof"] * 3, 8)
    s3, h3 = N.CGSolveMulti(A, C, dF, dU2, 3, True, 1e-14, False, 3)
    assert s3 == 4 and len(h3) == 4


def test_cg_multi_refuses_bad_input(device):
    say("[cg multi] bad input")
    P = problem(2, 4, 2)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    free = P["fes"].freedofs.astype(np.uint8)
    diag, _ = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    for K in (0, 9):
        f, u = N.DeviceVector(max(P["ndof"] * K, 1), 8), N.DeviceVector(max(P["ndof"] * K, 1), 8)
        with pytest.raises(RuntimeError):
            N.CGSolveMulti(A, C, f, u, K, True, 1e-10, False, 10)
    f, u = N.DeviceVector(P["ndof"] * 2, 8), N.DeviceVector(P["ndof"] * 2, 8)
    with pytest.raises(RuntimeError, match="CSR"):
        N.CGSolveMulti(C, None, f, u, 2, True, 1e-10, False, 10)  # Jacobi is not a CSR matrix
    with pytest.raises(RuntimeError, match="Jacobi"):
        N.CGSolveMulti(A, A, f, u, 2, True, 1e-10, False, 10)
    with pytest.raises(RuntimeError):
        N.CSRMultMulti(A, f, u, 3)  # sizes do not match 3 components


def test_cg_multi_zero_rhs(device):
    """All right-hand sides 0: w.d is reset to 1, <s, A s> = 0 ends the loop
    at once (steps 1), u stays 0, the recorded start residual is 0."""
    say("[cg multi] zero right-hand sides")
    P = _eliminated(problem(2, 4, 2))
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    C = N.CreateJacobi(diag, free)
    F = np.zeros((P["ndof"], 3))
    dF, dU = dvec(F.ravel()), N.DeviceVector(P["ndof"] * 3, 8)
    steps, hist = N.CGSolveMulti(A, C, dF, dU, 3, True, 1e-10, False, 100)
    Uo, so, ho = O.cg_multi(lambda V: _multi_spmv(P, V), inv, F, tol=1e-10, maxiter=100)
    assert steps == so == 1 and hist.tolist() == ho.tolist() == [0.0]
    assert np.all(dU.download() == 0.0)
