"""Adjoint gradient of J(kappa) = 0.5 (u-u_obs)^T M (u-u_obs), u = A(kappa)^{-1} f,
through torch.autograd: against the oracle's CPU adjoint gradient and against
central finite differences along random unit directions (the differentiation
configuration at 32x32 instead of 128x128)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fe_oracle as O
from helpers import say

pytestmark = pytest.mark.gpu


def _setup(nx, p):
    import ngsolve_amd as ngs
    m = ngs.MakeStructured2DMesh(nx=nx, ny=nx)
    fes = ngs.H1(m, order=p, dirichlet=".*")
    u, v = fes.TnT()
    f = ngs.LinearForm(1 * v * ngs.dx).Assemble().vec.NumPy().copy()
    mass = ngs.BilinearForm(u * v * ngs.dx).Assemble()
    pr = O.Problem(O.structured_2d(nx), p, intorder=2 * p - 2)
    prm = O.Problem(O.structured_2d(nx), p, intorder=2 * p)
    rp, c, vals = prm.assemble(stiffness=False, mass=True)
    M = sp.csr_matrix((vals, c, rp), shape=(pr.ndof, pr.ndof))
    return ngs, m, fes, f, mass, pr, M


@pytest.mark.parametrize("tdev", ["cuda", "cpu"])
def test_adjoint_gradient_matches_oracle_and_fd(device, tdev):
    import torch
    from ngsolve_amd.autograd import ApplySymmetric, PoissonProblem, SolvePoisson
    say(f"[autograd] 32x32 p=3 tensors on {tdev}")
    ngs, m, fes, f, mass, pr, M = _setup(32, 3)
    ne = m.ne
    rng = np.random.default_rng(3)
    kappa_true = np.exp(rng.standard_normal(ne))
    P = PoissonProblem(fes, f, tol=1e-13)
    u_true = SolvePoisson.apply(torch.tensor(kappa_true, device=tdev), P).cpu().numpy()
    u_obs = u_true + rng.normal(0.0, 1e-3, fes.ndof)
    kappa = np.exp(np.random.default_rng(4).normal(0.0, 0.5, ne))

    kt = torch.tensor(kappa, device=tdev, requires_grad=True)
    uo = torch.tensor(u_obs, device=tdev)
    u = SolvePoisson.apply(kt, P)
    r = u - uo
    J = 0.5 * torch.dot(r, ApplySymmetric.apply(r, mass.mat.op))
    J.backward()
    g = kt.grad.cpu().numpy()

    J_ref, g_ref, _ = O.adjoint_gradient(pr, kappa, f, u_obs, M)
    say(f"[autograd] J={J.item():.6e} ref={J_ref:.6e}")
    assert abs(J.item() - J_ref) <= 1e-9 * abs(J_ref)
    assert np.linalg.norm(g - g_ref) <= 1e-8 * np.linalg.norm(g_ref)

    # central differences (h = 1e-6) along 16 random unit directions
    drng = np.random.default_rng(6)
    h = 1e-6
    worst = 0.0
    for _ in range(16):
        d = drng.standard_normal(ne)
        d /= np.linalg.norm(d)
        Jp, _, _ = O.adjoint_gradient(pr, kappa + h * d, f, u_obs, M)
        Jm, _, _ = O.adjoint_gradient(pr, kappa - h * d, f, u_obs, M)
        fd = (Jp - Jm) / (2 * h)
        worst = max(worst, abs(fd - g @ d) / max(abs(fd), 1e-30))
    say(f"[autograd] worst FD relative error {worst:.2e}")
    # central differences at h = 1e-6 carry the solves' rounding,
    # ~eps * cond(A) * |J| / h (about 5e-5 of g.d here); a wrong sign, a
    # dropped term or a transposed operand is an O(1) error
    assert worst < 1e-3
