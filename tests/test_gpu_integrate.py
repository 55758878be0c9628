"""Integrate(c * grad(a) * grad(b), mesh, element_wise=True) on the device
(ngsamd_mf_element_grad) against the oracle's element matrices,
a_e^T K_e b_e (1e-12 of the largest), and the total against a^T A b with
the assembled matrix; a bare grad(gfu) refused."""
import numpy as np
import pytest

import ngsolve_amd as ngs
import oracle
from oracle import fe_oracle as O
from helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n,p", [(2, 6, 3), (3, 3, 3), (3, 2, 4)])
def test_element_energies_match_oracle(device, dim, n, p):
    say(f"[integrate] dim={dim} n={n} p={p}")
    m = ngs.MakeStructured2DMesh(nx=n, ny=n) if dim == 2 else ngs.MakeStructured3DMesh(nx=n)
    om = O.structured_2d(n, n) if dim == 2 else O.structured_3d(n, n, n)
    fes = ngs.H1(m, order=p)
    a, b = ngs.GridFunction(fes), ngs.GridFunction(fes)
    rng = np.random.default_rng(5)
    ua, ub = rng.uniform(-1, 1, fes.ndof), rng.uniform(-1, 1, fes.ndof)
    a.vec.NumPy()[:] = ua
    b.vec.NumPy()[:] = ub
    got = ngs.Integrate(ngs.grad(a) * ngs.grad(b), m, element_wise=True)
    pr = O.Problem(om, p, intorder=2 * p - 2)
    K = pr.element_matrices()
    ref = np.einsum("ei,eij,ej->e", ua[pr.dofs], K, ub[pr.dofs])
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    kappa = np.exp(rng.standard_normal(m.ne))
    got_k = ngs.Integrate(ngs.ElementCoefficient(kappa) * (ngs.grad(a) * ngs.grad(b)), m, element_wise=True)
    assert np.abs(got_k - kappa * ref).max() <= 1e-12 * np.abs(kappa * ref).max()
    u, v = fes.TnT()
    A = ngs.BilinearForm(ngs.grad(u) * ngs.grad(v) * ngs.dx).Assemble()
    total = ua @ oracle.spmv(A.mat.rowptr, A.mat.cols, A.mat.vals, ub)
    assert abs(ngs.Integrate(ngs.grad(a) * ngs.grad(b), m) - total) <= 1e-12 * abs(total) + 1e-14


def test_integrate_rejects_other_mesh(device):
    """Element-wise energies of grid functions on one mesh over another mesh
    (with fewer elements) are refused instead of written past the output."""
    m, m2 = ngs.MakeStructured2DMesh(nx=4, ny=4), ngs.MakeStructured2DMesh(nx=2, ny=2)
    a = ngs.GridFunction(ngs.H1(m, order=2))
    with pytest.raises(ValueError):
        ngs.Integrate(ngs.grad(a) * ngs.grad(a), m2, element_wise=True)


def test_integrate_rejects_vector_integrand(device):
    m = ngs.MakeStructured2DMesh(nx=2, ny=2)
    gfu = ngs.GridFunction(ngs.H1(m, order=1))
    with pytest.raises(TypeError):
        ngs.Integrate(ngs.grad(gfu), m)
