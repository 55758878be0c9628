"""Guards of the Python layer that need no device: the operators take one
coefficient value per element (the reference's fused path evaluates the
coefficient once per element, comp/matrixfree.cpp:1239-1297), so a
pointwise coefficient is refused, not collapsed; FreeDofs works as the
reference's method and as an attribute.  CPU only."""
import numpy as np
import pytest

import ngsolve_amd as ngs


def _setup():
    m = ngs.MakeStructured2DMesh(nx=3, ny=3)
    fes = ngs.H1(m, order=2, dirichlet=".*")
    return m, fes


@pytest.mark.parametrize("mf", [True, False])
def test_pointwise_coefficient_refused(mf):
    m, fes = _setup()
    u, v = fes.TnT()
    kappa = 1 + ngs.x * ngs.x   # varies inside every element
    a = ngs.BilinearForm(kappa * ngs.grad(u) * ngs.grad(v) * ngs.dx, mf=ngs.MFOpts() if mf else None)
    with pytest.raises(ValueError, match="element-wise"):
        a.Assemble()


def test_free_dofs_method_and_attribute():
    m, fes = _setup()
    a, b = fes.FreeDofs(), fes.FreeDofs
    assert np.array_equal(a.mask, b.mask)
    assert a.NumSet() == fes.ndof - int((~a.mask).sum()) and a.NumSet() < fes.ndof
