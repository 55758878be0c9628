"""Integrate of coefficient functions (host quadrature, weights |det F| w_q):
closed forms on the unit square and cube, element-wise sums."""
import numpy as np
import pytest

import ngsolve_amd as ngs


def test_integrate_closed_forms():
    m2 = ngs.MakeStructured2DMesh(nx=6, ny=5)
    assert abs(ngs.Integrate(1, m2) - 1.0) < 1e-13
    assert abs(ngs.Integrate(ngs.x * ngs.y, m2) - 0.25) < 1e-13
    assert abs(ngs.Integrate(ngs.x ** 3, m2, order=5) - 0.25) < 1e-13
    m3 = ngs.MakeStructured3DMesh(nx=3)
    assert abs(ngs.Integrate(1, m3) - 1.0) < 1e-13
    assert abs(ngs.Integrate(ngs.x * ngs.y * ngs.z, m3) - 0.125) < 1e-13
    ew = ngs.Integrate(1, m3, element_wise=True)
    assert ew.shape == (m3.ne,) and np.allclose(ew, 1.0 / m3.ne, rtol=1e-13)

