"""The colour-ordered scatter of the matrix-free operator
(ngsamd_mf_set_coloring; the reference's coloured element loop,
comp/fespace.cpp:518-760 colouring, comp/bilinearform.cpp:4669-4750 apply)
against the oracle: FP64 apply within 1e-13 of the oracle's CSR product,
FP32/BF16/FP16 within their bounds, 2D and 3D, p = 1..4; the apply equals
the gather scatter's to rounding; Mult / MultAdd semantics; CG on the
colour-scatter operator against the oracle's PCG (iterations within 1,
solution within 1e-10); a colouring whose elements of one colour share a dof
is refused."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, mf_operator, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


def _colors_table_order(P):
    classes = N.BuildClasses(P["fes"], P["ir"])
    el = np.concatenate([c["elements"] for c in classes]).astype(np.int64)
    return np.ascontiguousarray(N.ElementColoring(P["mesh"])[el], dtype=np.int32)


@pytest.mark.parametrize("dim,n,p", [(2, 9, 1), (2, 8, 3), (2, 6, 5), (3, 4, 2), (3, 4, 4)])
@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 1e-5), (2, 3e-2), (3, 4e-3), (4, 1e-5)])
def test_color_scatter_apply(device, dim, n, p, precision, tol):
    say(f"[coloring] dim {dim} n {n} p {p} precision {precision}")
    P = problem(dim, n, p, seed_kappa=2)
    A = mf_operator(P, precision)
    A.set_coloring(_colors_table_order(P))
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    dy.fill(5.0)
    A.mult(dx, dy)
    err = relerr_inf(dy.download(), ref)
    say(f"[coloring] apply vs oracle {err:.2e}")
    assert err < tol
    y0 = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
    dy.upload(y0)
    A.mult_add(2.0, dx, dy)
    assert relerr_inf(dy.download(), y0 + 2.0 * ref) < 2 * tol
    if precision == 0:
        G = mf_operator(P, 0)
        dg = N.DeviceVector(P["ndof"], 8)
        G.mult(dx, dg)
        A.mult(dx, dy)
        assert relerr_inf(dy.download(), dg.download()) < 1e-14


def test_color_scatter_cg(device):
    P = problem(3, 5, 3, seed_kappa=4)
    A = mf_operator(P, 0)
    A.set_coloring(_colors_table_order(P))
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, 1.0, 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 3000, 16)
    u_ref, its_ref, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-10, maxiter=3000)
    err = np.linalg.norm(du.download() - u_ref) / np.linalg.norm(u_ref)
    say(f"[coloring] CG its {its} oracle {its_ref}, solution {err:.2e}")
    assert abs(its - its_ref) <= 1 and err <= 1e-10


def test_bad_coloring_refused(device):
    P = problem(2, 4, 2)
    A = mf_operator(P, 0)
    with pytest.raises(RuntimeError, match="share a dof"):
        A.set_coloring(np.zeros(P["mesh"].ne, dtype=np.int32))
