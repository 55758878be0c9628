"""Fused matrix-free operator on the GPU against the oracle: FP64 within
1e-13 relative in the inf-norm of the assembled product, FP32 within 1e-5,
both scatter modes, the three MFMA kernel forms (factored reference
matrices, the default; the sum over the reference matrices; per quadrature
point), split FP16, the diagonal and the element gradient kernel."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, mf_operator, problem, relerr_inf, say

pytestmark = pytest.mark.gpu

CASES = [(2, 6, 1, None), (2, 6, 2, None), (2, 8, 3, 6), (2, 5, 4, None), (2, 5, 5, None),
         (3, 3, 1, None), (3, 3, 2, None), (3, 3, 3, None), (3, 3, 4, 8), (3, 2, 4, None), (3, 2, 5, None)]


KERNELS = ["fac", "ref", "quad"]


def _select(monkeypatch, kernel):
    monkeypatch.setenv("NGSAMD_MF_KERNEL", kernel)


@pytest.mark.parametrize("dim,n,p,order", CASES)
@pytest.mark.parametrize("scatter", [0, 1])
@pytest.mark.parametrize("kernel", KERNELS)
def test_mf_f64_matches_oracle(device, monkeypatch, dim, n, p, order, scatter, kernel):
    _select(monkeypatch, kernel)
    say(f"[mf f64 {kernel}] dim={dim} n={n} p={p} order={order} scatter={scatter}")
    P = problem(dim, n, p, order=order, seed_kappa=3)
    A = mf_operator(P, precision=0, scatter=scatter)
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    assert relerr_inf(dy.download(), ref) < 1e-13
    y0 = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
    dy.upload(y0)
    A.mult_add(2.5, dx, dy)
    assert relerr_inf(dy.download(), y0 + 2.5 * ref) < 1e-13


def test_mf_oracle_btdb_apply(device):
    """Device apply vs the oracle's own per-point F^-T D F^-1 restatement."""
    say("[mf f64] vs oracle BTDB")
    P = problem(3, 3, 4, order=8, seed_kappa=4)
    pr = O.Problem(P["omesh"], 4, intorder=8)
    cls = O.class_number(3, pr.mesh["els"])
    uniq = np.unique(cls)
    _, g = pr.shapes()
    bref = np.stack([g[np.where(cls == k)[0][0]] for k in uniq])
    x = np.random.default_rng(5).uniform(-1, 1, pr.ndof)
    ref = oracle.mf_apply(3, bref, np.searchsorted(uniq, cls), pr.wts, pr.dofs, pr.Finv, pr.detF, P["kappa"], x)
    A = mf_operator(P)
    dx, dy = dvec(x), N.DeviceVector(pr.ndof, 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), ref) < 1e-13


# FP32: 1e-5 relative in the inf-norm.  BF16 (8-bit significand operands,
# FP32 accumulation): 3e-2; FP16 (11-bit significand): 4e-3.  The refinement
# tests show both still drive the FP64 solve to 1e-10.
@pytest.mark.parametrize("precision,tol", [(1, 1e-5), (2, 3e-2), (3, 4e-3)])
@pytest.mark.parametrize("dim,n,p,order", [(2, 8, 3, 6), (3, 3, 4, 8), (3, 2, 2, None), (2, 5, 5, None),
                                           (3, 2, 3, None)])
@pytest.mark.parametrize("scatter", [0, 1])
@pytest.mark.parametrize("kernel", KERNELS)
def test_mf_low_precision(device, monkeypatch, precision, tol, dim, n, p, order, scatter, kernel):
    _select(monkeypatch, kernel)
    say(f"[mf prec={precision} {kernel}] dim={dim} n={n} p={p} scatter={scatter}")
    P = problem(dim, n, p, order=order, seed_kappa=3)
    A = mf_operator(P, precision=precision, scatter=scatter)
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), ref) < tol
    x32, y32 = dvec(x, 4), N.DeviceVector(P["ndof"], 4)
    A.mult_f32(x32, y32)
    assert relerr_inf(y32.download().astype(np.float64), ref) < tol


# Split FP16 (three FP16 products per contraction): FP32's tolerance, in the
# factored kernels (the only ones that take it).
@pytest.mark.parametrize("dim,n,p,order", [(2, 8, 3, 6), (3, 3, 4, 8), (3, 2, 2, None), (2, 5, 5, None),
                                           (2, 6, 1, None), (3, 3, 1, None), (3, 2, 3, None)])
@pytest.mark.parametrize("scatter", [0, 1])
def test_mf_split_f16(device, dim, n, p, order, scatter):
    say(f"[mf f16x3] dim={dim} n={n} p={p} scatter={scatter}")
    P = problem(dim, n, p, order=order, seed_kappa=3)
    A = mf_operator(P, precision=4, scatter=scatter)
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    err = relerr_inf(dy.download(), ref)
    x32, y32 = dvec(x, 4), N.DeviceVector(P["ndof"], 4)
    A.mult_f32(x32, y32)
    err32 = relerr_inf(y32.download().astype(np.float64), ref)
    say(f"  rel err {err:.2e} (FP32 vectors {err32:.2e})")
    assert err < 1e-5 and err32 < 1e-5
    # and far below plain FP16's error on the same operator
    A16 = mf_operator(P, precision=3, scatter=scatter)
    A16.mult(dx, dy)
    assert err < 0.05 * relerr_inf(dy.download(), ref)


def test_mf_split_f16_needs_factored_kernels(device, monkeypatch):
    monkeypatch.setenv("NGSAMD_MF_KERNEL", "ref")
    P = problem(2, 4, 2, seed_kappa=3)
    with pytest.raises(RuntimeError, match="F16X3"):
        mf_operator(P, precision=4)


def _mf_inputs(P):
    """The arrays mf_operator passes to CreateMatrixFree, per class."""
    fes, ir, dim = P["fes"], P["ir"], P["dim"]
    classes = N.BuildClasses(fes, ir)
    detj, ginv = N.ComputeGeometry(P["mesh"])
    out = []
    for c in classes:
        el = np.asarray(c["elements"], dtype=np.int32)
        G = detj[el, None, None] * ginv[el]
        if dim == 2:
            met = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 0, 1]], axis=1)
        else:
            met = np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 2, 2], G[:, 0, 1], G[:, 0, 2], G[:, 1, 2]], axis=1)
        out.append((el, np.asarray(c["bmat"]).ravel(), met, fes.el_dofs[el]))
    return out


def _create(P, parts, precision):
    fes, ir, dim = P["fes"], P["ir"], P["dim"]
    off = np.cumsum([0] + [len(p[0]) for p in parts]).astype(np.int64)
    el = np.ascontiguousarray(np.concatenate([p[0] for p in parts]).astype(np.int32))
    return N.CreateMatrixFree(dim, fes.nd_el, len(ir), fes.ndof, off,
                              np.ascontiguousarray(np.concatenate([p[1] for p in parts])),
                              np.ascontiguousarray(ir.weights),
                              np.ascontiguousarray(np.concatenate([p[3] for p in parts]).astype(np.int32)),
                              np.ascontiguousarray(np.concatenate([p[2] for p in parts])), None, el, precision, 0)


# Ragged inputs: a class without elements between two others gives the same
# bits (rounds and tiles skip it); an operator without elements applies 0.
@pytest.mark.parametrize("precision", [0, 2])
def test_mf_empty_class_and_no_elements(device, precision):
    say(f"[mf empty class] precision={precision}")
    P = problem(3, 2, 3, seed_kappa=None)
    parts = _mf_inputs(P)
    assert len(parts) >= 2
    A = _create(P, parts, precision)
    empty = (np.zeros(0, np.int32), parts[0][1], parts[0][2][:0], parts[0][3][:0])
    B = _create(P, parts[:1] + [empty] + parts[1:], precision)
    x = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
    dx, ya, yb = dvec(x), N.DeviceVector(P["ndof"], 8), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, ya)
    B.mult(dx, yb)
    assert np.array_equal(ya.download(), yb.download())
    Z = _create(P, [empty], precision)
    y0 = np.random.default_rng(3).uniform(-1, 1, P["ndof"])
    dz = dvec(y0)
    Z.mult(dx, dz)
    assert not np.any(dz.download())
    dz = dvec(y0)
    Z.mult_add(2.0, dx, dz)
    assert np.array_equal(dz.download(), y0)


@pytest.mark.parametrize("dim,n,p", [(2, 6, 3), (3, 3, 4)])
def test_mf_diagonal(device, dim, n, p):
    say(f"[mf diag] dim={dim} p={p}")
    P = problem(dim, n, p, seed_kappa=6)
    A = mf_operator(P)
    d = N.DeviceVector(P["ndof"], 8)
    A.diagonal(d)
    ref, _ = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], np.ones(P["ndof"], dtype=np.uint8))
    assert relerr_inf(d.download(), ref) < 1e-13


@pytest.mark.parametrize("dim,n,p", [(2, 6, 3), (3, 2, 4)])
def test_mf_element_gradient(device, dim, n, p):
    """g_e = -lam_e^T (K_e / kappa_e) u_e against the oracle's element matrices."""
    say(f"[mf grad] dim={dim} p={p}")
    P = problem(dim, n, p, seed_kappa=7)
    A = mf_operator(P)
    pr = O.Problem(P["omesh"], p, intorder=P["order"])
    K = pr.element_matrices()  # kappa = 1
    rng = np.random.default_rng(9)
    lam, u = rng.uniform(-1, 1, pr.ndof), rng.uniform(-1, 1, pr.ndof)
    ref = -np.einsum("ei,eij,ej->e", lam[pr.dofs], K, u[pr.dofs])
    g = N.DeviceVector(P["mesh"].ne, 8)
    A.element_grad(dvec(lam), dvec(u), g)
    assert relerr_inf(g.download(), ref) < 1e-12


def test_mf_set_kappa(device):
    say("[mf] set_kappa")
    P = problem(2, 6, 3, seed_kappa=None)
    A = mf_operator(P)
    kap = np.exp(np.random.default_rng(11).standard_normal(P["mesh"].ne))
    dk = dvec(kap)
    A.set_kappa(dk)
    vals = N.AssembleCSR(P["fes"], P["rowptr"], P["cols"], P["ir"], True, kap, False, None)
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(P["rowptr"], P["cols"], vals, x)) < 1e-13


def test_mf_coefficient_and_gradient_sizes_checked(device):
    """A coefficient or gradient vector whose length is not the element count
    is refused before any kernel reads or writes it; a coefficient handed over
    as a temporary stays alive while the operator holds it."""
    import gc
    say("[mf] coefficient / gradient sizes")
    P = problem(2, 6, 3, seed_kappa=None)
    A = mf_operator(P)
    ne = P["mesh"].ne
    assert A.ne == ne
    with pytest.raises(RuntimeError):
        A.set_kappa(N.DeviceVector(ne - 1, 8))
    with pytest.raises(RuntimeError):
        A.set_kappa(N.DeviceVector(ne, 4))
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx = dvec(x)
    with pytest.raises(RuntimeError):
        A.element_grad(dx, dx, N.DeviceVector(ne - 3, 8))
    kap = np.exp(np.random.default_rng(12).standard_normal(ne))
    A.set_kappa(dvec(kap))           # the only reference is the operator's
    gc.collect()
    filler = [dvec(np.full(ne, 1e300)) for _ in range(4)]  # reuse freed memory, if any
    dy = N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    vals = N.AssembleCSR(P["fes"], P["rowptr"], P["cols"], P["ir"], True, kap, False, None)
    assert relerr_inf(dy.download(), oracle.spmv(P["rowptr"], P["cols"], vals, x)) < 1e-13
    del filler


def test_mf_cg_matches_csr_cg(device):
    """Same operator, same recurrence: CG iteration counts agree within +-1."""
    say("[mf] CG vs CSR CG")
    P = problem(3, 3, 4, order=8, seed_kappa=2)
    A = mf_operator(P)
    Acsr = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    d = N.DeviceVector(P["ndof"], 8)
    A.diagonal(d)
    C = N.CreateJacobiDevice(d, P["fes"].freedofs)
    f = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    df = dvec(f)
    u1, u2 = N.DeviceVector(P["ndof"], 8), N.DeviceVector(P["ndof"], 8)
    its1, _ = N.CGSolve(A, C, df, u1, True, 1e-10, -1.0, 20000, 16)
    its2, _ = N.CGSolve(Acsr, C, df, u2, True, 1e-10, -1.0, 20000, 16)
    say(f"[mf] its mf={its1} csr={its2}")
    assert abs(its1 - its2) <= 1
    a, b = u1.download(), u2.download()
    assert np.linalg.norm(a - b) <= 1e-10 * np.linalg.norm(b)
