"""ngsamd_cg_solve_noalloc (the allocation-free CG) through the C ABI against
ngsamd_cg_solve on the same operators: two calls in a row, the second on a
smaller problem (the scratch array keeps the first call's contents), CSR and
matrix-free operators, Jacobi and block-Jacobi preconditioners; sizes beyond
the scratch array's limits are refused before any kernel runs."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from helpers import N, dvec, mf_operator, problem, say
from oracle import fe_oracle as O

pytestmark = pytest.mark.gpu

_lib = None


def lib():
    global _lib
    if _lib is None:
        # the library the _ngsamd module already loaded (same path, same handle)
        _lib = ctypes.CDLL(os.path.join(ROOT, "ngsolve_amd", "libngsamd.so"))
        f = _lib.ngsamd_cg_solve_noalloc
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double)]
        _lib.ngsamd_last_error.restype = ctypes.c_char_p
    return _lib


def noalloc(A, C, f, u, initialize=True, tol=1e-10, atol=-1.0, maxiter=5000):
    its = ctypes.c_int(0)
    res = np.zeros(maxiter)
    rc = lib().ngsamd_cg_solve_noalloc(A.handle, C.handle, f.ptr, u.ptr, 1 if initialize else 0, tol, atol,
                                       maxiter, ctypes.byref(its), res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return rc, its.value, res[:min(its.value, maxiter)]


def _case(dim, n, p, seed, op):
    say(f"[noalloc] dim {dim} n {n} p {p} seed {seed} op {op}")
    P = problem(dim, n, p, order=2 * p, seed_kappa=seed)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"]) if op == "csr" else mf_operator(P, 0)
    free = P["fes"].freedofs.astype(np.uint8)
    diag, _ = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, np.random.default_rng(seed + 10).uniform(-1, 1, P["ndof"]), 0.0)
    return P, A, C, f


@pytest.mark.parametrize("op", ["csr", "mf"])
def test_noalloc_matches_cg_solve_twice(device, op):
    # the larger problem first, then a smaller one on the same scratch
    for dim, n, p, seed in ((3, 6, 3, 2), (2, 9, 2, 3)):
        P, A, C, f = _case(dim, n, p, seed, op)
        df = dvec(f)
        u_ref, u = N.DeviceVector(P["ndof"], 8), N.DeviceVector(P["ndof"], 8)
        its_ref, res_ref = N.CGSolve(A, C, df, u_ref, True, 1e-10, -1.0, 5000, 16)
        u.fill(7.0)  # overwritten: initialize=1 starts from 0
        rc, its, res = noalloc(A, C, df, u)
        assert rc == 0, lib().ngsamd_last_error()
        ur, ua = u_ref.download(), u.download()
        err = np.abs(ua - ur).max() / np.abs(ur).max()
        say(f"[noalloc] {op} ndof {P['ndof']}: its {its} (cg_solve {its_ref}), solution diff {err:.2e}")
        assert abs(its - its_ref) <= 1
        assert err <= 1e-10
        assert np.allclose(res[:len(res_ref)], res_ref[:len(res)], rtol=1e-10, atol=0)


def test_noalloc_initial_guess_and_block_jacobi(device):
    """initialize = 0 continues from the given u (d = f - A u); a block-Jacobi
    preconditioner (one block per mesh node) takes the generic path."""
    P, A, C, f = _case(2, 8, 3, 4, "csr")
    free = P["fes"].freedofs.astype(bool)
    df = dvec(f)
    u0 = np.where(free, np.random.default_rng(5).uniform(-0.1, 0.1, P["ndof"]), 0.0)
    sp, nv = P["fes"], P["mesh"].nv
    starts = np.concatenate([np.arange(nv + 1), np.asarray(sp.first_edge_dof)[1:], np.asarray(sp.first_el_dof)[1:]])
    blocks = [b[free[b]] for b in np.split(np.arange(P["ndof"]), starts[1:-1])]
    blocks = [b for b in blocks if len(b)]
    bptr = np.zeros(len(blocks) + 1, dtype=np.int64)
    bptr[1:] = np.cumsum([len(b) for b in blocks])
    BJ = N.CreateBlockJacobi(P["ndof"], P["rowptr"], P["cols"], P["vals"], bptr,
                             np.concatenate(blocks).astype(np.int32))
    for pre, name in ((C, "jacobi"), (BJ, "block jacobi")):
        u_ref, u = dvec(u0), dvec(u0)
        its_ref, _ = N.CGSolve(A, pre, df, u_ref, False, 1e-10, -1.0, 5000, 16)
        rc, its, _ = noalloc(A, pre, df, u, initialize=False)
        assert rc == 0, lib().ngsamd_last_error()
        ur = u_ref.download()
        err = np.abs(u.download() - ur).max() / np.abs(ur).max()
        say(f"[noalloc] {name}, initial guess: its {its} (cg_solve {its_ref}), solution diff {err:.2e}")
        assert abs(its - its_ref) <= 1
        assert err <= 1e-10


def test_noalloc_refuses_sizes_beyond_scratch(device):
    P, A, C, f = _case(2, 4, 1, 6, "csr")
    df, u = dvec(f), N.DeviceVector(P["ndof"], 8)
    rc, _, _ = noalloc(A, C, df, u, maxiter=70000)
    assert rc == 2 and b"maxiter" in lib().ngsamd_last_error()
    # an operator of more than 4,194,304 dofs: a diagonal CSR matrix (no solve runs)
    n = (1 << 22) + 1
    rp = np.arange(n + 1, dtype=np.int64)
    big = N.CreateCSR(n, n, rp, np.arange(n, dtype=np.int32), np.ones(n))
    Cb = N.CreateJacobi(np.ones(n), None)
    fb, ub = N.DeviceVector(n, 8), N.DeviceVector(n, 8)
    rc, _, _ = noalloc(big, Cb, fb, ub)
    assert rc == 2 and b"4194304" in lib().ngsamd_last_error()
