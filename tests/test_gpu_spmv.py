"""CSR SpMV on the GPU against the oracle's CSR SpMV (FP64, 1e-13 relative
in the inf-norm)."""
import numpy as np
import scipy.sparse as sp
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n,p", [(2, 64, 3), (2, 1, 1), (2, 8, 5), (3, 4, 4), (3, 3, 2), (3, 1, 1)])
def test_spmv_matches_oracle(device, dim, n, p):
    say(f"[spmv] dim={dim} n={n} p={p}")
    P = problem(dim, n, p, seed_kappa=2)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    y0 = np.random.default_rng(7).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), dvec(y0)
    A.mult(dx, dy)
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    assert relerr_inf(dy.download(), ref) < 1e-13
    dy.upload(y0)
    A.mult_add(-0.5, dx, dy)
    ref2 = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x, y0.copy(), s=-0.5, accumulate=True)
    assert relerr_inf(dy.download(), ref2) < 1e-13


def test_spmv_oracle_assembled_matrix(device):
    """The device SpMV applied to the oracle's own assembly (64x64 trigs, p=3)."""
    say("[spmv] oracle-assembled 64x64 trigs p=3")
    pr = O.Problem(O.structured_2d(64), 3, intorder=6)
    rp, c, v = pr.assemble()
    A = N.CreateCSR(pr.ndof, pr.ndof, rp, c, v)
    x = np.random.default_rng(0).uniform(-1, 1, pr.ndof)
    dx, dy = dvec(x), N.DeviceVector(pr.ndof, 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(rp, c, v, x)) < 1e-13


def test_spmv_empty_rows_and_rectangular(device):
    say("[spmv] empty rows / rectangular")
    rp = np.array([0, 0, 2, 2, 5, 5], dtype=np.int64)
    cols = np.array([1, 6, 0, 3, 6], dtype=np.int32)
    vals = np.array([1.5, -2.0, 3.0, 4.0, 0.25])
    A = N.CreateCSR(5, 7, rp, cols, vals)
    x = np.arange(7, dtype=float) - 2
    dx, dy = dvec(x), dvec(np.full(5, 9.0))
    A.mult(dx, dy)
    assert np.array_equal(dy.download(), oracle.spmv(rp, cols, vals, x))


def test_spmv_long_rows(device):
    """Rows far longer than a lane group (dense 3 x 5000 block)."""
    say("[spmv] long rows")
    rng = np.random.default_rng(4)
    n, w = 3, 5000
    rp = np.arange(0, (n + 1) * w, w, dtype=np.int64)
    cols = np.tile(np.arange(w, dtype=np.int32), n)
    vals = rng.standard_normal(n * w)
    A = N.CreateCSR(n, w, rp, cols, vals)
    x = rng.standard_normal(w)
    dx, dy = dvec(x), N.DeviceVector(n, 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(rp, cols, vals, x)) < 1e-13


def test_spmv_rejects_bad_input(device):
    with pytest.raises(RuntimeError):
        N.CreateCSR(2, 2, np.array([0, 1, 2], dtype=np.int64), np.array([0, 5], dtype=np.int32),
                    np.array([1.0, 1.0]))


@pytest.mark.slow
def test_spmv_full_size_3d(device):
    """32^3 Kuhn tets, p=4 (2,146,689 dofs, 159,230,209 nonzeros), x ~ U[-1,1]
    seed 1: device SpMV vs the oracle's CSR SpMV on the same matrix."""
    say("[spmv] full size 32^3 p=4 (host assembly takes a few seconds)")
    P = problem(3, 32, 4, order=8)
    assert P["rowptr"][-1] == 159230209
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    y = dy.download()
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    assert relerr_inf(y, ref) < 1e-13
    # linearity at full size: A(2x) = 2 A x exactly
    dx.upload(2 * x)
    A.mult(dx, dy)
    assert np.array_equal(dy.download(), 2 * y)


ALGOS = ["runs", "stream", "rowgroup"]


def _select(monkeypatch, algo):
    monkeypatch.setenv("NGSAMD_SPMV", algo)


def _runs_matrix(rng):
    """Rows in runs of equal column lists: lengths 1..12 (runs longer than 8
    rows, or with more than 2048 values, are split), empty rows (9000 in a
    row: more rows than a block may hold), a row of 5000 nonzeros and runs
    of 1000-column rows."""
    w = 6000
    rows = []
    for R, L in [(1, 7), (3, 40), (12, 9), (0, 0), (5, 1000), (2, 5000), (4, 0), (6, 33), (1, 2049), (8, 256),
                 (9, 1), (3, 2048), (9000, 0), (1, 1)]:
        cols = np.sort(rng.choice(w, L, replace=False)).astype(np.int32)
        rows += [cols] * max(R, 1)
    rp = np.zeros(len(rows) + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    cols = np.concatenate(rows).astype(np.int32)
    return len(rows), w, rp, cols, rng.standard_normal(len(cols))


@pytest.mark.parametrize("algo", ALGOS)
def test_spmv_kernels_on_fe_matrices(device, monkeypatch, algo):
    """Each SpMV kernel on 3D p=4 and 2D p=6 (face and cell runs of 10 dofs
    are split) matrices, plus the fused <x, Ax> path through CG."""
    _select(monkeypatch, algo)
    for dim, n, p in [(3, 4, 4), (2, 6, 6)]:
        say(f"[spmv] kernel {algo} dim={dim} n={n} p={p}")
        P = problem(dim, n, p, seed_kappa=3)
        A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
        x = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
        y0 = np.random.default_rng(8).uniform(-1, 1, P["ndof"])
        dx, dy = dvec(x), dvec(y0)
        A.mult_add(0.75, dx, dy)
        ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x, y0.copy(), s=0.75, accumulate=True)
        assert relerr_inf(dy.download(), ref) < 1e-13
        d = N.DeviceVector(P["ndof"], 8)
        A.diagonal(d)
        dref = sp.csr_matrix((P["vals"], P["cols"], P["rowptr"]), shape=(P["ndof"], P["ndof"])).diagonal()
        assert np.array_equal(d.download(), dref)


@pytest.mark.parametrize("algo", ALGOS)
def test_spmv_kernels_on_runs_edge_cases(device, monkeypatch, algo):
    _select(monkeypatch, algo)
    say(f"[spmv] kernel {algo} run edge cases")
    rng = np.random.default_rng(11)
    n, w, rp, cols, vals = _runs_matrix(rng)
    A = N.CreateCSR(n, w, rp, cols, vals)
    x = rng.standard_normal(w)
    dx, dy = dvec(x), dvec(np.full(n, 3.0))
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), oracle.spmv(rp, cols, vals, x)) < 1e-13


@pytest.mark.parametrize("algo", ALGOS)
def test_spmv_zero_width(device, monkeypatch, algo):
    """A 3 x 0 matrix (all rows empty): y = A x = 0 with an empty x."""
    _select(monkeypatch, algo)
    say(f"[spmv] kernel {algo} zero width")
    A = N.CreateCSR(3, 0, np.zeros(4, dtype=np.int64), np.zeros(0, dtype=np.int32), np.zeros(0))
    dx, dy = N.DeviceVector(0, 8), dvec(np.full(3, 5.0))
    A.mult(dx, dy)
    assert np.array_equal(dy.download(), np.zeros(3))
