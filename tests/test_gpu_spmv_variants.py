"""The SpMV tuning switches: values per block (NGSAMD_SPMV_NNZ 1024, 2048,
4096) and workgroup counts (NGSAMD_SPMV_GRID, from 8, where each workgroup
walks many blocks, to 65536) of the runs and stream kernels, in the plain
apply, the accumulating apply and the CG variant with the fused <x, A x>
partials, against the oracle (1e-13); and the refinement solver without
graph capture (NGSAMD_CG_CHUNK=0) against its graph-captured run."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import N, dvec, mf_operator, problem, relerr_inf, say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("algo", ["runs", "stream"])
@pytest.mark.parametrize("nnz", ["1024", "2048", "4096"])
@pytest.mark.parametrize("grid", ["8", "1000", "65536"])
def test_spmv_block_and_grid_switches(device, monkeypatch, algo, nnz, grid):
    monkeypatch.setenv("NGSAMD_SPMV", algo)
    monkeypatch.setenv("NGSAMD_SPMV_NNZ", nnz)
    monkeypatch.setenv("NGSAMD_SPMV_GRID", grid)
    say(f"[spmv {algo}] nnz/block={nnz} grid={grid}")
    P = problem(3, 5, 3, seed_kappa=2)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(1).uniform(-1, 1, P["ndof"])
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), ref) < 1e-13
    y0 = np.random.default_rng(2).uniform(-1, 1, P["ndof"])
    dy.upload(y0)
    A.mult_add(-1.5, dx, dy)
    assert relerr_inf(dy.download(), y0 - 1.5 * ref) < 1e-13
    # the CG variant (fused partials of <s, A s>) through a Jacobi-PCG solve
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, 1.0, 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 2000, 16)
    u_ref, its_ref, _ = O.pcg(lambda t: oracle.spmv(P["rowptr"], P["cols"], P["vals"], t), inv, f, tol=1e-10,
                              maxiter=2000)
    assert abs(its - its_ref) <= 1
    assert np.abs(du.download() - u_ref).max() <= 1e-10 * np.abs(u_ref).max()


@pytest.mark.parametrize("case", ["runs_16cube_p4", "stream_64cube_p1"])
def test_spmv_size_dependent_defaults(device, monkeypatch, case):
    """The default block size and grid (no NGSAMD_SPMV* switch set) where they
    differ from a one-block-per-workgroup launch: the runs kernel on 16^3 p=4
    (about 10K blocks of 2048 values: a third as many workgroups as blocks)
    and the stream kernel on P1 over 64^3 cubes (above 1024 blocks: half as
    many workgroups as blocks), plain apply, accumulating apply and the CG
    variant against the oracle."""
    for k in ("NGSAMD_SPMV", "NGSAMD_SPMV_NNZ", "NGSAMD_SPMV_GRID"):
        monkeypatch.delenv(k, raising=False)
    n, p = (16, 4) if case.startswith("runs") else (64, 1)
    say(f"[spmv defaults] {case}")
    P = problem(3, n, p, seed_kappa=5)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(3).uniform(-1, 1, P["ndof"])
    ref = oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    A.mult(dx, dy)
    assert relerr_inf(dy.download(), ref) < 1e-13
    y0 = np.random.default_rng(4).uniform(-1, 1, P["ndof"])
    dy.upload(y0)
    A.mult_add(0.5, dx, dy)
    assert relerr_inf(dy.download(), y0 + 0.5 * ref) < 1e-13
    free = P["fes"].freedofs.astype(np.uint8)
    diag, inv = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, 1.0, 0.0)
    df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 5000, 16)
    u_ref, its_ref, _ = oracle.pcg_csr(P["rowptr"], P["cols"], P["vals"], inv, f, tol=1e-10, maxiter=5000)
    say(f"[spmv defaults] {case}: its {its} oracle {its_ref}")
    assert abs(its - its_ref) <= 1
    assert np.linalg.norm(du.download() - u_ref) <= 1e-10 * np.linalg.norm(u_ref)


def test_spmv_alternating_block_walk_is_bit_identical(device):
    """Consecutive applies of the runs kernel walk the blocks in alternating
    directions: every apply gives the same bits (each row is summed by the
    same code in the same order), for Mult and MultAdd."""
    P = problem(3, 8, 4, seed_kappa=5)
    A = N.CreateCSR(P["ndof"], P["ndof"], P["rowptr"], P["cols"], P["vals"])
    x = np.random.default_rng(3).uniform(-1, 1, P["ndof"])
    dx, dy = dvec(x), N.DeviceVector(P["ndof"], 8)
    outs = []
    for _ in range(3):
        A.mult(dx, dy)
        outs.append(dy.download())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])
    assert relerr_inf(outs[0], oracle.spmv(P["rowptr"], P["cols"], P["vals"], x)) < 1e-13
    y0 = np.random.default_rng(4).uniform(-1, 1, P["ndof"])
    adds = []
    for _ in range(2):
        dy.upload(y0)
        A.mult_add(0.25, dx, dy)
        adds.append(dy.download())
    assert np.array_equal(adds[0], adds[1])


def test_ir_without_graphs_matches_graphs(device, monkeypatch):
    say("[ir] NGSAMD_CG_CHUNK=0 vs graphs")
    P = problem(3, 4, 4, order=8, seed_kappa=3)
    free = P["fes"].freedofs.astype(np.uint8)
    diag, _ = O.jacobi_invdiag(P["rowptr"], P["cols"], P["vals"], free)
    C = N.CreateJacobi(diag, free)
    f = np.where(free != 0, np.random.default_rng(4).uniform(0, 1, P["ndof"]), 0.0)
    M64, M32 = mf_operator(P, 0), mf_operator(P, 1)
    out = []
    for chunk in (None, "0"):
        if chunk is None:
            monkeypatch.delenv("NGSAMD_CG_CHUNK", raising=False)
        else:
            monkeypatch.setenv("NGSAMD_CG_CHUNK", chunk)
        df, du = dvec(f), N.DeviceVector(P["ndof"], 8)
        outer, inner, rel = N.IRCGSolve(M64, M32, C, df, du, 1e-10, -1.0, 60, 100000)
        assert rel <= 1e-10
        out.append((outer, inner, du.download()))
    # the same kernels in the same order: identical iterates
    assert out[0][0] == out[1][0] and out[0][1] == out[1][1]
    assert np.array_equal(out[0][2], out[1][2])
