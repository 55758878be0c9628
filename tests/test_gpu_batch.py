"""Batched independent solves (the batching configuration at reduced size):
every problem's iteration count and solution against the oracle's CG on
that problem alone, FP64 and mixed precision (FP32 inner, FP64 refinement)."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import say

pytestmark = pytest.mark.gpu


def _problems():
    from ngsolve_amd.batch import make_problems
    return make_problems(24, seed=5, nx_range=(2, 20), p_range=(1, 5))


@pytest.mark.parametrize("runs", ["1", "0"])
def test_batch_fp64_matches_per_problem_cg(device, monkeypatch, runs):
    from ngsolve_amd.batch import BatchSolver
    monkeypatch.setenv("NGSAMD_BATCH_RUNS", runs)
    say(f"[batch] fp64 24 problems, runs layout {runs}")
    probs = _problems()
    bs = BatchSolver(probs)
    r = bs.solve_fp64(tol=1e-10)
    say(f"[batch] block_nnz={bs.block_nnz} its={r['iterations'].tolist()}")
    for k, pr in enumerate(probs):
        _, inv = O.jacobi_invdiag(pr.rowptr, pr.cols, pr.vals, pr.free)
        u_ref, its_ref, _ = oracle.pcg_csr(pr.rowptr, pr.cols, pr.vals, inv, pr.rhs, tol=1e-10, maxiter=100000)
        u = bs.solution(k)
        assert abs(int(r["iterations"][k]) - its_ref) <= 1, (k, r["iterations"][k], its_ref)
        assert np.linalg.norm(u - u_ref) <= 1e-9 * max(np.linalg.norm(u_ref), 1e-300)
        assert r["finish_ms"][k] >= 0


@pytest.mark.parametrize("runs", ["1", "0"])
def test_batch_mixed_reaches_1e10(device, monkeypatch, runs):
    from ngsolve_amd.batch import BatchSolver
    monkeypatch.setenv("NGSAMD_BATCH_RUNS", runs)
    say(f"[batch] mixed 24 problems, runs layout {runs}")
    probs = _problems()
    bs = BatchSolver(probs)
    r = bs.solve_mixed(tol=1e-10)
    say(f"[batch] outer={r['outer'].tolist()} relres max={r['relres'].max():.2e}")
    assert (r["relres"] <= 1e-10).all()
    for k, pr in enumerate(probs):
        u = bs.solution(k)
        res = pr.rhs - oracle.spmv(pr.rowptr, pr.cols, pr.vals, u)
        fr = pr.free.astype(bool)
        assert np.linalg.norm(res[fr]) <= 1.0001e-10 * np.linalg.norm(pr.rhs[fr])


@pytest.mark.parametrize("runs,nnz", [("1", None), ("0", None), ("1", "1024"), ("0", "4096")])
def test_batch_persistent_matches_graph(device, monkeypatch, runs, nnz):
    """All iterations in one cooperative launch (grid-wide barriers between
    the phases of an iteration) give the graph form's iterates bit for bit:
    iteration counts, last residuals and solutions, for both layouts and
    other block sizes, to convergence and with a maxiter stop."""
    from ngsolve_amd.batch import BatchSolver
    monkeypatch.setenv("NGSAMD_BATCH_RUNS", runs)
    if nnz:
        monkeypatch.setenv("NGSAMD_SPMV_NNZ", nnz)
    else:
        monkeypatch.delenv("NGSAMD_SPMV_NNZ", raising=False)
    say(f"[batch] persistent vs graph, runs layout {runs}, block_nnz {nnz}")
    bs = BatchSolver(_problems())
    for maxiter in (100000, 7):
        g = bs.solve_fp64(tol=1e-10, maxiter=maxiter)
        ug = bs.solution()
        q = bs.solve_fp64(tol=1e-10, maxiter=maxiter, persistent=True)
        uq = bs.solution()
        say(f"[batch] maxiter {maxiter}: its {q['iterations'].tolist()}")
        assert np.array_equal(g["iterations"], q["iterations"])
        assert np.array_equal(g["final_residual"], q["final_residual"])
        assert np.array_equal(ug, uq)
        assert (q["finish_ms"] >= 0).all()
        if maxiter == 7:
            assert (q["iterations"] <= 7).all()


def test_batch_rejects_bad_chunk(device):
    from ngsolve_amd.batch import BatchSolver
    bs = BatchSolver(_problems()[:3])
    with pytest.raises(RuntimeError, match="chunk"):
        bs.solve_fp64(chunk=-2)


def test_batch_rejects_coupled_problems(device):
    from ngsolve_amd import _ngsamd as N
    rp = np.array([0, 2, 3], dtype=np.int64)
    cols = np.array([0, 1, 1], dtype=np.int32)   # row 0 couples to problem 1
    with pytest.raises(RuntimeError):
        N.CreateBatchCSR(rp, cols, np.ones(3), np.array([0, 1, 2], dtype=np.int64), False)


def test_batch_mixed_device_outer_steps(device, monkeypatch):
    """Each problem takes its refinement steps on the device as soon as its
    inner solve stops: the same residual bound as the host-driven outer loop
    (NGSAMD_BATCH_IR_HOST=1, every problem finishing at the last outer
    check), and the problems finish at different times."""
    from ngsolve_amd.batch import BatchSolver
    probs = _problems()
    out = {}
    for mode in ("device", "host"):
        if mode == "host":
            monkeypatch.setenv("NGSAMD_BATCH_IR_HOST", "1")
        else:
            monkeypatch.delenv("NGSAMD_BATCH_IR_HOST", raising=False)
        bs = BatchSolver(probs)
        r = bs.solve_mixed(tol=1e-10)
        out[mode] = r
        for k, pr in enumerate(probs):
            u = bs.solution(k)
            res = pr.rhs - oracle.spmv(pr.rowptr, pr.cols, pr.vals, u)
            fr = pr.free.astype(bool)
            assert np.linalg.norm(res[fr]) <= 1.0001e-10 * np.linalg.norm(pr.rhs[fr]), (mode, k)
        say(f"[batch] {mode} outer steps: outer={r['outer'].tolist()} finish ms "
            f"{np.min(r['finish_ms']):.2f}..{np.max(r['finish_ms']):.2f}")
    ms = np.asarray(out["device"]["finish_ms"])
    assert (ms >= 0).all() and ms.min() < ms.max()
    assert (out["device"]["outer"] >= 1).all() and (out["device"]["inner"] >= 1).all()


def test_batch_autotune_runs_layout(device, monkeypatch):
    """The runs layout's block size is autotuned over 1024 / 2048 / 4096
    values per block (all three timed, the fastest kept); NGSAMD_SPMV_NNZ
    forces one.  The CG result is the same for every block size."""
    from ngsolve_amd import _ngsamd as N
    from ngsolve_amd.batch import BatchSolver
    probs = _problems()
    monkeypatch.delenv("NGSAMD_SPMV_NNZ", raising=False)
    bs = BatchSolver(probs)
    tune = N.BatchTuning(bs.op)
    say(f"[batch] autotune {tune} -> {bs.block_nnz}")
    assert sorted(tune) == [1024, 2048, 4096] and bs.block_nnz in tune
    assert bs.block_nnz == min(tune, key=tune.get)
    its = {}
    for bn in ("1024", "4096"):
        monkeypatch.setenv("NGSAMD_SPMV_NNZ", bn)
        b2 = BatchSolver(probs)
        assert b2.block_nnz == int(bn) and N.BatchTuning(b2.op) == {}
        its[bn] = b2.solve_fp64(tol=1e-10)["iterations"]
    assert np.array_equal(its["1024"], its["4096"])
