"""The batching configuration at its stated size: 512 independent 2D Poisson
problems from one generator (seed 5: nx ~ U{8..96}, p ~ U{1..5}, kappa_e =
exp(N(0,1)), rhs ~ U[-1,1]; 20,045,740 dofs in all), solved together on the
device to 1e-10 -- FP64 block CG (graph replays and one persistent launch)
and FP64 refinement over FP32 inner solves.
Every problem is checked against the oracle's Jacobi-PCG on that problem
alone (CGSolver recurrence, python/krylovspace.py:263-291): FP64 iteration
count within +-1 and solution within 1e-9; refinement: true FP64 residual
<= 1e-10 with the oracle's SpMV.  The problems are checked in four slices of
128 so that each test stays short."""
import numpy as np
import pytest

import oracle
from oracle import fe_oracle as O
from helpers import say

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

_state = {}


@pytest.fixture(scope="module", autouse=True)
def _release():
    yield
    _state.clear()  # the full-size problems hold several GB of host and device memory


def _solved():
    if "probs" in _state:
        return _state
    from ngsolve_amd.batch import BatchSolver, make_problems
    say("[full batch] building 512 problems")
    probs = make_problems(512, seed=5)
    assert sum(pr.ndof for pr in probs) == 20045740
    bs = BatchSolver(probs)
    say(f"[full batch] FP64 block CG, block_nnz={bs.block_nnz}")
    r64 = bs.solve_fp64(tol=1e-10)
    u64 = bs.solution()
    say("[full batch] FP64 block CG in one persistent launch")
    rps = bs.solve_fp64(tol=1e-10, persistent=True)
    ups = bs.solution()
    say("[full batch] mixed refinement")
    rmx = bs.solve_mixed(tol=1e-10)
    umx = bs.solution()
    _state.update(probs=probs, offsets=bs.offsets, r64=r64, u64=u64, rps=rps, ups=ups, rmx=rmx, umx=umx)
    del bs
    return _state


def test_512_problems_persistent_launch_matches_graph(device):
    """The persistent form (one cooperative launch) on all 512 problems: the
    graph form's iteration counts, residuals and solutions bit for bit."""
    S = _solved()
    assert np.array_equal(S["rps"]["iterations"], S["r64"]["iterations"])
    assert np.array_equal(S["rps"]["final_residual"], S["r64"]["final_residual"])
    assert np.array_equal(S["ups"], S["u64"])


@pytest.mark.parametrize("part", [0, 1, 2, 3])
def test_512_problems_every_problem_vs_oracle(device, part):
    S = _solved()
    probs, off = S["probs"], S["offsets"]
    worst_its, worst_err, worst_res = 0, 0.0, 0.0
    for k in range(128 * part, 128 * (part + 1)):
        pr = probs[k]
        _, inv = O.jacobi_invdiag(pr.rowptr, pr.cols, pr.vals, pr.free)
        u_ref, its_ref, _ = oracle.pcg_csr(pr.rowptr, pr.cols, pr.vals, inv, pr.rhs, tol=1e-10, maxiter=100000)
        its = int(S["r64"]["iterations"][k])
        u = S["u64"][off[k]:off[k + 1]]
        err = np.linalg.norm(u - u_ref) / max(np.linalg.norm(u_ref), 1e-300)
        assert abs(its - its_ref) <= 1, (k, pr.nx, pr.p, its, its_ref)
        assert err <= 1e-9, (k, pr.nx, pr.p, err)
        um = S["umx"][off[k]:off[k + 1]]
        fr = pr.free.astype(bool)
        res = pr.rhs - oracle.spmv(pr.rowptr, pr.cols, pr.vals, um)
        rel = np.linalg.norm(res[fr]) / np.linalg.norm(pr.rhs[fr])
        assert rel <= 1.0001e-10, (k, pr.nx, pr.p, rel)
        assert S["rmx"]["relres"][k] <= 1e-10
        worst_its = max(worst_its, abs(its - its_ref))
        worst_err, worst_res = max(worst_err, err), max(worst_res, rel)
    say(f"[full batch] problems {128 * part}..{128 * (part + 1) - 1}: FP64 its diff <= {worst_its}, "
        f"solution <= {worst_err:.1e}; mixed residual <= {worst_res:.1e}")
