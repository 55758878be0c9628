"""Many independent H1 Poisson problems solved together on the device.

make_problems() builds the batched configuration: independent 2D Poisson
problems on unit-square structured triangle meshes (homogeneous Dirichlet on
all sides) where one numpy generator draws, per problem and in this order,
nx ~ U{nx_min..nx_max}, the order p ~ U{p_min..p_max}, a per-element
coefficient kappa_e = exp(N(0,1)) and right-hand-side dof values ~ U[-1,1].
Matrices are assembled by the host FE layer (integration order 2p-2).

BatchSolver stacks them into one block-diagonal system (libngsamd's batched
operator) and solves all of them with one set of launches per iteration,
either in FP64 (ngsamd_batch_cg_solve) or with FP64 iterative refinement over
FP32 inner solves (ngsamd_batch_ir_cg_solve).
"""
import numpy as np

from ._native import N, ensure_device

ALL = 0xFFFFFFFFFFFFFFFF
PERSISTENT = -1  # NGSAMD_BATCH_PERSISTENT of ngsamd.h


class Problem:
    def __init__(self, nx, p, kappa_seed_values, rhs):
        self.nx, self.p = nx, p
        self.mesh = N.MakeStructured2DMesh(nx, nx)
        self.fes = N.H1Space(self.mesh, p, ALL)
        self.ir = N.TrigRule(max(2 * p - 2, 0))
        self.kappa = kappa_seed_values
        self.rhs = rhs
        self.rowptr, self.cols = N.BuildPattern(self.fes)
        self.vals = N.AssembleCSR(self.fes, self.rowptr, self.cols, self.ir, True, self.kappa, False, None)
        self.free = self.fes.freedofs

    @property
    def ndof(self):
        return self.fes.ndof


def make_problems(nprob, seed=5, nx_range=(8, 96), p_range=(1, 5)):
    rng = np.random.default_rng(seed)
    probs = []
    for _ in range(nprob):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        p = int(rng.integers(p_range[0], p_range[1] + 1))
        ne = 2 * nx * nx
        kappa = np.exp(rng.standard_normal(ne))
        ndof = N.H1Space(N.MakeStructured2DMesh(nx, nx), p, ALL).ndof
        rhs = rng.uniform(-1.0, 1.0, ndof)
        probs.append(Problem(nx, p, kappa, rhs))
    return probs


def stack(problems):
    """Block-diagonal CSR, concatenated rhs / free masks and problem offsets."""
    offs = np.cumsum([0] + [pr.ndof for pr in problems]).astype(np.int64)
    nnz_offs = np.cumsum([0] + [int(pr.rowptr[-1]) for pr in problems]).astype(np.int64)
    n = int(offs[-1])
    rowptr = np.empty(n + 1, dtype=np.int64)
    cols = np.empty(int(nnz_offs[-1]), dtype=np.int32)
    vals = np.empty(int(nnz_offs[-1]))
    for k, pr in enumerate(problems):
        rowptr[offs[k]:offs[k + 1]] = pr.rowptr[:-1] + nnz_offs[k]
        cols[nnz_offs[k]:nnz_offs[k + 1]] = pr.cols + offs[k]
        vals[nnz_offs[k]:nnz_offs[k + 1]] = pr.vals
    rowptr[n] = nnz_offs[-1]
    rhs = np.concatenate([pr.rhs for pr in problems])
    free = np.concatenate([pr.free for pr in problems]).astype(np.uint8)
    return offs, rowptr, cols, vals, rhs, free


def diagonal(rowptr, cols, vals):
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    d = np.zeros(n)
    m = cols == rows
    d[rows[m]] = vals[m]
    return d


class BatchSolver:
    def __init__(self, problems, with_f32=True):
        ensure_device()
        self.problems = problems
        self.offsets, rp, c, v, rhs, free = stack(problems)
        self.n = len(rhs)
        self.nnz = int(rp[-1])
        self.op = N.CreateBatchCSR(rp, c, v, self.offsets, with_f32)
        self.pre = N.CreateJacobi(diagonal(rp, c, v), free)
        self.rhs = N.DeviceVector(self.n, 8)
        self.rhs.upload(rhs)
        self.sol = N.DeviceVector(self.n, 8)
        self.nseg, self.block_nnz, self.nblocks = N.BatchInfo(self.op)

    def solve_fp64(self, tol=1e-10, maxiter=100000, chunk=16, persistent=False):
        """FP64 block CG.  chunk: iterations per replayed HIP graph (0: kernel
        by kernel); persistent: all iterations in one cooperative launch
        instead (NGSAMD_BATCH_PERSISTENT), with the same iterates."""
        if persistent:
            chunk = PERSISTENT
        its, res, ms = N.BatchCGSolve(self.op, self.pre, self.rhs, self.sol, tol, maxiter, chunk)
        return dict(iterations=its, final_residual=res, finish_ms=ms)

    def solve_mixed(self, tol=1e-10, inner_tol=1e-4, max_outer=50, max_inner=100000):
        outer, inner, rel, ms = N.BatchIRCGSolve(self.op, self.pre, self.rhs, self.sol, tol, inner_tol, max_outer,
                                                 max_inner)
        return dict(outer=outer, inner=inner, relres=rel, finish_ms=ms)

    def solution(self, k=None):
        u = self.sol.download()
        if k is None:
            return u
        return u[self.offsets[k]:self.offsets[k + 1]]
