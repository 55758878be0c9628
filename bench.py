"""Benchmark of the H1 Laplace hot path on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): 3D Poisson
on the unit cube, 32^3 cubes x 6 Kuhn tetrahedra (196,608 tets), H1 order 4
(2,146,689 dofs, 35 dofs per tet), integration order 8 (125 points),
x ~ U[-1,1] from numpy default_rng(1).

One step = one FP64 CSR operator application y = A x (the reference's
SparseMatrix<double>::MultAdd path).  `value` is its throughput in GDoF/s
(ndof / time per apply); reported beside it under "extra": the other CSR
kernels, the matrix-free applies in every precision, the 4-vector apply and
the lockstep 4-right-hand-side CG, the FP64 CSR Jacobi-CG and the
mixed-precision (FP32 / BF16 / FP16 matrix-free inner, FP64 refinement)
time-to-1e-10 on the lognormal-kappa variant, block-Jacobi CG, the device
set-up (pattern, assembly, topology and dof tables) against the host, the
adjoint gradient on 128^2 triangles, the 512 batched 2D problems and the 2D
64^2 p=3 case.  Every result states its error against the oracle (the CPU
restatement of the reference's algorithm, run on the host as the CPU
baseline), unless --no-cpu is given.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-cg]
                       [--no-batch] [--no-assembly] [--no-2d] [--no-adjoint]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed(N, fn, steps, warmup):
    """Average milliseconds per call of fn over `steps` calls, bracketed by
    HIP events on the library stream (synchronised on both sides)."""
    for _ in range(warmup):
        fn()
    N.synchronize()
    e0, e1 = N.event_create(), N.event_create()
    N.event_record(e0)
    for _ in range(steps):
        fn()
    N.event_record(e1)
    ms = N.event_elapsed_ms(e0, e1)
    N.synchronize()
    N.event_destroy(e0)
    N.event_destroy(e1)
    return ms / steps


def build_mf(N, fes, mesh, ir, kappa, precision, scatter=0):
    classes = N.BuildClasses(fes, ir)
    detj, ginv = N.ComputeGeometry(mesh)
    el = np.concatenate([c["elements"] for c in classes]).astype(np.int32)
    off = np.cumsum([0] + [len(c["elements"]) for c in classes]).astype(np.int64)
    G = detj[el, None, None] * ginv[el]
    metric = np.ascontiguousarray(np.stack([G[:, 0, 0], G[:, 1, 1], G[:, 2, 2], G[:, 0, 1], G[:, 0, 2], G[:, 1, 2]], 1))
    kap = None if kappa is None else np.ascontiguousarray(kappa[el])
    return N.CreateMatrixFree(3, fes.nd_el, len(ir), fes.ndof, off,
                              np.ascontiguousarray(np.concatenate([c["bmat"].ravel() for c in classes])),
                              np.ascontiguousarray(ir.weights), np.ascontiguousarray(fes.el_dofs[el]), metric, kap,
                              el, precision, scatter)


def host_inventory():
    """What the CPU baseline ran on: lscpu's model and core counts, nproc and
    the threads this process may use."""
    import subprocess
    info = {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU(s)", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "NUMA node(s)"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    return info


def _median_time(fn, budget_s, warmup=3, min_reps=3, max_reps=100000):
    """Median seconds per call: `warmup` untimed calls, then calls timed one
    by one (steady clock) until budget_s has passed and min_reps were made."""
    for _ in range(warmup):
        fn()
    ts = []
    t0 = time.perf_counter()
    while len(ts) < max_reps and (len(ts) < min_reps or time.perf_counter() - t0 < budget_s):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts)), len(ts)


def cpu_baseline_apply(rowptr, cols, vals, x, ndof, n):
    """CPU baseline of the apply: the oracle's row-parallel CSR SpMV and its
    per-quadrature-point matrix-free apply (the reference's MatrixFreeBTDTB
    formula, comp/matrixfree.cpp:681-1087) on the same 32^3 p=4 problem, on
    all the host threads this process has and on one thread; bounded samples
    (3 warm-ups, then the median of repeated calls).  Also returns the
    oracle's outputs, against which the device results' errors are stated."""
    import oracle
    from oracle import fe_oracle as O
    threads = oracle.num_threads()
    pr = O.Problem(O.structured_3d(n), 4, intorder=8)
    cls = O.class_number(3, pr.mesh["els"])
    uniq = np.unique(cls)
    reps = [int(np.where(cls == k)[0][0]) for k in uniq]
    _, bref = O.element_shapes(3, 4, pr.mesh["els"][reps], pr.pts)
    elcls = np.searchsorted(uniq, cls).astype(np.int32)
    y = np.empty(ndof)

    def spmv():
        oracle.spmv(rowptr, cols, vals, x, y)

    def mf():  # Mult = zero + MultAdd (the oracle's apply accumulates)
        y[:] = 0.0
        oracle.mf_apply(3, bref, elcls, pr.wts, pr.dofs, pr.Finv, pr.detF, None, x, y)

    out = {"threads_all": threads}
    t_all, r_all = _median_time(spmv, 8.0)
    y_csr = y.copy()
    tm_all, rm_all = _median_time(mf, 4.0, warmup=1)
    y_mf = y.copy()
    oracle.set_threads(1)
    try:
        t_one, r_one = _median_time(spmv, 3.0, warmup=1)
        tm_one, rm_one = _median_time(mf, 0.0, warmup=0, min_reps=1)
    finally:
        oracle.set_threads(threads)
    nnz = int(rowptr[-1])
    out["csr_spmv"] = {"threads": threads, "seconds_median": t_all, "reps": r_all,
                       "gdofs": round(ndof / t_all / 1e9, 4)}
    out["csr_spmv_one_thread"] = {"threads": 1, "seconds_median": t_one, "reps": r_one,
                                  "gdofs": round(ndof / t_one / 1e9, 4)}
    out["matrix_free"] = {"threads": threads, "seconds_median": tm_all, "reps": rm_all,
                          "gdofs": round(ndof / tm_all / 1e9, 4)}
    out["matrix_free_one_thread"] = {"threads": 1, "seconds_median": tm_one, "reps": rm_one,
                                     "gdofs": round(ndof / tm_one / 1e9, 4)}
    out["sample"] = (f"FP64 y=Ax of the 32^3 p=4 problem ({nnz} nonzeros): oracle C CSR SpMV, {r_all} calls on "
                     f"{threads} threads and {r_one} on 1; oracle C per-point matrix-free apply (125 points), "
                     f"{rm_all} calls on {threads} threads and {rm_one} on 1; median per call")
    return out, {"y_csr": y_csr, "y_mf": y_mf}


def cpu_baseline_pcg(rowptr, cols, vals, invdiag, f, tol=1e-10):
    """CPU baseline of a solve: the oracle's C Jacobi-PCG (CGSolver
    recurrence, OpenMP on all host threads) on the same system; its
    solution is the reference the device solutions are compared with."""
    import oracle
    t0 = time.perf_counter()
    u, its, _ = oracle.pcg_csr(rowptr, cols, vals, invdiag, f, tol=tol, maxiter=200000)
    return {"iterations": its, "seconds": round(time.perf_counter() - t0, 4), "threads": oracle.num_threads()}, u


def _rel(a, b, norm=None):
    if norm == "inf":
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def bench_batch(N, args):
    """Batching configuration: 512 independent 2D Poisson problems (seed 5:
    nx ~ U{8..96}, p ~ U{1..5}, kappa_e = exp(N(0,1)), rhs ~ U[-1,1]) solved
    together to 1e-10, FP64 block CG (graph replays of 16 iterations, and all
    iterations in one persistent cooperative launch) and mixed (FP32 inner,
    FP64 refinement), against solving them one after the other with the FP64
    device CG."""
    from ngsolve_amd.batch import BatchSolver, diagonal, make_problems
    t = time.perf_counter()
    probs = make_problems(args.batch_problems, seed=5)
    bs = BatchSolver(probs)
    log(f"[bench] batch setup {time.perf_counter() - t:.1f}s ndof={bs.n} nnz={bs.nnz} block_nnz={bs.block_nnz}")
    out = {"problems": len(probs), "total_dofs": bs.n, "total_nnz": bs.nnz, "autotuned_block_nnz": bs.block_nnz,
           "autotune_ms_per_apply": {str(k): round(v, 4) for k, v in N.BatchTuning(bs.op).items()}}
    sols = {}
    for name, fn in [("fp64_block_cg", lambda: bs.solve_fp64(tol=1e-10)),
                     ("fp64_block_cg_persistent", lambda: bs.solve_fp64(tol=1e-10, persistent=True)),
                     ("mixed_block_ir", lambda: bs.solve_mixed(tol=1e-10))]:
        N.synchronize()
        t0 = time.perf_counter()
        r = fn()
        el = time.perf_counter() - t0
        sols[name] = bs.solution()
        ms = r["finish_ms"]
        out[name] = {"seconds": round(el, 4), "problems_per_s": round(len(probs) / el, 1),
                     "p50_ms": round(float(np.percentile(ms, 50)), 3),
                     "p99_ms": round(float(np.percentile(ms, 99)), 3)}
        if "iterations" in r:
            out[name]["max_iterations"] = int(r["iterations"].max())
        else:
            out[name]["max_outer"] = int(r["outer"].max())
            out[name]["max_relres"] = float(r["relres"].max())
        log(f"[bench] batch {name}: {el:.3f}s p50 {out[name]['p50_ms']} ms p99 {out[name]['p99_ms']} ms")
    # the same problems one after the other (FP64 CSR + Jacobi device CG)
    seq = []
    for pr in probs:
        A = N.CreateCSR(pr.ndof, pr.ndof, pr.rowptr, pr.cols, pr.vals)
        C = N.CreateJacobi(diagonal(pr.rowptr, pr.cols, pr.vals), pr.free)
        f, u = N.DeviceVector(pr.ndof, 8), N.DeviceVector(pr.ndof, 8)
        f.upload(pr.rhs)
        N.synchronize()
        t0 = time.perf_counter()
        N.CGSolve(A, C, f, u, True, 1e-10, -1.0, 100000, 16)
        seq.append(time.perf_counter() - t0)
    seq = np.array(seq)
    out["sequential_fp64_cg"] = {"seconds": round(float(seq.sum()), 4),
                                 "problems_per_s": round(len(probs) / seq.sum(), 1),
                                 "p50_ms": round(float(np.percentile(seq, 50) * 1e3), 3),
                                 "p99_ms": round(float(np.percentile(seq, 99) * 1e3), 3)}
    log(f"[bench] sequential: {seq.sum():.3f}s")
    if not args.no_cpu:
        # error against the oracle's PCG on a sample of the problems (every
        # problem is checked in the tests): the 4 largest and 12 spread over
        # the batch; the oracle's time for the whole sample is its baseline
        order = np.argsort([pr.ndof for pr in probs])
        sample = sorted(set(order[-4:].tolist()) | set(range(0, len(probs), max(1, len(probs) // 12))))
        err = {k: 0.0 for k in sols}
        t_cpu = 0.0
        for k in sample:
            pr = probs[k]
            d = diagonal(pr.rowptr, pr.cols, pr.vals)
            inv = np.where(pr.free != 0, 1.0 / np.where(d == 0, 1.0, d), 0.0)
            res, u_ref = cpu_baseline_pcg(pr.rowptr, pr.cols, pr.vals, inv, pr.rhs)
            t_cpu += res["seconds"]
            for name in sols:
                err[name] = max(err[name], _rel(sols[name][bs.offsets[k]:bs.offsets[k + 1]], u_ref))
        for name in sols:
            out[name]["max_rel_err_vs_oracle_sample"] = err[name]
        out["oracle_sample"] = {"problems": len(sample), "cpu_oracle_pcg_seconds": round(t_cpu, 3),
                                "threads": res["threads"]}
        log(f"[bench] batch vs oracle on {len(sample)} problems: {err}")
    return out


def bench_2d(N, args):
    """The reference's CPU-runnable case: 2D Poisson on 64^2 squares x 2
    triangles, H1 p=3 (37,249 dofs, 628,609 nonzeros), kappa = 1, Dirichlet on
    every side, x ~ U[-1,1] (seed 0), f = 1, Jacobi-CG to 1e-10.  Its 8.3 MB
    per apply stay in the Infinity Cache, so the apply is bound by launch
    latency, not by HBM; the oracle's CSR SpMV and PCG time the same work on
    the host."""
    m = N.MakeStructured2DMesh(64, 64)
    fes = N.H1Space(m, 3, 0xFFFFFFFFFFFFFFFF)
    ir = N.TrigRule(6)  # the default integration order 2p
    rp, cols = N.BuildPattern(fes)
    vals = N.AssembleCSR(fes, rp, cols, ir, True, None, False, None)
    n = fes.ndof
    A = N.CreateCSR(n, n, rp, cols, vals)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    dx, dy = N.DeviceVector(n, 8), N.DeviceVector(n, 8)
    dx.upload(x)
    ms = timed(N, lambda: A.mult(dx, dy), 200, 20)
    out = {"ndof": n, "nnz": int(rp[-1]), "csr_apply_us": round(ms * 1e3, 2),
           "csr_gdofs": round(n / (ms * 1e-3) / 1e9, 3),
           "csr_algorithmic_gbs": round((int(rp[-1]) * 12 + (n + 1) * 4 + 2 * n * 8) / (ms * 1e-3) / 1e9, 1)}
    dd = N.DeviceVector(n, 8)
    A.diagonal(dd)
    diag = dd.download()
    C = N.CreateJacobi(diag, fes.freedofs)
    f = N.AssembleLinearForm(fes, N.TrigRule(6), np.ones((m.ne, len(N.TrigRule(6)))))
    df, du = N.DeviceVector(n, 8), N.DeviceVector(n, 8)
    df.upload(f)
    N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 100000, 16)  # warm-up (graph build)
    N.synchronize()
    t0 = time.perf_counter()
    its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 100000, 16)
    t_cg = time.perf_counter() - t0
    out.update({"cg_iterations": its, "cg_seconds": round(t_cg, 4),
                "cg_us_per_iteration": round(1e6 * t_cg / max(its - 1, 1), 2)})
    log(f"[bench] 2D 64^2 p=3: apply {ms * 1e3:.1f} us, CG {its} its {t_cg * 1e3:.1f} ms")
    if not args.no_cpu:
        import oracle
        dx.upload(x)
        A.mult(dx, dy)
        y_dev = dy.download()
        y = np.empty(n)
        t_sp, reps = _median_time(lambda: oracle.spmv(rp, cols, vals, x, y), 2.0)
        out["cpu_oracle_apply_us"] = round(1e6 * t_sp, 2)
        out["apply_rel_err_vs_oracle"] = _rel(y_dev, y, "inf")
        free = fes.freedofs.astype(bool)
        inv = np.where(free, 1.0 / np.where(diag == 0, 1.0, diag), 0.0)
        res, u_ref = cpu_baseline_pcg(rp, cols, vals, inv, f)
        out.update({"cpu_oracle_pcg_iterations": res["iterations"], "cpu_oracle_pcg_seconds": res["seconds"],
                    "cpu_threads": res["threads"], "cg_rel_err_vs_oracle": _rel(du.download(), u_ref)})
        log(f"[bench] 2D 64^2 p=3 CPU oracle: apply {out['cpu_oracle_apply_us']} us, PCG {res['iterations']} its "
            f"{res['seconds']} s; device errors {out['apply_rel_err_vs_oracle']:.1e} / "
            f"{out['cg_rel_err_vs_oracle']:.1e}")
    return out


def bench_adjoint(N, args):
    """Differentiation configuration: 2D unit square, 128 x 128 squares x 2
    triangles, H1 p=3 (148,225 dofs), f = 1, kappa_true = exp(N(0,1)) (seed 3),
    u_obs = u(kappa_true) + N(0, 1e-3^2) (same generator), J = 0.5 (u -
    u_obs)^T M (u - u_obs) at kappa = exp(N(0, 0.5)) (seed 4).  Times the
    forward solve and the backward pass (adjoint solve + element gradient
    kernel) through torch.autograd; checks the gradient against central
    differences (h = 1e-6) along 16 unit directions (seed 6, device solves)
    and, with the CPU baseline, against the oracle's FP64 adjoint gradient
    (sparse LU), whose time is reported beside it."""
    import torch
    import ngsolve_amd as ngs
    from ngsolve_amd.autograd import ApplySymmetric, PoissonProblem, SolvePoisson
    nx, p = 128, 3
    m = ngs.MakeStructured2DMesh(nx=nx, ny=nx)
    fes = ngs.H1(m, order=p, dirichlet=".*")
    u_, v_ = fes.TnT()
    f = ngs.LinearForm(1 * v_ * ngs.dx).Assemble().vec.NumPy().copy()
    mass = ngs.BilinearForm(u_ * v_ * ngs.dx).Assemble()
    ne = m.ne
    rng = np.random.default_rng(3)
    kappa_true = np.exp(rng.standard_normal(ne))
    P = PoissonProblem(fes, f, tol=1e-13)
    u_true = SolvePoisson.apply(torch.tensor(kappa_true, device="cuda"), P).cpu().numpy()
    u_obs = u_true + rng.normal(0.0, 1e-3, fes.ndof)
    kappa = np.exp(np.random.default_rng(4).normal(0.0, 0.5, ne))
    uo = torch.tensor(u_obs, device="cuda")

    def grad_once():
        kt = torch.tensor(kappa, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u = SolvePoisson.apply(kt, P)
        r = u - uo
        J = 0.5 * torch.dot(r, ApplySymmetric.apply(r, mass.mat.op))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        J.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return J.item(), kt.grad.cpu().numpy(), u.detach().cpu().numpy(), t1 - t0, t2 - t1

    grad_once()  # warm-up
    P.iterations.clear()
    J, g, u, t_fwd, t_bwd = grad_once()
    its = list(P.iterations)
    out = {"ndof": fes.ndof, "elements": ne, "J": J, "forward_seconds": round(t_fwd, 4),
           "backward_seconds": round(t_bwd, 4), "cg_iterations_forward_adjoint": its, "cg_tol": 1e-13}

    def J_of(k):
        uk = SolvePoisson.apply(torch.tensor(k, device="cuda"), P).cpu().numpy()
        rk = uk - u_obs
        return 0.5 * float(rk @ (Mh @ rk))

    import scipy.sparse as sp
    Mh = sp.csr_matrix((mass.mat.vals, mass.mat.cols, mass.mat.rowptr), shape=(fes.ndof, fes.ndof))
    drng = np.random.default_rng(6)
    worst = 0.0
    for _ in range(16):
        d = drng.standard_normal(ne)
        d /= np.linalg.norm(d)
        fd = (J_of(kappa + 1e-6 * d) - J_of(kappa - 1e-6 * d)) / 2e-6
        worst = max(worst, abs(fd - g @ d) / max(abs(fd), 1e-30))
    out["fd_worst_rel_err_16_directions"] = worst
    log(f"[bench] adjoint 128^2 p=3: forward {t_fwd * 1e3:.1f} ms backward {t_bwd * 1e3:.1f} ms its {its}, "
        f"FD worst {worst:.1e}")
    if not args.no_cpu:
        from oracle import fe_oracle as O
        pr = O.Problem(O.structured_2d(nx), p, intorder=2 * p - 2)
        prm = O.Problem(O.structured_2d(nx), p, intorder=2 * p)
        rp, c, vals = prm.assemble(stiffness=False, mass=True)
        M = sp.csr_matrix((vals, c, rp), shape=(pr.ndof, pr.ndof))
        t0 = time.perf_counter()
        J_ref, g_ref, u_ref = O.adjoint_gradient(pr, kappa, f, u_obs, M)
        t_cpu = time.perf_counter() - t0
        out.update({"gradient_rel_err_vs_oracle": _rel(g, g_ref), "u_rel_err_vs_oracle": _rel(u, u_ref),
                    "J_rel_err_vs_oracle": abs(J - J_ref) / abs(J_ref),
                    "cpu_oracle_adjoint_seconds": round(t_cpu, 3),
                    "cpu_oracle_note": "numpy/scipy oracle: sparse LU (SuperLU) of the free-dof block, two solves, "
                                       "element gradients"})
        log(f"[bench] adjoint vs oracle: gradient {out['gradient_rel_err_vs_oracle']:.1e}, CPU oracle {t_cpu:.2f}s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-assembly", action="store_true")
    ap.add_argument("--no-2d", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--profile-cg", action="store_true",
                    help="one FP64 CSR CG solve (kappa = 1, tol 1e-10) without graph capture, for a kernel trace")
    ap.add_argument("--profile-ir", action="store_true",
                    help="one refinement solve (lognormal kappa, tol 1e-10) without graph capture, for a kernel "
                         "trace; inner precision from --profile-ir-precision")
    ap.add_argument("--profile-ir-precision", default="f32", choices=["f32", "bf16", "f16", "f16x3"],
                    help="inner operator precision of --profile-ir (default f32)")
    ap.add_argument("--profile-mode", action="store_true",
                    help="short run for rocprofv3 counter passes: a few CSR and matrix-free applies, one "
                         "1 GiB device-to-device copy and one 1 GiB streaming read (the calibration kernels of "
                         "known size); no JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the operator and solver have no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    from ngsolve_amd import _ngsamd as N
    from ngsolve_amd._native import ensure_device
    ensure_device(0)

    n, p, order = args.n, 4, 8
    t = time.perf_counter()
    mesh = N.MakeStructured3DMesh(n, n, n)
    fes = N.H1Space(mesh, p, 0xFFFFFFFFFFFFFFFF)
    ir = N.TetRule(order)
    rowptr, cols = N.BuildPattern(fes)
    vals = N.AssembleCSR(fes, rowptr, cols, ir, True, None, False, None)
    ndof, nnz, ne = fes.ndof, int(rowptr[-1]), mesh.ne
    log(f"[bench] setup {time.perf_counter() - t:.1f}s ndof={ndof} nnz={nnz} ne={ne}")

    A = N.CreateCSR(ndof, ndof, rowptr, cols, vals)
    x = np.random.default_rng(1).uniform(-1, 1, ndof)
    dx, dy = N.DeviceVector(ndof, 8), N.DeviceVector(ndof, 8)
    dx.upload(x)
    # the CPU baseline (oracle) runs first: its outputs are the references
    # the device results' errors are stated against
    cpu, ref = None, None
    if not (args.no_cpu or args.profile_cg or args.profile_ir or args.profile_mode):
        log("[bench] cpu baseline (oracle CSR SpMV and matrix-free apply, all threads and 1 thread) ...")
        cpu, ref = cpu_baseline_apply(rowptr, cols, vals, x, ndof, n)
        log(f"[bench] cpu SpMV {cpu['csr_spmv']['gdofs']} GDoF/s on {cpu['threads_all']} threads, "
            f"{cpu['csr_spmv_one_thread']['gdofs']} on 1; matrix-free {cpu['matrix_free']['gdofs']} / "
            f"{cpu['matrix_free_one_thread']['gdofs']}")

    def err_apply(y):
        return None if ref is None else _rel(y, ref["y_mf"], "inf")

    if args.profile_cg:
        dd = N.DeviceVector(ndof, 8)
        A.diagonal(dd)
        C = N.CreateJacobi(dd.download(), fes.freedofs)
        f = N.AssembleLinearForm(fes, ir, np.ones((ne, len(ir))))
        df, du = N.DeviceVector(ndof, 8), N.DeviceVector(ndof, 8)
        df.upload(f)
        its, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 100000, 0)
        N.synchronize()
        log(f"[bench] profile CG done: {its} iterations")
        return

    if args.profile_ir:
        os.environ["NGSAMD_CG_CHUNK"] = "0"
        kappa = np.exp(np.random.default_rng(2).standard_normal(ne))
        vals_k = N.AssembleCSR(fes, rowptr, cols, ir, True, kappa, False, None)
        Ak = N.CreateCSR(ndof, ndof, rowptr, cols, vals_k)
        dd = N.DeviceVector(ndof, 8)
        Ak.diagonal(dd)
        Ck = N.CreateJacobi(dd.download(), fes.freedofs)
        f = N.AssembleLinearForm(fes, ir, np.ones((ne, len(ir))))
        df, du = N.DeviceVector(ndof, 8), N.DeviceVector(ndof, 8)
        df.upload(f)
        inner_prec = {"f32": 1, "bf16": 2, "f16": 3, "f16x3": 4}[args.profile_ir_precision]
        M64, M32 = build_mf(N, fes, mesh, ir, kappa, 0), build_mf(N, fes, mesh, ir, kappa, inner_prec)
        outer, inner, rel = N.IRCGSolve(M64, M32, Ck, df, du, 1e-10, -1.0, 60, 200000)
        N.synchronize()
        log(f"[bench] profile IR done: outer {outer} inner {inner} rel {rel:.1e}")
        return

    if args.profile_mode:
        for _ in range(3):
            A.mult(dx, dy)
        os.environ["NGSAMD_SPMV"] = "stream"
        As = N.CreateCSR(ndof, ndof, rowptr, cols, vals)
        del os.environ["NGSAMD_SPMV"]
        for _ in range(3):
            As.mult(dx, dy)
        del As
        x32, y32 = N.DeviceVector(ndof, 4), N.DeviceVector(ndof, 4)
        x32.upload(x.astype(np.float32))
        for prec in (0, 1, 2, 3, 4):
            M = build_mf(N, fes, mesh, ir, None, prec)
            for _ in range(3):
                M.mult(dx, dy)
                if prec:
                    M.mult_f32(x32, y32)
            del M
        big_a, big_b = N.DeviceVector(1 << 27, 8), N.DeviceVector(1 << 27, 8)  # 1 GiB each
        big_a.fill(1.0)
        big_b.copy_from(big_a)
        res = N.DeviceVector(1, 8)
        for _ in range(3):
            N.stream_read(big_a, res)
        N.synchronize()
        log("[bench] profile mode done")
        return

    # read ceiling: 1 GiB of non-temporal streaming loads (ngsamd_stream_read)
    big = N.DeviceVector(1 << 27, 8)
    big.fill(1.0)
    rsum = N.DeviceVector(1, 8)
    ms_read = timed(N, lambda: N.stream_read(big, rsum), 100, 10)
    read_gbs = (1 << 30) / (ms_read * 1e-3) / 1e9
    assert rsum.download()[0] == float(1 << 27)
    del big

    # ---------------------------------------------------- CSR SpMV (value)
    ms_csr = timed(N, lambda: A.mult(dx, dy), args.steps, args.warmup)
    y_csr = dy.download()
    csr_bytes = nnz * 12 + (ndof + 1) * 4 + 2 * ndof * 8
    gbs = csr_bytes / (ms_csr * 1e-3) / 1e9
    gdofs = ndof / (ms_csr * 1e-3) / 1e9
    log(f"[bench] csr spmv {ms_csr * 1e3:.1f} us  {gbs:.0f} GB/s  {gdofs:.2f} GDoF/s")

    extra = {}
    extra["csr_apply_rel_err_vs_oracle"] = err_apply(y_csr)
    extra["read_ceiling"] = {"gbs": round(read_gbs, 1), "frac_of_8tbs": round(read_gbs / HBM_PEAK_GBS, 4),
                             "kernel": "StreamReadKernel: 1 GiB FP64, 16-byte non-temporal loads"}
    log(f"[bench] streaming read ceiling {read_gbs:.0f} GB/s; csr spmv at {gbs / read_gbs:.3f} of it")
    # K = 4 vectors at once: the multi-vector apply reads the matrix once
    K = 4
    dXk, dYk = N.DeviceVector(ndof * K, 8), N.DeviceVector(ndof * K, 8)
    dXk.upload(np.random.default_rng(7).uniform(-1, 1, ndof * K))
    ms_multi = timed(N, lambda: N.CSRMultMulti(A, dXk, dYk, K), args.steps, args.warmup)
    extra["multi_rhs_4"] = {"apply_us": round(ms_multi * 1e3, 1), "single_apply_us": round(ms_csr * 1e3, 1),
                            "apply_speedup_vs_4_singles": round(4 * ms_csr / ms_multi, 2)}
    log(f"[bench] 4 vectors at once: {ms_multi * 1e3:.1f} us (single {ms_csr * 1e3:.1f} us)")
    del dXk, dYk
    # the plain-CSR kernels on the same matrix, for comparison (the default
    # above is the pattern-runs kernel: this matrix shares column lists
    # between the dofs of an edge, face or cell)
    for algo in ("stream", "rowgroup"):
        os.environ["NGSAMD_SPMV"] = algo
        Aalt = N.CreateCSR(ndof, ndof, rowptr, cols, vals)
        del os.environ["NGSAMD_SPMV"]
        ms_alt = timed(N, lambda: Aalt.mult(dx, dy), args.steps, args.warmup)
        extra[f"csr_{algo}_kernel"] = {
            "us_per_apply": round(ms_alt * 1e3, 1), "gbs": round(csr_bytes / (ms_alt * 1e-3) / 1e9, 1),
            "rel_err_vs_oracle": err_apply(dy.download())}
        del Aalt
        log(f"[bench] csr {algo} {ms_alt * 1e3:.1f} us")

    # ------------------------------------------------- matrix-free applies
    nq = len(ir)
    # flops the factored kernels execute per element: U = D^T x_e (nd x 3r),
    # V = G U (3 x 3 per basis function), y_e = D V; r = dim P_(p-1) = 20 for
    # tet p=4 (nd = 35): 2*35*60 + 18*20 + 2*35*60 = 8,760
    nd_el = fes.nd_el
    rank = 4 * 5 * 6 // 6
    flops_mf = (4.0 * nd_el * 3 * rank + 18.0 * rank) * ne
    classes_c = N.BuildClasses(fes, ir)
    el_c = np.concatenate([c["elements"] for c in classes_c]).astype(np.int64)
    colors = np.ascontiguousarray(N.ElementColoring(mesh)[el_c], dtype=np.int32)
    x32, y32 = N.DeviceVector(ndof, 4), N.DeviceVector(ndof, 4)
    x32.upload(x.astype(np.float32))
    for name, prec, peak in [("mf_f64", 0, FP64_PEAK_TFLOPS), ("mf_f32", 1, FP32_PEAK_TFLOPS),
                             ("mf_bf16", 2, BF16_PEAK_TFLOPS), ("mf_f16", 3, BF16_PEAK_TFLOPS),
                             ("mf_f16x3", 4, BF16_PEAK_TFLOPS)]:
        for scatter in (0, 1, 2):
            M = build_mf(N, fes, mesh, ir, None, prec, min(scatter, 1))
            if scatter == 2:  # colour-ordered direct scatter (33 colours at 32^3)
                M.set_coloring(colors)
            ms = timed(N, lambda: M.mult(dx, dy), args.steps, args.warmup)
            err = err_apply(dy.download())
            tf = flops_mf / (ms * 1e-3) / 1e12
            key = name + ("", "_atomic", "_color")[scatter]
            extra[key] = {"us_per_apply": round(ms * 1e3, 1), "gdofs": round(ndof / (ms * 1e-3) / 1e9, 3),
                          "tflops": round(tf, 2), "frac_of_peak": round(tf / peak, 4),
                          "rel_err_vs_oracle": err}
            if prec > 0:
                ms32 = timed(N, lambda: M.mult_f32(x32, y32), args.steps, args.warmup)
                extra[key]["us_per_apply_f32_vectors"] = round(ms32 * 1e3, 1)
                extra[key]["rel_err_vs_oracle_f32_vectors"] = err_apply(y32.download().astype(np.float64))
            log(f"[bench] {key} {ms * 1e3:.1f} us  {tf:.1f} TFLOP/s  err {err}")
            del M

    # ------------------------------------------- CSR assembly on the device
    if not args.no_assembly:
        t0 = time.perf_counter()
        rp_h, cols_h = N.BuildPattern(fes)
        t_pat_host = time.perf_counter() - t0
        t0 = time.perf_counter()
        vals_h = N.AssembleCSR(fes, rp_h, cols_h, ir, True, None, False, None)
        t_host = time.perf_counter() - t0
        M64 = build_mf(N, fes, mesh, ir, None, 0)
        N.synchronize()
        t0 = time.perf_counter()
        rp_d, cols_d = N.PatternDevice(M64)
        t_pat_dev = time.perf_counter() - t0
        same_pattern = bool(np.array_equal(rp_d, rp_h) and np.array_equal(cols_d, cols_h))
        t0 = time.perf_counter()
        Adev = N.AssembleCSRDevice(M64, rp_d, cols_d)
        N.synchronize()
        t_dev = time.perf_counter() - t0
        d_asm = float(np.abs(Adev.csr_values() - vals_h).max() / np.abs(vals_h).max())
        del Adev, M64, vals_h, rp_h, cols_h, rp_d, cols_d
        extra["csr_assembly"] = {
            "host_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
            "host_pattern_seconds": round(t_pat_host, 4), "host_values_seconds": round(t_host, 4),
            "device_pattern_seconds": round(t_pat_dev, 4), "device_pattern_identical": same_pattern,
            "device_values_seconds_with_host_structure": round(t_dev, 4), "rel_diff_vs_host": d_asm,
            "note": "device pattern: two sort passes on the device plus the host prefix sum and copies; device "
                    "values: pattern-runs structure built on the host, values by AssembleCsrKernel (kernel "
                    "times in the rocprof kernel trace)"}
        log(f"[bench] csr pattern host {t_pat_host:.3f}s device {t_pat_dev:.3f}s identical {same_pattern}; "
            f"values host {t_host:.3f}s device {t_dev:.3f}s diff {d_asm:.1e}")

        # topology + H1 dof tables: host setup vs ngsamd_h1_create (boundary
        # found as the faces of one element in both)
        els = np.ascontiguousarray(mesh.elements, dtype=np.int32)
        nob, nobi = np.zeros((0, 3), np.int32), np.zeros(0, np.int32)
        all_bits = (1 << 64) - 1
        t0 = time.perf_counter()
        hm = N.BuildTopologyHost(3, mesh.nv, els, nob, nobi)
        hf = N.H1Space(hm, p, all_bits)
        t_topo_host = time.perf_counter() - t0
        N.H1TablesDevice(3, mesh.nv, els, nob, nobi, p, all_bits)  # first call loads the kernels
        N.synchronize()
        t0 = time.perf_counter()
        D = N.H1TablesDevice(3, mesh.nv, els, nob, nobi, p, all_bits)
        t_topo_dev = time.perf_counter() - t0
        same_tables = bool(np.array_equal(D["el_dofs"], hf.el_dofs) and np.array_equal(D["freedofs"], hf.freedofs)
                           and np.array_equal(D["faces"], hm.faces) and np.array_equal(D["edges"], hm.edges))
        del D, hf, hm
        extra["h1_setup"] = {
            "host_seconds": round(t_topo_host, 4), "device_seconds": round(t_topo_dev, 4),
            "identical": same_tables,
            "note": "edge/face numbering, boundary, element dof tables, free dofs, element classes; device time "
                    "includes uploading the element table and downloading every table"}
        log(f"[bench] topology + H1 dofs host {t_topo_host:.3f}s device {t_topo_dev:.3f}s identical {same_tables}")

    # ----------------------------------------------------------------- CG
    if not args.no_cg:
        diag = np.empty(ndof)
        dd = N.DeviceVector(ndof, 8)
        A.diagonal(dd)
        diag[:] = dd.download()
        C = N.CreateJacobi(diag, fes.freedofs)
        f = N.AssembleLinearForm(fes, ir, np.ones((ne, nq)))
        df, du = N.DeviceVector(ndof, 8), N.DeviceVector(ndof, 8)
        df.upload(f)
        t0 = time.perf_counter()
        its, res = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 100000, 16)
        t_cg = time.perf_counter() - t0
        extra["cg_fp64_csr_kappa1"] = {"iterations": its, "seconds": round(t_cg, 4),
                                       "ms_per_iteration": round(1e3 * t_cg / max(its - 1, 1), 4)}
        log(f"[bench] FP64 CSR CG kappa=1: {its} its {t_cg:.3f}s")
        if cpu is not None:
            inv1 = np.where(fes.freedofs.astype(bool), 1.0 / np.where(diag == 0, 1.0, diag), 0.0)
            res1, u1_ref = cpu_baseline_pcg(rowptr, cols, vals, inv1, f)
            extra["cg_fp64_csr_kappa1"].update({"rel_err_vs_oracle": _rel(du.download(), u1_ref),
                                                "cpu_oracle_pcg": res1})

        # K right-hand sides of the same matrix: the lockstep CG
        # (CGSolver::MultiMult) against K single solves one after the other
        K = 4
        Fk = np.random.default_rng(8).uniform(0, 1, (ndof, K)) * fes.freedofs[:, None]
        dFk, dUk = N.DeviceVector(ndof * K, 8), N.DeviceVector(ndof * K, 8)
        dFk.upload(Fk.ravel())
        N.synchronize()
        t0 = time.perf_counter()
        steps_k, _ = N.CGSolveMulti(A, C, dFk, dUk, K, True, 1e-10, False, 100000)
        t_multi = time.perf_counter() - t0
        t_seq, its_seq = 0.0, []
        for j in range(K):
            df.upload(np.ascontiguousarray(Fk[:, j]))
            N.synchronize()
            t0 = time.perf_counter()
            its_j, _ = N.CGSolve(A, C, df, du, True, 1e-10, -1.0, 100000, 16)
            t_seq += time.perf_counter() - t0
            its_seq.append(its_j)
        df.upload(f)
        extra["multi_rhs_4"].update({
            "lockstep_cg_steps": steps_k, "lockstep_cg_seconds": round(t_multi, 4),
            "single_cg_iterations": its_seq, "single_cg_seconds_total": round(t_seq, 4),
            "cg_speedup": round(t_seq / t_multi, 2)})
        log(f"[bench] 4 right-hand sides: lockstep CG {steps_k} steps {t_multi:.3f}s vs 4 single solves "
            f"{t_seq:.3f}s")
        del dFk, dUk

        # lognormal kappa variant: FP64 CSR CG vs mixed-precision IR
        kappa = np.exp(np.random.default_rng(2).standard_normal(ne))
        vals_k = N.AssembleCSR(fes, rowptr, cols, ir, True, kappa, False, None)
        Ak = N.CreateCSR(ndof, ndof, rowptr, cols, vals_k)
        Ak.diagonal(dd)
        Ck = N.CreateJacobi(dd.download(), fes.freedofs)
        free = fes.freedofs.astype(bool)

        def true_res(u):
            r = f - _spmv_host(rowptr, cols, vals_k, u)
            return float(np.linalg.norm(r[free]) / np.linalg.norm(f[free]))

        u_oracle = None
        if cpu is not None:
            # the oracle's C Jacobi-PCG on the host: the CPU baseline of the
            # solve, and the solution every device result is compared with
            dk = np.empty(ndof)
            Ak.diagonal(dd)
            dk[:] = dd.download()
            invk = np.where(free, 1.0 / np.where(dk == 0, 1.0, dk), 0.0)
            res_cpu, u_oracle = cpu_baseline_pcg(rowptr, cols, vals_k, invk, f)
            log(f"[bench] CPU oracle PCG lognormal: {res_cpu['iterations']} its {res_cpu['seconds']:.2f}s on "
                f"{res_cpu['threads']} threads")

        def err_sol(u):
            return None if u_oracle is None else _rel(u, u_oracle)

        t0 = time.perf_counter()
        its_k, hist_k = N.CGSolve(Ak, Ck, df, du, True, 1e-10, -1.0, 200000, 16)
        t_k = time.perf_counter() - t0
        u64 = du.download()
        log(f"[bench] FP64 CSR CG lognormal: {its_k} its {t_k:.3f}s")
        M64 = build_mf(N, fes, mesh, ir, kappa, 0)
        res = {"fp64_csr_cg": {"iterations": its_k, "seconds": round(t_k, 4), "true_rel_residual": true_res(u64),
                               # the reference's stopping quantity sqrt|w.d|, last over first
                               "stopping_ratio": float(hist_k[-1] / hist_k[0]) if len(hist_k) > 1 else None,
                               "rel_err_vs_oracle": err_sol(u64)}}
        # the same FP64 CG on the FP64 matrix-free operator
        du_mf = N.DeviceVector(ndof, 8)
        t0 = time.perf_counter()
        its_mf, _ = N.CGSolve(M64, Ck, df, du_mf, True, 1e-10, -1.0, 200000, 16)
        t_mf = time.perf_counter() - t0
        u_mf = du_mf.download()
        log(f"[bench] FP64 matrix-free CG lognormal: {its_mf} its {t_mf:.3f}s")
        res["fp64_mf_cg"] = {"iterations": its_mf, "seconds": round(t_mf, 4), "true_rel_residual": true_res(u_mf),
                             "rel_err_vs_oracle": err_sol(u_mf), "speedup_vs_fp64_csr_cg": round(t_k / t_mf, 3)}
        del du_mf
        best = None
        for pname, prec in [("f32", 1), ("bf16", 2), ("f16", 3), ("f16x3", 4)]:
            Mlow = build_mf(N, fes, mesh, ir, kappa, prec)
            du2 = N.DeviceVector(ndof, 8)
            t0 = time.perf_counter()
            # inner tolerance -1: the solver's default for the inner precision
            outer, inner, rel = N.IRCGSolve(M64, Mlow, Ck, df, du2, 1e-10, -1.0, 60, 200000)
            t_ir = time.perf_counter() - t0
            u_ir = du2.download()
            log(f"[bench] mixed IR {pname}: outer {outer} inner {inner} rel {rel:.1e} {t_ir:.3f}s")
            res["mixed_ir_mf_" + pname] = {
                "outer": outer, "inner_total": inner, "seconds": round(t_ir, 4),
                "true_rel_residual": true_res(u_ir), "rel_err_vs_oracle": err_sol(u_ir),
                "speedup_vs_fp64_csr_cg": round(t_k / t_ir, 3)}
            if rel <= 1e-10 and (best is None or t_ir < best[1]):
                best = (pname, t_ir)
            del Mlow
        if best:
            res["best_mixed"] = best[0]
            res["speedup_best_mixed_vs_fp64_csr"] = round(t_k / best[1], 3)
        # FP64 CSR CG with block Jacobi on the reference's H1 smoothing blocks:
        # type 3 (one block per vertex, edge, face, cell) and type 4 (its
        # default: vertex-edge patches, faces, cells)
        from ngsolve_amd.comp import smoothing_blocks
        for btype in (3, 4):
            t0 = time.perf_counter()
            blocks = smoothing_blocks(mesh, fes, btype)
            t_blocks = time.perf_counter() - t0
            t0 = time.perf_counter()
            Cb = N.CreateBlockJacobi(ndof, rowptr, cols, vals_k, blocks.ptr, blocks.dofs)
            t_setup = time.perf_counter() - t0
            du3 = N.DeviceVector(ndof, 8)
            t0 = time.perf_counter()
            its_b, _ = N.CGSolve(Ak, Cb, df, du3, True, 1e-10, -1.0, 200000, 16)
            t_b = time.perf_counter() - t0
            u_b = du3.download()
            log(f"[bench] FP64 CSR CG block-Jacobi type {btype} lognormal: {its_b} its {t_b:.3f}s "
                f"(setup {t_setup:.2f}s)")
            res[f"fp64_csr_cg_blockjacobi_type{btype}"] = {
                "iterations": its_b, "seconds": round(t_b, 4), "setup_seconds": round(t_setup, 3),
                "blocks_seconds": round(t_blocks, 3),
                "blocks": len(blocks), "max_block": int(np.diff(blocks.ptr).max()),
                "true_rel_residual": true_res(u_b), "rel_err_vs_oracle": err_sol(u_b)}
            del Cb
        # two-level: Jacobi + exact coarse solve on the P1 (vertex) block by
        # the device sparse Cholesky, additively
        nv = mesh.nv
        keep = cols[:rowptr[nv]] < nv
        vrows = np.repeat(np.arange(nv), np.diff(rowptr[:nv + 1]))[keep]
        vrp = np.zeros(nv + 1, dtype=np.int64)
        np.add.at(vrp, vrows + 1, 1)
        vrp = np.cumsum(vrp)
        t0 = time.perf_counter()
        Cc = N.CreateSparseCholesky(nv, vrp, np.ascontiguousarray(cols[:rowptr[nv]][keep]),
                                    np.ascontiguousarray(vals_k[:rowptr[nv]][keep]),
                                    np.ascontiguousarray(fes.freedofs[:nv]))
        t_fact = time.perf_counter() - t0
        info = N.SparseCholeskyInfo(Cc)
        Ctl = N.CreateTwoLevel(Ck, Cc)
        dxv, dyv = N.DeviceVector(nv, 8), N.DeviceVector(nv, 8)
        dxv.upload(np.random.default_rng(7).uniform(-1, 1, nv))
        ms_coarse = timed(N, lambda: Cc.mult(dxv, dyv), 20, 3)
        du4 = N.DeviceVector(ndof, 8)
        t0 = time.perf_counter()
        its_t, _ = N.CGSolve(Ak, Ctl, df, du4, True, 1e-10, -1.0, 200000, 16)
        t_t = time.perf_counter() - t0
        u_t = du4.download()
        log(f"[bench] FP64 CSR CG two-level (Jacobi + P1 sparse Cholesky) lognormal: {its_t} its {t_t:.3f}s "
            f"(factor {t_fact:.2f}s, coarse solve {ms_coarse * 1e3:.1f} us, {info})")
        res["fp64_csr_cg_twolevel_p1"] = {
            "iterations": its_t, "seconds": round(t_t, 4), "factor_seconds": round(t_fact, 3),
            "coarse_solve_us": round(ms_coarse * 1e3, 1), "coarse": info,
            "true_rel_residual": true_res(u_t), "rel_err_vs_oracle": err_sol(u_t),
            "speedup_vs_fp64_csr_cg": round(t_k / t_t, 3)}
        # two-level with the multigrid V-cycle for the P1 block in place of
        # the exact coarse solve
        t0 = time.perf_counter()
        Cm = N.CreateMultiGrid(3, [n, n, n], nv, vrp, np.ascontiguousarray(cols[:rowptr[nv]][keep]),
                               np.ascontiguousarray(vals_k[:rowptr[nv]][keep]),
                               np.ascontiguousarray(fes.freedofs[:nv]), 2, 0.85, 32)
        t_mg = time.perf_counter() - t0
        levels = N.MultiGridInfo(Cm)
        Ctm = N.CreateTwoLevel(Ck, Cm)
        ms_vcycle = timed(N, lambda: Cm.mult(dxv, dyv), 20, 3)
        du5 = N.DeviceVector(ndof, 8)
        t0 = time.perf_counter()
        its_m, _ = N.CGSolve(Ak, Ctm, df, du5, True, 1e-10, -1.0, 200000, 16)
        t_m = time.perf_counter() - t0
        u_m = du5.download()
        log(f"[bench] FP64 CSR CG two-level (Jacobi + P1 multigrid V-cycle) lognormal: {its_m} its {t_m:.3f}s "
            f"(set-up {t_mg:.2f}s, levels {levels}, V-cycle {ms_vcycle * 1e3:.1f} us)")
        res["fp64_csr_cg_twolevel_p1_multigrid"] = {
            "iterations": its_m, "seconds": round(t_m, 4), "setup_seconds": round(t_mg, 3),
            "levels": levels, "vcycle_us": round(ms_vcycle * 1e3, 1), "smoothing": "2+2 damped Jacobi, 0.85",
            "true_rel_residual": true_res(u_m), "rel_err_vs_oracle": err_sol(u_m),
            "speedup_vs_fp64_csr_cg": round(t_k / t_m, 3)}
        del Ctl, Cc, Ctm, Cm, du4, du5, dxv, dyv
        if u_oracle is not None:
            res["cpu_oracle_pcg"] = dict(res_cpu, true_rel_residual=true_res(u_oracle),
                                         gpu_fp64_csr_cg_speedup=round(res_cpu["seconds"] / t_k, 1),
                                         gpu_best_mixed_speedup=round(res_cpu["seconds"] / best[1], 1) if best
                                         else None)
        extra["cg_lognormal_kappa"] = res

    if not args.no_adjoint:
        extra["adjoint_128sq_p3"] = bench_adjoint(N, args)
    if not args.no_batch:
        extra["batch_512_problems"] = bench_batch(N, args)
    if not args.no_2d:
        extra["poisson_2d_64sq_p3"] = bench_2d(N, args)

    cpu_line = None
    if cpu is not None:
        cpu_line = {"value": cpu["csr_spmv"]["gdofs"], "unit": "GDoF/s", "cores": cpu["threads_all"],
                    "kind": "port", "sample": cpu["sample"],
                    "one_core": cpu["csr_spmv_one_thread"], "matrix_free": cpu["matrix_free"],
                    "matrix_free_one_core": cpu["matrix_free_one_thread"], "host": host_inventory()}

    # HBM bytes per launch of the default CSR kernel, from the rocprofv3 PMC
    # passes summarised under profiles/ (only if they measured this kernel)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            tj = json.load(fh)
        if "CsrRunsKernel" in tj.get("kernel", ""):
            traffic = tj.get("csr_spmv_32cube_p4")

    out = {
        "metric": "GDoF/s apply + time-to-1e-10 CG, mixed vs FP64 on 1 GPU, % HBM/MFMA roofline",
        "value": round(gdofs, 4), "unit": "GDoF/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_csr, 5), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Poisson unit cube {n}^3 cubes x 6 Kuhn tets, H1 p=4, intorder 8 "
                               f"({ndof} dofs, {nnz} nnz), FP64 CSR y=Ax, x~U[-1,1] seed 1",
                   "ndof": ndof, "nnz": nnz, "elements": ne},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "CsrRunsKernel", "algorithmic_bytes": csr_bytes,
                     "read_ceiling_gbs": round(read_gbs, 1), "frac_of_read_ceiling": round(gbs / read_gbs, 4)},
        "cpu_baseline": cpu_line,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


def _spmv_host(rowptr, cols, vals, x):
    import scipy.sparse as sp
    n = len(rowptr) - 1
    return sp.csr_matrix((vals, cols, rowptr), shape=(n, n)) @ x


if __name__ == "__main__":
    main()
