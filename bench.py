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
mixed-precision (FP32 matrix-free inner, FP64 refinement) time-to-1e-10 on
the lognormal-kappa variant, block-Jacobi CG, the device set-up (pattern,
assembly, topology and dof tables) against the host, the 512 batched 2D
problems and the 2D 64^2 p=3 case.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-cg]
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


def cpu_baseline(rowptr, cols, vals, x, ndof, budget_s=10.0):
    """The oracle's row-parallel CSR SpMV (OpenMP, all host threads) on the
    same matrix; a bounded sample of about budget_s seconds."""
    import oracle
    y = np.empty(ndof)
    oracle.spmv(rowptr, cols, vals, x, y)  # warm-up
    t0 = time.perf_counter()
    reps = 0
    while True:
        oracle.spmv(rowptr, cols, vals, x, y)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    return {"value": round(ndof * reps / el / 1e9, 4), "unit": "GDoF/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"{reps} FP64 CSR applies y=Ax of the 32^3 p=4 matrix ({int(rowptr[-1])} nonzeros), "
                      f"oracle C SpMV with OpenMP, {el:.1f} s"}


def bench_batch(N, args):
    """Batching configuration: 512 independent 2D Poisson problems (seed 5:
    nx ~ U{8..96}, p ~ U{1..5}, kappa_e = exp(N(0,1)), rhs ~ U[-1,1]) solved
    together to 1e-10, FP64 block CG and mixed (FP32 inner, FP64 refinement),
    against solving them one after the other with the FP64 device CG."""
    from ngsolve_amd.batch import BatchSolver, diagonal, make_problems
    t = time.perf_counter()
    probs = make_problems(args.batch_problems, seed=5)
    bs = BatchSolver(probs)
    log(f"[bench] batch setup {time.perf_counter() - t:.1f}s ndof={bs.n} nnz={bs.nnz} block_nnz={bs.block_nnz}")
    out = {"problems": len(probs), "total_dofs": bs.n, "total_nnz": bs.nnz, "autotuned_block_nnz": bs.block_nnz}
    for name, fn in [("fp64_block_cg", lambda: bs.solve_fp64(tol=1e-10)),
                     ("mixed_block_ir", lambda: bs.solve_mixed(tol=1e-10))]:
        N.synchronize()
        t0 = time.perf_counter()
        r = fn()
        el = time.perf_counter() - t0
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
    ir = N.TrigRule(4)
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
        y = np.empty(n)
        oracle.spmv(rp, cols, vals, x, y)
        t0 = time.perf_counter()
        reps = 0
        while time.perf_counter() - t0 < 2.0:
            oracle.spmv(rp, cols, vals, x, y)
            reps += 1
        out["cpu_oracle_apply_us"] = round(1e6 * (time.perf_counter() - t0) / reps, 2)
        free = fes.freedofs.astype(bool)
        inv = np.where(free, 1.0 / np.where(diag == 0, 1.0, diag), 0.0)
        t0 = time.perf_counter()
        _, its_cpu, _ = oracle.pcg_csr(rp, cols, vals, inv, f, tol=1e-10, maxiter=100000)
        out.update({"cpu_oracle_pcg_iterations": its_cpu, "cpu_oracle_pcg_seconds": round(time.perf_counter() - t0, 4),
                    "cpu_threads": oracle.num_threads()})
        log(f"[bench] 2D 64^2 p=3 CPU oracle: apply {out['cpu_oracle_apply_us']} us, PCG {its_cpu} its "
            f"{out['cpu_oracle_pcg_seconds']} s")
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
    ap.add_argument("--batch-problems", type=int, default=512)
    ap.add_argument("--profile-cg", action="store_true",
                    help="one FP64 CSR CG solve (kappa = 1, tol 1e-10) without graph capture, for a kernel trace")
    ap.add_argument("--profile-ir", action="store_true",
                    help="one FP32-inner refinement solve (lognormal kappa, tol 1e-10) without graph capture, "
                         "for a kernel trace")
    ap.add_argument("--profile-mode", action="store_true",
                    help="short run for rocprofv3 counter passes: a few CSR and matrix-free applies and one "
                         "1 GiB device-to-device copy (the calibration kernel of known size); no JSON line")
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
        M64, M32 = build_mf(N, fes, mesh, ir, kappa, 0), build_mf(N, fes, mesh, ir, kappa, 1)
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
        for prec in (0, 1, 2, 3):
            M = build_mf(N, fes, mesh, ir, None, prec)
            for _ in range(3):
                M.mult(dx, dy)
                if prec:
                    M.mult_f32(x32, y32)
            del M
        big_a, big_b = N.DeviceVector(1 << 27, 8), N.DeviceVector(1 << 27, 8)  # 1 GiB each
        big_b.copy_from(big_a)
        N.synchronize()
        log("[bench] profile mode done")
        return

    # ---------------------------------------------------- CSR SpMV (value)
    ms_csr = timed(N, lambda: A.mult(dx, dy), args.steps, args.warmup)
    y_csr = dy.download()
    csr_bytes = nnz * 12 + (ndof + 1) * 4 + 2 * ndof * 8
    gbs = csr_bytes / (ms_csr * 1e-3) / 1e9
    gdofs = ndof / (ms_csr * 1e-3) / 1e9
    log(f"[bench] csr spmv {ms_csr * 1e3:.1f} us  {gbs:.0f} GB/s  {gdofs:.2f} GDoF/s")

    extra = {}
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
            "rel_err_vs_runs_kernel": float(np.abs(dy.download() - y_csr).max() / np.abs(y_csr).max())}
        del Aalt
        log(f"[bench] csr {algo} {ms_alt * 1e3:.1f} us")

    # ------------------------------------------------- matrix-free applies
    nq = len(ir)
    # flops of the form the kernels execute: y_e = sum_c S_c (g_ec x_e), six
    # symmetric nd x nd reference matrices per class (tet p=4: nd = 35)
    nd_el = fes.nd_el
    flops_mf = 2.0 * 6 * nd_el * nd_el * ne
    # the reference's per-quadrature-point form (435 Q flop per element),
    # reported as an equivalent rate only
    flops_mf_quad = 435.0 * nq * ne
    x32, y32 = N.DeviceVector(ndof, 4), N.DeviceVector(ndof, 4)
    x32.upload(x.astype(np.float32))
    for name, prec, peak in [("mf_f64", 0, FP64_PEAK_TFLOPS), ("mf_f32", 1, FP32_PEAK_TFLOPS),
                             ("mf_bf16", 2, BF16_PEAK_TFLOPS), ("mf_f16", 3, BF16_PEAK_TFLOPS)]:
        for scatter in (0, 1):
            M = build_mf(N, fes, mesh, ir, None, prec, scatter)
            ms = timed(N, lambda: M.mult(dx, dy), args.steps, args.warmup)
            err = float(np.abs(dy.download() - y_csr).max() / np.abs(y_csr).max())
            tf = flops_mf / (ms * 1e-3) / 1e12
            key = name + ("" if scatter == 0 else "_atomic")
            extra[key] = {"us_per_apply": round(ms * 1e3, 1), "gdofs": round(ndof / (ms * 1e-3) / 1e9, 3),
                          "tflops": round(tf, 2), "frac_of_peak": round(tf / peak, 4),
                          "per_point_equivalent_tflops": round(flops_mf_quad / (ms * 1e-3) / 1e12, 2),
                          "rel_err_vs_fp64_csr": err}
            if prec > 0:
                ms32 = timed(N, lambda: M.mult_f32(x32, y32), args.steps, args.warmup)
                extra[key]["us_per_apply_f32_vectors"] = round(ms32 * 1e3, 1)
            log(f"[bench] {key} {ms * 1e3:.1f} us  {tf:.1f} TFLOP/s  err {err:.1e}")
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

        t0 = time.perf_counter()
        its_k, hist_k = N.CGSolve(Ak, Ck, df, du, True, 1e-10, -1.0, 200000, 16)
        t_k = time.perf_counter() - t0
        u64 = du.download()
        log(f"[bench] FP64 CSR CG lognormal: {its_k} its {t_k:.3f}s")
        M64 = build_mf(N, fes, mesh, ir, kappa, 0)
        res = {"fp64_csr_cg": {"iterations": its_k, "seconds": round(t_k, 4), "true_rel_residual": true_res(u64),
                               # the reference's stopping quantity sqrt|w.d|, last over first
                               "stopping_ratio": float(hist_k[-1] / hist_k[0]) if len(hist_k) > 1 else None}}
        # the same FP64 CG on the FP64 matrix-free operator
        du_mf = N.DeviceVector(ndof, 8)
        t0 = time.perf_counter()
        its_mf, _ = N.CGSolve(M64, Ck, df, du_mf, True, 1e-10, -1.0, 200000, 16)
        t_mf = time.perf_counter() - t0
        u_mf = du_mf.download()
        log(f"[bench] FP64 matrix-free CG lognormal: {its_mf} its {t_mf:.3f}s")
        res["fp64_mf_cg"] = {"iterations": its_mf, "seconds": round(t_mf, 4), "true_rel_residual": true_res(u_mf),
                             "rel_diff_vs_csr_solution": float(np.linalg.norm(u_mf - u64) / np.linalg.norm(u64)),
                             "speedup_vs_fp64_csr_cg": round(t_k / t_mf, 3)}
        del du_mf
        best = None
        for pname, prec in [("f32", 1), ("bf16", 2), ("f16", 3)]:
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
                "true_rel_residual": true_res(u_ir),
                "rel_diff_vs_fp64_solution": float(np.linalg.norm(u_ir - u64) / np.linalg.norm(u64)),
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
                "true_rel_residual": true_res(u_b),
                "rel_diff_vs_jacobi_solution": float(np.linalg.norm(u_b - u64) / np.linalg.norm(u64))}
            del Cb
        if not args.no_cpu:
            # the same solve with the oracle's C Jacobi-PCG (OpenMP, all host threads)
            import oracle
            dk = np.empty(ndof)
            Ak.diagonal(dd)
            dk[:] = dd.download()
            invk = np.where(free, 1.0 / np.where(dk == 0, 1.0, dk), 0.0)
            t0 = time.perf_counter()
            u_cpu, its_cpu, _ = oracle.pcg_csr(rowptr, cols, vals_k, invk, f, tol=1e-10, maxiter=200000)
            t_cpu = time.perf_counter() - t0
            log(f"[bench] CPU oracle PCG lognormal: {its_cpu} its {t_cpu:.2f}s on {oracle.num_threads()} threads")
            res["cpu_oracle_pcg"] = {
                "iterations": its_cpu, "seconds": round(t_cpu, 3), "threads": oracle.num_threads(),
                "true_rel_residual": true_res(u_cpu),
                "gpu_fp64_csr_cg_speedup": round(t_cpu / t_k, 1),
                "gpu_best_mixed_speedup": round(t_cpu / best[1], 1) if best else None}
        extra["cg_lognormal_kappa"] = res

    if not args.no_batch:
        extra["batch_512_problems"] = bench_batch(N, args)
    extra["poisson_2d_64sq_p3"] = bench_2d(N, args)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(rowptr, cols, vals, x, ndof)
        log(f"[bench] cpu {cpu['value']} GDoF/s on {cpu['cores']} threads")

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
                     "kernel": "CsrRunsKernel", "algorithmic_bytes": csr_bytes},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


def _spmv_host(rowptr, cols, vals, x):
    import scipy.sparse as sp
    n = len(rowptr) - 1
    return sp.csr_matrix((vals, cols, rowptr), shape=(n, n)) @ x


if __name__ == "__main__":
    main()
