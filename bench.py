"""Benchmark: time-to-rtol of the preconditioned Krylov hot path on one MI355X.

Workload (one "step" = one complete solve of one system with the matrix and
right-hand side already resident in HBM): 3D 7-point convection-diffusion on a
128^3 grid (n = 2,097,152, nnz = 14,581,760; diffusion 1, velocity direction
(1, 0.5, 0.25) scaled to cell Peclet 0.5, central differences), b = A x* with
x* ~ U[-1, 1] (seed 1), x0 = 0, BiCGSTAB + ParILU(0) (20 sweeps, factorisation
inside the step), FP64, CSR SpMV, to relative residual 1e-10.  The ELL format
is timed beside it.

Printed: one JSON line (value = seconds per solve, lower is better) with the
dominant kernel's roofline (algorithmic bytes / mean kernel time from HIP
events recorded on the library's stream during the timed steps) and a CPU
baseline (the oracle's C restatement, one core, on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-rtol 1e-10 (s) + SpMV/SpTRSV HBM GB/s; mixed-prec vs FP64 on GPU"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n, nnz, nnz_l, nnz_u, ell_w):
    return {
        "spmv_csr": 12 * nnz + 4 * (n + 1) + 16 * n,
        "spmv_ell": 12 * n * ell_w + 16 * n,
        "trsv_lower": 12 * nnz_l + 4 * (n + 1) + 16 * n,
        "trsv_upper": 12 * nnz_u + 4 * (n + 1) + 16 * n,
    }


def run_gpu(A, b, fmt, steps, warmup, profile=True, precision=0, factor_storage=0, sweeps=None):
    import torch
    from xsolver_amd import XSolver

    dev = torch.device("cuda:0")
    rp, ci, v = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A)
    bd = torch.from_numpy(b).to(dev)
    xd = torch.zeros_like(bd)
    S = XSolver()
    S.mat_setup_device(rp, ci, v, use_ell=(fmt == "ell"))
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 2000)
    S.pc_settype("ilu0")
    S.set_option("parilu_sweeps", 20)
    S.set_option("parilu_async", 1)
    S.set_option("precision", precision)
    S.set_option("factor_storage", factor_storage)
    if sweeps is not None:   # the reference's iterative ILU apply (Jacobi sweeps)
        S.set_option("lsolve_sweeps", sweeps[0])
        S.set_option("usolve_sweeps", sweeps[1])

    def step():
        S.set_option("refactor", 1)
        S.pc_setup()
        xd.zero_()
        st = S.solve_device(bd, xd)
        assert st == 0, f"solve status {st}"

    for _ in range(warmup):
        step()
    S.set_option("profile", (1 << 1) | (1 << 2) | (1 << 3) if profile else 0)
    pc_ms, solve_ms, its = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        pc_ms.append(S.time_ms("pc_setup"))
        solve_ms.append(S.time_ms("solve"))
        its.append(S.iterations)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kern = {}
    for k in ("spmv", "trsv_lower", "trsv_upper"):
        cnt = S.time_ms("count:" + k)
        tot = S.time_ms("total:" + k)
        if cnt > 0:
            kern[k] = {"count": int(cnt), "total_ms": tot, "avg_us": 1000.0 * tot / cnt}
    x = xd.cpu().numpy()
    S.close()
    return {
        "sec_per_step": (t1 - t0) / steps,
        "pc_setup_ms": float(np.mean(pc_ms)),
        "solve_ms": float(np.mean(solve_ms)),
        "iterations": its,
        "kernels": kern,
        "x": x,
    }


# kernel families timed by the library -> kernel names in the rocprofv3 summary
PROFILE_NAMES = {"spmv": "k_spmv_csr<2>", "trsv_lower": "k_trsv<false>",
                 "trsv_upper": "k_trsv<true>", "spmv_ell": "k_spmv_ell<2>"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the 128^3 workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "convdiff3d_128_hbm_traffic.md")
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        return json.loads(block).get(PROFILE_NAMES.get(kernel, ""))
    except (OSError, IndexError, ValueError):
        return None


def cpu_baseline(A, b, iters_needed, sample_iters=8):
    """Oracle (C restatement, 1 core): ILU(0) set-up plus `sample_iters`
    BiCGSTAB iterations on the same system, scaled to `iters_needed`."""
    import oracle

    oracle.lib()
    t0 = time.perf_counter()
    oracle.solve(A, b, "bicgstab", "ilu0", tol=1e-10, maxiter=0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, st, it, _ = oracle.solve(A, b, "bicgstab", "ilu0", tol=1e-10, maxiter=sample_iters)
    t_run = time.perf_counter() - t0
    per_iter = max(t_run - t_setup, 0.0) / max(it, 1)
    est = t_setup + per_iter * iters_needed
    return {
        "value": round(est, 4),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle ILU(0) set-up ({t_setup:.2f} s) + {it} BiCGSTAB iterations "
                   f"({per_iter:.3f} s/iter) on the same 128^3 system, scaled to "
                   f"{iters_needed} iterations"),
    }


def bench_batch(args):
    """512 independent 2D variable-coefficient diffusion systems (s ~ U{32..160},
    kappa = exp(N(0, 0.5)) per cell, harmonic faces, seed 3), CG + Jacobi to
    rtol 1e-8 in one persistent launch, FP64 and FP32-inner/FP64-refine."""
    import torch
    from xsolver_amd import XBatch, batch_concat, gen

    sizes, systems = gen.batch_systems(512, 32, 160, seed=3)
    sys_ptr, rp, ci, v = batch_concat(systems)
    n = int(sys_ptr[-1])
    rng = np.random.default_rng(3)
    xs = rng.uniform(-1, 1, n)
    b = gen.spmv((rp, ci, v), xs)
    dev = torch.device("cuda:0")
    B = XBatch(sys_ptr, rp, ci, v)
    bd = torch.from_numpy(b).to(dev)
    out = {}
    for prec in (0, 1):
        xd = torch.zeros_like(bd)
        for _ in range(args.warmup):
            xd.zero_()
            B.solve_device(bd, xd, tol=1e-8, maxiter=5000, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            xd.zero_()
            B.solve_device(bd, xd, tol=1e-8, maxiter=5000, precision=prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        its, rel = B.results()
        out[prec] = {"sec": dt, "max_iters": int(its.max()), "max_relres": float(rel.max())}
    B.close()
    print(json.dumps({
        "metric": "batched solve time, 512 systems (s)", "value": round(out[0]["sec"], 5),
        "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * out[0]["sec"], 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, xsolver_amd/gen.py)",
        "config": {"workload": "batch512 vardiff2d s~U{32..160} CG+Jacobi rtol 1e-8",
                   "rows": n, "nnz": int(len(v)), "systems_per_s": round(512 / out[0]["sec"], 1),
                   "max_iterations": out[0]["max_iters"], "max_relres": out[0]["max_relres"],
                   "mixed_sec": round(out[1]["sec"], 5),
                   "mixed_max_relres": out[1]["max_relres"],
                   "mixed_speedup_vs_fp64": round(out[0]["sec"] / out[1]["sec"], 3)}}), flush=True)


def bench_rhs32(args):
    """32 right-hand sides on the 128^3 convection-diffusion matrix (RHS j:
    x*_j with seed 1 + j), BiCGSTAB + ParILU(0) to rtol 1e-10, one block solve
    (SpMM + 32-column SpTRSV) against 32 single solves."""
    import torch
    from xsolver_amd import XSolver, gen

    N, k = args.grid, 32
    A = gen.convdiff3d(N)
    n = len(A[0]) - 1
    B = np.stack([gen.manufactured(A, 1 + j)[1] for j in range(k)], 1)
    dev = torch.device("cuda:0")
    rp, ci, v = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A)
    S = XSolver()
    S.mat_setup_device(rp, ci, v)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 2000)
    S.pc_settype("ilu0")
    S.set_option("parilu_sweeps", 20)
    S.pc_setup()
    Bd = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    Xd = torch.zeros_like(Bd)
    for _ in range(args.warmup):
        Xd.zero_()
        S.solve_multi_device(Bd, Xd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Xd.zero_()
        S.solve_multi_device(Bd, Xd)
    torch.cuda.synchronize()
    block = (time.perf_counter() - t0) / args.steps
    its, r0, rl = S.multi_results()
    X = Xd.cpu().numpy()
    rel = max(float(np.linalg.norm(B[:, j] - gen.spmv(A, X[:, j])) / np.linalg.norm(B[:, j]))
              for j in range(k))
    # single solves of the same columns, back to back (4 sampled, scaled to 32)
    bj = torch.empty(n, dtype=torch.float64, device=dev)
    xj = torch.empty_like(bj)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in range(4):
        bj.copy_(Bd[:, j])
        xj.zero_()
        S.solve_device(bj, xj)
    torch.cuda.synchronize()
    single = (time.perf_counter() - t0) / 4
    # SpMM alone (k = 32): algorithmic bytes 12 nnz + 4(n+1) + 16 n k
    nnz = len(A[2])
    Y = torch.empty_like(Xd)
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        S.spmm(Xd, Y, st)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        S.spmm(Xd, Y, st)
    e1.record()
    torch.cuda.synchronize()
    spmm_us = e0.elapsed_time(e1) * 100.0
    spmm_bytes = 12 * nnz + 4 * (n + 1) + 16 * n * k
    S.close()
    print(json.dumps({
        "metric": "32-RHS block solve time (s)", "value": round(block, 5), "unit": "s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000 * block, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded generator, xsolver_amd/gen.py)",
        "config": {"workload": f"convdiff3d_{N} x 32 RHS BiCGSTAB+ParILU(0) FP64 rtol 1e-10",
                   "n": n, "k": k, "sec_per_rhs": round(block / k, 5),
                   "single_solve_sec": round(single, 5),
                   "speedup_vs_single_solves": round(single * k / block, 3),
                   "max_iterations": int(its.max()), "min_iterations": int(its.min()),
                   "max_true_relres": rel, "spmm_us": round(spmm_us, 2),
                   "spmm_GBs": round(spmm_bytes / spmm_us / 1e3, 1),
                   "spmm_bytes": spmm_bytes}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ell", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-sweeps", action="store_true",
                    help="skip the runs with the reference's Jacobi-sweep ILU apply")
    ap.add_argument("--workload", default="convdiff3d",
                    choices=["convdiff3d", "batch512", "rhs32"],
                    help="convdiff3d (default, the headline), batch512, rhs32")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        n = args.copy_calib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calib_bytes_read": 8 * n, "copy_calib_bytes_written": 8 * n}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (single-GPU kernels; the "
              "multi-GPU decomposition is design-only) -- run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch512":
        return bench_batch(args)
    if args.workload == "rhs32":
        return bench_rhs32(args)

    from xsolver_amd import gen

    N = args.grid
    A = gen.convdiff3d(N)
    xs, b = gen.manufactured(A, 1)
    n, nnz = len(A[0]) - 1, len(A[2])
    nnz_l = (nnz - n) // 2
    nnz_u = nnz - nnz_l
    by = algorithmic_bytes(n, nnz, nnz_l, nnz_u, int(np.diff(A[0]).max()))

    csr = run_gpu(A, b, "csr", args.steps, args.warmup)
    relres = float(np.linalg.norm(b - gen.spmv(A, csr["x"])) / np.linalg.norm(b))
    err_xs = float(np.linalg.norm(csr["x"] - xs) / np.linalg.norm(xs))
    ell = None if args.no_ell else run_gpu(A, b, "ell", args.steps, args.warmup)
    mixed = None if args.no_mixed else run_gpu(A, b, "csr", args.steps, args.warmup,
                                               precision=1, factor_storage=0)
    mixed_relres = (float(np.linalg.norm(b - gen.spmv(A, mixed["x"])) / np.linalg.norm(b))
                    if mixed else None)
    # the reference GPU backend's ILU apply: 20 Jacobi sweeps per factor
    # (hip/Precond/precon_gpuilu0.cpp:17-18), FP64, and FP32 inside refinement
    sw = None if args.no_sweeps else run_gpu(A, b, "csr", args.steps, args.warmup,
                                             profile=False, sweeps=(20, 20))
    swm = None if args.no_sweeps or args.no_mixed else run_gpu(
        A, b, "csr", args.steps, args.warmup, profile=False, precision=1, sweeps=(20, 20))
    def relres_of(r):
        return float(np.linalg.norm(b - gen.spmv(A, r["x"])) / np.linalg.norm(b)) if r else None
    mixed_err = (float(np.linalg.norm(mixed["x"] - csr["x"]) / np.linalg.norm(csr["x"]))
                 if mixed else None)

    # dominant kernel by total time in the timed CSR steps
    k = csr["kernels"]
    dom = max(k, key=lambda name: k[name]["total_ms"]) if k else None
    roof = None
    if dom:
        bytes_per = by["spmv_csr"] if dom == "spmv" else by[dom]
        achieved = bytes_per / (k[dom]["avg_us"] * 1e-6) / 1e9
        traffic = measured_traffic(dom) if N == 128 else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": dom, "bytes_per_launch": bytes_per,
                "avg_us": round(k[dom]["avg_us"], 2)}
    kernels_gbs = {}
    for name, rec in k.items():
        bp = by["spmv_csr"] if name == "spmv" else by[name]
        kernels_gbs[name] = {"avg_us": round(rec["avg_us"], 2),
                             "GB/s": round(bp / (rec["avg_us"] * 1e-6) / 1e9, 1),
                             "launches": rec["count"]}
    if ell and "spmv" in ell["kernels"]:
        r = ell["kernels"]["spmv"]
        kernels_gbs["spmv_ell"] = {"avg_us": round(r["avg_us"], 2),
                                   "GB/s": round(by["spmv_ell"] / (r["avg_us"] * 1e-6) / 1e9, 1),
                                   "launches": r["count"]}

    iters = int(np.median(csr["iterations"]))
    cpu = None if args.no_cpu else cpu_baseline(A, b, iters)

    out = {
        "metric": METRIC,
        "value": round(csr["sec_per_step"], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * csr["sec_per_step"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, xsolver_amd/gen.py)",
        "config": {
            "workload": f"convdiff3d_{N} BiCGSTAB+ParILU(0) FP64 rtol 1e-10 (CSR)",
            "n": n, "nnz": nnz, "format": "csr", "parilu_sweeps": 20,
            "iterations": iters, "pc_setup_ms": round(csr["pc_setup_ms"], 3),
            "solve_ms": round(csr["solve_ms"], 3), "true_relres": relres,
            "rel_err_vs_xstar": err_xs,
            "ell_sec_per_step": round(ell["sec_per_step"], 5) if ell else None,
            "ell_iterations": int(np.median(ell["iterations"])) if ell else None,
            "mixed_sec_per_step": round(mixed["sec_per_step"], 5) if mixed else None,
            "mixed_inner_iterations": int(np.median(mixed["iterations"])) if mixed else None,
            "mixed_true_relres": mixed_relres,
            "mixed_rel_err_vs_fp64_gpu": mixed_err,
            "mixed_speedup_vs_fp64": (round(csr["sec_per_step"] / mixed["sec_per_step"], 3)
                                      if mixed else None),
            "sweep20_sec_per_step": round(sw["sec_per_step"], 5) if sw else None,
            "sweep20_iterations": int(np.median(sw["iterations"])) if sw else None,
            "sweep20_true_relres": relres_of(sw),
            "sweep20_mixed_sec_per_step": round(swm["sec_per_step"], 5) if swm else None,
            "sweep20_mixed_inner_iterations": int(np.median(swm["iterations"])) if swm else None,
            "sweep20_mixed_true_relres": relres_of(swm),
        },
        "kernels": kernels_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
