"""Benchmark: time-to-rtol of the preconditioned Krylov hot path on one MI355X.

Workload (one "step" = one complete solve of one system with the matrix and
right-hand side already resident in HBM): 3D 7-point convection-diffusion on a
128^3 grid (n = 2,097,152, nnz = 14,581,760; diffusion 1, velocity direction
(1, 0.5, 0.25) scaled to cell Peclet 0.5, central differences), b = A x* with
x* ~ U[-1, 1] (seed 1), x0 = 0, CSR input, BiCGSTAB + ParILU(0) (20
asynchronous sweeps, factorisation inside the step), FP64, to relative
residual 1e-10.  The ILU factors are applied the way the reference's GPU
backend applies them -- Jacobi sweeps per factor -- with 5 sweeps on L and 2
on U in the headline configuration (the reference hard-codes 20 / 20; 5 / 2
gave the shortest time-to-rtol of the sweep counts tried on this system);
the exact sync-free SpTRSV apply, the reference's 20/20 sweeps, ELL input and
the FP32-inner mixed solves (FP32- and BF16-stored factors) are timed beside
it.

The same JSON line carries, under "workloads", every other configuration of
the hot path, each timed on the GPU with the matrix resident in HBM: the
27-point lognormal 100^3 GMRES(30) + ParILUT system (FP64 and mixed
precision, set-up included), 512 independent systems in one persistent
launch, 32 right-hand sides on the 128^3 matrix, the 512^2 adjoint
(forward + transposed solve + gradient gather), and SpMV kernel timings
(ELL and row-binned vector CSR, with warm and with flushed last-level cache,
beside a flushed-cache device copy that sets the practical bandwidth
ceiling).

Printed: one JSON line (value = seconds per solve, lower is better) with the
dominant kernel's roofline (algorithmic bytes / mean time from HIP events
recorded on the library's stream during the timed steps) and a CPU baseline
(the oracle's C restatement of the reference CPU path -- ILU(0) set-up, then
BiCGSTAB with the exact LU apply -- run to convergence on the host cores,
SpMV and BLAS-1 on OpenMP threads, triangular solves serial as in the
reference).
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


APPLY = (5, 2)   # Jacobi sweeps per factor (L, U) in the headline configuration


def algorithmic_bytes(n, nnz, nnz_l, nnz_u, sweeps, idx_a=4, idx_f=4):
    """Bytes each kernel (or apply sequence) must move at least: 8 B of value
    plus the stored column (idx_a for A's ELL copy, idx_f for the factors'
    sweep copies: 2 with 16-bit offsets, else 4) per entry, 8 B per vector
    entry read or written.  nnz_l / nnz_u: strictly lower / upper entries of
    the factors.  An apply of k >= 2 sweeps is one fused launch (sweeps 1-2,
    no previous iterate) plus k - 2 plain ones; U also reads 1/u_ii.
    Of BiCGSTAB's two L applies per iteration the second forms s = r - gama u
    in its fused launch (reads r and u, writes s: 16n more than the plain
    fused launch's 16n); the first is the plain one (p has its own update
    kernel by default), so the L apply is counted at their mean, 8n above the
    plain fused launch.  The exact SpTRSV reads CSR factors (4-byte
    columns)."""
    kl, ku = sweeps
    ea, ef = 8 + idx_a, 8 + idx_f
    l_fused, l_plain = ef * nnz_l + 24 * n, ef * nnz_l + 24 * n
    u_fused, u_plain = ef * nnz_u + 24 * n, ef * nnz_u + 32 * n
    return {
        "spmv": ea * nnz + 16 * n,                  # ELL values + columns, x, y
        "ilu_lower_apply": l_fused + (kl - 2) * l_plain,
        "ilu_upper_apply": u_fused + (ku - 2) * u_plain,
        "trsv_lower_exact": 12 * nnz_l + 4 * (n + 1) + 4 * n + 16 * n,
        "trsv_upper_exact": 12 * nnz_u + 4 * (n + 1) + 4 * n + 24 * n,
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
    idx = (S.time_ms("spmv:idx16"), S.time_ms("pc:idx16"), S.time_ms("spmv:dia"),
           S.time_ms("pc:dia"))
    S.close()
    return {
        "idx16": idx,
        "sec_per_step": (t1 - t0) / steps,
        "pc_setup_ms": float(np.mean(pc_ms)),
        "solve_ms": float(np.mean(solve_ms)),
        "iterations": its,
        "kernels": kern,
        "x": x,
    }


# library timing families -> (kernel name in the rocprofv3 summary, launches)
PROFILE_NAMES = {
    "spmv": [("k_spmv_dia<2>", 1)],
    "ilu_lower_apply": [("k_tri_sweep_dia<false, true, false>", 0.5),
                        ("k_tri_sweep_dia<false, true, true>", 0.5),
                        ("k_tri_sweep_dia<false, false, false>", APPLY[0] - 2)],
    "ilu_upper_apply": [("k_tri_sweep_dia<true, true, false>", 1),
                        ("k_tri_sweep_dia<true, false, false>", APPLY[1] - 2)],
    "trsv_lower_exact": [("k_trsv<false, 4>", 1)],
    "trsv_upper_exact": [("k_trsv<true, 4>", 1)],
}


def measured_traffic(family):
    """HBM bytes per launch (per apply sequence) of `family` on the 128^3
    workload from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "convdiff3d_128_hbm_traffic.md")
    try:
        text = open(path).read()
        block = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        parts = [block.get(k) for k, m in PROFILE_NAMES.get(family, []) if m > 0]
        if not parts or any(p is None for p in parts):
            return None
        return int(sum(block[k] * m for k, m in PROFILE_NAMES[family] if m > 0))
    except (OSError, IndexError, ValueError):
        return None


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, b, sweeps):
    """The oracle (the C restatement of the reference CPU path) running the
    whole solve on the host: ILU(0) set-up (seqilu0,
    hip/HostMatrix/HostMatrixCSR.cpp:164-197), then BiCGSTAB
    (hip/Solver/xsolver_bicgstab.cpp:20-124) with the exact LU apply
    (LUsolve, :385-409) to rtol 1e-10.  SpMV and BLAS-1 run on OMP_NUM_THREADS
    OpenMP threads (the reference uses threaded MKL there); the triangular
    solves are serial as in the reference.  Also timed: the same solve with
    the headline's Jacobi-sweep apply."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    oracle.lib()
    t0 = time.perf_counter()
    _, st, it, _ = oracle.solve(A, b, "bicgstab", "ilu0", tol=1e-10, maxiter=2000)
    t_exact = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, st2, it2, _ = oracle.solve(A, b, "bicgstab", "ilu0", tol=1e-10, maxiter=2000,
                                  lsweeps=sweeps[0], usweeps=sweeps[1])
    t_sw = time.perf_counter() - t0
    return {
        "value": round(t_exact, 4),
        "unit": "s",
        "cores": threads,
        "kind": "port",
        "cpu": _cpu_model(),
        "sample": (f"the whole solve, no extrapolation: ILU(0) set-up + {it} BiCGSTAB iterations "
                   f"with the exact LU apply to rtol 1e-10 (status {st}); SpMV/BLAS-1 on {threads} "
                   f"OpenMP threads, triangular solves serial"),
        "sweeps_apply_s": round(t_sw, 4),
        "sweeps_apply_iterations": it2,
    }


def bench_batch(steps=3, warmup=1):
    """512 independent 2D variable-coefficient diffusion systems (s ~ U{32..160},
    kappa = exp(N(0, 0.5)) per cell, harmonic faces, seed 3), CG + Jacobi to
    rtol 1e-8 in one persistent launch, FP64 and FP32 CG with FP64 reliable
    updates."""
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
        for _ in range(warmup):
            xd.zero_()
            B.solve_device(bd, xd, tol=1e-8, maxiter=5000, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            xd.zero_()
            B.solve_device(bd, xd, tol=1e-8, maxiter=5000, precision=prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        its, rel = B.results()
        out[prec] = {"sec": dt, "max_iters": int(its.max()), "max_relres": float(rel.max())}
    B.close()
    return {"workload": "batch512 vardiff2d s~U{32..160} CG+Jacobi rtol 1e-8, one persistent launch",
            "sec": round(out[0]["sec"], 5), "rows": n, "nnz": int(len(v)),
            "systems_per_s": round(512 / out[0]["sec"], 1),
            "max_iterations": out[0]["max_iters"], "max_relres": out[0]["max_relres"],
            "mixed_sec": round(out[1]["sec"], 5), "mixed_max_relres": out[1]["max_relres"],
            "mixed_max_iterations": out[1]["max_iters"],
            "mixed_speedup_vs_fp64": round(out[0]["sec"] / out[1]["sec"], 3)}


def bench_rhs32(grid=128, steps=2, warmup=1):
    """32 right-hand sides on the 128^3 convection-diffusion matrix (RHS j:
    x*_j with seed 1 + j), BiCGSTAB + ParILU(0) with the headline's sweep
    apply to rtol 1e-10, one block solve (SpMM + 32-column sweeps) against 32
    single solves."""
    import torch
    from xsolver_amd import XSolver, gen

    N, k = grid, 32
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
    S.set_option("lsolve_sweeps", APPLY[0])
    S.set_option("usolve_sweeps", APPLY[1])
    S.pc_setup()
    Bd = torch.from_numpy(np.ascontiguousarray(B)).to(dev)
    Xd = torch.zeros_like(Bd)
    for _ in range(warmup):
        Xd.zero_()
        S.solve_multi_device(Bd, Xd)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        Xd.zero_()
        S.solve_multi_device(Bd, Xd)
    torch.cuda.synchronize()
    block = (time.perf_counter() - t0) / steps
    its, r0, rl = S.multi_results()
    X = Xd.cpu().numpy()
    rel = max(float(np.linalg.norm(B[:, j] - gen.spmv(A, X[:, j])) / np.linalg.norm(B[:, j]))
              for j in range(k))
    # single solves of the same columns, back to back (4 sampled, scaled to 32)
    bj = torch.empty(n, dtype=torch.float64, device=dev)
    xj = torch.empty_like(bj)
    bj.copy_(Bd[:, 0])
    xj.zero_()
    S.solve_device(bj, xj)        # untimed: the single solve's first-use set-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j in range(4):
        bj.copy_(Bd[:, j])
        xj.zero_()
        S.solve_device(bj, xj)
    torch.cuda.synchronize()
    single = (time.perf_counter() - t0) / 4
    # SpMM alone (k = 32): algorithmic bytes (8 + index bytes) nnz + 16 n k,
    # + 4(n+1) row pointers for the CSR kernel
    nnz = len(A[2])
    spmm_ell = S.time_ms("spmv:kernel") == 2
    idx_bytes = 2 if (spmm_ell and S.time_ms("spmv:idx16")) else 4
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
    spmm_bytes = (8 + idx_bytes) * nnz + (0 if spmm_ell else 4 * (n + 1)) + 16 * n * k
    S.close()
    return {"workload": (f"convdiff3d_{N} x 32 RHS BiCGSTAB+ParILU(0) FP64 rtol 1e-10, "
                         f"ILU apply {APPLY[0]}/{APPLY[1]} Jacobi sweeps"),
            "sec": round(block, 5), "n": n, "k": k, "sec_per_rhs": round(block / k, 5),
            "single_solve_sec": round(single, 5),
            "speedup_vs_single_solves": round(single * k / block, 3),
            "max_iterations": int(its.max()), "min_iterations": int(its.min()),
            "max_true_relres": rel, "spmm_us": round(spmm_us, 2),
            "spmm_GBs": round(spmm_bytes / spmm_us / 1e3, 1), "spmm_bytes": spmm_bytes,
            "spmm_kernel": ("ELL, 16-bit offsets" if idx_bytes == 2 else "ELL")
            if spmm_ell else "CSR"}


def bench_c3(steps=3, warmup=1, exact=True):
    """27-point lognormal system on 100^3 (n = 1,000,000, nnz = 26,463,592,
    seed 2), GMRES(30) + ParILUT (fill 2, drop 1e-4) to rtol 1e-10: FP64
    against FP32-inner / FP64-outer with BF16- and FP32-stored factors, the
    factors applied by the reference's ParILUT sweeps (10 on L, 15 on U) and
    by the exact sync-free solve.  Time-to-rtol = ParILUT set-up + solve."""
    import torch
    from xsolver_amd import XSolver, gen

    A = gen.lognormal27(100, seed=2)
    xs, b = gen.manufactured(A, 2)
    n, nnz = len(A[0]) - 1, len(A[2])
    dev = torch.device("cuda:0")
    rp, ci, v = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A)
    bd = torch.from_numpy(b).to(dev)
    out = {}
    runs = [("fp64", 0, 0, (10, 15)), ("mixed_bf16", 1, 1, (10, 15)), ("mixed_fp32", 1, 0, (10, 15))]
    if exact:
        runs += [("fp64_exact", 0, 0, None), ("mixed_bf16_exact", 1, 1, None)]
    for name, prec, fs, sw in runs:
        S = XSolver()
        S.mat_setup_device(rp, ci, v)
        S.ksp_settype("gmres")
        S.ksp_setoption(1e-10, 3000, 30)
        S.pc_settype("ilut")
        S.pc_setilut(1e-4, 2)
        S.set_option("precision", prec)
        S.set_option("factor_storage", fs)
        if sw:
            S.set_option("lsolve_sweeps", sw[0])
            S.set_option("usolve_sweeps", sw[1])
        setups, ts = [], []
        xd = torch.zeros_like(bd)
        for rep in range(warmup + steps):
            S.set_option("refactor", 1)
            xd.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            S.pc_setup()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st = S.solve_device(bd, xd)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            assert st == 0, f"{name}: status {st}"
            if rep >= warmup:
                setups.append(t1 - t0)
                ts.append(t2 - t1)
        x = xd.cpu().numpy()
        su, so = float(np.mean(setups)), float(np.mean(ts))
        out[name] = {"time_to_rtol_s": round(su + so, 5), "setup_s": round(su, 5),
                     "solve_s": round(so, 5), "iterations": S.iterations,
                     "true_relres": float(np.linalg.norm(b - gen.spmv(A, x)) / np.linalg.norm(b))}
        S.close()
    f64 = out["fp64"]
    return {"workload": "lognormal27_100 GMRES(30)+ParILUT(fill 2, drop 1e-4) rtol 1e-10, "
                        "ILU apply 10/15 sweeps; time-to-rtol = ParILUT set-up + solve",
            "n": n, "nnz": nnz, "runs": out,
            "mixed_bf16_speedup_vs_fp64": round(f64["time_to_rtol_s"] / out["mixed_bf16"]["time_to_rtol_s"], 3),
            "mixed_bf16_solve_speedup_vs_fp64": round(f64["solve_s"] / out["mixed_bf16"]["solve_s"], 3),
            "mixed_fp32_speedup_vs_fp64": round(f64["time_to_rtol_s"] / out["mixed_fp32"]["time_to_rtol_s"], 3)}


def bench_mixed(grid=128, steps=3, warmup=1):
    """The headline system's precision comparison alone (for profiling the
    per-iteration kernel budget): FP64 and the FP32 BiCGSTAB recurrence with
    FP64 reliable updates, FP32- and BF16-stored sweep copies, 5/2 and 5/3
    sweeps, re-factorisation inside every step as in the headline."""
    from xsolver_amd import gen

    A = gen.convdiff3d(grid)
    _, b = gen.manufactured(A, 1)
    out = {}
    for name, prec, fs, sw in (("fp64", 0, 0, APPLY), ("mixed_fp32", 1, 0, APPLY),
                               ("mixed_bf16", 1, 1, APPLY), ("mixed_bf16_5_3", 1, 1, (5, 3))):
        r = run_gpu(A, b, "csr", steps, warmup, profile=False, precision=prec,
                    factor_storage=fs, sweeps=sw)
        out[name] = {"sec_per_step": round(r["sec_per_step"], 5),
                     "pc_setup_ms": round(r["pc_setup_ms"], 3),
                     "iterations": int(np.median(r["iterations"])),
                     "true_relres": float(np.linalg.norm(b - gen.spmv(A, r["x"])) / np.linalg.norm(b))}
        print(f"bench: mixed {name} done", file=sys.stderr, flush=True)
    f64 = out["fp64"]["sec_per_step"]
    for name in out:
        out[name]["speedup_vs_fp64"] = round(f64 / out[name]["sec_per_step"], 3)
    return {"workload": f"convdiff3d_{grid} BiCGSTAB+ParILU(0) rtol 1e-10, FP64 vs mixed",
            "runs": out}


def bench_adjoint(steps=3, warmup=1):
    """Adjoint on the 512^2 variable-coefficient diffusion system (kappa =
    exp of a squared-exponential Gaussian random field, length 0.1, seed 4),
    CG + ParILU(0) FP64: forward solve to rtol 1e-10, transposed solve
    A^T lam = 2 (x - x_t) (transposed SpMV, transposed ILU triangular
    solves), gradient gather dL/dA_ij = -lam_i x_j on the pattern."""
    import torch
    from xsolver_amd import XSolver, gen

    A = gen.vardiff2d(gen.grf_kappa(512, 0.1, 4))
    rp, ci, v = A
    n, nnz = len(rp) - 1, len(v)
    xt = np.random.default_rng([4, 1]).uniform(-1, 1, n)
    b = gen.spmv(A, np.random.default_rng([4, 2]).uniform(-1, 1, n))
    dev = torch.device("cuda:0")
    S = XSolver()
    S.mat_setup(rp, ci, v)
    S.ksp_settype("cg")
    S.ksp_setoption(1e-10, 20000)
    S.pc_settype("ilu0")
    S.pc_setup()
    bd = torch.from_numpy(b).to(dev)
    xtd = torch.from_numpy(xt).to(dev)
    xd = torch.zeros_like(bd)
    lam = torch.zeros_like(bd)
    gv = torch.empty(nnz, dtype=torch.float64, device=dev)
    tf, ta = [], []
    for rep in range(warmup + steps):
        xd.zero_()
        lam.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        assert S.solve_device(bd, xd) == 0
        itf = S.iterations
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        g = 2.0 * (xd - xtd)
        assert S.solve_transpose_device(g, lam) == 0
        ita = S.iterations
        S.grad_values(lam, xd, gv, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if rep >= warmup:
            tf.append(t1 - t0)
            ta.append(t2 - t1)
    S.close()
    return {"workload": "grf2d_512 CG+ParILU(0) FP64 forward + adjoint (A^T lam = dL/dx) + "
                        "gradient gather",
            "n": n, "nnz": nnz, "forward_s": round(float(np.mean(tf)), 5),
            "adjoint_and_gradient_s": round(float(np.mean(ta)), 5),
            "time_to_gradient_s": round(float(np.mean(tf) + np.mean(ta)), 5),
            "forward_iterations": itf, "adjoint_iterations": ita}


def bench_spmv_kernels(cases=("convdiff3d_128", "lognormal27_100"), reps=20):
    """SpMV y = A x alone through xs_spmv on HIP events of the launching
    stream: the ELL kernel and the row-binned vector-CSR kernel on each
    matrix, back to back (warm: the 256 MB last-level cache keeps part of
    the operands) and with the cache flushed before every launch (a 1 GiB
    buffer written in between).  Beside them the practical ceiling measured
    the same way: a flushed-cache device copy (torch copy_, 1 GiB)."""
    import torch
    from xsolver_amd import XSolver, gen

    dev = torch.device("cuda:0")
    st = torch.cuda.current_stream().cuda_stream
    flush = torch.empty(1 << 27, dtype=torch.float64, device=dev)   # 1 GiB

    def timed(fn, cold):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        fn()
        for e0, e1 in ev:
            if cold:
                flush.fill_(1.0)
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        return float(np.median([e0.elapsed_time(e1) * 1000.0 for e0, e1 in ev]))

    out = {}
    src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(2.0)
    dst = torch.empty_like(src)
    cu = timed(lambda: dst.copy_(src), True)
    wu = timed(lambda: dst.copy_(src), False)
    out["copy_1GiB"] = {"cold_us": round(cu, 1), "cold_GBs": round(2 * 8 * src.numel() / cu / 1e3, 1),
                        "warm_us": round(wu, 1), "warm_GBs": round(2 * 8 * src.numel() / wu / 1e3, 1)}
    del src, dst
    for name in cases:
        A = gen.convdiff3d(128) if name == "convdiff3d_128" else gen.lognormal27(100, seed=2)
        n, nnz = len(A[0]) - 1, len(A[2])
        rp, ci, v = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in A)
        x = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, n)).to(dev)
        y = torch.empty_like(x)
        rec = {"n": n, "nnz": nnz}
        for kname, ell, vec, dia in (("diagonal_slots", True, None, 1), ("ell", True, None, 0),
                                     ("vector_csr", False, 1, 0)):
            S = XSolver()
            S.set_option("spmv_dia", dia)
            S.mat_setup_device(rp, ci, v, use_ell=ell)
            if vec is not None:
                S.set_option("spmv_kernel", vec)
            kind = S.time_ms("spmv:kernel")
            if (ell and kind != 2) or (dia and S.time_ms("spmv:dia") == 0):
                S.close()
                continue
            idx = 0 if dia else 2 if (ell and S.time_ms("spmv:idx16") == 1) else 4
            # bytes the kernel must move: values + stored columns (+ row
            # pointers for CSR), x once, y once
            by = (8 + idx) * nnz + (0 if ell else 4 * (n + 1)) + 16 * n
            warm = timed(lambda: S.spmv(x, y, st), False)
            cold = timed(lambda: S.spmv(x, y, st), True)
            rec[kname] = {"bytes": by, "warm_us": round(warm, 2),
                          "warm_GBs": round(by / warm / 1e3, 1), "cold_us": round(cold, 2),
                          "cold_GBs": round(by / cold / 1e3, 1),
                          "cold_frac_of_peak": round(by / cold / 1e3 / HBM_PEAK_GBS, 3)}
            S.close()
        out[name] = rec
        del rp, ci, v, x, y
    return out


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
                    choices=["convdiff3d", "batch512", "rhs32", "c3", "adjoint", "spmv", "mixed"],
                    help="convdiff3d (default: the headline line, with every other workload "
                         "under 'workloads'), or one workload alone (for profiling): batch512, "
                         "rhs32, c3 (27-point 100^3 GMRES+ParILUT, mixed precision vs FP64), "
                         "adjoint (512^2), spmv (SpMV kernels warm / flushed-cache, copy ceiling), "
                         "mixed (the headline system, FP64 vs mixed precision only)")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the other workloads of the default line")
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

    single = {"batch512": lambda: bench_batch(args.steps, args.warmup),
              "rhs32": lambda: bench_rhs32(args.grid, args.steps, args.warmup),
              "c3": lambda: bench_c3(args.steps, args.warmup),
              "adjoint": lambda: bench_adjoint(args.steps, args.warmup),
              "spmv": lambda: bench_spmv_kernels(),
              "mixed": lambda: bench_mixed(args.grid, args.steps, args.warmup)}
    if args.workload in single:
        rec = single[args.workload]()
        print(json.dumps({"workload": args.workload, "steps": args.steps,
                          "warmup": args.warmup, "result": rec}), flush=True)
        return

    from xsolver_amd import gen

    N = args.grid
    A = gen.convdiff3d(N)
    xs, b = gen.manufactured(A, 1)
    n, nnz = len(A[0]) - 1, len(A[2])
    nnz_l = (nnz - n) // 2          # symmetric pattern: strictly lower = strictly upper

    def relres_of(r):
        return float(np.linalg.norm(b - gen.spmv(A, r["x"])) / np.linalg.norm(b)) if r else None

    def its(r):
        return int(np.median(r["iterations"])) if r else None

    def sec(r):
        return round(r["sec_per_step"], 5) if r else None

    # timed without per-kernel events (they add gaps between launches); the
    # kernel figures come from separate profiled runs of the same solves
    head = run_gpu(A, b, "csr", args.steps, args.warmup, profile=False, sweeps=APPLY)
    relres = relres_of(head)
    err_xs = float(np.linalg.norm(head["x"] - xs) / np.linalg.norm(xs))
    headp = run_gpu(A, b, "csr", 1, 1, sweeps=APPLY)
    exact = run_gpu(A, b, "csr", args.steps, args.warmup, profile=False)   # sync-free SpTRSV
    exactp = run_gpu(A, b, "csr", 1, 1)
    mixed = None if args.no_mixed else run_gpu(A, b, "csr", args.steps, args.warmup,
                                               profile=False, precision=1, sweeps=APPLY)
    mixed_bf16 = None if args.no_mixed else run_gpu(A, b, "csr", args.steps, args.warmup,
                                                    profile=False, precision=1,
                                                    factor_storage=1, sweeps=APPLY)
    # the mixed solve at its own best apply (5/3: the FP32 recurrence gains
    # more from a third U sweep than FP64 does, measured)
    mixed_bf16_53 = None if args.no_mixed else run_gpu(A, b, "csr", args.steps, args.warmup,
                                                       profile=False, precision=1,
                                                       factor_storage=1, sweeps=(5, 3))
    mixed_exact = None if args.no_mixed else run_gpu(A, b, "csr", args.steps, args.warmup,
                                                     profile=False, precision=1)
    # the reference GPU backend's own apply: 20 Jacobi sweeps per factor
    # (hip/Precond/precon_gpuilu0.cpp:17-18)
    sw = None if args.no_sweeps else run_gpu(A, b, "csr", args.steps, args.warmup,
                                             profile=False, sweeps=(20, 20))
    ell = None if args.no_ell else run_gpu(A, b, "ell", args.steps, args.warmup,
                                           profile=False, sweeps=APPLY)

    ia, ifac, dia, fdia = headp["idx16"]
    # A's and the factors' copies: diagonal slots (no index bytes), else ELL
    # with 16- or 32-bit columns
    by = algorithmic_bytes(n, nnz, nnz_l, nnz_l, APPLY, 0 if dia > 0 else 2 if ia == 1 else 4,
                           0 if fdia > 0 else 2 if ifac == 1 else 4)
    # kernel families: the headline run's SpMV and ILU applies, the exact run's SpTRSV
    fam = {"spmv": headp["kernels"].get("spmv"),
           "ilu_lower_apply": headp["kernels"].get("trsv_lower"),
           "ilu_upper_apply": headp["kernels"].get("trsv_upper"),
           "trsv_lower_exact": exactp["kernels"].get("trsv_lower"),
           "trsv_upper_exact": exactp["kernels"].get("trsv_upper")}
    fam = {k: v for k, v in fam.items() if v}
    kernels_gbs = {name: {"avg_us": round(rec["avg_us"], 2),
                          "GB/s": round(by[name] / (rec["avg_us"] * 1e-6) / 1e9, 1),
                          "launches": rec["count"]} for name, rec in fam.items()}
    # dominant family by total time in the profiled headline solve
    headfam = {k: v for k, v in fam.items() if not k.endswith("_exact")}
    dom = max(headfam, key=lambda name: headfam[name]["total_ms"]) if headfam else None
    roof = None
    if dom:
        rec = headfam[dom]
        achieved = by[dom] / (rec["avg_us"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(dom) if N == 128 else None,
                "kernel": dom, "launches": PROFILE_NAMES[dom],
                "bytes_per_launch": by[dom], "avg_us": round(rec["avg_us"], 2)}

    iters = its(head)
    cpu = None if args.no_cpu else cpu_baseline(A, b, APPLY)
    # every other workload of the hot path, each bounded to a few timed runs
    workloads = {}
    if not args.headline_only:
        for name, fn in (("spmv_kernels", bench_spmv_kernels), ("lognormal27_100", bench_c3),
                         ("batch512", bench_batch), ("rhs32", bench_rhs32),
                         ("grf2d_512_adjoint", bench_adjoint)):
            t0 = time.perf_counter()
            workloads[name] = fn()
            workloads[name]["bench_wall_s"] = round(time.perf_counter() - t0, 1)
            print(f"bench: {name} done ({workloads[name]['bench_wall_s']} s)", file=sys.stderr,
                  flush=True)

    out = {
        "metric": METRIC,
        "value": round(head["sec_per_step"], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000 * head["sec_per_step"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, xsolver_amd/gen.py)",
        "config": {
            "workload": (f"convdiff3d_{N} BiCGSTAB+ParILU(0) FP64 rtol 1e-10, CSR input, "
                         f"ILU apply {APPLY[0]}/{APPLY[1]} Jacobi sweeps"),
            "n": n, "nnz": nnz, "format": "csr", "parilu_sweeps": 20,
            "ilu_apply_sweeps": list(APPLY), "iterations": iters,
            "pc_setup_ms": round(head["pc_setup_ms"], 3), "solve_ms": round(head["solve_ms"], 3),
            "true_relres": relres, "rel_err_vs_xstar": err_xs,
            "exact_trsv_sec_per_step": sec(exact), "exact_trsv_iterations": its(exact),
            "exact_trsv_true_relres": relres_of(exact),
            "mixed_sec_per_step": sec(mixed), "mixed_inner_iterations": its(mixed),
            "mixed_true_relres": relres_of(mixed),
            "mixed_speedup_vs_fp64": (round(head["sec_per_step"] / mixed["sec_per_step"], 3)
                                      if mixed else None),
            "mixed_bf16_factors_sec_per_step": sec(mixed_bf16),
            "mixed_bf16_factors_iterations": its(mixed_bf16),
            "mixed_bf16_factors_true_relres": relres_of(mixed_bf16),
            "mixed_bf16_factors_speedup_vs_fp64": (
                round(head["sec_per_step"] / mixed_bf16["sec_per_step"], 3) if mixed_bf16 else None),
            "mixed_bf16_factors_apply_5_3_sec_per_step": sec(mixed_bf16_53),
            "mixed_bf16_factors_apply_5_3_iterations": its(mixed_bf16_53),
            "mixed_bf16_factors_apply_5_3_true_relres": relres_of(mixed_bf16_53),
            "mixed_bf16_factors_apply_5_3_speedup_vs_fp64": (
                round(head["sec_per_step"] / mixed_bf16_53["sec_per_step"], 3)
                if mixed_bf16_53 else None),
            "mixed_exact_trsv_sec_per_step": sec(mixed_exact),
            "mixed_exact_trsv_speedup_vs_fp64": (
                round(exact["sec_per_step"] / mixed_exact["sec_per_step"], 3)
                if mixed_exact else None),
            "sweep20_sec_per_step": sec(sw), "sweep20_iterations": its(sw),
            "sweep20_true_relres": relres_of(sw),
            "ell_input_sec_per_step": sec(ell), "ell_input_iterations": its(ell),
        },
        "kernels": kernels_gbs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "workloads": workloads,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
