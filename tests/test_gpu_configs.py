"""Parity at the benchmarked configurations' full sizes, through the C-ABI.

* 128^3 convection-diffusion, BiCGSTAB + ParILU(0), rtol 1e-10: CSR input
  through the row-binned vector-CSR SpMV and ELL input, each with the 5/2
  and the reference's 20/20 Jacobi-sweep ILU apply and with the exact
  sync-free SpTRSV, against the oracle running the same apply (oracle:
  ILU(0) = seqilu0, hip/HostMatrix/HostMatrixCSR.cpp:164-197; exact apply =
  LUsolve, :385-409; BiCGSTAB = hip/Solver/xsolver_bicgstab.cpp:20-124).
  The GPU factorisation runs deterministic ParILU sweeps on the unscaled
  matrix to ILU(0)'s fixed point.  Acceptance: iterations within +-2, FP64
  true relative residual below 1e-10, ||x - x_oracle|| / ||x_oracle|| <=
  1e-10, the residual history within 1e-6.  The benchmark's own setting (20
  asynchronous sweeps) is checked beside it.
* 512 independent systems (2D variable-coefficient diffusion, s ~ U{32..160}),
  CG + Jacobi to rtol 1e-8 in one persistent launch, each against the
  oracle's CG on that system.
* 512^2 adjoint: gradients of L = ||x - x_t||^2 w.r.t. b and every nonzero
  against the FP64 CPU adjoint (the oracle solving A^T lam = dL/dx) and
  against central finite differences on 200 random nonzeros (h = 1e-6).
* 27-point lognormal 100^3, GMRES(30) + ParILUT (fill 2, drop 1e-4): the
  FP64 solve's true residual, the ParILUT keep counts, and the mixed
  (FP32 inner, BF16 factors) solve's error against the FP64 solution.
"""
import ctypes
import functools

import numpy as np
import pytest

import oracle
import xsolver_amd
from tests.helpers import check_fp64_parity, say, true_relres
from xsolver_amd import XBatch, XSolver, batch_concat, gen

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------------------
# 128^3 convection-diffusion, BiCGSTAB + ParILU(0)
# ---------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def convdiff128():
    A = gen.convdiff3d(128)
    xs, b = gen.manufactured(A, 1)
    return A, xs, b


@functools.lru_cache(maxsize=None)
def convdiff128_oracle(apply):
    A, _, b = convdiff128()
    kl, ku = apply if apply else (0, 0)
    return oracle.solve(A, b, "bicgstab", "ilu0", tol=1e-10, maxiter=2000, lsweeps=kl, usweeps=ku)


APPLIES = [(5, 2), (20, 20), None]
APPLY_IDS = ["sweeps5_2", "sweeps20_20", "exact"]


def _c2_solver(A, fmt, apply, **opts):
    S = XSolver()
    S.mat_setup(*A, advanced_pam=(0, 1 if fmt == "ell" else 0, 1))
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 2000)
    S.pc_settype("ilu0")
    if fmt == "csr_vec":
        S.set_option("spmv_kernel", 1)
    if apply:
        S.set_option("lsolve_sweeps", apply[0])
        S.set_option("usolve_sweeps", apply[1])
    for k, v in opts.items():
        S.set_option(k, v)
    return S


@pytest.mark.parametrize("apply", APPLIES, ids=APPLY_IDS)
@pytest.mark.parametrize("fmt", ["csr_vec", "ell"])
def test_convdiff128_bicgstab_parilu0_matches_oracle(dev, fmt, apply):
    A, xs, b = convdiff128()
    n = len(b)
    say(f"[c2] convdiff3d_128 n={n} nnz={len(A[2])} fmt={fmt} apply={apply}")
    xo, sto, ito, reso = convdiff128_oracle(apply)
    assert sto == 0
    S = _c2_solver(A, fmt, apply, parilu_async=0, parilu_sweeps=100, parilu_scale=0)
    if fmt == "csr_vec":
        assert S.time_ms("spmv:kernel") == 1        # the binned vector-CSR kernel, no ELL copy
    x, st = S.solve(b)
    assert st == xsolver_amd.XS_OK
    if apply == (5, 2) and fmt == "ell":
        # the GPU's ILU(0) equals the oracle's to rounding
        rp, ci, lu, dpos = S.factors(n, len(A[2]))
        lo, uo = oracle.ilu0(A)
        dlu = float(np.abs(lu - lo).max() / np.abs(lo).max())
        say(f"   factors vs oracle ILU(0): max rel {dlu:.2e}")
        assert np.array_equal(dpos, uo) and dlu < 1e-14
    kl, ku = apply if apply else (0, 0)
    check_fp64_parity(f"c2 {fmt} {APPLY_IDS[APPLIES.index(apply)]}", A, b, x, S.iterations,
                      S.resvec(), xo, ito, reso, 1e-10,
                      oracle_rerun=lambda bp: oracle.solve(A, bp, "bicgstab", "ilu0", tol=1e-10,
                                                           maxiter=2000, lsweeps=kl,
                                                           usweeps=ku)[::3])
    S.close()


def test_convdiff128_benchmark_setting(dev):
    """The benchmark's factorisation (20 asynchronous sweeps on the scaled
    matrix, the reference's default) with the 5/2 apply: FP64 residual below
    1e-10 and the iteration count of the oracle's exact ILU(0) within +-2;
    the asynchronous factors are not ILU(0) to rounding, so the solution
    agrees with the oracle's to what rtol 1e-10 leaves."""
    A, xs, b = convdiff128()
    xo, sto, ito, reso = convdiff128_oracle((5, 2))
    S = _c2_solver(A, "csr", (5, 2), parilu_async=1, parilu_sweeps=20)
    x, st = S.solve(b)
    rr = true_relres(A, b, x)
    dx = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    say(f"[c2 bench] gpu it={S.iterations} oracle it={ito} relres={rr:.3e} x-diff={dx:.2e}")
    assert st == 0 and abs(S.iterations - ito) <= 2 and rr < 1e-10 and dx < 1e-7
    S.close()


# ---------------------------------------------------------------------------
# 512 independent systems, CG + Jacobi
# ---------------------------------------------------------------------------
@functools.lru_cache(maxsize=None)
def batch512():
    sizes, systems = gen.batch_systems(512, 32, 160, seed=3)
    sys_ptr, rp, ci, v = batch_concat(systems)
    xs = np.random.default_rng(3).uniform(-1, 1, int(sys_ptr[-1]))
    b = gen.spmv((rp, ci, v), xs)
    return sys_ptr, (rp, ci, v), b


def test_batch512_matches_per_system_oracle(dev):
    sys_ptr, A, b = batch512()
    say(f"[c4] 512 systems, rows={sys_ptr[-1]} nnz={len(A[2])}")
    xo, ito, relo, bad = oracle.batch_cg(sys_ptr, A, b, tol=1e-8, maxiter=5000)
    assert bad == 0
    B = XBatch(sys_ptr, *A)
    x, st = B.solve(b, tol=1e-8, maxiter=5000, precision=0)
    it, rr = B.results()
    B.close()
    assert st == 0
    dits = np.abs(it.astype(np.int64) - ito)
    dx = np.array([np.linalg.norm(x[lo:hi] - xo[lo:hi]) / np.linalg.norm(xo[lo:hi])
                   for lo, hi in zip(sys_ptr[:-1], sys_ptr[1:])])
    say(f"   iterations {it.min()}..{it.max()} (oracle {ito.min()}..{ito.max()}), "
        f"max |d it| {dits.max()}, max true relres {rr.max():.3e}, max x-diff {dx.max():.2e}")
    assert dits.max() <= 2
    assert rr.max() < 1e-8
    assert dx.max() <= 1e-10


# ---------------------------------------------------------------------------
# 512^2 adjoint
# ---------------------------------------------------------------------------
def test_adjoint_grf512_cpu_adjoint_and_fd(dev):
    import scipy.sparse as sp
    import torch
    from xsolver_amd.autograd import make_solver, sparse_solve
    A = gen.vardiff2d(gen.grf_kappa(512, 0.1, 4))
    rp, ci, v = A
    n, nnz = len(rp) - 1, len(v)
    xt = np.random.default_rng([4, 1]).uniform(-1, 1, n)
    b = gen.spmv(A, np.random.default_rng([4, 2]).uniform(-1, 1, n))
    say(f"[c5] grf2d_512 n={n} nnz={nnz}")
    S = make_solver(rp, ci, v, "cg", "ilu0", tol=1e-13, maxiter=20000, parilu_async=0,
                    parilu_sweeps=200)
    vals = torch.from_numpy(v.copy()).to(dev).requires_grad_(True)
    bt = torch.from_numpy(b.copy()).to(dev).requires_grad_(True)
    x = sparse_solve(vals, bt, S)
    loss = ((x - torch.from_numpy(xt).to(dev)) ** 2).sum()
    loss.backward()
    gv = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    xg = x.detach().cpu().numpy()
    # FP64 CPU adjoint: A^T lam = 2 (x - x_t), dL/dA_ij = -lam_i x_j
    xo, st, _, _ = oracle.solve(A, b, "cg", "ilu0", tol=1e-14, maxiter=20000)
    At = sp.csr_matrix((v, ci, rp), shape=(n, n)).T.tocsr()
    At.sort_indices()
    Atc = (At.indptr.astype(np.int32), At.indices.astype(np.int32), At.data)
    lam_o, st2, _, _ = oracle.solve(Atc, 2 * (xo - xt), "cg", "ilu0", tol=1e-14, maxiter=20000)
    assert st == 0 and st2 == 0
    rows = np.repeat(np.arange(n), np.diff(rp))
    gv_o = -lam_o[rows] * xo[ci]
    rel_v = np.linalg.norm(gv - gv_o) / np.linalg.norm(gv_o)
    rel_b = np.linalg.norm(gb - lam_o) / np.linalg.norm(lam_o)
    say(f"   vs FP64 CPU adjoint: dL/dA rel {rel_v:.2e}, dL/db rel {rel_b:.2e}")
    assert rel_v < 1e-6 and rel_b < 1e-6
    # central finite differences, 200 random nonzeros, h = 1e-6.  Each
    # perturbed system A +- h e_i e_j^T is solved from the unperturbed
    # solution, so the solver only has to resolve the O(h) change, with the
    # unperturbed ILU(0) kept as preconditioner (it sets the convergence rate,
    # not the solution) and BiCGSTAB with the 5/2 Jacobi-sweep apply.  The
    # loss difference is formed without cancellation:
    #   L+ - L- = sum (x+ - x-) (x+ + x- - 2 x_t).
    idx = np.random.default_rng(4).choice(nnz, size=200, replace=False)
    h = 1e-6
    Sf = XSolver()
    Sf.mat_setup(rp, ci, v)
    Sf.ksp_settype("bicgstab")
    Sf.ksp_setoption(1e-11, 20000)
    Sf.pc_settype("ilu0")
    Sf.set_option("parilu_async", 0)
    Sf.set_option("parilu_sweeps", 200)
    Sf.set_option("lsolve_sweeps", 5)
    Sf.set_option("usolve_sweeps", 2)
    Sf.pc_setup()
    Sf.set_option("pc_reuse", 1)
    lib = xsolver_amd.lib()

    def x_at(vv):
        assert lib.xs_mat_update_values(Sf.h, vv.ctypes.data_as(ctypes.c_void_p), 0) == 0
        xx, stt = Sf.solve(b, x0=xg)
        assert stt == 0
        return xx

    fd = np.zeros(len(idx))
    its = []
    for q, k in enumerate(idx):
        vp = v.copy(); vp[k] += h
        vm = v.copy(); vm[k] -= h
        xp = x_at(vp); its.append(Sf.iterations)
        xm = x_at(vm); its.append(Sf.iterations)
        fd[q] = float(np.dot(xp - xm, xp + xm - 2 * xt)) / (2 * h)
        if q % 50 == 49:
            say(f"   FD {q + 1}/200 done (perturbed solves {min(its)}..{max(its)} iterations)")
    Sf.close()
    rel_fd = np.linalg.norm(gv[idx] - fd) / np.linalg.norm(fd)
    say(f"   vs central FD (200 entries, h=1e-6): rel {rel_fd:.2e}")
    # FD accuracy: the truncation error is O(h^2) relative; each perturbed
    # solution carries O(eps |x|) rounding per entry from its updates, which
    # the loss difference sums over n = 262,144 entries with random signs:
    # sqrt(n) eps / (2 h) ~ 1e-7 absolute per derivative -- the bound below
    # leaves a decade above that for the smallest of the 200 derivatives
    assert rel_fd < 1e-5
    S.close()


# ---------------------------------------------------------------------------
# 27-point lognormal 100^3, GMRES(30) + ParILUT
# ---------------------------------------------------------------------------
def test_lognormal27_100_gmres_parilut_fp64_and_mixed(dev):
    A = gen.lognormal27(100, seed=2)
    rp, ci, v = A
    n, nnz = len(rp) - 1, len(v)
    xs, b = gen.manufactured(A, 2)
    say(f"[c3] lognormal27_100 n={n} nnz={nnz}")
    out = {}
    for name, prec, fs in (("fp64", 0, 0), ("mixed_bf16", 1, 1)):
        S = XSolver()
        S.mat_setup(*A)
        S.ksp_settype("gmres")
        S.ksp_setoption(1e-10, 3000, 30)
        S.pc_settype("ilut")
        S.pc_setilut(1e-4, 2)
        S.set_option("precision", prec)
        S.set_option("factor_storage", fs)
        S.set_option("lsolve_sweeps", 10)
        S.set_option("usolve_sweeps", 15)
        x, st = S.solve(b)
        assert st == 0
        out[name] = (x, S.iterations)
        if prec == 0:
            # the final keep: 2 x each factor's count in A with its diagonal
            frp, fci, flu, fdp = S.factors(n, 3 * nnz + 2 * n)
            rows = np.repeat(np.arange(n), np.diff(rp))
            nl0, nu0 = int(np.sum(ci < rows)), int(np.sum(ci > rows))
            nlf = int(np.sum(fdp - frp[:-1]))
            nuf = int(frp[-1]) - nlf
            say(f"   ParILUT factor: strict L {nlf} (keep {2 * (nl0 + n) - n}), "
                f"U {nuf} (keep {2 * (nu0 + n)})")
            assert nlf + n <= 2 * (nl0 + n) + n // 100 and nuf <= 2 * (nu0 + n) + n // 100
        S.close()
    x64, it64 = out["fp64"]
    xm, itm = out["mixed_bf16"]
    rr64, rrm = true_relres(A, b, x64), true_relres(A, b, xm)
    err = np.linalg.norm(xm - x64) / np.linalg.norm(x64)
    say(f"   fp64 it={it64} relres={rr64:.3e}; mixed bf16 it={itm} relres={rrm:.3e} "
        f"err vs fp64={err:.2e}")
    assert rr64 < 1e-10 and rrm < 1e-10
    assert err < 1e-8
