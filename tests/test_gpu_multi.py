"""Block of right-hand sides through the C-ABI (xs_spmm, xs_pc_apply_multi,
xs_solve_multi[_device]) against the oracle column by column: SpMM and the
k-column triangular solves to rounding, and the block BiCGSTAB/CG solve per
column to the single right-hand side bar (iterations within +-2 of the
oracle, FP64 relative residual below rtol, solution close to the oracle's)."""
import numpy as np
import pytest

import oracle
import xsolver_amd
from tests.helpers import check_fp64_parity, random_sparse, rel_err, say, split_factors, true_relres
from xsolver_amd import XSolver, gen

pytestmark = pytest.mark.gpu


def torch_stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def dev_tensor(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


SPMM_MATS = {
    "convdiff3d_20": lambda: gen.convdiff3d(20),
    "lognormal27_10": lambda: gen.lognormal27(10, seed=2),
    "longrows_1500": lambda: random_sparse(1500, 4, 22, long_rows=[(3, 1400), (900, 700)]),
}


@pytest.mark.parametrize("k", [1, 2, 5, 6, 12, 32, 64])
@pytest.mark.parametrize("name", list(SPMM_MATS))
def test_spmm_matches_oracle(dev, name, k):
    import torch
    A = SPMM_MATS[name]()
    n = len(A[0]) - 1
    say(f"[spmm] {name} n={n} k={k}")
    S = XSolver()
    S.mat_setup(*A)
    X = np.random.default_rng(k).uniform(-1, 1, (n, k))
    Yd = torch.empty((n, k), dtype=torch.float64, device=dev)
    S.spmm(dev_tensor(X, dev), Yd, torch_stream())
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy()
    scale = np.abs(A[2]).max() * np.diff(A[0]).max()
    for j in range(k):
        assert np.abs(Y[:, j] - oracle.spmv(A, X[:, j])).max() <= 1e-14 * scale
    S.close()


@pytest.mark.parametrize("k", [2, 6, 12, 32, 64])
def test_spmm_ell_matches_csr_bitwise(dev, k):
    """On a matrix with an automatic ELL copy, even k runs the ELL SpMM (two
    or four columns per thread, 16-bit offsets); every column sums in the
    row's entry order, so it equals the CSR SpMM bit for bit.  A block that
    is not 16-byte aligned takes the CSR kernel."""
    import torch
    A = gen.convdiff3d(20)
    n = len(A[0]) - 1
    X = np.random.default_rng(10 + k).uniform(-1, 1, (n, k))
    outs = []
    for kernel in (None, 1):
        S = XSolver()
        S.mat_setup(*A)
        if kernel is not None:
            S.set_option("spmv_kernel", kernel)     # vector CSR, no ELL copy
        assert S.time_ms("spmv:kernel") == (2 if kernel is None else 1)
        Yd = torch.empty((n, k), dtype=torch.float64, device=dev)
        S.spmm(dev_tensor(X, dev), Yd, torch_stream())
        torch.cuda.synchronize()
        outs.append(Yd.cpu().numpy())
        if kernel is None:
            # blocks shifted by one double: 8-byte aligned only
            xb = torch.empty(n * k + 1, dtype=torch.float64, device=dev)
            yb = torch.empty(n * k + 1, dtype=torch.float64, device=dev)
            xb[1:].copy_(dev_tensor(X.reshape(-1), dev))
            S.spmm(xb[1:].view(n, k), yb[1:].view(n, k), torch_stream())
            torch.cuda.synchronize()
            outs.append(yb[1:].cpu().numpy().reshape(n, k))
        S.close()
    say(f"[spmm-ell] k={k} max |ell - csr| = {np.abs(outs[0] - outs[2]).max():.3e}")
    assert np.array_equal(outs[0], outs[2])
    assert np.array_equal(outs[1], outs[2])


TRSV_MATS = {
    "poisson2d_128": lambda: gen.poisson2d(128),     # only r-1 inside a wave: scan path
    "convdiff3d_12": lambda: gen.convdiff3d(12),     # r-1, r-12 inside a wave: loop path
    "lognormal27_10": lambda: gen.lognormal27(10, seed=2),
    "random_3000": lambda: random_sparse(3000, 7, 23),
}


@pytest.mark.parametrize("k", [1, 3, 16, 21, 32])
@pytest.mark.parametrize("name", list(TRSV_MATS))
def test_pc_apply_multi_matches_oracle(dev, name, k):
    import torch
    A = TRSV_MATS[name]()
    n = len(A[0]) - 1
    say(f"[mtrsv] {name} n={n} k={k}")
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.pc_settype("ilu0")
    S.pc_setup()
    rp, ci, lu, dpos = S.factors(n, len(A[2]))
    L, U, dinv = split_factors(rp, ci, lu, dpos)
    B = np.random.default_rng(5 + k).uniform(-1, 1, (n, k))
    Bd = dev_tensor(B, dev)
    Yd = torch.empty((n, k), dtype=torch.float64, device=dev)
    for rep in range(2):  # the tickets reset themselves between calls
        S.pc_apply_multi(Bd, Yd, torch_stream())
        torch.cuda.synchronize()
        Y = Yd.cpu().numpy()
        for j in range(k):
            ref = oracle.upper_solve(U, dinv, oracle.lower_solve(L, B[:, j]))
            assert rel_err(Y[:, j], ref) < 1e-12, (rep, j)
    # one column against the single right-hand side kernel
    yd = torch.empty(n, dtype=torch.float64, device=dev)
    S.pc_apply(dev_tensor(B[:, 0], dev), yd, torch_stream())
    torch.cuda.synchronize()
    assert rel_err(Y[:, 0], yd.cpu().numpy()) < 1e-13
    S.close()


def test_pc_apply_multi_parilut_factors(dev):
    import torch
    A = gen.lognormal27(10, seed=2)
    n = len(A[0]) - 1
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("gmres")
    S.pc_settype("ilut")
    S.pc_setilut(1e-4, 2)
    S.pc_setup()
    rp, ci, lu, dpos = S.factors(n, 4 * len(A[2]))
    L, U, dinv = split_factors(rp, ci, lu, dpos)
    B = np.random.default_rng(11).uniform(-1, 1, (n, 8))
    Yd = torch.empty((n, 8), dtype=torch.float64, device=dev)
    S.pc_apply_multi(dev_tensor(B, dev), Yd, torch_stream())
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy()
    for j in range(8):
        ref = oracle.upper_solve(U, dinv, oracle.lower_solve(L, B[:, j]))
        assert rel_err(Y[:, j], ref) < 1e-12
    S.close()


@pytest.mark.parametrize("kl,ku", [(0, 1), (2, 2), (5, 2), (3, 6)])
@pytest.mark.parametrize("k", [1, 6, 7, 12, 32])
def test_pc_apply_multi_sweeps(dev, k, kl, ku):
    """k-column Jacobi-sweep apply (ELL copies, first two sweeps in one pass)
    against the oracle's sweeps on the GPU's own factors, column by column,
    and bit-identical to the single right-hand side apply."""
    import torch
    A = gen.convdiff3d(16)
    n = len(A[0]) - 1
    say(f"[msweeps] k={k} kl={kl} ku={ku}")
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.pc_settype("ilu0")
    S.set_option("lsolve_sweeps", kl)
    S.set_option("usolve_sweeps", ku)
    S.pc_setup()
    rp, ci, lu, dpos = S.factors(n, len(A[2]))
    B = np.random.default_rng(13 + k).uniform(-1, 1, (n, k))
    Yd = torch.empty((n, k), dtype=torch.float64, device=dev)
    S.pc_apply_multi(dev_tensor(B, dev), Yd, torch_stream())
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy()
    for j in range(k):
        ref = oracle.ilu_sweeps(rp, ci, lu, dpos, kl, ku, B[:, j])
        assert rel_err(Y[:, j], ref) < 1e-13, j
    yd = torch.empty(n, dtype=torch.float64, device=dev)
    S.pc_apply(dev_tensor(B[:, k - 1], dev), yd, torch_stream())
    torch.cuda.synchronize()
    assert np.array_equal(Y[:, k - 1], yd.cpu().numpy())
    S.close()


def test_block_solve_with_sweep_apply(dev):
    """Block BiCGSTAB with the 5/2 sweep apply: every column converges with
    the iteration count of the oracle running the same apply (+-2)."""
    A = gen.convdiff3d(24)
    k = 6
    Xs, B = rhs_block(A, k, 21)
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 1000)
    S.pc_settype("ilu0")
    S.set_option("parilu_async", 0)
    S.set_option("parilu_sweeps", 60)
    S.set_option("lsolve_sweeps", 5)
    S.set_option("usolve_sweeps", 2)
    X, st = S.solve_multi(B)
    its, r0, rl = S.multi_results()
    assert st == xsolver_amd.XS_OK
    for j in range(k):
        xo, sto, ito, _ = oracle.solve(A, B[:, j], "bicgstab", "ilu0", tol=1e-10, maxiter=1000,
                                       lsweeps=5, usweeps=2)
        rr = true_relres(A, B[:, j], X[:, j])
        say(f"   col {j}: oracle it={ito} gpu it={its[j]} relres={rr:.2e}")
        assert sto == 0 and abs(int(its[j]) - ito) <= 2 and rr < 1e-10
    S.close()


BLOCK_CASES = [
    # name, builder, ksp, pc, tol, maxiter, k
    ("convdiff3d_24_bicgstab_ilu0", lambda: gen.convdiff3d(24), "bicgstab", "ilu0", 1e-10, 1000, 8),
    ("convdiff3d_24_bicgstab_jacobi", lambda: gen.convdiff3d(24), "bicgstab", "jacobi", 1e-10, 2000, 5),
    ("poisson2d_96_cg_jacobi", lambda: gen.poisson2d(96), "cg", "jacobi", 1e-8, 5000, 4),
    ("grf2d_48_cg_ilu0", lambda: gen.vardiff2d(gen.grf_kappa(48, 0.1, 4)), "cg", "ilu0", 1e-10, 2000, 3),
    ("poisson2d_48_cg_none", lambda: gen.poisson2d(48), "cg", "none", 1e-9, 2000, 2),
]


def rhs_block(A, k, seed0):
    xs = []
    bs = []
    for j in range(k):
        x, b = gen.manufactured(A, seed0 + j)
        xs.append(x)
        bs.append(b)
    return np.stack(xs, 1), np.stack(bs, 1)


@pytest.mark.parametrize("case", BLOCK_CASES, ids=[c[0] for c in BLOCK_CASES])
def test_block_solve_matches_oracle_per_column(dev, case):
    name, build, ksp, pc, tol, maxiter, k = case
    A = build()
    Xs, B = rhs_block(A, k, 1)
    say(f"[block] {name} n={len(Xs)} k={k}")
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype(ksp)
    S.ksp_setoption(tol, maxiter)
    S.pc_settype(pc)
    if pc == "ilu0":
        S.set_option("parilu_async", 0)
        S.set_option("parilu_sweeps", 200)
        S.set_option("parilu_scale", 0)
    X, st = S.solve_multi(B)
    its, r0, rl = S.multi_results()
    assert st == xsolver_amd.XS_OK and len(its) == k
    for j in range(k):
        xo, sto, ito, reso = oracle.solve(A, B[:, j], ksp, pc, tol=tol, maxiter=maxiter)
        assert sto == 0
        assert r0[j] == pytest.approx(np.linalg.norm(B[:, j]), rel=1e-12)
        res = S.multi_resvec(j)
        assert res[-1] == pytest.approx(rl[j])
        check_fp64_parity(f"col {j}", A, B[:, j], X[:, j], int(its[j]), res, xo, ito, reso, tol,
                          oracle_rerun=lambda bp: oracle.solve(A, bp, ksp, pc, tol=tol,
                                                               maxiter=maxiter)[::3])
    S.close()


def test_block_device_layouts_and_single_column(dev):
    """Column-major and interleaved device blocks give the same answer, and a
    1-column block matches the single right-hand side solve's iterations."""
    import torch
    A = gen.convdiff3d(20)
    Xs, B = rhs_block(A, 6, 3)
    n = B.shape[0]
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 1000)
    S.pc_settype("ilu0")
    Bi = dev_tensor(B, dev)
    Xi = torch.zeros_like(Bi)
    torch.cuda.synchronize()
    assert S.solve_multi_device(Bi, Xi, xsolver_amd.XS_RHS_INTERLEAVED) == 0
    it_i = S.multi_results()[0].copy()
    Bc = dev_tensor(B.T, dev)
    Xc = torch.zeros_like(Bc)
    torch.cuda.synchronize()
    assert S.solve_multi_device(Bc, Xc, xsolver_amd.XS_RHS_COLMAJOR) == 0
    it_c = S.multi_results()[0].copy()
    assert np.array_equal(it_i, it_c)
    assert np.array_equal(Xi.cpu().numpy(), Xc.cpu().numpy().T)   # same kernels, same order
    x1, st = S.solve(B[:, 2])
    it1 = S.iterations
    X1, st1 = S.solve_multi(B[:, 2:3])
    assert st == 0 and st1 == 0
    assert abs(int(S.multi_results()[0][0]) - it1) <= 1
    assert rel_err(X1[:, 0], x1) < 1e-8
    S.close()


def test_block_per_column_masks(dev):
    """A zero column and an already-solved column finish at once; the others
    iterate to their own counts."""
    A = gen.convdiff3d(16)
    Xs, B = rhs_block(A, 4, 5)
    B[:, 1] = 0.0
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 1000)
    S.pc_settype("jacobi")
    X0 = np.zeros_like(B)
    X0[:, 3] = Xs[:, 3]                       # exact start: r0 = 0 to rounding
    S.ksp_set_absolute_tol(1e-12 * np.linalg.norm(B[:, 3]))
    X, st = S.solve_multi(B, x0=X0)
    its, r0, rl = S.multi_results()
    say(f"[masks] its={its}")
    assert st == 0
    assert its[1] == 0 and its[3] == 0 and its[0] > 0 and its[2] > 0
    assert np.array_equal(X[:, 1], np.zeros(B.shape[0]))
    for j in (0, 2):
        assert true_relres(A, B[:, j], X[:, j]) < 1e-10
    S.close()


def test_block_errors(dev):
    A = gen.poisson2d(16)
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("gmres")
    S.ksp_setoption(1e-8, 100)
    S.pc_settype("jacobi")
    B = np.ones((256, 3))
    with pytest.raises(xsolver_amd.XSolverError) as e:
        S.solve_multi(B)
    assert e.value.status == xsolver_amd.XS_ERR_UNSUPPORTED
    S.ksp_settype("cg")
    with pytest.raises(xsolver_amd.XSolverError) as e:
        S.solve_multi(np.ones((256, 65)))
    assert e.value.status == xsolver_amd.XS_ERR_ARG
    S.close()
