"""Adjoint through the C-ABI: A^T lam = dL/dx with transposed SpMV and
transposed ILU solves; gradients w.r.t. b and the nonzeros checked against
central finite differences and against an FP64 CPU adjoint (the oracle's
solve of A^T lam = g)."""
import numpy as np
import pytest

import oracle
from tests.helpers import say
from xsolver_amd import XSolver, gen

pytestmark = pytest.mark.gpu


def system(N=48):
    A = gen.vardiff2d(gen.grf_kappa(N, 0.1, 4))
    n = len(A[0]) - 1
    xt = np.random.default_rng([4, 1]).uniform(-1, 1, n)
    b = gen.spmv(A, np.random.default_rng([4, 2]).uniform(-1, 1, n))
    return A, b, xt


def test_transposed_kernels_match_oracle(dev):
    import torch
    A = gen.convdiff3d(14)           # nonsymmetric: A^T differs from A
    n = len(A[0]) - 1
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab"); S.ksp_setoption(1e-12, 500); S.pc_settype("ilu0")
    S.set_option("parilu_async", 0); S.set_option("parilu_sweeps", 60)
    S.pc_setup()
    x = np.random.default_rng(1).uniform(-1, 1, n)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    S.spmv_t(xd, yd, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.abs(yd.cpu().numpy() - oracle.spmv_t(A, x)).max() < 1e-13 * np.abs(x).max() * 8
    # M^{-T} applied to x: check (LU)^T y = x with the GPU's own factors
    S.pc_apply_t(xd, yd, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    rp, ci, lu, dpos = S.factors(n, len(A[2]))
    import scipy.sparse as sp
    Lm = sp.csr_matrix((np.where(np.arange(len(lu)) < np.repeat(dpos, np.diff(rp)), lu, 0.0), ci, rp), shape=(n, n)) + sp.eye(n)
    Um = sp.csr_matrix((np.where(np.arange(len(lu)) >= np.repeat(dpos, np.diff(rp)), lu, 0.0), ci, rp), shape=(n, n))
    y = yd.cpu().numpy()
    assert np.abs((Lm @ Um).T @ y - x).max() < 1e-10
    # adjoint solve A^T lam = g
    g = np.random.default_rng(2).uniform(-1, 1, n)
    gd = torch.from_numpy(g).to(dev)
    lam = torch.zeros_like(gd)
    st = S.solve_transpose_device(gd, lam)
    assert st == 0
    import scipy.sparse as sp2
    At = sp2.csr_matrix((A[2], A[1], A[0]), shape=(n, n)).T.tocsr()
    lam_np = lam.cpu().numpy()
    assert np.linalg.norm(At @ lam_np - g) / np.linalg.norm(g) < 1e-11
    S.close()


def test_adjoint_gradients_fd_and_cpu_adjoint(dev):
    import torch
    from xsolver_amd.autograd import make_solver, sparse_solve
    A, b, xt = system(48)
    rp, ci, v = A
    n, nnz = len(rp) - 1, len(v)
    say(f"[adjoint] grf2d_48 n={n} nnz={nnz}")
    S = make_solver(rp, ci, v, "cg", "ilu0", tol=1e-13, maxiter=5000)
    vals = torch.from_numpy(v.copy()).to(dev).requires_grad_(True)
    bt = torch.from_numpy(b.copy()).to(dev).requires_grad_(True)
    xtt = torch.from_numpy(xt).to(dev)
    x = sparse_solve(vals, bt, S)
    loss = ((x - xtt) ** 2).sum()
    loss.backward()
    gv = vals.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    # FP64 CPU adjoint (oracle): A^T lam = 2 (x - x_t), grad_A = -lam_i x_j
    xo, st, _, _ = oracle.solve(A, b, "cg", "ilu0", tol=1e-14, maxiter=5000)
    import scipy.sparse as sp
    At = sp.csr_matrix((v, ci, rp), shape=(n, n)).T.tocsr()
    At.sort_indices()
    Atc = (At.indptr.astype(np.int32), At.indices.astype(np.int32), At.data)
    lam_o, st2, _, _ = oracle.solve(Atc, 2 * (xo - xt), "cg", "ilu0", tol=1e-14, maxiter=5000)
    rows = np.repeat(np.arange(n), np.diff(rp))
    gv_o = -lam_o[rows] * xo[ci]
    rel_v = np.linalg.norm(gv - gv_o) / np.linalg.norm(gv_o)
    rel_b = np.linalg.norm(gb - lam_o) / np.linalg.norm(lam_o)
    say(f"   vs FP64 CPU adjoint: dL/dA rel {rel_v:.2e}, dL/db rel {rel_b:.2e}")
    assert rel_v < 1e-6 and rel_b < 1e-6
    # central finite differences on 200 random nonzeros (h = 1e-6)
    rng = np.random.default_rng(4)
    idx = rng.choice(nnz, size=200, replace=False)
    h = 1e-6
    fd = np.zeros(len(idx))
    Sf = XSolver()
    Sf.mat_setup(rp, ci, v); Sf.ksp_settype("cg"); Sf.ksp_setoption(1e-14, 5000); Sf.pc_settype("ilu0")

    def L(vv):
        Sf.mat_setup(rp, ci, vv)
        xx, stt = Sf.solve(b)
        assert stt == 0
        return float(((xx - xt) ** 2).sum())

    for q, k in enumerate(idx):
        vp = v.copy(); vp[k] += h
        vm = v.copy(); vm[k] -= h
        fd[q] = (L(vp) - L(vm)) / (2 * h)
    Sf.close()
    rel_fd = np.linalg.norm(gv[idx] - fd) / np.linalg.norm(fd)
    say(f"   vs central FD (200 entries): rel {rel_fd:.2e}")
    assert rel_fd < 1e-4
    S.close()
