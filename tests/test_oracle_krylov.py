"""The oracle's restatement of the rest of the intel Krylov family
(intel/mpisolver-C-v1.cpp: cpu_cg_solve, cpu_icg_solve, cpu_icgm_solve,
cpu_ibicgstab_solve, cpu_gcr_solve, cpu_igcr_solve) pinned by what the
mathematics fixes:

* CG on M^-1 A with a constant-diagonal Jacobi is preconditioned CG:
  residual norms equal the PCG's scaled by the constant.
* ICG / ICGM are CG / PCG with the inner products regrouped: the same
  iterates to rounding (ICGM judges convergence one residual late, so it
  performs one iteration more).  Their recurrence for r.r keeps an absolute
  error of about eps ||r0||^2, so it stops tracking the residual near
  ||r|| / ||r0|| ~ sqrt(eps) and then diverges: as in the reference, they are
  compared at rtol 1e-6, and a 1e-10 request is shown to break down.
* Improved BiCGSTAB is BiCGSTAB: the same residual history to rounding.
* Left-preconditioned GCR without restart is the minimal-residual method on
  M^-1 A, so its residual norms equal left GMRES's (no restart); CGS and MGS
  agree; with restart it still converges.
* Each reaches the true tolerance on its monitored residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from xsolver_amd import gen


def dense(A):
    rp, ci, v = A
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()


def test_cg_left_with_constant_jacobi_is_pcg():
    A = gen.poisson2d(20)
    xs, b = gen.manufactured(A, 0)
    x1, s1, i1, r1 = oracle.solve(A, b, "cg_left", "jacobi", tol=1e-10, maxiter=500)
    x2, s2, i2, r2 = oracle.solve(A, b, "cgm", "jacobi", tol=1e-10, maxiter=500)
    assert s1 == 0 and s2 == 0
    # M = 4 I: ||M^-1 r|| = ||r|| / 4 at every step
    m = min(len(r1), len(r2))
    assert np.allclose(r1[:m] * 4.0, r2[:m], rtol=1e-7)
    assert abs(i1 - i2) <= 1
    assert np.linalg.norm(x1 - x2) / np.linalg.norm(x2) < 1e-8


@pytest.mark.parametrize("A,pc", [(gen.vardiff2d(gen.grf_kappa(24, 0.2, 4)), "jacobi"),
                                  (gen.poisson2d(24), "ic0")])
def test_icg_is_cg_left_and_icgm_is_pcg(A, pc):
    xs, b = gen.manufactured(A, 3)
    tol = 1e-6
    xa, sa, ia, ra = oracle.solve(A, b, "cg_left", pc, tol=tol, maxiter=2000)
    xb, sb, ib, rb = oracle.solve(A, b, "icg", pc, tol=tol, maxiter=2000)
    # the first iterations agree to rounding (a wrong term or sign would
    # show there); later CG on the nonsymmetric M^-1 A amplifies rounding
    assert sa == 0 and sb == 0 and abs(ia - ib) <= 2
    assert np.allclose(ra[:10], rb[:10], rtol=1e-9)
    xc, sc, ic, rc = oracle.solve(A, b, "cgm", pc, tol=tol, maxiter=2000)
    xd, sd, id_, rd = oracle.solve(A, b, "icgm", pc, tol=tol, maxiter=2000)
    assert sc == 0 and sd == 0
    assert id_ == ic + 1                      # the test sees ||r|| one step late
    assert np.allclose(rd[1:11], rc[:10], rtol=1e-9)
    Ad = dense(A)
    assert np.linalg.norm(b - Ad @ xd) / np.linalg.norm(b) < tol
    Dinv = None
    if pc == "jacobi":   # ICG's monitored residual is M^-1 (b - A x)
        Dinv = 1.0 / np.diag(Ad)
        rb_true = np.linalg.norm(Dinv * (b - Ad @ xb))
        assert rb_true / rb[0] < 1.01 * tol
    # far below sqrt(eps) the recurrence no longer tracks r.r
    _, sbad, _, _ = oracle.solve(A, b, "icg", pc, tol=1e-13, maxiter=3000)
    assert sbad != 0


@pytest.mark.parametrize("pc", ["jacobi", "ilu0"])
def test_ibicgstab_is_bicgstab(pc):
    A = gen.convdiff3d(10)
    xs, b = gen.manufactured(A, 1)
    x1, s1, i1, r1 = oracle.solve(A, b, "bicgstab", pc, tol=1e-10, maxiter=500)
    x2, s2, i2, r2 = oracle.solve(A, b, "ibicgstab", pc, tol=1e-10, maxiter=500)
    assert s1 == 0 and s2 == 0 and abs(i1 - i2) <= 1
    m = min(len(r1), len(r2), 12)
    assert np.allclose(r1[:m], r2[:m], rtol=1e-5)
    assert np.linalg.norm(x1 - x2) / np.linalg.norm(x1) < 1e-8
    assert np.linalg.norm(b - dense(A) @ x2) / np.linalg.norm(b) < 1e-10


@pytest.mark.parametrize("pc", ["jacobi", "ilu0"])
def test_gcr_left_is_left_minimal_residual(pc):
    A = gen.convdiff3d(8)
    xs, b = gen.manufactured(A, 2)
    x1, s1, i1, r1 = oracle.solve(A, b, "gcr_left", pc, tol=1e-10, maxiter=400, restart=400)
    x2, s2, i2, r2 = oracle.solve(A, b, "igcr_left", pc, tol=1e-10, maxiter=400, restart=400)
    x3, s3, i3, r3 = oracle.solve(A, b, "gmres", pc, tol=1e-10, maxiter=400, restart=400)
    assert s1 == 0 and s2 == 0 and s3 == 0
    m = min(len(r1), len(r2), len(r3), 25)
    assert np.allclose(r1[:m], r3[:m], rtol=1e-6)
    assert np.allclose(r2[:m], r3[:m], rtol=1e-6)
    assert abs(i1 - i3) <= 1 and abs(i2 - i3) <= 1
    # restarted: still converges on its monitored residual
    x4, s4, i4, r4 = oracle.solve(A, b, "igcr_left", pc, tol=1e-10, maxiter=2000, restart=6)
    assert s4 == 0 and r4[-1] / r4[0] < 1e-10 and i4 > i2
