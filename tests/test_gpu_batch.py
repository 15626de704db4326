"""Batched persistent CG + Jacobi through the C-ABI (xs_batch_*), per system
against the oracle's CG on the same seeded systems."""
import numpy as np
import pytest

import oracle
from tests.helpers import X_TOL, say
from xsolver_amd import XBatch, batch_concat, gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32inner"])
def test_batch_matches_per_system_oracle(dev, precision):
    sizes, systems = gen.batch_systems(count=40, smin=8, smax=48, seed=3)
    sys_ptr, rp, ci, v = batch_concat(systems)
    rng = np.random.default_rng(7)
    xs = rng.uniform(-1, 1, sys_ptr[-1])
    b = gen.spmv((rp, ci, v), xs)
    say(f"[batch] {len(systems)} systems, n={sys_ptr[-1]}, precision={precision}")
    B = XBatch(sys_ptr, rp, ci, v)
    x, st = B.solve(b, tol=1e-8, maxiter=3000, precision=precision)
    it, rr = B.results()
    B.close()
    assert st == 0
    for s, A in enumerate(systems):
        lo, hi = sys_ptr[s], sys_ptr[s + 1]
        xo, sto, ito, _ = oracle.solve(A, b[lo:hi], "cg", "jacobi", tol=1e-8, maxiter=3000)
        true = np.linalg.norm(b[lo:hi] - gen.spmv(A, x[lo:hi])) / np.linalg.norm(b[lo:hi])
        assert sto == 0
        assert true < 1.05e-8 and rr[s] == pytest.approx(true, rel=1e-6, abs=1e-14)
        if precision == 0:
            assert abs(int(it[s]) - ito) <= 2, (s, it[s], ito)
            assert np.linalg.norm(x[lo:hi] - xo) / np.linalg.norm(xo) <= X_TOL


def test_batch_uneven_sizes_and_zero_rhs(dev):
    sizes, systems = gen.batch_systems(count=12, smin=4, smax=70, seed=5)
    sys_ptr, rp, ci, v = batch_concat(systems)
    b = np.ones(sys_ptr[-1])
    b[sys_ptr[3]:sys_ptr[4]] = 0.0            # one system with a zero right-hand side
    B = XBatch(sys_ptr, rp, ci, v)
    x, st = B.solve(b, tol=1e-10, maxiter=5000)
    it, rr = B.results()
    B.close()
    assert st == 0 and it[3] == 0 and np.all(x[sys_ptr[3]:sys_ptr[4]] == 0)
    assert len(set(it.tolist())) > 3          # systems stop at their own counts


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "fp32inner"])
def test_batch_nonzero_x0_uses_initial_residual_base(dev, precision):
    """Both precisions stop on ||b - A x|| < tol ||b - A x0|| and report relres
    on that base (the reference's convention, hip/Solver/xsolver_cg.cpp:16-74),
    as the oracle's CG started from the same x0 does."""
    sizes, systems = gen.batch_systems(count=10, smin=8, smax=40, seed=8)
    sys_ptr, rp, ci, v = batch_concat(systems)
    rng = np.random.default_rng(9)
    b = gen.spmv((rp, ci, v), rng.uniform(-1, 1, sys_ptr[-1]))
    x0 = rng.uniform(-1, 1, sys_ptr[-1])
    B = XBatch(sys_ptr, rp, ci, v)
    x, st = B.solve(b, tol=1e-9, maxiter=3000, precision=precision, x0=x0)
    it, rr = B.results()
    B.close()
    assert st == 0
    for s, A in enumerate(systems):
        lo, hi = sys_ptr[s], sys_ptr[s + 1]
        r0 = np.linalg.norm(b[lo:hi] - gen.spmv(A, x0[lo:hi]))
        true = np.linalg.norm(b[lo:hi] - gen.spmv(A, x[lo:hi])) / r0
        assert true < 1e-9 and rr[s] == pytest.approx(true, rel=1e-6, abs=1e-15)
        if precision == 0:
            xo, sto, ito, _ = oracle.solve(A, b[lo:hi], "cg", "jacobi", tol=1e-9, maxiter=3000,
                                           x0=x0[lo:hi])
            assert sto == 0 and abs(int(it[s]) - ito) <= 2
            assert np.linalg.norm(x[lo:hi] - xo) / np.linalg.norm(xo) <= X_TOL
