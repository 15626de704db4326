"""ParIC(0) on the GPU through the C-ABI against the oracle's IC(0) (the
reference's hip/Precond/ic0.cpp restated, pinned by golden vectors).

* Factors: deterministic ParIC(0) sweeps to the fixed point, written as
  combined L~ | D L~^T factors on A's pattern, equal the oracle's IC(0) in
  that form to rounding, with and without the symmetric scaling.
* Solves: CG and BiCGSTAB with IC(0) (exact sync-free triangular solves)
  against the oracle's solve with seqic0 + LUsolve
  (hip/HostMatrix/HostMatrixCSR.cpp:269-326, :385-409), to the plain-path
  acceptance; the 512^2 lognormal-field system at full size; the
  reference's asynchronous default; the Jacobi-sweep apply; the FP32-inner
  solve.
* A structurally nonsymmetric pattern is refused (XS_ERR_UNSUPPORTED).
"""
import numpy as np
import pytest

import oracle
import xsolver_amd
from tests.helpers import check_fp64_parity, say, true_relres
from xsolver_amd import XSolver, XSolverError, gen

pytestmark = pytest.mark.gpu

MATS = [
    ("poisson2d_32", lambda: gen.poisson2d(32)),
    ("grf2d_48", lambda: gen.vardiff2d(gen.grf_kappa(48, 0.1, 4))),
    ("lognormal27_12", lambda: gen.lognormal27(12, seed=2)),
]
DET = {"parilu_async": 0, "parilu_sweeps": 600}


def _solver(A, ksp, tol, maxiter, **opts):
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype(ksp)
    S.ksp_setoption(tol, maxiter)
    S.pc_settype("ic0")
    for k, v in opts.items():
        S.set_option(k, v)
    return S


@pytest.mark.parametrize("scale", [0, 1], ids=["unscaled", "scaled"])
@pytest.mark.parametrize("name,build", MATS, ids=[m[0] for m in MATS])
def test_paric0_factors_match_oracle_ic0(dev, name, build, scale):
    A = build()
    n, nnz = len(A[0]) - 1, len(A[2])
    S = _solver(A, "cg", 1e-10, 10, parilu_scale=scale, **DET)
    S.pc_setup()
    rp, ci, lu, dpos = S.factors(n, nnz)
    S.close()
    lo = oracle.ic_combined(A, *oracle.ic0(A))
    d = float(np.abs(lu - lo).max() / np.abs(lo).max())
    say(f"[ic0] {name} scale={scale}: combined factors vs oracle IC(0) max rel {d:.2e}")
    assert np.array_equal(rp, A[0]) and np.array_equal(ci, A[1])
    assert d < 1e-13


CASES = [
    ("poisson2d_64_cg", lambda: gen.poisson2d(64), "cg", 1e-10, 1000, 0),
    ("grf2d_64_cg", lambda: gen.vardiff2d(gen.grf_kappa(64, 0.1, 4)), "cg", 1e-10, 1000, 4),
    ("lognormal27_16_cg", lambda: gen.lognormal27(16, seed=2), "cg", 1e-10, 500, 2),
    # BiCGSTAB on the diagonally dominant 27-point system (on the 64^2
    # lognormal-field system a one-ulp change of b moves the oracle's own
    # BiCGSTAB solution by 1.4e-8: that case pins nothing)
    ("lognormal27_16_bicgstab", lambda: gen.lognormal27(16, seed=2), "bicgstab", 1e-10, 500, 2),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_ic0_solve_matches_oracle(dev, case):
    name, build, ksp, tol, maxiter, seed = case
    A = build()
    xs, b = gen.manufactured(A, seed)
    say(f"[ic0-solve] {name} n={len(b)}")
    xo, sto, ito, reso = oracle.solve(A, b, ksp, "ic0", tol=tol, maxiter=maxiter)
    assert sto == 0
    S = _solver(A, ksp, tol, maxiter, **DET)
    x, st = S.solve(b)
    assert st == xsolver_amd.XS_OK
    check_fp64_parity(name, A, b, x, S.iterations, S.resvec(), xo, ito, reso, tol,
                      oracle_rerun=lambda bp: oracle.solve(A, bp, ksp, "ic0", tol=tol,
                                                           maxiter=maxiter)[::3])
    S.close()


def test_ic0_grf512_full_size(dev):
    """The 512^2 lognormal-field diffusion matrix (n = 262,144), CG + IC(0)."""
    A = gen.vardiff2d(gen.grf_kappa(512, 0.1, 4))
    xs, b = gen.manufactured(A, 4)
    say(f"[ic0-512] n={len(b)} nnz={len(A[2])}")
    xo, sto, ito, reso = oracle.solve(A, b, "cg", "ic0", tol=1e-10, maxiter=5000)
    assert sto == 0
    S = _solver(A, "cg", 1e-10, 5000, parilu_async=0, parilu_sweeps=2200)
    x, st = S.solve(b)
    assert st == 0
    check_fp64_parity("grf2d_512 cg ic0", A, b, x, S.iterations, S.resvec(), xo, ito, reso, 1e-10,
                      oracle_rerun=lambda bp: oracle.solve(A, bp, "cg", "ic0", tol=1e-10,
                                                           maxiter=5000)[::3])
    S.close()


def test_ic0_reference_default_async(dev):
    """20 asynchronous in-place sweeps on the scaled matrix (the default):
    converges to rtol with the oracle's iteration count within a few."""
    A = gen.vardiff2d(gen.grf_kappa(64, 0.1, 4))
    xs, b = gen.manufactured(A, 4)
    xo, sto, ito, _ = oracle.solve(A, b, "cg", "ic0", tol=1e-10, maxiter=1000)
    S = _solver(A, "cg", 1e-10, 1000)
    x, st = S.solve(b)
    rr = true_relres(A, b, x)
    say(f"[ic0-async] gpu it={S.iterations} oracle it={ito} relres={rr:.2e}")
    assert st == 0 and rr < 1e-10 and abs(S.iterations - ito) <= max(2, ito // 10)
    S.close()


def test_ic0_sweep_apply_and_mixed(dev):
    """Jacobi-sweep apply of the IC factors equals the oracle's sweeps on the
    GPU's own factors; BiCGSTAB with FP32 inner iterations reaches an FP64
    residual below 1e-10."""
    import torch
    A = gen.lognormal27(12, seed=2)
    n = len(A[0]) - 1
    S = _solver(A, "bicgstab", 1e-10, 500, lsolve_sweeps=4, usolve_sweeps=3, **DET)
    S.pc_setup()
    rp, ci, lu, dpos = S.factors(n, len(A[2]))
    x = np.random.default_rng(1).uniform(-1, 1, n)
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    S.pc_apply(xd, yd, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    yo = oracle.ilu_sweeps(rp, ci, lu, dpos, 4, 3, x)
    assert np.abs(yd.cpu().numpy() - yo).max() < 1e-13 * np.abs(yo).max()
    xs, b = gen.manufactured(A, 2)
    xg, st = S.solve(b)
    assert st == 0 and true_relres(A, b, xg) < 1e-10
    S.set_option("precision", 1)
    S.set_option("factor_storage", 1)
    xm, st = S.solve(b)
    rr = true_relres(A, b, xm)
    say(f"[ic0-mixed] fp32 inner, bf16 factors: relres={rr:.2e} "
        f"err vs fp64 {np.linalg.norm(xm - xg) / np.linalg.norm(xg):.2e}")
    assert st == 0 and rr < 1e-10
    S.close()


def test_ic0_refuses_nonsymmetric_pattern(dev):
    A = gen.convdiff3d(6)
    rp, ci, v = (a.copy() for a in A)
    # drop one entry above the diagonal: the pattern loses its symmetry
    i = 10
    e = [k for k in range(rp[i], rp[i + 1]) if ci[k] > i][-1]
    keep = np.ones(len(ci), bool)
    keep[e] = False
    rp2 = np.r_[0, np.cumsum([keep[rp[r]:rp[r + 1]].sum() for r in range(len(rp) - 1)])]
    S = XSolver()
    S.mat_setup(rp2.astype(np.int32), ci[keep], v[keep])
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-8, 10)
    S.pc_settype("ic0")
    with pytest.raises(XSolverError) as err:
        S.pc_setup()
    assert err.value.status == xsolver_amd.XS_ERR_UNSUPPORTED
    S.close()
