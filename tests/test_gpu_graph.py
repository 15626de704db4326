"""HIP-graph replay of the iteration batches: a solve with use_graph = 1
(batches of check_every iterations captured once per solve and replayed)
gives bit-for-bit the result of the same solve with kernels launched one by
one (same kernels, same order), on repeated solves with eager work and
synchronisations in between, for the single, block and mixed-precision
drivers."""
import numpy as np
import pytest

from tests.helpers import say, true_relres
from xsolver_amd import XSolver, gen

pytestmark = pytest.mark.gpu


def make(A, ksp, pc, tol, maxiter, graph, **opts):
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype(ksp)
    S.ksp_setoption(tol, maxiter)
    S.pc_settype(pc)
    S.set_option("parilu_async", 0)
    S.set_option("use_graph", graph)
    for k, v in opts.items():
        S.set_option(k, v)
    return S


CASES = [
    ("convdiff3d_20_bicgstab_ilu0", lambda: gen.convdiff3d(20), "bicgstab", "ilu0", 1e-10, 1000, {}),
    ("poisson2d_128_cg_jacobi", lambda: gen.poisson2d(128), "cg", "jacobi", 1e-8, 5000, {}),
    ("grf2d_40_cg_ilu0", lambda: gen.vardiff2d(gen.grf_kappa(40, 0.1, 4)), "cg", "ilu0", 1e-10, 2000, {}),
    ("convdiff3d_20_mixed_bicgstab", lambda: gen.convdiff3d(20), "bicgstab", "ilu0", 1e-10, 1000,
     {"precision": 1}),
    ("convdiff3d_20_bicgstab_sweeps_5_2", lambda: gen.convdiff3d(20), "bicgstab", "ilu0", 1e-10,
     1000, {"lsolve_sweeps": 5, "usolve_sweeps": 2}),
    ("convdiff3d_20_mixed_sweeps_5_2", lambda: gen.convdiff3d(20), "bicgstab", "ilu0", 1e-10,
     1000, {"precision": 1, "lsolve_sweeps": 5, "usolve_sweeps": 2}),
    ("lognormal27_10_gmres_ilut_sweeps", lambda: gen.lognormal27(10, seed=2), "gmres", "ilut",
     1e-10, 500, {"lsolve_sweeps": 10, "usolve_sweeps": 15}),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_graph_replay_matches_eager(dev, case):
    import torch
    name, build, ksp, pc, tol, maxiter, opts = case
    A = build()
    _, b = gen.manufactured(A, 1)
    say(f"[graph] {name}")
    E = make(A, ksp, pc, tol, maxiter, 0, **opts)
    G = make(A, ksp, pc, tol, maxiter, 1, **opts)
    xe, ste = E.solve(b)
    for rep in range(3):
        xg, stg = G.solve(b)
        assert ste == stg == 0
        assert G.iterations == E.iterations
        assert np.array_equal(xg, xe), rep
        # eager work and a synchronisation between replays
        t = torch.arange(1000, dtype=torch.float64, device=dev)
        assert float(t.sum()) == 499500.0
        torch.cuda.synchronize()
    assert true_relres(A, b, xg) < tol
    E.close()
    G.close()


def test_graph_replay_block_solve(dev):
    A = gen.convdiff3d(16)
    B = np.stack([gen.manufactured(A, s)[1] for s in range(1, 7)], 1)
    E = make(A, "bicgstab", "ilu0", 1e-10, 1000, 0)
    G = make(A, "bicgstab", "ilu0", 1e-10, 1000, 1)
    Xe, ste = E.solve_multi(B)
    for rep in range(2):
        Xg, stg = G.solve_multi(B)
        assert ste == stg == 0
        assert np.array_equal(G.multi_results()[0], E.multi_results()[0])
        assert np.array_equal(Xg, Xe), rep
    E.close()
    G.close()
