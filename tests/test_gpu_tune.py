"""SpMV autotuner (xs_autotune): the chosen format/row-block size is one of
the candidates, the SpMV and a solve with it still match the oracle, and the
per-matrix cache file is written and then reused."""
import numpy as np
import pytest

import oracle
from tests.helpers import random_sparse, say, true_relres
from xsolver_amd import XSolver, gen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,build", [
    ("convdiff3d_48", lambda: gen.convdiff3d(48)),
    ("longrows_4000", lambda: random_sparse(4000, 5, 31, long_rows=[(10, 3000), (2000, 1500)])),
])
def test_autotune_choice_and_parity(dev, name, build, tmp_path, monkeypatch):
    import torch
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("XS_TUNE_CACHE", str(cache))
    A = build()
    n = len(A[0]) - 1
    S = XSolver()
    S.mat_setup(*A)
    chunk, ell, us = S.autotune()
    say(f"[tune] {name}: chunk={chunk} ell={ell} {us:.2f} us")
    # chunk 0 = the row-binned vector-CSR kernel
    assert chunk in (0, 512, 1024, 1536, 2048, 3072, 4096) and ell in (0, 1) and us > 0
    lines = cache.read_text().split("\n")
    assert len([l for l in lines if l.strip()]) == 1
    x = np.random.default_rng(2).uniform(-1, 1, n)
    yd = torch.empty(n, dtype=torch.float64, device=dev)
    S.spmv(torch.from_numpy(x).to(dev), yd, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    scale = np.abs(A[2]).max() * np.diff(A[0]).max()
    assert np.abs(yd.cpu().numpy() - oracle.spmv(A, x)).max() <= 1e-14 * scale
    _, b = gen.manufactured(A, 1)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 1000)
    S.pc_settype("jacobi")
    xg, st = S.solve(b)
    assert st == 0 and true_relres(A, b, xg) < 1e-10
    S.close()
    # a second context on the same matrix takes the cached choice
    T = XSolver()
    T.mat_setup(*A)
    assert T.autotune() == (chunk, ell, pytest.approx(us, rel=1e-3))
    assert len([l for l in cache.read_text().split("\n") if l.strip()]) == 1
    T.close()
