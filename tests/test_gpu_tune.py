"""Autotuner (xs_autotune): the chosen SpMV format / row-block size /
vector-CSR lanes and long-row threshold / k-RHS SpMM columns per thread, and
after a factorisation the triangular solves' workgroups / gather width /
group mode, are among the candidates; the SpMV, the SpMM, the triangular
solves and a solve with them still match the oracle; the per-matrix cache
file is written and then reused."""
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
    lanes, thr, cw = (int(S.time_ms(k)) for k in ("tune:vec_lanes", "tune:long_thr", "tune:spmm_cw"))
    say(f"   vector CSR {lanes} lanes, long rows past {thr}; SpMM {cw} columns per thread")
    assert lanes in (2, 4, 8, 16, 32, 64) and thr // lanes in (4, 8, 16) and thr % lanes == 0
    # the SpMM is tuned with the ELL copy (0: no copy for these row lengths)
    assert cw in ((2, 4) if ell else (0, 2, 4))
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


def test_autotune_triangular_solves_and_spmm(dev, tmp_path, monkeypatch):
    """After ILU(0) set-up the tuner also picks the SpTRSV configuration; the
    exact triangular solves with it equal the oracle's, the k-RHS SpMM with
    the tuned columns per thread equals per-column SpMVs, and a second
    context takes both choices from the cache."""
    import torch
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("XS_TUNE_CACHE", str(cache))
    A = gen.convdiff3d(24)
    n = len(A[0]) - 1
    S = XSolver()
    S.mat_setup(*A)
    S.ksp_settype("bicgstab")
    S.ksp_setoption(1e-10, 500)
    S.pc_settype("ilu0")
    S.set_option("parilu_async", 0)
    S.set_option("parilu_sweeps", 400)
    S.set_option("parilu_scale", 0)
    S.pc_setup()
    S.autotune()
    wg, ga, pr, tus = (S.time_ms(k) for k in ("tune:trsv_wg", "tune:trsv_gather", "tune:trsv_prog",
                                               "tune:trsv_us"))
    say(f"[tune] convdiff3d_24 SpTRSV: {int(wg)} workgroups, gather {int(ga)}, mode {int(pr)}, "
        f"{tus:.1f} us per lower + upper solve")
    assert wg in (128, 256, 512, 1024) and ga in (4, 16) and pr in (0, 1) and tus > 0
    assert len([l for l in cache.read_text().split("\n") if l.strip()]) == 2
    # exact triangular solves with the tuned configuration: L U z = b for the
    # device factors (L unit lower, U with the pivots, one combined CSR)
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, n)
    bd = torch.from_numpy(b).to(dev)
    y = torch.empty_like(bd)
    z = torch.empty_like(bd)
    st = torch.cuda.current_stream().cuda_stream
    S.trsv(0, bd, y, st)
    S.trsv(1, y, z, st)
    torch.cuda.synchronize()
    rp, ci, lu, dpos = S.factors(n, len(A[2]))
    zz = z.cpu().numpy()
    u_z = np.zeros(n)
    for i in range(n):
        u_z[i] = sum(lu[e] * zz[ci[e]] for e in range(dpos[i], rp[i + 1]))
    lu_z = u_z.copy()
    for i in range(n):
        lu_z[i] += sum(lu[e] * u_z[ci[e]] for e in range(rp[i], dpos[i]))
    assert np.abs(lu_z - b).max() <= 1e-10 * np.abs(b).max()
    # the k-RHS SpMM with the tuned columns per thread
    k = 8
    X = rng.uniform(-1, 1, (n, k))
    Xd = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
    Yd = torch.empty_like(Xd)
    S.spmm(Xd, Yd, st)
    torch.cuda.synchronize()
    Y = Yd.cpu().numpy()
    scale = np.abs(A[2]).max() * np.diff(A[0]).max()
    for j in range(k):
        assert np.abs(Y[:, j] - oracle.spmv(A, X[:, j])).max() <= 1e-14 * scale
    xs, bb = gen.manufactured(A, 1)
    S.set_option("lsolve_sweeps", -1)
    S.set_option("usolve_sweeps", -1)
    xg, stt = S.solve(bb)
    assert stt == 0 and true_relres(A, bb, xg) < 1e-10
    S.close()
    T = XSolver()
    T.mat_setup(*A)
    T.ksp_settype("bicgstab")
    T.pc_settype("ilu0")
    T.set_option("parilu_async", 0)
    T.set_option("parilu_sweeps", 400)
    T.set_option("parilu_scale", 0)
    T.pc_setup()
    T.autotune()
    assert (T.time_ms("tune:trsv_wg"), T.time_ms("tune:trsv_gather"),
            T.time_ms("tune:trsv_prog")) == (wg, ga, pr)
    assert len([l for l in cache.read_text().split("\n") if l.strip()]) == 2
    T.close()
