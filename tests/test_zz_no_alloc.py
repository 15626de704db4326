"""The allocation-free batched solve (xs_batch_solve_noalloc) through the
C-ABI against the allocating entry point (xs_batch_create +
xs_batch_solve_device) on the same systems: two calls in a row, the second on
a smaller batch, then the size limit and a pattern that leaves its block.
Both entry points run the same per-system kernel with workgroup-deterministic
reductions, so iterations, solutions and residuals agree bit for bit."""
import ctypes

import numpy as np
import pytest

import xsolver_amd
from tests.helpers import say
from xsolver_amd import XBatch, batch_concat, gen

pytestmark = pytest.mark.gpu


def _systems(count, smin, smax, seed):
    sizes, systems = gen.batch_systems(count=count, smin=smin, smax=smax, seed=seed)
    sys_ptr, rp, ci, v = batch_concat(systems)
    xs = np.random.default_rng(seed + 100).uniform(-1, 1, sys_ptr[-1])
    b = gen.spmv((rp, ci, v), xs)
    return sys_ptr, rp, ci, v, b


def _reference(sys_ptr, rp, ci, v, b, tol, maxiter):
    B = XBatch(sys_ptr, rp, ci, v)
    x, st = B.solve(b, tol=tol, maxiter=maxiter, precision=0)
    it, rr = B.results()
    B.close()
    return x, st, it, rr


def _noalloc(torch, dev, sys_ptr, rp, ci, v, b, tol, maxiter, x0=None):
    lib = xsolver_amd.lib()
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
    d_sp, d_rp, d_ci = t(sys_ptr, torch.int32), t(rp, torch.int32), t(ci, torch.int32)
    d_v, d_b = t(v, torch.float64), t(b, torch.float64)
    d_x = torch.zeros_like(d_b) if x0 is None else t(x0, torch.float64)
    nsys = len(sys_ptr) - 1
    d_it = torch.zeros(nsys, dtype=torch.int32, device=dev)
    d_rr = torch.zeros(nsys, dtype=torch.float64, device=dev)
    p = lambda a: ctypes.c_void_p(a.data_ptr())
    stream = torch.cuda.current_stream()
    rc = lib.xs_batch_solve_noalloc(nsys, int(sys_ptr[-1]), len(v), p(d_sp), p(d_rp), p(d_ci),
                                    p(d_v), p(d_b), p(d_x), tol, maxiter, p(d_it), p(d_rr),
                                    ctypes.c_void_p(stream.cuda_stream))
    stream.synchronize()
    return rc, d_x.cpu().numpy(), d_it.cpu().numpy(), d_rr.cpu().numpy(), (d_sp, d_rp, d_ci, d_v,
                                                                           d_b, d_x, d_it, d_rr)


def test_noalloc_matches_allocating_entry_point(dev):
    import torch
    tol, maxiter = 1e-8, 5000
    for call, (count, smin, smax, seed) in enumerate([(60, 8, 96, 11), (17, 6, 40, 12)]):
        sys_ptr, rp, ci, v, b = _systems(count, smin, smax, seed)
        say(f"[noalloc] call {call + 1}: {count} systems, n={sys_ptr[-1]}, nnz={len(v)}")
        x_ref, st_ref, it_ref, rr_ref = _reference(sys_ptr, rp, ci, v, b, tol, maxiter)
        assert st_ref == 0
        rc, x, it, rr, _ = _noalloc(torch, dev, sys_ptr, rp, ci, v, b, tol, maxiter)
        assert rc == xsolver_amd.XS_OK
        dx = float(np.abs(x - x_ref).max())
        say(f"   iterations {it.min()}..{it.max()}, max |x - x_ref| {dx:.1e}, "
            f"max relres {rr.max():.2e}")
        assert np.array_equal(it, it_ref)
        assert np.array_equal(x, x_ref)
        assert np.array_equal(rr, rr_ref)
        assert rr.max() < tol


def test_noalloc_nonzero_x0_matches(dev):
    import torch
    sys_ptr, rp, ci, v, b = _systems(9, 10, 30, 13)
    x0 = np.random.default_rng(5).uniform(-1, 1, sys_ptr[-1])
    B = XBatch(sys_ptr, rp, ci, v)
    x_ref, st = B.solve(b, tol=1e-9, maxiter=3000, precision=0, x0=x0)
    it_ref, rr_ref = B.results()
    B.close()
    rc, x, it, rr, _ = _noalloc(torch, dev, sys_ptr, rp, ci, v, b, 1e-9, 3000, x0=x0)
    assert st == 0 and rc == 0
    assert np.array_equal(it, it_ref) and np.array_equal(x, x_ref)


def test_noalloc_limits_and_bad_pattern(dev):
    import torch
    lib = xsolver_amd.lib()
    sys_ptr, rp, ci, v, b = _systems(6, 8, 20, 14)
    rc, x, it, rr, bufs = _noalloc(torch, dev, sys_ptr, rp, ci, v, b, 1e-8, 2000)
    assert rc == 0 and (it > 0).all()
    p = lambda a: ctypes.c_void_p(a.data_ptr())
    d_sp, d_rp, d_ci, d_v, d_b, d_x, d_it, d_rr = bufs
    # beyond the workspace: rejected on the host before any device work
    big = xsolver_amd.XS_BATCH_NOALLOC_MAX_ROWS + 1
    rc = lib.xs_batch_solve_noalloc(6, big, len(v), p(d_sp), p(d_rp), p(d_ci), p(d_v), p(d_b),
                                    p(d_x), 1e-8, 2000, p(d_it), p(d_rr), None)
    assert rc == xsolver_amd.XS_ERR_TOO_LARGE
    assert lib.xs_batch_solve_noalloc(0, 10, 10, p(d_sp), p(d_rp), p(d_ci), p(d_v), p(d_b),
                                      p(d_x), 1e-8, 2000, p(d_it), p(d_rr), None) == \
        xsolver_amd.XS_ERR_ARG
    # one column of system 2 moved into system 3 (still inside the batch):
    # system 2 is reported unsolved, the others solve as before
    ci_bad = ci.copy()
    e = int(rp[sys_ptr[2]])
    ci_bad[e] = int(sys_ptr[3])
    rc, x2, it2, rr2, _ = _noalloc(torch, dev, sys_ptr, rp, ci_bad, v, b, 1e-8, 2000)
    assert rc == 0
    assert it2[2] == -1 and np.isnan(rr2[2])
    ok = [s for s in range(6) if s != 2]
    assert np.array_equal(it2[ok], it[ok]) and (rr2[ok] < 1e-8).all()
