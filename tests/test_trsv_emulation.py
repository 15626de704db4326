"""CPU emulation of the sync-free triangular-solve kernel's per-wavefront
algorithm (xsolver_amd/csrc/xs_kernels.hip, k_trsv): 64 rows per ticket,
dependencies before the wave's first row (lower) / after its last row (upper)
read from the solution vector, dependencies inside the wave resolved by the
log-step affine scan when every row depends at most on its neighbour, by the
64-step broadcast loop otherwise.  Waves are emulated in ticket order, which
is the order the kernel's waits resolve in.  Checked against the oracle's
sequential solves (oracle/xs_oracle.c)."""
import numpy as np
import pytest

import oracle
from tests.helpers import random_sparse
from xsolver_amd import gen

W = 64


def emulate(rp, ci, lu, dpos, dinv, b, upper):
    n = len(rp) - 1
    x = np.full(n, np.nan)
    nw = (n + W - 1) // W
    paths = set()
    for w in range(nw):
        if not upper:
            r0, r1 = w * W, min(n, w * W + W)
        else:
            r1 = n - w * W
            r0 = max(0, r1 - W)
        acc = np.zeros(W)
        di = np.ones(W)
        eb = np.zeros(W, int)
        ee = np.zeros(W, int)
        act = np.zeros(W, bool)
        for lane in range(W):
            r = r0 + lane
            if r >= r1:
                continue
            act[lane] = True
            acc[lane] = b[r]
            if not upper:
                e, end = rp[r], dpos[r]
                xe = e
                while xe < end and ci[xe] < r0:
                    xe += 1
                for q in range(e, xe):
                    assert not np.isnan(x[ci[q]])   # produced by an earlier wave
                    acc[lane] -= lu[q] * x[ci[q]]
                eb[lane], ee[lane] = xe, end
            else:
                di[lane] = dinv[r]
                beg = dpos[r] + 1
                e = rp[r + 1] - 1
                xb = e + 1
                while xb > beg and ci[xb - 1] >= r1:
                    xb -= 1
                for q in range(e, xb - 1, -1):
                    assert not np.isnan(x[ci[q]])
                    acc[lane] -= lu[q] * x[ci[q]]
                eb[lane], ee[lane] = beg, xb
        nint = ee - eb
        chain = all((not act[l]) or nint[l] == 0 or
                    (nint[l] == 1 and ci[eb[l]] == (r0 + l + (1 if upper else -1)))
                    for l in range(W))
        if chain:
            paths.add("scan")
            if not upper:
                c = acc.copy()
                m = np.array([-lu[eb[l]] if act[l] and nint[l] == 1 else 0.0 for l in range(W)])
            else:
                c = np.where(act, di * acc, 0.0)
                m = np.array([-di[l] * lu[eb[l]] if act[l] and nint[l] == 1 else 0.0 for l in range(W)])
            d = 1
            while d < W:
                if not upper:
                    cj = np.concatenate([c[:d], c[:-d]]); mj = np.concatenate([m[:d], m[:-d]])
                    has = np.arange(W) >= d
                else:
                    cj = np.concatenate([c[d:], c[-d:]]); mj = np.concatenate([m[d:], m[-d:]])
                    has = np.arange(W) + d < W
                c = np.where(has, c + m * cj, c)
                m = np.where(has, m * mj, m)
                d <<= 1
            val = c
        else:
            paths.add("loop")
            nrows = r1 - r0
            if not upper:
                e = eb.copy()
                for k in range(nrows):
                    xk = acc[k]
                    for l in range(W):
                        if act[l] and e[l] < ee[l] and ci[e[l]] == r0 + k:
                            acc[l] -= lu[e[l]] * xk
                            e[l] += 1
                val = acc
            else:
                e = ee - 1
                for k in range(nrows - 1, -1, -1):
                    xk = acc[k] * di[k]
                    for l in range(W):
                        if act[l] and e[l] >= eb[l] and ci[e[l]] == r0 + k:
                            acc[l] -= lu[e[l]] * xk
                            e[l] -= 1
                val = acc * di
        for l in range(W):
            if act[l]:
                x[r0 + l] = val[l]
    return x, paths


CASES = {
    "poisson2d_64": lambda: gen.poisson2d(64),         # only r-1 inside a wave: scan path
    "convdiff3d_9": lambda: gen.convdiff3d(9),         # r-1, r-9 inside a wave: loop path
    "poisson2d_12": lambda: gen.poisson2d(12),         # r-1 and r-12 inside a wave: loop path
    "lognormal27_6": lambda: gen.lognormal27(6, seed=2),
    "random_300": lambda: random_sparse(300, 6, 5),
}


@pytest.mark.parametrize("name", list(CASES))
def test_wave_algorithm_matches_sequential_solves(name):
    A = CASES[name]()
    rp, ci, _ = A
    lu, dpos = oracle.ilu0(A)
    n = len(rp) - 1
    dinv = 1.0 / lu[dpos]
    b = np.random.default_rng(9).uniform(-1, 1, n)
    from tests.helpers import split_factors
    L, U, dv = split_factors(rp, ci, lu, dpos)
    xl, pl = emulate(rp, ci, lu, dpos, dinv, b, upper=False)
    assert np.abs(xl - oracle.lower_solve(L, b)).max() < 1e-12 * np.abs(xl).max()
    xu, pu = emulate(rp, ci, lu, dpos, dinv, b, upper=True)
    assert np.abs(xu - oracle.upper_solve(U, dv, b)).max() < 1e-12 * np.abs(xu).max()
    if name == "poisson2d_64":
        assert pl == {"scan"} and pu == {"scan"}
    if name == "poisson2d_12":
        assert "loop" in pl and "loop" in pu
