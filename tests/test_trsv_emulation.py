"""CPU emulation of the sync-free triangular-solve kernel's per-wavefront
algorithm (xsolver_amd/csrc/xs_kernels.hip, k_trsv): 64 rows per group, lanes
along the solve order (upper solves: lane l holds row r1-1-l), dependencies
outside the group read from the solution vector, dependencies inside the
group resolved by the DPP affine scan (row shifts 1/2/4/8, then the
row_bcast:15 / row_bcast:31 steps) when every row depends at most on the
previous one, otherwise lane by lane (each row final and published at its
turn, its in-wave products summed apart from the external ones).  Waves are emulated in ticket order, which
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
        nrows = r1 - r0
        # lanes along the solve order: upper solves run rows backwards
        row = [(r1 - 1 - l) if upper else (r0 + l) for l in range(W)]
        acc = np.zeros(W)
        di = np.ones(W)
        eb = np.zeros(W, int)
        ee = np.zeros(W, int)
        act = np.arange(W) < nrows
        for lane in range(nrows):
            r = row[lane]
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
                    (nint[l] == 1 and ci[eb[l]] == row[l] + (1 if upper else -1))
                    for l in range(W))
        if chain:
            # inclusive affine scan in lane order (the kernel's DPP steps:
            # shifts 1, 2, 4, 8 inside 16-lane rows, then row broadcasts)
            paths.add("scan")
            c = np.where(act, di * acc, 0.0)
            m = np.array([-di[l] * lu[eb[l]] if act[l] and nint[l] == 1 else 0.0 for l in range(W)])
            lanes = np.arange(W)
            for d in (1, 2, 4, 8):
                src = lanes - d
                ok = (lanes % 16) >= d
                cj = np.where(ok, c[np.clip(src, 0, W - 1)], 0.0)
                mj = np.where(ok, m[np.clip(src, 0, W - 1)], 1.0)
                c, m = c + m * cj, m * mj
            for srclane, rows_in in ((15, (1,)), (47, (3,)), (31, (2, 3))):
                ok = np.isin(lanes // 16, rows_in)
                if srclane == 15:
                    ok = ok & (lanes // 16 == 1)
                cj = np.where(ok, c[srclane], 0.0)
                mj = np.where(ok, m[srclane], 1.0)
                c, m = c + m * cj, m * mj
            val = c
        else:
            # one lane at a time: lane k's row is (external part + in-wave
            # part) * d once the rows of lanes < k are broadcast, published
            # at once; the in-wave products are summed apart, in solve order
            paths.add("loop")
            e = np.where(upper, ee - 1, eb)
            ain = np.zeros(W)
            val = np.zeros(W)
            for k in range(nrows):
                val[k] = (acc[k] + ain[k]) * di[k]
                x[row[k]] = val[k]
                rk = row[k]
                for l in range(nrows):
                    if nint[l] == 0:
                        continue
                    inside = (e[l] >= eb[l]) if upper else (e[l] < ee[l])
                    if inside and ci[e[l]] == rk:
                        ain[l] -= lu[e[l]] * val[k]
                        e[l] += -1 if upper else 1
        for l in range(nrows):
            x[row[l]] = val[l]
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
