"""Shared helpers for the GPU parity tests."""
import sys

import numpy as np

from xsolver_amd import gen


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_sparse(n, per_row, seed, dom=1.5, long_rows=()):
    """Nonsymmetric random pattern, diagonal present and dominant; rows listed
    in long_rows get (row_len) entries."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        k = dict(long_rows).get(i, per_row)
        cols = set(rng.choice(n, size=min(k, n), replace=False).tolist())
        cols.add(i)
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], np.int32)
    v = rng.standard_normal(len(ci))
    for i in range(n):
        seg = slice(rp[i], rp[i + 1])
        d = rp[i] + rows[i].index(i)
        v[d] = dom * np.abs(v[seg]).sum() + 1.0
    return rp, ci, v


def split_factors(rp, ci, lu, dpos):
    """Combined factors -> L strict lower CSR, U strict upper CSR, dinv."""
    n = len(rp) - 1
    lrp = [0]; lci = []; lv = []; urp = [0]; uci = []; uv = []
    for i in range(n):
        for e in range(rp[i], dpos[i]):
            lci.append(ci[e]); lv.append(lu[e])
        for e in range(dpos[i] + 1, rp[i + 1]):
            uci.append(ci[e]); uv.append(lu[e])
        lrp.append(len(lci)); urp.append(len(uci))
    L = (np.array(lrp, np.int32), np.array(lci, np.int32), np.array(lv))
    U = (np.array(urp, np.int32), np.array(uci, np.int32), np.array(uv))
    return L, U, 1.0 / lu[dpos]


def rel_err(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def true_relres(A, b, x):
    return float(np.linalg.norm(b - gen.spmv(A, x)) / np.linalg.norm(b))


# north-star acceptance of the plain FP64 path: the solution within 1e-10
# relative of the oracle's FP64 solution on the same seeded input
X_TOL = 1e-10
# residual histories: every entry within this relative deviation of the
# oracle's (absolute norms, entry 0 = ||r0||)
RES_RTOL = 1e-6


def resvec_dev(res, reso):
    m = min(len(res), len(reso))
    if m == 0:
        return 0.0
    den = np.maximum(np.abs(reso[:m]), 1e-300)
    return float(np.max(np.abs(res[:m] - reso[:m]) / den))


def check_fp64_parity(tag, A, b, x, it, res, xo, ito, reso, tol, it_slack=2,
                      true_res=True):
    """The acceptance of the plain FP64 path against the oracle run on the
    same input: iterations within +-it_slack, true relative residual below
    tol (recomputed in FP64 on the host), ||x - x_oracle|| / ||x_oracle|| <=
    X_TOL, and the whole residual history within RES_RTOL of the oracle's."""
    rr = true_relres(A, b, x)
    dx = float(np.linalg.norm(x - xo) / np.linalg.norm(xo))
    dres = resvec_dev(res, reso)
    say(f"   {tag}: oracle it={ito} gpu it={it} true relres={rr:.3e} "
        f"x-diff={dx:.2e} resvec dev={dres:.2e}")
    assert abs(it - ito) <= it_slack
    if true_res:
        assert rr < tol
    assert dx <= X_TOL
    assert len(res) == it + 1
    assert dres <= RES_RTOL
