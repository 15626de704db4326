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
