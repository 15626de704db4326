"""The oracle's incomplete Cholesky (or_ic0, or_ict, or_ic_split) pinned.

* Bit-exact against golden vectors written by the reference's own
  hip/Precond/ic0.cpp and ict_l.cpp compiled in place (tests/golden/).
* Closed forms: IC(0) of a tridiagonal SPD matrix and ICT without dropping
  are the exact Cholesky factor (L L^T = A); ICT keeping no off-diagonal
  entry is 1/sqrt|a_ii|; the IC(0) invariant (L L^T)_ij = a_ij holds on the
  lower pattern of A.
* The LUsolve form (seqic0's split) applies (L~ D L~^T)^{-1}: checked with
  dense algebra.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from xsolver_amd import gen

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ic_golden.npz")
ICT_PARAMS = [(1e-8, 100), (1e-2, 3), (1e-3, 1), (0.0, 50)]


def names():
    G = np.load(GOLD)
    return sorted({k.split("/")[0] for k in G.files})


def load(name):
    G = np.load(GOLD)
    return G, (G[name + "/rowptr"], G[name + "/colidx"], G[name + "/val"])


def lower_dense(lp, lc, lv, n):
    """Cholesky-form L (diagonal = l_ii, not its reciprocal) as a dense array."""
    L = np.zeros((n, n))
    for i in range(n):
        for e in range(lp[i], lp[i + 1]):
            L[i, lc[e]] = 1.0 / lv[e] if lc[e] == i else lv[e]
    return L


@pytest.mark.parametrize("name", names())
def test_ic0_matches_reference_golden(name):
    G, A = load(name)
    lp, lc, lv = oracle.ic0(A)
    assert np.array_equal(lp, G[name + "/ic0/lptr"])
    assert np.array_equal(lc, G[name + "/ic0/lcol"])
    assert np.array_equal(lv, G[name + "/ic0/lval"])


@pytest.mark.parametrize("name", names())
@pytest.mark.parametrize("tol,maxfil", ICT_PARAMS)
def test_ict_matches_reference_golden(name, tol, maxfil):
    G, A = load(name)
    n = len(A[0]) - 1
    mf = min(maxfil, n - 1)
    k = f"{name}/ict_{tol:g}_{mf}"
    lp, lc, lv = oracle.ict(A, tol, mf)
    assert np.array_equal(lp, G[k + "/lptr"])
    assert np.array_equal(lc, G[k + "/lcol"])
    assert np.array_equal(lv, G[k + "/lval"])


def test_ic0_tridiagonal_is_exact_cholesky():
    n = 40
    rng = np.random.default_rng(3)
    off = -rng.uniform(0.2, 1.0, n - 1)
    d = np.abs(np.r_[off, 0]) + np.abs(np.r_[0, off]) + rng.uniform(0.1, 1.0, n)
    Ad = np.diag(d) + np.diag(off, 1) + np.diag(off, -1)
    S = sp.csr_matrix(Ad)
    A = (S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data)
    lp, lc, lv = oracle.ic0(A)
    L = lower_dense(lp, lc, lv, n)
    assert np.abs(L @ L.T - Ad).max() < 1e-14 * np.abs(Ad).max()
    assert np.allclose(L, np.linalg.cholesky(Ad), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("A", [gen.poisson2d(9), gen.vardiff2d(gen.grf_kappa(10, 0.2, 1)),
                               gen.lognormal27(4, seed=5)])
def test_ic0_pattern_invariant(A):
    rp, ci, v = A
    n = len(rp) - 1
    lp, lc, lv = oracle.ic0(A)
    L = lower_dense(lp, lc, lv, n)
    P = L @ L.T
    Ad = sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    rows = np.repeat(np.arange(n), np.diff(rp))
    low = ci <= rows
    err = np.abs(P[rows[low], ci[low]] - v[low]).max()
    assert err < 1e-13 * np.abs(v).max()
    # and the factor has exactly A's lower pattern
    assert np.array_equal(lp, np.r_[0, np.cumsum([np.sum(ci[rp[i]:rp[i + 1]] <= i)
                                                  for i in range(n)])])


@pytest.mark.parametrize("A", [gen.poisson2d(7), gen.lognormal27(3, seed=2)])
def test_ict_without_dropping_is_exact_cholesky(A):
    rp, ci, v = A
    n = len(rp) - 1
    lp, lc, lv = oracle.ict(A, 0.0, n - 1)
    L = lower_dense(lp, lc, lv, n)
    Ad = sp.csr_matrix((v, ci, rp), shape=(n, n)).toarray()
    assert np.abs(L @ L.T - Ad).max() < 1e-13 * np.abs(Ad).max()


def test_ict_keeping_nothing_is_inverse_sqrt_diagonal():
    A = gen.vardiff2d(gen.grf_kappa(6, 0.3, 2))
    rp, ci, v = A
    n = len(rp) - 1
    lp, lc, lv = oracle.ict(A, 1e6, 5)       # every off-diagonal is below the drop level
    assert np.array_equal(lp, np.arange(n + 1))
    diag = v[[e for i in range(n) for e in range(rp[i], rp[i + 1]) if ci[e] == i]]
    assert np.array_equal(lv, 1.0 / np.sqrt(np.abs(diag)))


@pytest.mark.parametrize("factor", ["ic0", "ict"])
def test_ic_split_lusolve_applies_the_factor_inverse(factor):
    A = gen.vardiff2d(gen.grf_kappa(8, 0.2, 3))
    n = len(A[0]) - 1
    lp, lc, lv = oracle.ic0(A) if factor == "ic0" else oracle.ict(A, 1e-3, 4)
    (lrp, lci, lvs), (urp, uci, uvs), dinv = oracle.ic_split(lp, lc, lv)
    Lt = sp.csr_matrix((lvs, lci, lrp), shape=(n, n)).toarray() + np.eye(n)
    U = sp.csr_matrix((uvs, uci, urp), shape=(n, n)).toarray() + np.eye(n)
    assert np.array_equal(U, Lt.T)
    L = lower_dense(lp, lc, lv, n)
    M = Lt @ np.diag(1.0 / dinv) @ U
    assert np.abs(M - L @ L.T).max() < 1e-13 * np.abs(M).max()
    # the solve through the preconditioned driver matches the dense one
    b = np.random.default_rng(0).uniform(-1, 1, n)
    x, st, it, _ = oracle.solve(A, b, "cg", factor, tol=1e-12, maxiter=500, ilut_tol=1e-3,
                                ilut_fill=4)
    assert st == 0
    Ad = sp.csr_matrix((A[2], A[1], A[0]), shape=(n, n)).toarray()
    assert np.linalg.norm(Ad @ x - b) / np.linalg.norm(b) < 1e-11
