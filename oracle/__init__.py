"""ctypes bindings for the CPU oracle (test infrastructure only).

The oracle is the checker: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module.  The product library
(xsolver_amd/libxsolver_amd.so) never calls it.

``liboracle.so`` is the C restatement in xs_oracle.c.  ``_ref/libxsref.so`` is
the reference's own hip/Precond/ilut_diag_in_csr.cpp compiled in place (it
exists only where the reference tree was available at build time).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int_p = ctypes.POINTER(ctypes.c_int)


def build(ref=None):
    """Compile liboracle.so (and _ref/libxsref.so when the reference exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference")
    if ref:
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        L.or_spmv.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_bmax.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.or_spmv_t.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_lower_solve.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp]
        L.or_upper_solve.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.or_jacobi_setup.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_double, _dp]
        L.or_ilu0.argtypes = [ctypes.c_int, _ip, _ip, _dp, _dp, _ip]
        L.or_ilut.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, _dp, _ip, _ip, _ip, _dp]
        L.or_ilu_split.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _ip, _ip, _dp,
                                   _ip, _ip, _dp, _dp]
        L.or_parilu_sweep.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _ip, _dp,
                                      _ip, _ip, _dp, _dp, _dp]
        L.or_solve.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _ip, _ip, _dp,
                               _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _dp, _c_int_p]
        L.or_ilu_sweeps.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, ctypes.c_int,
                                    ctypes.c_int, _dp, _dp]
        L.or_parilut.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _ip]
        L.or_batch_cg.argtypes = [ctypes.c_int, _ip, _ip, _ip, _dp, _dp, _dp, ctypes.c_double,
                                  ctypes.c_int, _ip, _dp]
        L.or_ic0.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_double, ctypes.c_double,
                             _ip, _ip, _dp]
        L.or_ict.argtypes = [ctypes.c_int, _ip, _ip, _dp, ctypes.c_double, ctypes.c_double,
                             ctypes.c_int, _ip, _ip, _dp]
        L.or_ic_split.argtypes = [ctypes.c_int, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip, _dp, _dp]
        for f in ("or_jacobi_setup", "or_ilu0", "or_ilut", "or_solve", "or_parilut",
                  "or_batch_cg", "or_ic0", "or_ict"):
            getattr(L, f).restype = ctypes.c_int
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's compiled ILU code, or None where it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libxsref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_ilu0_simple.argtypes = [_dp, _dp, _ip, _ip, ctypes.c_int, _ip]
        R.ref_ilutp_final.argtypes = [_dp, _dp, _ip, _ip, ctypes.c_int, _dp, ctypes.c_int,
                                      _ip, _ip, _ip, ctypes.c_double]
        R.ref_ic0_csr_A.argtypes = [ctypes.c_int, _ip, _dp, _ip, _ip, _dp, _ip, ctypes.c_double,
                                    ctypes.c_double]
        R.ref_ict_csr_A_l.argtypes = [ctypes.c_int, _ip, _dp, _ip, _ip, _dp, _ip,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int]
        _REF = R
    return _REF


def _csr(A):
    rp, ci, v = A
    return (np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32),
            np.ascontiguousarray(v, np.float64))


def spmv(A, x):
    rp, ci, v = _csr(A)
    y = np.empty(len(rp) - 1)
    lib().or_spmv(len(rp) - 1, rp, ci, v, np.ascontiguousarray(x, np.float64), y)
    return y


def bmax(A, b, x):
    rp, ci, v = _csr(A)
    y = np.empty(len(rp) - 1)
    lib().or_bmax(len(rp) - 1, rp, ci, v, np.ascontiguousarray(b, np.float64),
                  np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_t(A, x):
    rp, ci, v = _csr(A)
    y = np.empty(len(rp) - 1)
    lib().or_spmv_t(len(rp) - 1, rp, ci, v, np.ascontiguousarray(x, np.float64), y)
    return y


def lower_solve(L, b):
    rp, ci, v = _csr(L)
    x = np.empty(len(rp) - 1)
    lib().or_lower_solve(len(rp) - 1, rp, ci, v, np.ascontiguousarray(b, np.float64), x)
    return x


def upper_solve(U, dinv, b):
    rp, ci, v = _csr(U)
    x = np.empty(len(rp) - 1)
    lib().or_upper_solve(len(rp) - 1, rp, ci, v, np.ascontiguousarray(dinv, np.float64),
                         np.ascontiguousarray(b, np.float64), x)
    return x


def ilu_sweeps(rp, ci, luval, uptr, kl, ku, x):
    """Reference GPU ILU apply: kl Jacobi sweeps on L, ku on U (combined factors)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    y = np.empty(len(rp) - 1)
    lib().or_ilu_sweeps(len(rp) - 1, rp, ci, np.ascontiguousarray(luval, np.float64),
                        np.ascontiguousarray(uptr, np.int32), kl, ku,
                        np.ascontiguousarray(x, np.float64), y)
    return y


def jacobi_dinv(A, small=0.0):
    rp, ci, v = _csr(A)
    d = np.empty(len(rp) - 1)
    bad = lib().or_jacobi_setup(len(rp) - 1, rp, ci, v, small, d)
    return d, bad


def ilu0(A):
    """Combined ILU(0) values on A's pattern and the pivot positions."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    lu = np.empty(len(v))
    uptr = np.empty(n, np.int32)
    rc = lib().or_ilu0(n, rp, ci, v, lu, uptr)
    if rc:
        raise ValueError("row %d has no diagonal entry" % (rc - 1))
    return lu, uptr


def ilut(A, lfil, droptol):
    """Combined ILUT rows (lurows, lucols, luval), pivot positions, 1/pivot."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    cap = (2 * lfil + 2) * n
    luval = np.zeros(cap)
    lucols = np.zeros(cap, np.int32)
    lurows = np.zeros(n + 1, np.int32)
    uptr = np.zeros(n + 1, np.int32)
    diag = np.zeros(n)
    rc = lib().or_ilut(n, rp, ci, v, lfil, droptol, cap, luval, lucols, lurows, uptr, diag)
    if rc:
        raise ValueError("ILUT capacity exceeded")
    nnz = lurows[n]
    return lurows, lucols[:nnz].copy(), luval[:nnz].copy(), uptr[:n].copy(), diag


def ic0(A, smallnum=1e-4, droptol=1e-4):
    """IC(0) as the reference stores it: lower CSR (lp, lc, lv) with 1/l_ii
    last in each row."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    cap = max(len(v), 1)
    lp = np.zeros(n + 1, np.int32); lc = np.zeros(cap, np.int32); lv = np.zeros(cap)
    rc = lib().or_ic0(n, rp, ci, v, smallnum, droptol, lp, lc, lv)
    if rc:
        raise ValueError("row %d: lower part does not end at the diagonal" % (rc - 1))
    return lp, lc[:lp[n]].copy(), lv[:lp[n]].copy()


def ict(A, droptol, maxfil, smallnum=1e-4):
    """ICT (the reference's ict_csr_A_l) in the same layout as ic0()."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    cap = (maxfil + 1) * n
    lp = np.zeros(n + 1, np.int32); lc = np.zeros(cap, np.int32); lv = np.zeros(cap)
    rc = lib().or_ict(n, rp, ci, v, smallnum, droptol, maxfil, lp, lc, lv)
    if rc:
        raise ValueError("row %d has no diagonal entry" % (rc - 1))
    return lp, lc[:lp[n]].copy(), lv[:lp[n]].copy()


def ic_split(lp, lc, lv):
    """seqic0 / seqict's conversion: (L~ strict lower CSR, U = L~^T strict
    upper CSR, dinv = 1/l_ii^2)."""
    lp = np.ascontiguousarray(lp, np.int32); lc = np.ascontiguousarray(lc, np.int32)
    lv = np.ascontiguousarray(lv, np.float64)
    n = len(lp) - 1
    cap = max(int(lp[n]) - n, 1)
    lrp = np.zeros(n + 1, np.int32); urp = np.zeros(n + 1, np.int32)
    lci = np.zeros(cap, np.int32); uci = np.zeros(cap, np.int32)
    lvs = np.zeros(cap); uvs = np.zeros(cap); dinv = np.zeros(n)
    lib().or_ic_split(n, lp, lc, lv, lrp, lci, lvs, urp, uci, uvs, dinv)
    m = lrp[n]
    return (lrp, lci[:m].copy(), lvs[:m].copy()), (urp, uci[:m].copy(), uvs[:m].copy()), dinv


def ic_combined(A, lp, lc, lv):
    """An IC factor (ic0() layout, on a lower pattern contained in A's) as
    combined LU values on A's pattern: L~_ij = l_ij / l_jj below the
    diagonal, U = D L~^T (D = diag(l_ii^2)) on and above it; entries of A
    outside the factor's pattern are 0."""
    rp, ci, _ = _csr(A)
    n = len(rp) - 1
    d = np.empty(n)
    vals = {}
    for i in range(n):
        for e in range(lp[i], lp[i + 1]):
            if lc[e] == i:
                d[i] = 1.0 / lv[e]
    for i in range(n):
        for e in range(lp[i], lp[i + 1]):
            j = int(lc[e])
            if j != i:
                vals[(i, j)] = lv[e] / d[j]
    lu = np.zeros(len(ci))
    for i in range(n):
        for e in range(rp[i], rp[i + 1]):
            j = int(ci[e])
            if j < i:
                lu[e] = vals.get((i, j), 0.0)
            elif j == i:
                lu[e] = d[i] * d[i]
            else:
                lu[e] = d[i] * d[i] * vals.get((j, i), 0.0)
    return lu


def ilu_split(rp, ci, luval, uptr):
    """Split combined factors into L (strict lower), U (strict upper scaled by
    1/u_ii) and dinv, as the reference's seqilu0/seqilut do."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    luval = np.ascontiguousarray(luval, np.float64)
    uptr = np.ascontiguousarray(uptr, np.int32)
    n = len(rp) - 1
    cap = max(int(rp[n]), 1)
    lrp = np.zeros(n + 1, np.int32); urp = np.zeros(n + 1, np.int32)
    lci = np.zeros(cap, np.int32); uci = np.zeros(cap, np.int32)
    lv = np.zeros(cap); uv = np.zeros(cap); dinv = np.zeros(n)
    lib().or_ilu_split(n, rp, ci, luval, uptr, lrp, lci, lv, urp, uci, uv, dinv)
    nl, nu = lrp[n], urp[n]
    return ((lrp, lci[:nl].copy(), lv[:nl].copy()),
            (urp, uci[:nu].copy(), uv[:nu].copy()), dinv)


def parilu_sweep(A, L, Ucsc, n):
    rp, ci, a = _csr(A)
    lrp, lci, lv = _csr(L)
    ucp, uri, uv = _csr(Ucsc)
    lo = np.empty_like(lv); uo = np.empty_like(uv)
    lib().or_parilu_sweep(n, rp, ci, a, lrp, lci, lv, ucp, uri, uv, lo, uo)
    return lo, uo


def parilut(A, sweeps=10, fill=2.0, droptol=1e-4, trace=False):
    """ParILUT factors: L (strict lower CSR), U (upper CSR, diagonal first).
    With trace=True also an (sweeps, 4) array per outer step: strict-L
    candidates, U candidates, kept strict-L entries, kept U entries (U counts
    include the pivots)."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    cap = int((fill + 3.0) * len(v)) + 2 * n + 16
    lrp = np.zeros(n + 1, np.int32); urp = np.zeros(n + 1, np.int32)
    lci = np.zeros(cap, np.int32); uci = np.zeros(cap, np.int32)
    lv = np.zeros(cap); uv = np.zeros(cap)
    tr = np.zeros(4 * max(sweeps, 1), np.int32)
    rc = lib().or_parilut(n, rp, ci, v, sweeps, fill, droptol, cap, lrp, lci, lv, urp, uci, uv, tr)
    if rc:
        raise ValueError("ParILUT capacity exceeded")
    out = ((lrp, lci[:lrp[n]].copy(), lv[:lrp[n]].copy()),
           (urp, uci[:urp[n]].copy(), uv[:urp[n]].copy()))
    return out + (tr.reshape(-1, 4)[:sweeps].copy(),) if trace else out


def batch_cg(sys_ptr, A, b, tol=1e-8, maxiter=5000, x0=None):
    """CG + Jacobi on every diagonal block of a block-diagonal system.
    Returns (x, iterations per system, final true relative residual per
    system, number not converged)."""
    rp, ci, v = _csr(A)
    sp_ = np.ascontiguousarray(sys_ptr, np.int32)
    nsys = len(sp_) - 1
    x = np.zeros(len(rp) - 1) if x0 is None else np.array(x0, np.float64, copy=True)
    its = np.zeros(nsys, np.int32)
    rel = np.zeros(nsys)
    bad = lib().or_batch_cg(nsys, sp_, rp, ci, v, np.ascontiguousarray(b, np.float64), x, tol,
                            maxiter, its, rel)
    return x, its, rel, bad


def solve(A, b, ksp="cg", pc="jacobi", tol=1e-8, abstol=0.0, maxiter=1000,
          restart=0, x0=None, ilut_tol=1e-4, ilut_fill=10, lsweeps=0, usweeps=0):
    """Run one oracle solve.  Returns (x, status, usediter, resvec)."""
    rp, ci, v = _csr(A)
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64, copy=True)
    resvec = np.zeros(maxiter + 2)
    it = ctypes.c_int(0)
    st = lib().or_solve(ksp.encode(), pc.encode(), n, rp, ci, v,
                        np.ascontiguousarray(b, np.float64), x, tol, abstol, maxiter,
                        restart, ilut_tol, ilut_fill, lsweeps, usweeps, resvec,
                        ctypes.byref(it))
    return x, st, it.value, resvec[: it.value + 1].copy()
