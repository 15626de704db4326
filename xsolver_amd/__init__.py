"""MI355X (gfx950) preconditioned Krylov solver -- Python host mirror.

The product is the C-ABI library ``libxsolver_amd.so`` (include/xsolver_c.h).
This module binds it with ctypes and mirrors the reference's API
(hip/xsolver_c.h:37-49: mat_setup -> ksp_settype -> ksp_setoption ->
pc_settype -> solve -> free) as a class, so tests read like the reference's
call sequence (hip/solve_base.h:4-24).  Device vectors are passed as torch
tensors (``data_ptr()``) for the kernel-level entry points; torch is only
used for device memory and streams.

There is no CPU fallback: if the library is missing, importing the solver
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxsolver_amd.so")
_LIB = None

XS_OK = 0
XS_NOT_CONVERGED = 1
XS_ERR_ARG = -1
XS_ERR_STATE = -2
XS_ERR_UNSUPPORTED = -3
XS_ERR_DEVICE = -4
XS_ERR_BREAKDOWN = -5
XS_ERR_PIVOT = -6
XS_ERR_MEMORY = -7
XS_ERR_TOO_LARGE = -8
XS_BATCH_NOALLOC_MAX_ROWS = 1048576

EXPORTS = [
    "xsolver_mat_setup", "xsolver_ksp_settype", "xsolver_ksp_set_absolute_tol",
    "xsolver_ksp_setoption", "xsolver_pc_settype", "xsolver_pc_setilut",
    "xsolver_pc_setict", "xsolver_solve", "xsolver_free", "xsolver_last_error",
    "xsolver_get_iterations", "xsolver_get_residual",
    "xs_create", "xs_destroy", "xs_status_string", "xs_mat_setup", "xs_mat_setup_device",
    "xs_mat_update_values", "xs_ksp_settype", "xs_ksp_set_absolute_tol", "xs_ksp_setoption",
    "xs_pc_settype", "xs_pc_setilut", "xs_set_option", "xs_pc_setup", "xs_solve",
    "xs_solve_device", "xs_get_iterations", "xs_get_residual", "xs_get_initial_residual",
    "xs_get_resvec", "xs_get_time_ms", "xs_spmv", "xs_bmax", "xs_spmv_t", "xs_pc_apply",
    "xs_pc_apply_t", "xs_trsv", "xs_pc_get_factors",
    "xs_batch_create", "xs_batch_solve", "xs_batch_solve_device", "xs_batch_get_results",
    "xs_batch_get_time_ms", "xs_batch_destroy", "xs_solve_transpose_device",
    "xs_grad_values_device", "xs_solve_multi_device", "xs_solve_multi", "xs_get_multi_results",
    "xs_get_multi_resvec", "xs_spmm", "xs_pc_apply_multi", "xs_autotune", "xs_mtx_read",
    "xs_batch_solve_noalloc",
    "cpu_mat_setup", "cpu_mat_update", "gpu_mat_setup", "gpu_mat_format", "gpu_mat_update",
    "gpu_mat_free", "gpu_preconditioner_setup", "gpu_preconditioner_update",
    "gpu_preconditioner_free", "gpu_cg_solve", "gpu_icg_solve", "gpu_cgm_solve",
    "gpu_icgm_solve", "gpu_bicgstab_solve", "gpu_ibicgstab_solve", "gpu_gcr_solve",
    "gpu_igcr_solve", "gpu_gmres_solve", "gpu_fgmres_solve", "gpu_last_error",
]

XS_RHS_COLMAJOR = 0
XS_RHS_INTERLEAVED = 1
MAX_RHS = 64


class XSolverError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_cp = ctypes.c_char_p


def lib():
    """Load libxsolver_amd.so; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `python -m xsolver_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "xs_create": (_i, [ctypes.POINTER(_vp)]),
        "xs_destroy": (_i, [_vp]),
        "xs_status_string": (_cp, [_i]),
        "xs_mat_setup": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
        "xs_mat_setup_device": (_i, [_vp, _i, _i, _vp, _vp, _vp, _i]),
        "xs_mat_update_values": (_i, [_vp, _vp, _i]),
        "xs_ksp_settype": (_i, [_vp, _cp]),
        "xs_ksp_set_absolute_tol": (_i, [_vp, _d]),
        "xs_ksp_setoption": (_i, [_vp, _d, _i, _i]),
        "xs_pc_settype": (_i, [_vp, _cp]),
        "xs_pc_setilut": (_i, [_vp, _d, _i]),
        "xs_set_option": (_i, [_vp, _cp, _d]),
        "xs_pc_setup": (_i, [_vp]),
        "xs_solve": (_i, [_vp, _vp, _vp]),
        "xs_solve_device": (_i, [_vp, _vp, _vp]),
        "xs_get_iterations": (_i, [_vp]),
        "xs_get_residual": (_d, [_vp]),
        "xs_get_initial_residual": (_d, [_vp]),
        "xs_get_resvec": (_i, [_vp, _vp, _i]),
        "xs_get_time_ms": (_d, [_vp, _cp]),
        "xs_spmv": (_i, [_vp, _vp, _vp, _vp]),
        "xs_bmax": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "xs_spmv_t": (_i, [_vp, _vp, _vp, _vp]),
        "xs_pc_apply": (_i, [_vp, _vp, _vp, _vp]),
        "xs_pc_apply_t": (_i, [_vp, _vp, _vp, _vp]),
        "xs_trsv": (_i, [_vp, _i, _vp, _vp, _vp]),
        "xs_pc_get_factors": (_i, [_vp, _vp, _vp, _vp, _vp, _i]),
        "xs_batch_create": (_i, [ctypes.POINTER(_vp), _i, _vp, _vp, _vp, _vp]),
        "xs_batch_solve": (_i, [_vp, _vp, _vp, _d, _i, _i]),
        "xs_batch_solve_device": (_i, [_vp, _vp, _vp, _d, _i, _i]),
        "xs_batch_get_results": (_i, [_vp, _vp, _vp]),
        "xs_batch_solve_noalloc": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _vp,
                                        _vp, _vp]),
        "xs_batch_get_time_ms": (_d, [_vp]),
        "xs_batch_destroy": (_i, [_vp]),
        "xs_solve_transpose_device": (_i, [_vp, _vp, _vp]),
        "xs_grad_values_device": (_i, [_vp, _vp, _vp, _vp, _vp]),
        "xs_solve_multi_device": (_i, [_vp, _i, _vp, _vp, _i]),
        "xs_solve_multi": (_i, [_vp, _i, _vp, _vp, _i]),
        "xs_get_multi_results": (_i, [_vp, _vp, _vp, _vp, _i]),
        "xs_get_multi_resvec": (_i, [_vp, _i, _vp, _i]),
        "xs_spmm": (_i, [_vp, _i, _vp, _vp, _vp]),
        "xs_autotune": (_i, [_vp]),
        "xs_mtx_read": (_i, [_cp, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
        "xs_pc_apply_multi": (_i, [_vp, _i, _vp, _vp, _vp]),
        "xsolver_mat_setup": (None, [_i, _vp, _vp, _vp, _vp]),
        "xsolver_ksp_settype": (None, [_cp]),
        "xsolver_ksp_set_absolute_tol": (None, [_d]),
        "xsolver_ksp_setoption": (None, [_d, _i, _vp, _vp, _i]),
        "xsolver_pc_settype": (None, [_cp]),
        "xsolver_pc_setilut": (None, [_d, _i]),
        "xsolver_pc_setict": (None, [_d, _i]),
        "xsolver_solve": (None, [_vp, _vp]),
        "xsolver_free": (None, []),
        "xsolver_last_error": (_i, []),
        "xsolver_get_iterations": (_i, []),
        "xsolver_get_residual": (_d, []),
        # Fortran-style MCSR surface: every argument by address
        "cpu_mat_setup": (None, [_vp] * 7),
        "cpu_mat_update": (None, [_vp, _vp]),
        "gpu_mat_setup": (None, []),
        "gpu_mat_format": (None, [_cp, _vp]),
        "gpu_mat_update": (None, []),
        "gpu_mat_free": (None, []),
        "gpu_preconditioner_setup": (None, [_cp]),
        "gpu_preconditioner_update": (None, [_cp]),
        "gpu_preconditioner_free": (None, [_cp]),
        "gpu_last_error": (_i, []),
    }
    for name in ("gpu_cg_solve", "gpu_icg_solve", "gpu_cgm_solve", "gpu_icgm_solve",
                 "gpu_bicgstab_solve", "gpu_ibicgstab_solve"):
        sig[name] = (None, [_vp] * 7)
    for name in ("gpu_gcr_solve", "gpu_igcr_solve", "gpu_gmres_solve", "gpu_fgmres_solve"):
        sig[name] = (None, [_vp] * 8)
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def _check(rc, what):
    if rc < 0:
        msg = lib().xs_status_string(rc).decode()
        raise XSolverError(rc, f"{what}: {msg} ({rc})")
    return rc


def _ptr(a):
    """Address of a numpy array or torch tensor."""
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


class XSolver:
    """One solver context (handle API of include/xsolver_c.h)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().xs_create(ctypes.byref(h)), "xs_create")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib().xs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference call sequence ------------------------------------------
    def mat_setup(self, rowptr, colidx, val, advanced_pam=(0, 0, 1)):
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(colidx, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        pam = np.ascontiguousarray(advanced_pam, np.int32)
        return _check(lib().xs_mat_setup(self.h, len(rp) - 1, _ptr(rp), _ptr(ci), _ptr(v),
                                         _ptr(pam)), "xs_mat_setup")

    def mat_setup_device(self, rowptr, colidx, val, use_ell=False):
        """rowptr/colidx (int32) and val (float64) are device tensors kept alive here."""
        self._keep = [rowptr, colidx, val]
        return _check(lib().xs_mat_setup_device(self.h, rowptr.numel() - 1, val.numel(),
                                                _ptr(rowptr), _ptr(colidx), _ptr(val),
                                                int(use_ell)), "xs_mat_setup_device")

    def ksp_settype(self, name):
        return _check(lib().xs_ksp_settype(self.h, name.encode()), "xs_ksp_settype")

    def ksp_set_absolute_tol(self, tol):
        return _check(lib().xs_ksp_set_absolute_tol(self.h, tol), "xs_ksp_set_absolute_tol")

    def ksp_setoption(self, tol, maxiter, restart=0):
        return _check(lib().xs_ksp_setoption(self.h, tol, maxiter, restart), "xs_ksp_setoption")

    def pc_settype(self, name):
        return _check(lib().xs_pc_settype(self.h, name.encode()), "xs_pc_settype")

    def pc_setilut(self, tol, max_fill):
        return _check(lib().xs_pc_setilut(self.h, tol, max_fill), "xs_pc_setilut")

    def set_option(self, name, value):
        return _check(lib().xs_set_option(self.h, name.encode(), float(value)), "xs_set_option")

    def autotune(self):
        """Pick the SpMV format / row-block size by timing; returns
        (chunk, ell, microseconds per SpMV)."""
        _check(lib().xs_autotune(self.h), "xs_autotune")
        return (int(self.time_ms("tune:chunk")), int(self.time_ms("tune:ell")),
                self.time_ms("tune:us"))

    def pc_setup(self):
        return _check(lib().xs_pc_setup(self.h), "xs_pc_setup")

    def solve(self, rhs, x0=None):
        """Host solve; returns (x, status)."""
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True)
        st = _check(lib().xs_solve(self.h, _ptr(b), _ptr(x)), "xs_solve")
        return x, st

    def solve_device(self, d_rhs, d_sol):
        return _check(lib().xs_solve_device(self.h, _ptr(d_rhs), _ptr(d_sol)), "xs_solve_device")

    # -- k right-hand sides -------------------------------------------------
    def solve_multi(self, rhs, x0=None):
        """Host block solve: rhs is (n, k); returns (X (n, k), status)."""
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True, order="C")
        st = _check(lib().xs_solve_multi(self.h, b.shape[1], _ptr(b), _ptr(x),
                                         XS_RHS_INTERLEAVED), "xs_solve_multi")
        return x, st

    def solve_multi_device(self, d_rhs, d_sol, layout=XS_RHS_INTERLEAVED):
        """Device block solve.  Interleaved: (n, k) row-major tensors;
        column-major: (k, n) row-major tensors (k vectors back to back)."""
        k = d_rhs.shape[1] if layout == XS_RHS_INTERLEAVED else d_rhs.shape[0]
        return _check(lib().xs_solve_multi_device(self.h, k, _ptr(d_rhs), _ptr(d_sol), layout),
                      "xs_solve_multi_device")

    def spmm(self, X, Y, stream=None):
        """Y = A X for (n, k) row-major device tensors."""
        return _check(lib().xs_spmm(self.h, X.shape[1], _ptr(X), _ptr(Y), stream), "xs_spmm")

    def pc_apply_multi(self, X, Y, stream=None):
        """Y = M^{-1} X for (n, k) row-major device tensors."""
        return _check(lib().xs_pc_apply_multi(self.h, X.shape[1], _ptr(X), _ptr(Y), stream),
                      "xs_pc_apply_multi")

    def multi_results(self):
        """(iterations, res0, resl) per column of the last block solve."""
        cap = MAX_RHS
        it = np.zeros(cap, np.int32)
        r0 = np.zeros(cap)
        rl = np.zeros(cap)
        k = _check(lib().xs_get_multi_results(self.h, _ptr(it), _ptr(r0), _ptr(rl), cap),
                   "xs_get_multi_results")
        return it[:k], r0[:k], rl[:k]

    def multi_resvec(self, col):
        it, _, _ = self.multi_results()
        out = np.zeros(int(it[col]) + 1)
        m = _check(lib().xs_get_multi_resvec(self.h, col, _ptr(out), len(out)),
                   "xs_get_multi_resvec")
        return out[:m]

    # -- queries ------------------------------------------------------------
    @property
    def iterations(self):
        return lib().xs_get_iterations(self.h)

    @property
    def residual(self):
        return lib().xs_get_residual(self.h)

    @property
    def initial_residual(self):
        return lib().xs_get_initial_residual(self.h)

    def resvec(self):
        n = self.iterations + 1
        out = np.zeros(max(n, 1))
        m = lib().xs_get_resvec(self.h, _ptr(out), len(out))
        return out[:m]

    def time_ms(self, phase):
        return lib().xs_get_time_ms(self.h, phase.encode())

    # -- kernel-level (device tensors) ---------------------------------------
    def spmv(self, x, y, stream=None):
        return _check(lib().xs_spmv(self.h, _ptr(x), _ptr(y), stream), "xs_spmv")

    def bmax(self, b, x, y, stream=None):
        return _check(lib().xs_bmax(self.h, _ptr(b), _ptr(x), _ptr(y), stream), "xs_bmax")

    def spmv_t(self, x, y, stream=None):
        return _check(lib().xs_spmv_t(self.h, _ptr(x), _ptr(y), stream), "xs_spmv_t")

    def pc_apply(self, x, y, stream=None):
        return _check(lib().xs_pc_apply(self.h, _ptr(x), _ptr(y), stream), "xs_pc_apply")

    def pc_apply_t(self, x, y, stream=None):
        return _check(lib().xs_pc_apply_t(self.h, _ptr(x), _ptr(y), stream), "xs_pc_apply_t")

    def trsv(self, which, b, x, stream=None):
        return _check(lib().xs_trsv(self.h, int(which), _ptr(b), _ptr(x), stream), "xs_trsv")

    def solve_transpose_device(self, d_rhs, d_sol):
        return _check(lib().xs_solve_transpose_device(self.h, _ptr(d_rhs), _ptr(d_sol)),
                      "xs_solve_transpose_device")

    def grad_values(self, d_lambda, d_x, d_grad, stream=None):
        return _check(lib().xs_grad_values_device(self.h, _ptr(d_lambda), _ptr(d_x), _ptr(d_grad),
                                                  stream), "xs_grad_values_device")

    def factors(self, n, cap):
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        dp = np.zeros(n, np.int32)
        nnz = _check(lib().xs_pc_get_factors(self.h, _ptr(rp), _ptr(ci), _ptr(v), _ptr(dp), cap),
                     "xs_pc_get_factors")
        return rp, ci[:nnz], v[:nnz], dp


class XBatch:
    """Many independent systems solved in one persistent launch
    (block-diagonal CSR, sys_ptr = row offsets of the systems)."""

    def __init__(self, sys_ptr, rowptr, colidx, val):
        self.sys_ptr = np.ascontiguousarray(sys_ptr, np.int32)
        rp = np.ascontiguousarray(rowptr, np.int32)
        ci = np.ascontiguousarray(colidx, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        h = ctypes.c_void_p()
        _check(lib().xs_batch_create(ctypes.byref(h), len(self.sys_ptr) - 1, _ptr(self.sys_ptr),
                                     _ptr(rp), _ptr(ci), _ptr(v)), "xs_batch_create")
        self.h = h
        self.nsys = len(self.sys_ptr) - 1

    def solve(self, rhs, tol=1e-8, maxiter=5000, precision=0, x0=None):
        b = np.ascontiguousarray(rhs, np.float64)
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True)
        st = _check(lib().xs_batch_solve(self.h, _ptr(b), _ptr(x), tol, maxiter, precision),
                    "xs_batch_solve")
        return x, st

    def solve_device(self, d_rhs, d_sol, tol=1e-8, maxiter=5000, precision=0):
        return _check(lib().xs_batch_solve_device(self.h, _ptr(d_rhs), _ptr(d_sol), tol, maxiter,
                                                  precision), "xs_batch_solve_device")

    def results(self):
        it = np.zeros(self.nsys, np.int32)
        rr = np.zeros(self.nsys)
        lib().xs_batch_get_results(self.h, _ptr(it), _ptr(rr))
        return it, rr

    @property
    def time_ms(self):
        return lib().xs_batch_get_time_ms(self.h)

    def close(self):
        if self.h:
            lib().xs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


XS_MTX_NEGATE_GENERAL = 1


def read_mtx(path, negate_general=False):
    """Matrix Market file -> (rowptr, colidx, val, diag), 0-based CSR with
    sorted columns (xs_mtx_read; host only).  negate_general reproduces the
    reference reader's sign flip for general files."""
    L = lib()
    flags = XS_MTX_NEGATE_GENERAL if negate_general else 0
    n = ctypes.c_int(0)
    nnz = ctypes.c_int(0)
    p = str(path).encode()
    _check(L.xs_mtx_read(p, flags, ctypes.byref(n), ctypes.byref(nnz), None, None, None, None),
           "xs_mtx_read")
    rp = np.zeros(n.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    v = np.zeros(max(nnz.value, 1))
    d = np.zeros(n.value)
    _check(L.xs_mtx_read(p, flags, ctypes.byref(n), ctypes.byref(nnz), _ptr(rp), _ptr(ci),
                         _ptr(v), _ptr(d)), "xs_mtx_read")
    return rp, ci[:nnz.value], v[:nnz.value], d


def batch_concat(systems):
    """Stack CSR systems into one block-diagonal CSR; returns (sys_ptr, rp, ci, v)."""
    sys_ptr = [0]
    rps, cis, vs = [], [], []
    nnz_off = 0
    for rp, ci, v in systems:
        n = len(rp) - 1
        rps.append(np.asarray(rp[:-1], np.int64) + nnz_off)
        cis.append(np.asarray(ci, np.int64) + sys_ptr[-1])
        vs.append(np.asarray(v, np.float64))
        nnz_off += len(v)
        sys_ptr.append(sys_ptr[-1] + n)
    rp = np.concatenate(rps + [np.array([nnz_off])]).astype(np.int32)
    return (np.array(sys_ptr, np.int32), rp, np.concatenate(cis).astype(np.int32),
            np.concatenate(vs))


def legacy():
    """The reference-style global API functions (ctypes)."""
    return lib()
