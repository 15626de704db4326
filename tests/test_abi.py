"""The C-ABI library builds for gfx950, loads, and exports every function that
include/xsolver_c.h declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

import xsolver_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xsolver_c.h")


def declared_functions():
    text = open(HEADER).read()
    text = text.split("#ifdef __cplusplus\n} /* extern \"C\" */")[0]
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M)
    return sorted({n for n in names if n.startswith(("xs_", "xsolver_"))})


def test_library_built():
    assert os.path.exists(xsolver_amd.LIB_PATH), "run python -m xsolver_amd.build"


def test_header_declares_reference_api():
    names = declared_functions()
    for ref in ["xsolver_mat_setup", "xsolver_ksp_settype", "xsolver_ksp_set_absolute_tol",
                "xsolver_ksp_setoption", "xsolver_pc_settype", "xsolver_pc_setilut",
                "xsolver_pc_setict", "xsolver_solve", "xsolver_free"]:
        assert ref in names


def test_every_declared_symbol_is_exported():
    lib = xsolver_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(xsolver_amd.EXPORTS)


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", xsolver_amd.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if " T " in l}
    for n in declared_functions():
        assert n in syms, n  # unmangled: real C linkage


def test_code_object_is_gfx950():
    # the embedded device code object is built for gfx950 only
    blob = open(xsolver_amd.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"amdgcn-amd-amdhsa--gfx942" not in blob


def test_status_strings_without_device():
    lib = xsolver_amd.lib()
    assert lib.xs_status_string(xsolver_amd.XS_ERR_UNSUPPORTED) == b"unsupported"
    assert lib.xs_status_string(xsolver_amd.XS_NOT_CONVERGED) == b"maximum iterations reached"


def test_cpp_header_compiles_with_reference_signature(tmp_path):
    # Code written against the reference header (int &usediter, default restart)
    # compiles against include/xsolver_c.h and links to the library.
    src = tmp_path / "caller.cpp"
    src.write_text(
        '#include "xsolver_c.h"\n'
        'int main(){ int n=1,rp[2]={0,1},ci[1]={0}; double v[1]={2.0}; int pam[3]={0,0,1};\n'
        ' int used=0; double res[11]; char cg[]="cg", jac[]="Jacobi";\n'
        ' xsolver_mat_setup(n,rp,ci,v,pam); xsolver_ksp_settype(cg);\n'
        ' xsolver_ksp_setoption(1e-8,10,used,res); xsolver_pc_settype(jac);\n'
        ' double b[1]={1.0}, x[1]={0.0}; xsolver_solve(b,x); xsolver_free();\n'
        ' return xsolver_last_error()==0 ? 0 : 1; }\n')
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"), str(src),
                    xsolver_amd.LIB_PATH, "-o", str(exe)], check=True)
    # the CUDA backend's Fortran-style surface, called as the reference's
    # C++ declarations allow (references, default resvec)
    src2 = tmp_path / "caller_mcsr.cpp"
    src2.write_text(
        '#include "xsolver_c.h"\n'
        'int main(){ int n=2,h=0,ns=2,ptr[3]={0,1,2},col[2]={1,0},nz=7,used=0,maxit=10,rs=5;\n'
        ' double d[2]={4.0,4.0},off[2]={-1.0,-1.0},b[2]={1.0,1.0},x[2]={0,0},tol=1e-10;\n'
        ' char jac[]="Jacobi", ell[]="ELL";\n'
        ' cpu_mat_setup(n,h,ns,d,off,ptr,col); gpu_mat_setup(); gpu_mat_format(ell,nz);\n'
        ' gpu_preconditioner_setup(jac); gpu_cg_solve(b,x,tol,maxit,NULL,used);\n'
        ' gpu_gcr_solve(b,x,rs,tol,maxit,x,used); gpu_mat_free();\n'
        ' return gpu_last_error()==0 ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"), str(src2),
                    xsolver_amd.LIB_PATH, "-o", str(tmp_path / "caller_mcsr")], check=True)
    src_c = tmp_path / "caller.c"
    src_c.write_text('#include "xsolver_c.h"\nint main(void){ xs_ctx *c = 0; int rc = xs_create(&c);\n'
                     ' if (rc == 0) xs_destroy(c); return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src_c),
                    xsolver_amd.LIB_PATH, "-o", str(tmp_path / "caller_c")], check=True)


def test_no_device_fails_loudly_on_cpu():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    h = ctypes.c_void_p()
    rc = xsolver_amd.lib().xs_create(ctypes.byref(h))
    assert rc == xsolver_amd.XS_ERR_DEVICE
    with pytest.raises(xsolver_amd.XSolverError):
        xsolver_amd.XSolver()
