"""Build libxsolver_amd.so (gfx950) from xsolver_amd/csrc with hipcc.

Usage: python -m xsolver_amd.build   (or __graft_entry__.build())
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxsolver_amd.so")
ARCH = os.environ.get("XS_OFFLOAD_ARCH", "gfx950")
SOURCES = ["xs_kernels.hip", "xs_krylov.hip", "xs_parilut.hip", "xs_split.hip", "xs_ic.hip", "xs_intel.hip", "xs_mixed.hip", "xs_batch.hip", "xs_multi.hip", "xs_host.cpp", "xs_solvers.cpp", "xs_batch_host.cpp", "xs_multi_host.cpp", "xs_io.cpp", "xs_mcsr.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed-FP32 instructions are disabled for all device code (-packed-fp32-ops).
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
         "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", "hip", *FLAGS, "-I", os.path.join(HERE, "..", "include"),
           "-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msgs}")
    return out, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, msgs in results:
        if msgs and verbose:
            print(msgs, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
