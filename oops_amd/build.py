"""Build liboops_amd.so (HIP kernels + C-ABI runtime) in-tree for gfx950.

Every HIP translation unit is compiled with hipcc for gfx950 and packed-FP32
instruction generation switched off.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liboops_amd.so")
ARCH = os.environ.get("OOPS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["cce_krylov.hip", "cce_lanczos.hip", "cce_dense64.hip", "cce_medium.hip", "expv_large.hip", "term_builder.hip", "cce_factored.hip",
           "capi.cpp", "cce_host.cpp", "large_host.cpp", "serve.cpp"]
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True, jobs=4):
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmds.append(["hipcc"] + HIPFLAGS + lang + ["-c", path, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
