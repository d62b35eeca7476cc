"""Build the oracle (test infrastructure only).

* liborc_expv.so   — the C++ restatement in expv_oracle.cpp (g++).
* _ref/libdlaran_ref.so — the reference's own DLARAN.f compiled from its
  source where it lies under /root/reference (amdflang), used only to pin
  orc_dlaran.  Built only when /root/reference is present; never committed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DLARAN = "/root/reference/src/expv/source/expokit/DLARAN.f"


def build(verbose=True):
    out = os.path.join(HERE, "liborc_expv.so")
    cmd = ["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fPIC", "-shared", "-pthread",
           os.path.join(HERE, "expv_oracle.cpp"), "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    flang = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if os.path.exists(REF_DLARAN) and flang:
        os.makedirs(os.path.join(HERE, "_ref"), exist_ok=True)
        ref = os.path.join(HERE, "_ref", "libdlaran_ref.so")
        cmd = [flang, "-O2", "-fPIC", "-shared", REF_DLARAN, "-o", ref]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)


if __name__ == "__main__":
    build()
    sys.exit(0)
