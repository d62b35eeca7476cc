"""Build the C++ test programs of tests/cpp (run by __graft_entry__.build()).

test_classes: compiled with g++ against include/oops_classes.hpp and linked
with oops_amd/liboops_amd.so (found at run time through its rpath)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ROCM_LIB = os.environ.get("ROCM_PATH", "/opt/rocm") + "/lib"


def build(verbose=True):
    out = os.path.join(HERE, "test_classes")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", os.path.join(HERE, "test_classes.cpp"), "-o", out + ".tmp",
           "-L" + os.path.join(ROOT, "oops_amd"), "-loops_amd", "-Wl,-rpath,$ORIGIN/../../oops_amd",
           "-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
