"""The reference's C++ class interfaces (include/oops_classes.hpp: KronProd,
SumKronProd, MatExpVector, cClusterIndex, cSpinCollection,
cDepthFirstPathTracing, cSpinCluster) through a C++ program compiled against
the header and linked with liboops_amd.so (tests/cpp/test_classes.cpp).

* host: the Kronecker superoperators against their defining identities
  (Flat(m) vec(rho) = vec(m rho), Sharp(m) vec(rho) = vec(rho m)), KronProd
  element by element, and the cluster classes on the reference's RoyCoord.xyz
  (counts of orders 1-4 at 6 and 8 A; sub-clusters against brute force);
* gpu: MatExpVector with every method and constructor against the scipy
  expm fixtures (tests/golden/expv_cases.npz) to 1e-10.

Reference: include/math/MatExp.h:45-87, include/kron/KronProd.h:32-104,
include/spin/SpinCluster.h:31-68, include/spin/SpinClusterAlgorithm.h:68-82."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_classes")
GOLD = os.path.join(ROOT, "tests", "golden")


def _binary():
    if not os.path.exists(BIN):
        raise RuntimeError(f"{BIN} missing: run __graft_entry__.build() (tests/cpp/build_cpp_tests.py)")
    return BIN


def test_class_layer_host_checks():
    r = subprocess.run([_binary(), "host", os.path.join(GOLD, "RoyCoord.xyz")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout


def _write_cases(path):
    g = np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)
    with open(path, "wb") as f:
        nc = int(g["ncases"])
        f.write(struct.pack("<i", nc))
        for c in range(nc):
            p = f"c{c}_"
            sd = g[p + "spin_dim"]
            coeff, nbody, poff = g[p + "coeff"], g[p + "nbody"], g[p + "pos_off"]
            pos, dims, moff, mats = g[p + "pos"], g[p + "dims"], g[p + "mat_off"], g[p + "mats"]
            t, v, w = g[p + "t"], g[p + "v"], g[p + "w"]
            f.write(struct.pack("<iii", len(sd), len(coeff), len(t)))
            f.write(np.asarray(sd, "<i8").tobytes())
            for k in range(len(coeff)):
                nb = int(nbody[k])
                b0 = int(poff[k])
                f.write(struct.pack("<di", float(coeff[k]), nb))
                f.write(np.asarray(pos[b0:b0 + nb], "<i8").tobytes())
                f.write(np.asarray(dims[b0:b0 + nb], "<i8").tobytes())
                for b in range(b0, b0 + nb):
                    f.write(np.asarray(mats[moff[b]:moff[b + 1]], "<c16").tobytes())
            f.write(np.asarray(v, "<c16").tobytes())
            f.write(np.asarray(t, "<f8").tobytes())
            f.write(np.asfortranarray(np.asarray(w, "<c16")).tobytes(order="F"))


@pytest.mark.gpu
def test_matexpvector_against_scipy_fixtures(tmp_path):
    cases = str(tmp_path / "cases.bin")
    _write_cases(cases)
    print("MatExpVector through the class layer", flush=True)
    r = subprocess.run([_binary(), "gpu", cases], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all ok" in r.stdout and r.stdout.count("vs scipy expm") == 12
