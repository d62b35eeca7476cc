"""GPU: the device-side term builder (oops_amd/csrc/term_builder.hip) against
the host compiler it replaces on the batch path.

* the packed tables of every generator (lvA, -lvB^dagger, lvB, -lvA^dagger)
  of clusters of 1-4 spins are identical, field by field and bit for bit:
  local operators, ELL values, offsets, flip-mask slots and masks; clusters
  from the 5 nm bath, lattice geometries whose dipolar tensors have exact
  zeros (pairs along a cube edge and a face diagonal), and with dephasing;
* batches built on the device give the same L(t) as batches built on the
  host (OOPS_TERM_BUILDER=host), FP64 and mixed, Hahn echo and CPMG-N.

Reference of the terms: source/misc/misc.cpp:11-57,
source/app/ensemble_cce.cpp:41-91, source/kron/KronProd.cpp:9-107."""
import numpy as np
import pytest

import oops_amd as O

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)
A = 3.567


def _geometries():
    X = O.bath_diamond(A, 50.0, 0.011, seed=0)
    C = O.Clusters(X, 8.0, 4)
    out = []
    for o in range(4):
        cl = C.get(o)
        for i in np.linspace(0, len(cl) - 1, 6).astype(int):
            out.append(X[list(cl[i])])
    base = np.array([A * 1.0, A * 1.0, 0.0])
    out.append(np.array([base, base + [0.0, 0.0, A]]))                      # along z: J_xz = J_yz = 0
    out.append(np.array([base, base + [A / 2, A / 2, 0.0]]))                # face diagonal: nz = 0
    out.append(np.array([base, base + [0.0, 0.0, A], base + [A / 2, A / 2, 0.0],
                         base + [A / 4, A / 4, A / 4]]))
    return out


@pytest.mark.gpu
def test_device_tables_equal_host_tables(ctx):
    geos = _geometries()
    params = [O.cce_params(B=B111), O.cce_params(B=np.array([0.0, 0.0, 0.2])),
              O.cce_params(B=B111, dephasing_rate=2e3, dephasing_axis=(0.3, 0.1, 1.0))]
    nops = 0
    for par in params:
        for X in geos:
            for which in range(4):
                h = O.op_tables(ctx, X, par, which, device=False)
                d = O.op_tables(ctx, X, par, which, device=True)
                for key in ("counts", "locs", "off", "eslot", "masks"):
                    assert np.array_equal(h[key], d[key]), (len(X), which, key)
                assert np.array_equal(h["val"].view(np.float64), d["val"].view(np.float64)), (len(X), which)
                nops += 1
    print(f"{nops} generators: device tables identical to the host compiler's", flush=True)


def _batches(ctx, C, par, precision, tol, npulse, t):
    import os
    out = {}
    # both builds on the Krylov kernels (the factored CPMG path takes no tables)
    os.environ["OOPS_FACTORED"] = "0"
    for mode in ("host", "device"):
        os.environ["OOPS_TERM_BUILDER"] = mode
        try:
            if npulse == 1 and t is None:
                b = C.hahn_batch(ctx, par, m=30, tol=tol, precision=precision)
                b.run(np.linspace(0.0, 1e-3, 41) / 2.0)
                L, st = b.results()
            else:
                b = C.echo_batch(ctx, par, npulse, t, m=30, tol=tol, precision=precision)
                b.run()
                Ls, st = b.results()
                L = np.stack(Ls)
            out[mode] = (L, st)
        finally:
            os.environ.pop("OOPS_TERM_BUILDER", None)
    os.environ.pop("OOPS_FACTORED", None)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol,npulse", [("fp64", 1e-12, 1), ("mixed", 3e-3, 1), ("mixed", 1e-4, 4)])
def test_device_built_batch_equals_host_built_batch(ctx, precision, tol, npulse):
    X = O.bath_diamond(A, 25.0, 0.011, seed=5)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=B111)
    t = None if npulse == 1 else np.linspace(0.0, 1e-3, 21)
    print(f"{precision} N={npulse}: clusters {C.counts()}", flush=True)
    r = _batches(ctx, C, par, precision, tol, npulse, t)
    (Lh, sh), (Ld, sd) = r["host"], r["device"]
    assert ((sh[:, 8] & 0xFF) == 0).all()
    assert np.array_equal(sh[:, :8], sd[:, :8])      # the same steps, products and rejections
    assert np.array_equal(Lh, Ld)
