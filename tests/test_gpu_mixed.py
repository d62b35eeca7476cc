"""GPU: the mixed-precision batched path against the FP64 path.

Mixed precision (complex FP32 couplings and basis, FP64 recurrence,
re-orthogonalisation and small exponential) must give every L_c(t) within
1e-6 absolute of the FP64 path (the FP64 path is itself pinned to the oracle
at 1e-10 in test_gpu_hahn.py).  Hermitian generators run the Lanczos kernel,
non-Hermitian ones (dephasing) the mixed Arnoldi kernel; both are covered."""
import numpy as np
import pytest

import oops_amd as O

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)
ABS_TOL = 1e-6


def _sample(C, per):
    counts = C.counts()
    sel, base = [], 0
    for c in counts:
        sel += list(base + np.linspace(0, c - 1, min(per, c)).astype(np.int64))
        base += c
    return np.array(sorted(set(sel)), np.int64)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,m,tol", [("mixed", 20, 3e-4), ("mixed", 30, 1e-6), ("mixed-arnoldi", 20, 3e-4)])
def test_mixed_matches_fp64(ctx, precision, m, tol):
    X = O.bath_diamond(3.567, 30.0, 0.011, seed=5)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=B111)
    sel = _sample(C, 24)
    tau = np.linspace(0.0, 2e-3, 81) / 2
    print(f"{precision} m={m} tol={tol} clusters={len(sel)}", flush=True)
    ref = C.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64", select=sel)
    ref.run(tau)
    L64, st64 = ref.results()
    b = C.hahn_batch(ctx, par, m=m, tol=tol, precision=precision, select=sel)
    b.run(tau)
    L, st = b.results()
    assert (st64[:, 8] & 0xFF).max() == 0 and (st[:, 8] & 0xFF).max() == 0
    err = np.abs(L - L64).max()
    print(f"max |dL| = {err:.3e}", flush=True)
    assert err <= ABS_TOL
    assert np.all(np.abs(L[:, 0] - 1.0) < 1e-6)


@pytest.mark.gpu
def test_mixed_with_dephasing_uses_arnoldi_and_matches_fp64(ctx):
    X = O.bath_diamond(3.567, 25.0, 0.011, seed=6)
    C = O.Clusters(X, 8.0, 3)
    par = O.cce_params(B=B111, dephasing_rate=300.0)
    sel = _sample(C, 16)
    tau = np.linspace(0.0, 1e-3, 41) / 2
    ref = C.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64", select=sel)
    ref.run(tau)
    L64, _ = ref.results()
    b = C.hahn_batch(ctx, par, m=20, tol=3e-4, precision="mixed", select=sel)
    b.run(tau)
    L, st = b.results()
    assert (st[:, 8] & 0xFF).max() == 0
    assert np.abs(L - L64).max() <= ABS_TOL
