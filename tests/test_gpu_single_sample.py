"""GPU parity of the single-sample CCE batch (pure bath state, Hilbert space;
source/app/single_sample_cce.cpp:48-128) against the oracle's dense
restatement: every cluster of orders 1-4 of a small bath (Hilbert dims 2-16,
the dense path) in FP64 to 1e-10 under a Hahn echo and CPMG-2/4, and 5-spin
clusters (n = 32) on the flip-mask Lanczos kernel in FP64 (1e-7 at tol 1e-12)
and mixed precision (1e-6).  The CCE product of the batch equals the host
reduction of the oracle's L_c(t)."""
import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _oracle(X, C, states, order_max, npulse, tl):
    nv = P.NVCenter((0.0, 0.0, 0.0), B111)
    out = []
    for o in range(order_max):
        for cl in C.get(o):
            prob = P.single_sample_problem(X, list(cl), B111, nv, states)
            out.append(P.single_sample_dense(prob, tl, npulse))
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [1, 2, 4])
def test_single_sample_batch_matches_oracle(ctx, npulse):
    X = O.bath_diamond(3.567, 14.0, 0.011, seed=3)
    C = O.Clusters(X, 8.0, 4)
    states = O.bath_states(11, len(X))
    tl = np.linspace(0.0, 2e-4, 41)
    print(f"single-sample N={npulse}: clusters {C.counts()}", flush=True)
    b = C.single_sample_batch(ctx, O.cce_params(B=B111), states, npulse, tl, m=30, tol=1e-12, precision="fp64")
    b.run()
    Ls, st = b.results()
    L = np.stack(Ls)
    assert ((st[:, 8] & 0xFF) == 0).all()
    assert {t["kernel"] for t in b.timing()} == {"dense_pade"}
    want = _oracle(X, C, states, 4, npulse, tl)
    err = np.abs(L - want).max()
    print(f"  max |L - oracle| {err:.2e}; L(t_end) range {L[:, -1].min():.4f}..{L[:, -1].max():.4f}", flush=True)
    assert err <= 1e-10
    assert np.ptp(want) > 1e-3
    # the CCE reduction of the device L equals that of the oracle's
    got = C.reduce(L.reshape(-1), len(tl))["final"]
    ref = C.reduce(want.reshape(-1), len(tl))["final"]
    assert np.abs(got - ref).max() <= 1e-9


def _local_bath():
    """The 9 spins nearest to the centre of one 5-spin cluster of a dense bath:
    clusters [9, 17, 35, 62, 78] of orders 1-5 at cutoff 8 A."""
    X = O.bath_diamond(3.567, 14.0, 0.05, seed=2)
    C = O.Clusters(X, 8.0, 5)
    c0 = X[list(C.get(4)[0])].mean(axis=0)
    keep = np.sort(np.argsort(np.linalg.norm(X - c0, axis=1))[:9])
    return X[keep]


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [1, 2])
def test_single_sample_five_spin_clusters_on_the_lanczos_kernel(ctx, npulse):
    X = _local_bath()
    C = O.Clusters(X, 8.0, 5)
    assert C.count(4) > 0
    states = O.bath_states(5, len(X))
    tl = np.linspace(0.0, 1e-4, 21)
    want = _oracle(X, C, states, 5, npulse, tl)
    print(f"single-sample 5 spins N={npulse}: clusters {C.counts()}", flush=True)
    # FP64 with the mixed path's Lanczos / Chebyshev scheme at tol 1e-12, and
    # mixed precision at the bar of the mixed path
    for precision, tol, bar in (("fp64-lanczos", 1e-12, 1e-7), ("mixed", 1e-4, 1e-6)):
        b = C.single_sample_batch(ctx, O.cce_params(B=B111), states, npulse, tl, m=30, tol=tol, precision=precision)
        b.run()
        Ls, st = b.results()
        L = np.stack(Ls)
        kinds = {t["n"]: t["kernel"] for t in b.timing()}
        err = np.abs(L - want).max()
        print(f"  {precision}: kernels {kinds}, max |L - oracle| {err:.2e}", flush=True)
        assert ((st[:, 8] & 0xFF) == 0).all()
        assert kinds[32] == ("lanczos_slot_fp64" if precision == "fp64-lanczos" else "lanczos_slot")
        assert err <= bar
