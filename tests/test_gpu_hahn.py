"""GPU parity of the batched Hahn-echo Krylov expv against the oracle.

The oracle (oracle/expv_oracle.cpp, a CPU restatement of the reference's
MAIN_MKL / HAMVEC_ZGEXPV_W_MKL_PROFILE Krylov driver) runs the same two
sweeps per cluster on the same term tables; the FP64 device path must agree
to 1e-10 relative (tolerance of the FP64 path, stated by the workload)."""
import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)
REL_TOL = 1e-10


def _problems(X, clusters, B=B111):
    par = O.cce_params(B=B)
    items, oA, oB, rhos, Ds = [], [], [], [], []
    for cl in clusters:
        c = X[list(cl)]
        A, rho, D = O.cluster_terms(c, par, 0)
        Bd, _, _ = O.cluster_terms(c, par, 2)
        items.append((A, Bd, rho, rho, D))
        oA.append(P.TermTable(len(cl), A.coeff, A.nbody, A.pos_offset, A.pos, A.dims, A.mat_offset, A.mats,
                              A.spin_dim, A.n))
        oB.append(P.TermTable(len(cl), Bd.coeff, Bd.nbody, Bd.pos_offset, Bd.pos, Bd.dims, Bd.mat_offset,
                              Bd.mats, Bd.spin_dim, Bd.n))
        rhos.append(rho)
        Ds.append(D)
    return items, oA, oB, rhos, Ds


@pytest.mark.gpu
@pytest.mark.parametrize("m", [30, 20])
def test_hahn_batch_matches_oracle(ctx, m):
    X = O.bath_diamond(3.567, 20.0, 0.011, seed=1)
    C = O.Clusters(X, 8.0, 4)
    picks = []
    for order in range(4):
        cl = C.get(order)
        picks += [tuple(r) for r in cl[: 3]]
    items, oA, oB, rhos, Ds = _problems(X, picks)
    tlist = np.linspace(0.0, 4e-5, 9)
    tau = tlist / 2
    batch = O.HahnBatch.from_items(ctx, items, m=m)
    batch.run(tau)
    L, st = batch.results()
    assert (st[:, 8] & 0xF).max() == 0 and (st[:, 8] >> 4 & 0xF).max() == 0
    ref = np.zeros_like(L)
    for i in range(len(items)):
        n = oA[i].n
        if n > m:
            Lr, _, rc = P.hahn_batch([oA[i]], [oB[i]], [rhos[i]], [rhos[i]], [Ds[i]], tau, m=m, nthreads=1)
            assert rc == 0
            ref[i] = Lr[0]
        else:
            # n <= m: exact dense propagation (the reference stops there)
            import scipy.linalg as sla
            A = P.full_from_table(oA[i]) if hasattr(P, "full_from_table") else None
            Ad = _dense(oA[i])
            Bd = _dense(oB[i])
            for k, t in enumerate(tau):
                u = sla.expm(-1j * Ad * t) @ rhos[i]
                y = sla.expm(-1j * Bd * t) @ rhos[i]
                ref[i, k] = np.real(Ds[i] * np.vdot(y, u))
    err = np.abs(L - ref) / np.maximum(np.abs(ref), 1e-300)
    assert err.max() < REL_TOL, (err.max(), np.unravel_index(err.argmax(), err.shape))


def _dense(t):
    n = t.n
    M = np.zeros((n, n), complex)
    for j in range(n):
        e = np.zeros(n, complex)
        e[j] = 1.0
        M[:, j] = P.hamvec(t, e)
    return M
