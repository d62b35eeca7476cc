"""GPU: factored propagator powers (oops_amd/csrc/cce_factored.hip) for the
4-spin clusters of unpolarised CCE batches on uniform grids.

Without dephasing exp(-i lvA s) rho = e^{i H1 s} rho e^{-i H0 s}, so the
reference's dense Liouville propagator (PiecewiseFullMatrixVectorEvolution,
QuantumEvolutionAlgorithm.cpp:60-83) factors into two 16 x 16 Hilbert
propagators.  Checks:

* CPMG-N (N = 2, 3, 4) on 4-spin clusters sampled from a bath against the
  oracle's dense 256 x 256 piecewise propagator of riffle(lvA, lvB, N)
  (ensemble_cce.cpp:33-55, misc.cpp:59-84) to 1e-10;
* N = 8, 16 on 41 points and N = 16 on the 401-point 0-5 ms grid against the
  oracle's Hilbert-space chains (single_sample_dense, riffle(H0, H1, N) and
  riffle(H1, H0, N)) averaged over the 16 basis states of the cluster, which
  for an unpolarised cluster is the same coherence (the first check pins
  that identity on the Liouville propagator) to 1e-10 / 1e-9;
* the same batch with OOPS_FACTORED=0 (flip-mask Lanczos, tol 1e-6) within
  1e-6 of the factored result;
* the Hahn echo (OOPS_FACTORED=all) equal to the FP64 Hahn batch (the
  reference's Arnoldi algorithm) to 1e-10;
* L(0) = 1 and |L| <= 1 on the 401-point 0-5 ms grid for N = 16.
"""
import os

import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _bath():
    X = O.bath_diamond(3.567, 22.0, 0.02, seed=4)
    C = O.Clusters(X, 7.0, 4)
    return X, C


def _oracle(coords, npulse, t):
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    n = int(np.prod(prob["ldims"]))
    A = P.full(prob["lvA"], n)
    Bm = P.full(prob["lvB"], n)
    st = P.piecewise_dense(P.riffle(A, Bm, npulse), P.pulse_interval(npulse), t[1] - t[0], prob["rho"], len(t))
    return np.real(prob["D"] * (prob["rho"].conj() @ st))


def _oracle_hilbert(coords, npulse, t):
    """(1/16) sum over the basis states of Re <psi_1(t), psi_2(t)>."""
    nv = P.NVCenter((0, 0, 0), B111)
    k = len(coords)
    out = np.zeros(len(t))
    for j in range(2 ** k):
        states = [(j >> (k - 1 - s)) & 1 for s in range(k)]
        prob = P.single_sample_problem(coords, list(range(k)), B111, nv, states)
        out += P.single_sample_dense(prob, t, npulse)
    return out / 2 ** k


def _run(ctx, C, npulse, t, mode, precision="mixed", tol=1e-6):
    old = os.environ.get("OOPS_FACTORED")
    if mode is None:
        os.environ.pop("OOPS_FACTORED", None)
    else:
        os.environ["OOPS_FACTORED"] = mode
    try:
        b = C.echo_batch(ctx, O.cce_params(B=B111), npulse, t, m=30, tol=tol, precision=precision)
        b.run()
        Ls, st = b.results()
        kinds = {x["n"]: x["kernel"] for x in b.timing()}
    finally:
        if old is None:
            os.environ.pop("OOPS_FACTORED", None)
        else:
            os.environ["OOPS_FACTORED"] = old
    return np.stack(Ls), st, kinds


def _check(ctx, npulse, t, oracle, nclusters, bar):
    X, C = _bath()
    counts = C.counts()
    assert counts[3] >= nclusters
    L, st, kinds = _run(ctx, C, npulse, t, None)
    print(f"factored CPMG-{npulse}, {len(t)} points: clusters {counts}, kernels {kinds}", flush=True)
    assert kinds[256] == "factored_powers"
    assert ((st[:, 8] & 0xFF) == 0).all()
    base = sum(counts[:3])
    four = C.get(3)
    worst = 0.0
    for i in np.linspace(0, len(four) - 1, nclusters).astype(int):
        want = oracle(X[list(four[i])], npulse, t)
        assert np.ptp(want) > 1e-4
        err = np.abs(L[base + i] - want).max()
        worst = max(worst, err)
        assert err <= bar, (i, err)
    print(f"  max |L - oracle| {worst:.2e}", flush=True)
    return L[base:]


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [2, 3, 4])
def test_factored_cpmg_matches_reference_liouville_propagator(ctx, npulse):
    _check(ctx, npulse, np.linspace(0.0, 8e-4, 9), _oracle, 2, 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [8, 16])
def test_factored_cpmg_matches_hilbert_chains(ctx, npulse):
    _check(ctx, npulse, np.linspace(0.0, 1e-3, 41), _oracle_hilbert, 4, 1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [2, 8])
def test_factored_equals_lanczos_path(ctx, npulse):
    X, C = _bath()
    t = np.linspace(0.0, 6e-4, 25)
    Lf, sf, kf = _run(ctx, C, npulse, t, None)
    Ll, sl, kl = _run(ctx, C, npulse, t, "0")
    assert kf[256] == "factored_powers" and kl[256] == "lanczos_slot"
    assert ((sl[:, 8] & 0xFF) == 0).all()
    base = sum(C.counts()[:3])
    err = np.abs(Lf[base:] - Ll[base:]).max()
    print(f"CPMG-{npulse}: factored vs Lanczos max |dL| {err:.2e}", flush=True)
    assert err <= 1e-6
    assert np.array_equal(Lf[:base], Ll[:base])    # the other buckets are untouched


@pytest.mark.gpu
def test_factored_hahn_equals_fp64_hahn_batch(ctx):
    X, C = _bath()
    t = np.linspace(0.0, 1e-3, 41)
    Lf, sf, kf = _run(ctx, C, 1, t / 2.0, "all")
    assert kf[256] == "factored_powers"
    ref = C.hahn_batch(ctx, O.cce_params(B=B111), m=30, tol=1e-12, precision="fp64")
    ref.run(t / 2.0)
    L64, st64 = ref.results()
    base = sum(C.counts()[:3])
    err = np.abs(Lf[base:] - L64[base:]).max()
    print(f"Hahn: factored vs FP64 Arnoldi max |dL| {err:.2e}", flush=True)
    assert err <= 1e-10


@pytest.mark.gpu
def test_factored_long_grid(ctx):
    t = np.linspace(0.0, 5e-3, 401)
    four = _check(ctx, 16, t, _oracle_hilbert, 2, 1e-9)
    assert np.all(np.abs(four[:, 0] - 1.0) <= 1e-13)
    assert np.all(np.abs(four) <= 1.0 + 1e-9)
