"""GPU: pulse sequences with per-item time grids (oops_echo_batch_*).

CPMG-N (Pulse_Interval / riffle, misc.cpp:59-84) against the oracle's
restatement of the reference's production dense propagator
(PiecewiseFullMatrixVectorEvolution, QuantumEvolutionAlgorithm.cpp:60-83):
the dense paths (n <= m VALU kernel, n = 64 FP64-MFMA kernel on uniform
grids) to 1e-10, the mixed Lanczos path (n = 256, or non-uniform grids) to
1e-6.
Items of different sizes, pulse counts and time grids give the same L(t)
alone and mixed in one batch (the kernels compute each item independently),
and an echo batch with npulse = 1 equals the Hahn batch on the same times."""
import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _clusters(k, count, seed=3):
    X = O.bath_diamond(3.567, 25.0, 0.02, seed=seed)
    C = O.Clusters(X, 7.0, k)
    cl = C.get(k - 1)
    return [X[list(r)] for r in cl[np.linspace(0, len(cl) - 1, count).astype(int)]]


def _item(coords, npulse, t):
    par = O.cce_params(B=B111)
    A, rho, D = O.cluster_terms(coords, par, 0)
    Bd, _, _ = O.cluster_terms(coords, par, 2)
    it = {"opA": A, "opB": Bd, "u0": rho, "y0": rho, "scale": D, "npulse": npulse, "t": t}
    if npulse > 1:
        it["opC"], _, _ = O.cluster_terms(coords, par, 1)
        it["opD"], _, _ = O.cluster_terms(coords, par, 5)
    return it


def _oracle_cpmg(coords, npulse, t):
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    n = int(np.prod(prob["ldims"]))
    A = P.full(prob["lvA"], n)
    Bm = P.full(prob["lvB"], n)
    ops = P.riffle(A, Bm, npulse)
    seg = P.pulse_interval(npulse)
    st = P.piecewise_dense(ops, seg, t[1] - t[0], prob["rho"], len(t))
    return np.real(prob["D"] * (prob["rho"].conj() @ st))


@pytest.mark.gpu
@pytest.mark.parametrize("k,npulse", [(1, 2), (2, 4), (2, 16), (1, 8)])
def test_cpmg_dense_path_matches_reference_propagator(ctx, k, npulse):
    t = np.linspace(0.0, 1e-3, 41)
    for coords in _clusters(k, 3):
        print(f"dense cpmg k={k} N={npulse}", flush=True)
        b = O.EchoBatch.from_items(ctx, [_item(coords, npulse, t)], m=30, precision="fp64")
        b.run()
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        assert np.abs(L[0] - _oracle_cpmg(coords, npulse, t)).max() < 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [1, 2, 4, 16])
def test_dense64_mfma_path_matches_reference_propagator(ctx, npulse):
    t = np.linspace(0.0, 1e-3, 33)
    for coords in _clusters(3, 3):
        print(f"dense64 k=3 N={npulse}", flush=True)
        it = _item(coords, npulse, t / 2 if npulse == 1 else t)
        for prec in ("fp64", "mixed"):
            b = O.EchoBatch.from_items(ctx, [it], m=30, precision=prec)
            b.run()
            L, st = b.results()
            assert (st[:, 8] & 0xFF).max() == 0
            assert np.abs(L[0] - _oracle_cpmg(coords, npulse, t)).max() < 1e-10


@pytest.mark.gpu
def test_dense64_hahn_long_grid_matches_reference_propagator(ctx):
    """Hahn echo on the n = 64 path advances each side by one matrix-vector
    product per output time; over the 401-point 0-5 ms grid of the CCE-4
    workload it stays within 1e-10 of the oracle's powers of the step
    propagator (PiecewiseFullMatrixVectorEvolution)."""
    t = np.linspace(0.0, 5e-3, 401)
    for coords in _clusters(3, 2):
        it = _item(coords, 1, t / 2)
        b = O.EchoBatch.from_items(ctx, [it], m=30, precision="fp64")
        b.run()
        L, st = b.results()
        assert b.timing()[0]["kernel"] == "dense64_mfma"
        assert (st[:, 8] & 0xFF).max() == 0
        e = np.abs(L[0] - _oracle_cpmg(coords, 1, t)).max()
        print(f"dense64 hahn 401 points: max |dL| {e:.2e}", flush=True)
        assert e < 1e-10


@pytest.mark.gpu
def test_nonuniform_grid_n64_uses_krylov_and_matches_fp64(ctx):
    t = np.sort(np.concatenate([[0.0], np.random.default_rng(1).uniform(0, 5e-4, 12)]))
    for coords in _clusters(3, 2):
        it = _item(coords, 1, t)
        ref = O.EchoBatch.from_items(ctx, [it], m=30, tol=1e-12, precision="fp64")
        ref.run()
        mix = O.EchoBatch.from_items(ctx, [it], m=30, tol=1e-4, precision="mixed")
        mix.run()
        assert np.abs(ref.results()[0][0] - mix.results()[0][0]).max() < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("k,npulse", [(4, 2), (4, 4)])
def test_cpmg_lanczos_path_matches_reference_propagator(ctx, k, npulse):
    t = np.linspace(0.0, 4e-4, 17)
    for coords in _clusters(k, 2):
        print(f"lanczos cpmg k={k} N={npulse}", flush=True)
        b = O.EchoBatch.from_items(ctx, [_item(coords, npulse, t)], m=30, tol=1e-6, precision="mixed")
        b.run()
        L, st = b.results()
        assert b.timing()[0]["kernel"] == "lanczos_slot"
        assert (st[:, 8] & 0xFF).max() == 0
        assert np.abs(L[0] - _oracle_cpmg(coords, npulse, t)).max() < 1e-6


@pytest.mark.gpu
def test_items_alone_and_in_the_mix_agree(ctx):
    rng = np.random.default_rng(0)
    items = []
    for k in (1, 2, 3, 4):
        for coords in _clusters(k, 2):
            npulse = int(rng.choice([1, 2, 4]))
            nt = int(rng.integers(5, 30))
            t = np.linspace(0.0, float(rng.uniform(2e-4, 8e-4)), nt)
            items.append(_item(coords, npulse, t / 2 if npulse == 1 else t))
    mix = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    mix.run()
    Lmix, _ = mix.results()
    for i, it in enumerate(items):
        print(f"item {i} npulse={it['npulse']} nt={len(it['t'])}", flush=True)
        alone = O.EchoBatch.from_items(ctx, [it], m=30, tol=1e-4, precision="mixed")
        alone.run()
        La, _ = alone.results()
        assert np.abs(La[0] - Lmix[i]).max() <= 1e-12


@pytest.mark.gpu
def test_echo_npulse1_equals_hahn_batch(ctx):
    tau = np.linspace(0.0, 5e-4, 21)
    coords = _clusters(3, 2) + _clusters(4, 2)
    items = [_item(c, 1, tau) for c in coords]
    e = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    e.run()
    Le, _ = e.results()
    h = O.HahnBatch.from_items(ctx, [(it["opA"], it["opB"], it["u0"], it["y0"], it["scale"]) for it in items],
                               m=30, tol=1e-4, precision="mixed")
    h.run(tau)
    Lh, _ = h.results()
    for i in range(len(items)):
        assert np.abs(Le[i] - Lh[i]).max() <= 1e-13


@pytest.mark.gpu
def test_graph_replay_matches_eager_launch(ctx):
    """A batch whose bucket launches are captured in a HIP graph gives, on
    every replay, the eager batch's L(t); eager runs of another batch and
    synchronisations sit between the replays."""
    t = np.linspace(0.0, 6e-4, 25)
    items = []
    for k, npulse in ((1, 1), (2, 2), (3, 4), (3, 1), (4, 2)):
        for coords in _clusters(k, 2):
            items.append(_item(coords, npulse, t / 2 if npulse == 1 else t))
    eager = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    graph = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    graph.capture()
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        eager.run()
        Le, _ = eager.results()
        te = sum(r["ms"] for r in eager.timing())
        graph.run()
        Lg, st = graph.results()
        tg = sum(r["ms"] for r in graph.timing())
        assert (st[:, 8] & 0xFF).max() == 0
        for i in range(len(items)):
            assert np.array_equal(Lg[i], Le[i]), (rep, i)
        # the replay did the work (not stale results of an earlier replay)
        assert tg > 0.3 * te, (tg, te)


@pytest.mark.gpu
def test_graph_replay_of_medium_bucket_matches_eager_launch(ctx):
    """Five-spin items (n = 1024): a captured batch holds the medium kernel's
    sweep units and the inner-product kernel (Hahn) and a two-sided unit
    (CPMG-N); every replay, with eager runs between them, equals the eager
    batch's L(t)."""
    t = np.linspace(0.0, 4e-4, 17)
    items = [_item(c, 1, t / 2) for c in _clusters(5, 2)] + [_item(_clusters(5, 3)[2], 2, t)]
    eager = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    graph = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-4, precision="mixed")
    graph.capture()
    for rep in range(3):
        print(f"medium graph replay {rep}", flush=True)
        eager.run()
        Le, _ = eager.results()
        assert eager.timing()[0]["kernel"] == "lanczos_medium"
        graph.run()
        Lg, st = graph.results()
        assert (st[:, 8] & 0xFF).max() == 0
        assert (st[:, 14:16] == 0).all()
        for i in range(len(items)):
            assert np.array_equal(Lg[i], Le[i]), (rep, i)


@pytest.mark.gpu
def test_union_of_baths_batch_equals_single_bath_batch(ctx):
    """One batch over several baths (Clusters.union) gives bath 0's clusters
    the same L(t), bit for bit, as a batch over bath 0 alone: every cluster
    is evolved on its own whatever else the batch holds."""
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    Xs = [O.bath_diamond(3.567, 22.0, 0.02, seed=s) for s in range(3)]
    CU, first = O.Clusters.union(Xs, 7.0, 4)
    C0 = O.Clusters(Xs[0], 7.0, 4)
    t = np.linspace(0.0, 1e-3, 21)
    bu = CU.echo_batch(ctx, par, 1, t / 2.0, m=30, tol=1e-3, precision="mixed")
    bu.run()
    LU, su = bu.results()
    b0 = C0.echo_batch(ctx, par, 1, t / 2.0, m=30, tol=1e-3, precision="mixed")
    b0.run()
    L0, s0 = b0.results()
    cu, c0 = CU.counts(), C0.counts()
    base = np.concatenate([[0], np.cumsum(cu)])
    rows = np.concatenate([np.arange(base[o], base[o] + c0[o]) for o in range(4)])
    print(f"union {cu} vs bath 0 {c0}", flush=True)
    assert np.array_equal(np.stack(LU)[rows], np.stack(L0))
