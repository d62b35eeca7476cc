"""GPU: the flip-mask (slot) form of the mixed Lanczos kernel.

For a 4-spin spin-1/2 Liouvillian every merged local operator entry moves
element i to i ^ mask; the slot kernel sums the entries per mask into
per-element coefficients (9 FP64 one-spin masks, 12 FP32 two-spin masks) and
applies 21 gathers per element instead of the 64 ELL entries.  It must give
the same L_c(t) as the FP64 path to 1e-6 (the mixed-precision bar) and agree
with the ELL form of the same Lanczos kernel (precision "mixed-ell"), for
Hahn echoes and CPMG-N, on n = 256 and n = 64 buckets.  (CPMG-N on n = 256
against the oracle's dense propagator: test_gpu_echo.py, which runs this
kernel too.)"""
import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _sel(C, order, count):
    counts = C.counts()
    base = int(sum(counts[:order - 1]))
    return base + np.linspace(0, counts[order - 1] - 1, count).astype(np.int64)


def _kinds(batch):
    return {r["n"]: r["kernel"] for r in batch.timing()}


@pytest.mark.gpu
@pytest.mark.parametrize("m,tol", [(30, 1e-3), (20, 3e-4)])
def test_slot_kernel_matches_fp64_and_ell(ctx, m, tol):
    X = O.bath_diamond(3.567, 30.0, 0.011, seed=5)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=B111)
    sel = _sel(C, 4, 48)
    tau = np.linspace(0.0, 2.5e-3, 101) / 2
    print(f"slot m={m} tol={tol} clusters={len(sel)}", flush=True)
    out = {}
    for prec in ("fp64", "mixed", "mixed-ell"):
        b = C.hahn_batch(ctx, par, m=30 if prec == "fp64" else m, tol=1e-12 if prec == "fp64" else tol,
                         precision=prec, select=sel)
        b.run(tau)
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        out[prec] = (L, _kinds(b))
        b.close()
    assert out["mixed"][1][256] == "lanczos_slot"
    assert out["mixed-ell"][1][256] == "lanczos_mixed"
    e64 = np.abs(out["mixed"][0] - out["fp64"][0]).max()
    eell = np.abs(out["mixed"][0] - out["mixed-ell"][0]).max()
    print(f"max |dL| vs fp64 {e64:.3e}, vs ell {eell:.3e}", flush=True)
    assert e64 <= 1e-6
    assert eell <= 1e-6


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
    """L(t_k) from the oracle's dense piecewise propagator, one time at a time
    (any grid)."""
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    n = int(np.prod(prob["ldims"]))
    ops = P.riffle(P.full(prob["lvA"], n), P.full(prob["lvB"], n), npulse)
    seg = P.pulse_interval(npulse)
    if t[0] == 0.0 and np.allclose(np.diff(t), t[1], rtol=1e-12, atol=0.0):
        st = P.piecewise_dense(ops, seg, t[1], prob["rho"], len(t))   # uniform grid: powers of one step
        return np.real(prob["D"] * (prob["rho"].conj() @ st))
    out = []
    for tk in t:
        x = P.piecewise_dense(ops, seg, tk, prob["rho"], 2)[:, 1] if tk > 0 else prob["rho"]
        out.append(np.real(prob["D"] * (prob["rho"].conj() @ x)))
    return np.array(out)


@pytest.mark.gpu
@pytest.mark.parametrize("k,npulse", [(4, 2), (4, 8), (3, 4)])
def test_slot_kernel_cpmg_matches_reference_propagator(ctx, k, npulse):
    items, coords = [], _clusters(k, 2 if k == 4 else 3)
    # a non-uniform grid keeps n = 64 items off the dense propagator path
    tt = 1e-3 * np.linspace(0.0, 1.0, 21) ** (1.0 if k == 4 else 1.5)
    for c in coords:
        items.append(_item(c, npulse, tt))
    print(f"slot cpmg k={k} N={npulse}", flush=True)
    res = {}
    for prec in ("mixed", "mixed-ell"):
        b = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-5, precision=prec)
        b.run()
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        res[prec] = (L, _kinds(b))
    n = 4 ** k
    assert res["mixed"][1][n] == "lanczos_slot"
    assert res["mixed-ell"][1][n] == "lanczos_mixed"
    for i, c in enumerate(coords):
        assert np.abs(res["mixed"][0][i] - res["mixed-ell"][0][i]).max() < 1e-6
        if k == 3:   # n = 64: the oracle's dense propagator per output time is cheap
            ref = _oracle_cpmg(c, npulse, tt)
            assert np.abs(res["mixed"][0][i] - ref).max() < 1e-6


@pytest.mark.gpu
def test_slot_hahn_outputs_inside_steps(ctx):
    """Mixed Hahn sweeps evaluate every output time that falls inside a Krylov
    step from that step's basis (sweep A stores its states, sweep B takes the
    inner products).  On a grid finer than the natural step they take fewer
    steps than output times, while the FP64 path (the reference's schedule,
    zgexpv_cache.f:440-661) ends a step at every output time; L(t) agrees to
    1e-6 on fine, irregular, repeated, single and offset grids."""
    rng = np.random.default_rng(11)
    coords = _clusters(4, 4)
    irregular = np.sort(np.concatenate([[0.0, 0.0], rng.uniform(0.0, 1.25e-3, 58)]))
    irregular[30] = irregular[29]
    grids = [np.linspace(0.0, 2.5e-3, 801) / 2, irregular, np.array([7e-4]), np.linspace(1e-4, 1.2e-3, 37)]
    items = []
    for c, g in zip(coords, grids):
        it = _item(c, 1, g)
        items.append(it)
    print("slot hahn grids", [len(g) for g in grids], flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        b = O.EchoBatch.from_items(ctx, items, m=30, tol=1e-12 if prec == "fp64" else 1e-3, precision=prec)
        b.run()
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        res[prec] = (L, st, _kinds(b))
    assert res["mixed"][2][256] == "lanczos_slot"
    for i, g in enumerate(grids):
        e = np.abs(np.asarray(res["mixed"][0][i]) - np.asarray(res["fp64"][0][i])).max()
        print(f"grid {i} ({len(g)} times): max |dL| {e:.3e}, mixed steps A/B {res['mixed'][1][i, 0]}/"
              f"{res['mixed'][1][i, 4]}, fp64 steps A {res['fp64'][1][i, 0]}", flush=True)
        assert e <= 1e-6
    # the fine grid: 800 intervals of 1.56 us; the FP64 schedule takes at least
    # one step per interval, the mixed sweeps far fewer
    assert res["fp64"][1][0, 0] >= 800
    assert res["mixed"][1][0, 0] < 600 and res["mixed"][1][0, 4] < 600


@pytest.mark.gpu
def test_slot_hahn_rerun_with_longer_grid(ctx):
    """A uniform-time batch run again with more output times grows its sweep
    state store and still matches the FP64 path."""
    X = O.bath_diamond(3.567, 30.0, 0.011, seed=5)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=B111)
    sel = _sel(C, 4, 24)
    b = C.hahn_batch(ctx, par, m=30, tol=1e-3, precision="mixed", select=sel)
    b.run(np.linspace(0.0, 1e-3, 51) / 2)
    tau = np.linspace(0.0, 2.5e-3, 251) / 2
    b.run(tau)
    L, st = b.results()
    assert (st[:, 8] & 0xFF).max() == 0
    assert _kinds(b)[256] == "lanczos_slot"
    b.close()
    r = C.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64", select=sel)
    r.run(tau)
    L64, _ = r.results()
    r.close()
    e = np.abs(L - L64).max()
    print(f"rerun max |dL| {e:.3e}", flush=True)
    assert e <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("field", [0.01, 0.45])
def test_slot_hahn_low_and_high_field(ctx, field):
    """At 0.01 T one Krylov step spans many output times; at 0.45 T the steps
    are shorter than the output spacing and clusters far from the centre reach
    a happy breakdown (the step then covers every remaining output time).
    Both must stay within 1e-6 of the FP64 path."""
    X = O.bath_diamond(3.567, 40.0, 0.011, seed=8)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=np.ones(3) * field / np.sqrt(3.0))
    sel = _sel(C, 4, 32)
    tau = np.linspace(0.0, 2e-3, 151) / 2
    out = {}
    for prec in ("fp64", "mixed"):
        b = C.hahn_batch(ctx, par, m=30, tol=1e-12 if prec == "fp64" else 1e-3, precision=prec, select=sel)
        b.run(tau)
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        out[prec] = (L, st, _kinds(b))
        b.close()
    assert out["mixed"][2][256] == "lanczos_slot"
    st = out["mixed"][1]
    e = np.abs(out["mixed"][0] - out["fp64"][0]).max()
    print(f"B={field}: max |dL| {e:.3e}, mixed steps/sweep {st[:, 0].mean():.1f}, breakdowns "
          f"{int(((st[:, 8] >> 8) & 3).astype(bool).sum())}, fp64 steps/sweep {out['fp64'][1][:, 0].mean():.1f}",
          flush=True)
    assert e <= 1e-6


@pytest.mark.gpu
def test_hahn_without_state_store_matches_with_store(ctx, monkeypatch):
    """OOPS_SWEEP_STORE=0 (also the fallback when the store would exceed
    32 GB or cannot be allocated): the Hahn sweeps interleave per output time
    with a step cut at every output, on the n = 256 and n = 1024 kernels; L(t)
    agrees with the store-based sweeps to 1e-6."""
    X = O.bath_diamond(3.567, 30.0, 0.011, seed=5)
    C = O.Clusters(X, 8.0, 4)
    par = O.cce_params(B=B111)
    sel = _sel(C, 4, 16)
    tau = np.linspace(0.0, 1e-3, 81) / 2
    out = {}
    for store in ("1", "0"):
        monkeypatch.setenv("OOPS_SWEEP_STORE", store)
        b = C.hahn_batch(ctx, par, m=30, tol=1e-3, precision="mixed", select=sel)
        b.run(tau)
        L, st = b.results()
        assert (st[:, 8] & 0xFF).max() == 0
        out[store] = (L, st[:, 0].sum())
        b.close()
    e = np.abs(out["1"][0] - out["0"][0]).max()
    print(f"store vs no store: max |dL| {e:.2e}, steps {out['1'][1]} vs {out['0'][1]}", flush=True)
    assert e <= 1e-6
    assert out["1"][1] < out["0"][1]
    # n = 1024 (medium kernel units vs both sweeps in one workgroup)
    X5 = O.bath_diamond(3.567, 25.0, 0.011, seed=2)
    C5 = O.Clusters(X5, 6.0, 5)
    cl = C5.get(4)
    items = []
    for r in cl[np.linspace(0, len(cl) - 1, 2).astype(int)]:
        A, rho, D = O.cluster_terms(X5[list(r)], par, 0)
        Bd, _, _ = O.cluster_terms(X5[list(r)], par, 2)
        items.append((A, Bd, rho, rho, D))
    res = {}
    for store in ("1", "0"):
        monkeypatch.setenv("OOPS_SWEEP_STORE", store)
        b = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-3, precision="mixed")
        b.run(tau)
        L, st = b.results()
        assert b.timing()[0]["kernel"] == "lanczos_medium"
        assert (st[:, 8] & 0xFF).max() == 0
        res[store] = L
    e5 = np.abs(res["1"] - res["0"]).max()
    print(f"medium store vs no store: max |dL| {e5:.2e}", flush=True)
    assert e5 <= 1e-6
