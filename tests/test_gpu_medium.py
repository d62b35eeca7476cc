"""GPU: the batched mixed Lanczos kernel for n = 1024 and 4096 (5- and 6-spin
Liouvillians, CCE-5/6): one 1024-thread workgroup per cluster with the product
input and the active generator's tables in LDS.  Its L(t) must match the FP64
single-state path (itself pinned to the oracle at 1e-10 in
test_gpu_large.py) to 1e-6, and CPMG-N a segment-by-segment scipy
expm_multiply of the oracle's Liouvillians to 1e-6."""
import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _clusters(k, count, seed=2):
    X = O.bath_diamond(3.567, 25.0, 0.011, seed=seed)
    C = O.Clusters(X, 6.0, k)
    cl = C.get(k - 1)
    return [X[list(r)] for r in cl[np.linspace(0, len(cl) - 1, count).astype(int)]]


def _hahn_items(coords):
    par = O.cce_params(B=B111)
    A, rho, D = O.cluster_terms(coords, par, 0)
    Bd, _, _ = O.cluster_terms(coords, par, 2)
    return (A, Bd, rho, rho, D)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [5, 6])
def test_medium_kernel_matches_fp64(ctx, k):
    items = [_hahn_items(c) for c in _clusters(k, 3)]
    tau = np.linspace(0.0, 1e-3, 41) / 2
    print(f"medium k={k} n={items[0][0].n}", flush=True)
    mixed = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-3, precision="mixed")
    mixed.run(tau)
    Lm, st = mixed.results()
    assert mixed.timing()[0]["kernel"] == "lanczos_medium"
    assert (st[:, 8] & 0xFF).max() == 0
    ref = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-12, precision="fp64")
    ref.run(tau)
    L64, _ = ref.results()
    assert ref.timing()[0]["kernel"] == "single_state"
    err = np.abs(Lm - L64).max()
    print(f"max |dL| {err:.2e}", flush=True)
    assert err <= 1e-6


@pytest.mark.gpu
def test_medium_kernel_cpmg_matches_reference_propagator(ctx):
    coords = _clusters(5, 1)[0]
    par = O.cce_params(B=B111)
    A, rho, D = O.cluster_terms(coords, par, 0)
    Bd, _, _ = O.cluster_terms(coords, par, 2)
    Cc, _, _ = O.cluster_terms(coords, par, 1)
    Dd, _, _ = O.cluster_terms(coords, par, 5)
    t = np.linspace(0.0, 4e-4, 9)
    it = {"opA": A, "opB": Bd, "opC": Cc, "opD": Dd, "u0": rho, "y0": rho, "scale": D, "npulse": 2, "t": t}
    print("medium cpmg k=5 N=2", flush=True)
    b = O.EchoBatch.from_items(ctx, [it], m=30, tol=1e-6, precision="mixed")
    b.run()
    L, st = b.results()
    assert b.timing()[0]["kernel"] == "lanczos_medium"
    assert (st[:, 8] & 0xFF).max() == 0
    # reference: the segments applied with scipy's expm_multiply on the
    # oracle's explicit Liouvillians (a dense 1024 x 1024 propagator per time
    # point would take minutes)
    import scipy.sparse as sp
    from scipy.sparse.linalg import expm_multiply
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    n = int(np.prod(prob["ldims"]))
    ops = P.riffle(sp.csr_matrix(P.full(prob["lvA"], n)), sp.csr_matrix(P.full(prob["lvB"], n)), 2)
    seg = P.pulse_interval(2)
    Lr = []
    for tk in t:
        v = prob["rho"].astype(complex)
        for j, op in enumerate(ops):
            v = expm_multiply(-1j * seg[j] * tk * op, v)
        Lr.append(np.real(prob["D"] * (prob["rho"].conj() @ v)))
    assert np.abs(L[0] - np.array(Lr)).max() < 1e-6


@pytest.mark.gpu
def test_medium_hahn_units_and_cpmg_in_one_bucket(ctx):
    """Hahn items with a state store run as one work unit per sweep (output
    times inside a Krylov step evaluated from its basis, L(t) formed by a
    second kernel); a CPMG-N item in the same bucket runs both sides in one
    unit.  Hahn items on fine and irregular grids match the FP64 single-state
    path to 1e-6, and the CPMG item gives the same L as when it runs alone."""
    coords = _clusters(5, 3)
    par = O.cce_params(B=B111)
    rng = np.random.default_rng(4)
    grids = [np.linspace(0.0, 1e-3, 161) / 2, np.sort(np.concatenate([[0.0, 0.0], rng.uniform(0.0, 5e-4, 30)]))]
    hahn = []
    for c, g in zip(coords[:2], grids):
        A, rho, D = O.cluster_terms(c, par, 0)
        Bd, _, _ = O.cluster_terms(c, par, 2)
        hahn.append({"opA": A, "opB": Bd, "u0": rho, "y0": rho, "scale": D, "npulse": 1, "t": g})
    A, rho, D = O.cluster_terms(coords[2], par, 0)
    Bd, _, _ = O.cluster_terms(coords[2], par, 2)
    Cc, _, _ = O.cluster_terms(coords[2], par, 1)
    Dd, _, _ = O.cluster_terms(coords[2], par, 5)
    cp = {"opA": A, "opB": Bd, "opC": Cc, "opD": Dd, "u0": rho, "y0": rho, "scale": D, "npulse": 2,
          "t": np.linspace(0.0, 4e-4, 5)}
    print("medium units: hahn grids", [len(g) for g in grids], "+ cpmg N=2", flush=True)
    mixed = O.EchoBatch.from_items(ctx, hahn + [cp], m=30, tol=1e-3, precision="mixed")
    mixed.run()
    Lm, st = mixed.results()
    assert mixed.timing()[0]["kernel"] == "lanczos_medium"
    assert (st[:, 8] & 0xFF).max() == 0
    assert (st[:, 14:16] == 0).all()
    ref = O.EchoBatch.from_items(ctx, hahn, m=30, tol=1e-12, precision="fp64")
    ref.run()
    L64, st64 = ref.results()
    for i in range(2):
        e = np.abs(Lm[i] - L64[i]).max()
        print(f"hahn item {i}: max |dL| {e:.2e}, mixed steps {st[i, 0]}/{st[i, 4]}, fp64 steps {st64[i, 0]}",
              flush=True)
        assert e <= 1e-6
    alone = O.EchoBatch.from_items(ctx, [cp], m=30, tol=1e-3, precision="mixed")
    alone.run()
    La, _ = alone.results()
    assert np.abs(Lm[2] - La[0]).max() <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("k", [5, 6])
def test_medium_flip_mask_form_matches_ell_form(ctx, k):
    """The n = 1024 / 4096 kernel's flip-mask product (per-element coefficient
    tables, 11 + 20 masks at k = 5, 13 + 30 at k = 6) against its merged-ELL
    product (precision "mixed-ell"): L(t) within 1e-6."""
    items = [_hahn_items(c) for c in _clusters(k, 2)]
    tau = np.linspace(0.0, 1e-3, 41) / 2
    out = {}
    for prec in ("mixed", "mixed-ell"):
        b = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-3, precision=prec)
        b.run(tau)
        L, st = b.results()
        assert b.timing()[0]["kernel"] == "lanczos_medium"
        assert (st[:, 8] & 0xFF).max() == 0
        out[prec] = L
    e = np.abs(out["mixed"] - out["mixed-ell"]).max()
    print(f"medium k={k} flip-mask vs ELL: max |dL| {e:.2e}", flush=True)
    assert e <= 1e-6


@pytest.mark.gpu
def test_many_units_take_the_ell_form_and_agree_with_the_flip_mask_form(ctx):
    """A medium bucket with more work units than 2/3 of the CUs runs the
    merged ELL product from LDS (the flip-mask coefficient stream would
    saturate HBM); its L(t) agrees with the same clusters run as a small
    bucket on the flip-mask form to the mixed path's bar."""
    ncu = ctx.info()["num_cu"]
    many = ncu // 3 + 8                     # 2 units per Hahn item: > 2/3 of the CUs
    coords = _clusters(5, many)
    items = [_hahn_items(c) for c in coords]
    tau = np.linspace(0.0, 6e-4, 25) / 2
    big = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-3, precision="mixed")
    big.run(tau)
    Lb, sb = big.results()
    small = O.HahnBatch.from_items(ctx, items[:8], m=30, tol=1e-3, precision="mixed")
    small.run(tau)
    Ls, ss = small.results()
    assert big.timing()[0]["kernel"] == small.timing()[0]["kernel"] == "lanczos_medium"
    assert (sb[:, 8] & 0xFF).max() == 0 and (ss[:, 8] & 0xFF).max() == 0
    err = np.abs(Lb[:8] - Ls).max()
    print(f"{many} items (ELL form) vs 8 items (flip-mask form): max |dL| {err:.2e}", flush=True)
    assert err <= 1e-6
