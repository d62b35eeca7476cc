"""GPU parity of the single-state path (oops_hamvec, oops_expv) against the
golden fixtures (explicit Kronecker products and scipy expm) and against the
oracle's restatement of the reference's multi-output Krylov driver
(hamvec_zgexpv_w_mkl_profile.f), on the same term tables.

Tolerances: the product H v is a fixed sum of the same terms, so it agrees to
1e-12 relative; Krylov results agree to 1e-10 relative of max |w| (the FP64
path's bar), the oracle and the device differ only in rounding (modified vs
classical-twice Gram-Schmidt)."""
import os

import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")
B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _case(g, c):
    pre = f"c{c}_"
    fields = [g[pre + k] for k in ("coeff", "nbody", "pos_off", "pos", "dims", "mat_off", "mats", "spin_dim")]
    return O.TermTable(*fields), P.TermTable(len(g[pre + "spin_dim"]), *fields, int(np.prod(g[pre + "spin_dim"])))


@pytest.mark.gpu
def test_hamvec_against_explicit_kron(ctx):
    g = np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)
    for c in range(int(g["ncases"])):
        T, _ = _case(g, c)
        got = O.hamvec(ctx, T, g[f"c{c}_x"])
        want = g[f"c{c}_Hx"]
        print(f"hamvec case {c} n={T.n}", flush=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), c


@pytest.mark.gpu
@pytest.mark.parametrize("axis", ["1", "0"])
def test_tiled_product_matches_flat_product(ctx, axis, monkeypatch):
    """The large path's products of a 10-spin Liouvillian (n = 2^20): the
    two-axis product (ket and bra sides, 2 passes) and, with OOPS_LARGE_AXIS=0,
    the tiled product (6 passes), against the flat ELL product evaluated on
    the host from the same compiled tables (pinned by the Kronecker fixtures
    above); -lvB^dagger and an 8-spin lvA through the two-axis product too."""
    monkeypatch.setenv("OOPS_LARGE_AXIS", axis)
    rng = np.random.default_rng(0)
    X = np.array([p for p in rng.uniform(-15, 15, (400, 3)) if np.linalg.norm(p) <= 15.0][:10])
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    cases = [(10, 0)] if axis == "0" else [(10, 0), (10, 2), (8, 0)]
    for k, which in cases:
        T, _, _ = O.cluster_terms(X[:k], par, which)
        if k == 10:
            assert O.tile_plan(T) == 6
        v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
        print(f"{'two-axis' if axis == '1' else 'tiled'} hamvec k={k} which={which} n={T.n}", flush=True)
        got = O.hamvec(ctx, T, v)
        _, want = O.tile_plan(T, v, flat=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.gpu
def test_expv_against_scipy_fixture(ctx):
    g = np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)
    for c in range(int(g["ncases"])):
        T, _ = _case(g, c)
        if T.n <= 8:
            continue  # the driver needs m < n
        m = min(30, T.n - 1)
        print(f"expv case {c} n={T.n} m={m}", flush=True)
        W, st = O.expv(ctx, T, g[f"c{c}_v"], g[f"c{c}_t"], m=m, tol=1e-12)
        want = g[f"c{c}_w"]
        err = np.abs(W - want).max() / np.abs(want).max()
        assert err < 1e-10, (c, err)
        assert st["nstep"] > 0 and st["nmult"] > 0


def _liouville_cluster(k, seed):
    X = O.bath_diamond(3.567, 14.0, 0.05, seed=seed)
    C = O.Clusters(X, 8.0, k)
    cl = C.get(k - 1)
    assert len(cl) > 0
    return X[list(cl[0])]


@pytest.mark.gpu
@pytest.mark.parametrize("k,m", [(4, 30), (5, 30), (5, 40)])
def test_expv_matches_oracle_on_cluster_liouvillian(ctx, k, m):
    coords = _liouville_cluster(k, seed=2)
    par = O.cce_params(B=B111)
    A, rho, D = O.cluster_terms(coords, par, 0)
    oA = P.TermTable(k, A.coeff, A.nbody, A.pos_offset, A.pos, A.dims, A.mat_offset, A.mats, A.spin_dim, A.n)
    tl = np.linspace(0.0, 2e-5, 6)
    print(f"cluster k={k} n={A.n} m={m}", flush=True)
    W, st = O.expv(ctx, A, rho, tl, m=m, tol=1e-12)
    Wr, sr, rc = P.expv(oA, rho, tl, m=m, tol=1e-12)
    assert rc == 0
    err = np.abs(W - Wr).max() / np.abs(Wr).max()
    assert err < 1e-10, err
    assert st["nstep"] == sr.nstep and st["nreject"] == sr.nreject
    # the norm estimate (10 Lanczos steps without reorthogonalisation, then a
    # power iteration) amplifies rounding; it only sets the first step size,
    # rounded to 2 significant digits
    assert abs(st["anorm"] - sr.anorm) <= 1e-4 * sr.anorm
    # the overlap output equals <v, w(t)> of the state output
    ov, _ = O.expv(ctx, A, rho, tl, m=m, tol=1e-12, overlaps=True)
    assert np.abs(ov - rho.conj() @ W).max() < 1e-12 * np.abs(rho).sum() * np.abs(W).max()


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [[3, 2, 3], [2, 3, 4], [9, 9, 9], [3] * 7])
def test_padded_local_dimensions_on_device(ctx, dims):
    """Spin-1, its Liouville space and mixed kinds (local dimensions 3, 9 and
    3/2/4 embedded in the next power of two on the device; the reference's
    hamvec_func3.cpp:307-389 takes any local dimension): H v against the
    explicit Kronecker matrix, exp(-i H t) v against scipy expm and against
    the oracle's Krylov driver on the unpadded table, with the same step
    schedule (the norm estimate's random draws follow the caller's indices)."""
    import scipy.linalg as sla
    from mixed_dims import random_table
    T, H = random_table(dims, seed=len(dims))
    oT = P.TermTable(len(dims), T.coeff, T.nbody, T.pos_offset, T.pos, T.dims, T.mat_offset, T.mats,
                     T.spin_dim, T.n)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
    print(f"padded dims {dims} n={T.n}", flush=True)
    got = O.hamvec(ctx, T, v)
    assert np.abs(got - H @ v).max() <= 1e-12 * np.abs(H @ v).max()
    v /= np.linalg.norm(v)
    ta = np.array([0.0, 2.0, 7.0, 20.0]) / np.abs(H).sum(0).max()
    m = min(30, T.n - 1)
    W, st = O.expv(ctx, T, v, ta, m=m, tol=1e-12)
    want = np.stack([sla.expm(-1j * H * t) @ v for t in ta], axis=1)
    assert np.abs(W - want).max() < 1e-10
    Wr, sr, rc = P.expv(oT, v, ta, m=m, tol=1e-12)
    assert rc == 0
    assert np.abs(W - Wr).max() < 1e-10
    assert st["nstep"] == sr.nstep and st["nreject"] == sr.nreject
    assert abs(st["anorm"] - sr.anorm) <= 1e-4 * sr.anorm


@pytest.mark.gpu
def test_expv_rejects_bad_arguments(ctx):
    g = np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)
    T, _ = _case(g, 0)  # n = 4
    with pytest.raises(O.OopsError):
        O.expv(ctx, T, g["c0_v"], [0.0, 1e-5], m=4)
    T, _ = _case(g, 3)
    with pytest.raises(O.OopsError):
        O.expv(ctx, T, g["c3_v"], [1e-5, 0.0], m=10)
    with pytest.raises(O.OopsError):
        O.expv(ctx, T, g["c3_v"], [0.0, 1e-5], m=41)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [5, 6])
def test_hahn_batch_large_clusters_match_oracle(ctx, k):
    """Clusters above 256 Liouville elements (CCE-5/6: n = 1024, 4096) run on
    the single-state path inside a batch; L(t) against the oracle's two-sweep
    Krylov restatement to 1e-10."""
    coords = _liouville_cluster(k, seed=2)
    par = O.cce_params(B=B111)
    A, rho, D = O.cluster_terms(coords, par, 0)
    Bd, _, _ = O.cluster_terms(coords, par, 2)
    tau = np.linspace(0.0, 2e-5, 5)
    print(f"large hahn k={k} n={A.n}", flush=True)
    b = O.HahnBatch.from_items(ctx, [(A, Bd, rho, rho, D)], m=30, tol=1e-12, precision="fp64")
    b.run(tau)
    L, st = b.results()
    assert (st[:, 8] & 0xFF).max() == 0 and st[0, 0] > 0
    tA = P.TermTable(k, A.coeff, A.nbody, A.pos_offset, A.pos, A.dims, A.mat_offset, A.mats, A.spin_dim, A.n)
    tB = P.TermTable(k, Bd.coeff, Bd.nbody, Bd.pos_offset, Bd.pos, Bd.dims, Bd.mat_offset, Bd.mats, Bd.spin_dim,
                     Bd.n)
    Lr, _, rc = P.hahn_batch([tA], [tB], [rho], [rho], [D], tau, m=30, nthreads=1)
    assert rc == 0
    assert np.abs(L[0] - Lr[0]).max() < 1e-10 * max(1.0, np.abs(Lr[0]).max())


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_expv_outputs_inside_steps_match_reference_schedule(ctx, precision):
    """steps="inside" (bit 1 of oops_expv_ex's options): the steps run past the
    output times and evaluate those inside from their own basis.  On a
    12-spin Hilbert cluster (n = 4096) with 201 outputs 1 us apart (finer than
    the natural step) the overlaps and the full states agree with the
    reference's schedule (a step ends at every output time) to 1e-10 with FP64
    and 1e-6 with the FP32 basis (the states have unit norm), in fewer Krylov
    steps."""
    rng = np.random.default_rng(3)
    X = np.array([p for p in rng.uniform(-15, 15, (600, 3)) if np.linalg.norm(p) <= 15.0][:12])
    T, _, _ = O.cluster_terms(X, O.cce_params(B=B111), 3)
    v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
    v /= np.linalg.norm(v)
    tl = np.arange(201) * 1e-6
    bar = 1e-10 if precision == "fp64" else 1e-6
    ref, st_ref = O.expv(ctx, T, v, tl, m=30, tol=1e-12, overlaps=True, precision=precision)
    ins, st_ins = O.expv(ctx, T, v, tl, m=30, tol=1e-12, overlaps=True, precision=precision, steps="inside")
    print(f"inside steps n={T.n} {precision}: steps {st_ins['nstep']:.0f} vs {st_ref['nstep']:.0f}, "
          f"max |d<v,w>| {np.abs(ins - ref).max():.2e}", flush=True)
    assert np.abs(ins - ref).max() <= bar
    assert st_ins["nstep"] < st_ref["nstep"]
    # full states (out_mode 0) on a shorter grid with a repeated time
    t2 = np.concatenate([[0.0], np.arange(1, 41) * 1e-6, [40e-6]])
    wr, _ = O.expv(ctx, T, v, t2, m=30, tol=1e-12, precision=precision)
    wi, _ = O.expv(ctx, T, v, t2, m=30, tol=1e-12, precision=precision, steps="inside")
    assert np.abs(wi - wr).max() <= bar   # unit-norm states


@pytest.mark.gpu
def test_expv_repeated_calls_are_identical(ctx):
    """Each oops_expv call builds and frees its own device state and stream;
    back-to-back calls on FP64 and FP32 bases give bit-identical results."""
    rng = np.random.default_rng(5)
    X = np.array([p for p in rng.uniform(-15, 15, (600, 3)) if np.linalg.norm(p) <= 15.0][:11])
    T, _, _ = O.cluster_terms(X, O.cce_params(B=B111), 3)
    v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
    v /= np.linalg.norm(v)
    tl = np.arange(31) * 1e-6
    for precision in ("fp64", "fp32"):
        runs = [O.expv(ctx, T, v, tl, m=30, tol=1e-12, overlaps=True, precision=precision)[0] for _ in range(3)]
        print(f"repeated expv n={T.n} {precision}", flush=True)
        assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])


@pytest.mark.gpu
def test_hahn_batch_large_items_on_threads_equal_single_items(ctx):
    """An FP64 batch of three 5-spin clusters (n = 1024) runs its items on
    host threads, each on its own stream; every item's L(t) equals the same
    item run alone (1e-12)."""
    X = O.bath_diamond(3.567, 14.0, 0.05, seed=2)
    C = O.Clusters(X, 8.0, 5)
    cl = C.get(4)
    assert len(cl) >= 3
    par = O.cce_params(B=B111)
    items = []
    for r in cl[:3]:
        A, rho, D = O.cluster_terms(X[list(r)], par, 0)
        Bd, _, _ = O.cluster_terms(X[list(r)], par, 2)
        items.append((A, Bd, rho, rho, D))
    tau = np.linspace(0.0, 2e-5, 5)
    print("threaded large batch of 3 items", flush=True)
    b = O.HahnBatch.from_items(ctx, items, m=30, tol=1e-12, precision="fp64")
    b.run(tau)
    L, st = b.results()
    assert (st[:, 8] & 0xFF).max() == 0
    for i, it in enumerate(items):
        one = O.HahnBatch.from_items(ctx, [it], m=30, tol=1e-12, precision="fp64")
        one.run(tau)
        L1, _ = one.results()
        assert np.abs(L[i] - L1[0]).max() <= 1e-12, i


# --------------------------------------------------------------- single large cluster at full size
def _sphere(k, radius=15.0, seed=0):
    """k positions uniform in a sphere of `radius` A (the large-cluster workload)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < k:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) <= radius:
            out.append(p)
    return np.array(out)


def _unit(rng, n):
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return v / np.linalg.norm(v)


def _oracle_table(T, sign=1.0):
    return P.TermTable(T.spin_dim.size, sign * np.asarray(T.coeff), T.nbody, T.pos_offset, T.pos, T.dims,
                       T.mat_offset, T.mats, T.spin_dim, T.n)


def _vec_outer(a, b, k):
    """vec(|a><b|) in the Liouville ordering: spin s contributes the local
    index ket_s + 2 bra_s, spin 0 the slowest digit."""
    D = 1 << k
    I, J = np.meshgrid(np.arange(D), np.arange(D), indexing="ij")
    idx = np.zeros((D, D), np.int64)
    for s in range(k):
        sh = k - 1 - s
        idx = idx * 4 + ((I >> sh) & 1) + 2 * ((J >> sh) & 1)
    v = np.zeros(D * D, np.complex128)
    v[idx.ravel()] = np.outer(a, b.conj()).ravel()
    return v


@pytest.mark.gpu
def test_config4_liouville10_rank1_state_equals_two_hilbert_evolutions(ctx):
    """10 spins in a 1.5 nm sphere, lvA (n = 2^20), m = 40.  With
    lvA vec(rho) = vec(rho H0 - H1 rho), a rank-1 start |a><b| evolves into
    exp(i H1 t)|a><b|exp(-i H0 t): the whole 2^20-element state must equal the
    outer product of two 1024-element Hilbert evolutions, which the oracle
    computes.  Also: unit norm kept to 1e-10 and the FP32 basis within 1e-6."""
    k = 10
    X = _sphere(k)
    par = O.cce_params(B=B111)
    A, _, _ = O.cluster_terms(X, par, 0)
    H0, _, _ = O.cluster_terms(X, par, 3)
    H1, _, _ = O.cluster_terms(X, par, 4)
    assert A.n == 1 << 20 and O.tile_plan(A) == 6
    rng = np.random.default_rng(1)
    a, b = _unit(rng, 1 << k), _unit(rng, 1 << k)
    v = _vec_outer(a, b, k)
    tl = np.arange(11) * 1e-6
    full_t = tl[[0, 5, 10]]
    print(f"liouville-10 n={A.n} m=40: full states at {len(full_t)} times", flush=True)
    W, st = O.expv(ctx, A, v, full_t, m=40, tol=1e-12)
    at, _, rc1 = P.expv(_oracle_table(H1, -1.0), a, full_t, m=40, tol=1e-12)
    bt, _, rc2 = P.expv(_oracle_table(H0, -1.0), b, full_t, m=40, tol=1e-12)
    assert rc1 == 0 and rc2 == 0
    for q in range(len(full_t)):
        want = _vec_outer(at[:, q], bt[:, q], k)
        err = np.abs(W[:, q] - want).max()
        print(f"  t={full_t[q]:.0e}: max |w - outer| {err:.2e}, |norm - 1| {abs(np.linalg.norm(W[:, q]) - 1):.1e}",
              flush=True)
        assert err <= 1e-10
        assert abs(np.linalg.norm(W[:, q]) - 1.0) <= 1e-10
    ov64, s64 = O.expv(ctx, A, v, tl, m=40, tol=1e-12, overlaps=True)
    ov32, s32 = O.expv(ctx, A, v, tl, m=40, tol=1e-12, overlaps=True, precision="fp32")
    print(f"  fp32 basis vs fp64: max |d<v,w>| {np.abs(ov32 - ov64).max():.2e} "
          f"({s64['nstep']:.0f} / {s32['nstep']:.0f} Krylov steps)", flush=True)
    assert np.abs(ov32 - ov64).max() <= 1e-6
    assert np.abs(ov64[[0, 5, 10]] - v.conj() @ W).max() <= 1e-12


@pytest.mark.gpu
def test_config4_hilbert14_matches_oracle(ctx):
    """14 spins in a 1.5 nm sphere, the m_s = 0 conditional Hamiltonian
    (n = 16384), m = 40: states against the oracle's Krylov driver to 1e-10,
    unit norm to 1e-10, and the FP32 basis within 1e-6 over 11 outputs 1 us
    apart."""
    X = _sphere(14)
    T, _, _ = O.cluster_terms(X, O.cce_params(B=B111), 3)
    assert T.n == 16384
    v = _unit(np.random.default_rng(1), T.n)
    tl = np.array([0.0, 1e-6, 2e-6])
    print(f"hilbert-14 n={T.n} m=40", flush=True)
    W, st = O.expv(ctx, T, v, tl, m=40, tol=1e-12)
    Wr, sr, rc = P.expv(_oracle_table(T), v, tl, m=40, tol=1e-12)
    assert rc == 0
    err = np.abs(W - Wr).max()
    print(f"  max |w - oracle| {err:.2e}; steps {st['nstep']:.0f} vs {sr.nstep}", flush=True)
    assert err <= 1e-10
    assert np.abs(np.linalg.norm(W, axis=0) - 1.0).max() <= 1e-10
    t11 = np.arange(11) * 1e-6
    ov64, _ = O.expv(ctx, T, v, t11, m=40, tol=1e-12, overlaps=True)
    ov32, _ = O.expv(ctx, T, v, t11, m=40, tol=1e-12, overlaps=True, precision="fp32")
    print(f"  fp32 basis vs fp64: max |d<v,w>| {np.abs(ov32 - ov64).max():.2e}", flush=True)
    assert np.abs(ov32 - ov64).max() <= 1e-6
