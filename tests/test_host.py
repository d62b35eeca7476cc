"""Product host logic (liboops_amd.so, no GPU needed) against the oracle:
term tables, NV states, enumeration order and parents, CCE reduction, the
synthetic bath, and the C-ABI export list."""
import os
import re

import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def test_library_exports_every_declared_symbol():
    lib = O.lib()
    names = set()
    for h in ("oops_expv.h", "oops_cce.h", "oops_serve.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:int|double|const char\*)\s+(oops_\w+)\s*\(", txt, re.M))
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # link-compatible symbols of the reference's expv library
    txt = open(os.path.join(ROOT, "include", "oops_expv.h")).read()
    ref = set(re.findall(r"^\s*(?:int|void)\s+((?:krylov_\w+)|(?:\w+_))\s*\(", txt, re.M))
    assert {"main_cache_", "main_mkl_", "krylov_zgexpv", "krylov_zcooexpv", "hamvec_func3_", "hamvec_cuda3_",
            "hamvec_cuda3_init_", "hamvec_cuda3_term_", "zgpadm_"} <= ref
    missing = [n for n in sorted(ref) if not hasattr(lib, n)]
    assert not missing, missing


def test_nv_states_match_oracle():
    ev, vec = O.nv_states([0, 0, 0], B111)
    nv = P.NVCenter((0, 0, 0), B111)
    assert np.allclose(ev, nv.evals, rtol=1e-14, atol=1e-5)
    for k in range(3):
        assert np.allclose(P.spin_vector(3, vec[:, k]), P.spin_vector(3, nv.state(k)), atol=1e-12)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_cluster_terms_match_oracle(k):
    a = 3.567
    rng = np.random.default_rng(k)
    coords = np.array([[a * x for x in rng.integers(-4, 5, 3)] for _ in range(k)]) + \
        np.array([[0.25 * a * (i % 2)] * 3 for i in range(k)])
    par = O.cce_params(B=B111, dephasing_rate=2.0 if k == 2 else 0.0)
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv, dephasing_rate=2.0 if k == 2 else 0.0)
    for which, key in [(0, "lvA"), (1, "lvB"), (2, "lvBdag")]:
        tt, rho, D = O.cluster_terms(coords, par, which)
        ot = P.flatten(prob[key], prob["ldims"])
        assert len(tt.coeff) == len(ot.coeff)
        assert np.array_equal(tt.nbody, ot.nbody)
        assert np.array_equal(tt.pos[:len(ot.pos)], ot.pos[:len(tt.pos)])
        assert np.allclose(tt.coeff, ot.coeff, rtol=1e-13, atol=1e-9)
        assert np.abs(tt.mats - ot.mats).max() < 1e-15
        assert np.abs(rho - prob["rho"]).max() < 1e-15 and D == prob["D"]


def test_enumeration_bit_exact_against_oracle():
    X = O.bath_diamond(3.567, 22.0, 0.011, seed=9)
    C = O.Clusters(X, 8.0, 4)
    cl, par = P.enumerate_clusters(P.connection_matrix(X, 8.0), 4)
    for o in range(4):
        assert np.array_equal(C.get(o), np.array(cl[o], np.int32).reshape(-1, o + 1))
        assert C.parents(o) == par[o]


def test_enumeration_counts_roycoord():
    path = os.path.join(ROOT, "tests", "golden", "RoyCoord.xyz")
    lines = open(path).read().split("\n")
    n = int(lines[0])
    X = np.array([[float(v) for v in l.split()[1:4]] for l in lines[1:n + 1]])
    assert O.Clusters(X, 6.0, 4).counts() == [548, 504, 672, 1107]
    assert O.Clusters(X, 8.0, 4).counts() == [548, 1179, 3530, 12662]


def test_reduction_bit_exact_against_oracle():
    X = O.bath_diamond(3.567, 18.0, 0.02, seed=4)
    C = O.Clusters(X, 7.0, 3)
    cl, par = P.enumerate_clusters(P.connection_matrix(X, 7.0), 3)
    rng = np.random.default_rng(2)
    nt = 5
    L = [rng.uniform(0.7, 1.0, size=(len(c), nt)) for c in cl]
    tilde, final, each = P.cce_reduce(L, par)
    r = C.reduce(np.concatenate(L), nt)
    assert np.array_equal(r["tilde"], np.concatenate(tilde))
    assert np.array_equal(r["final"], final)
    assert np.array_equal(r["each"], each)
    assert np.allclose(np.exp(r["log_abs"]) * r["sign"], final, rtol=1e-12)


def test_bath_diamond_properties():
    a = 3.567
    X = O.bath_diamond(a, 30.0, 0.011, seed=0)
    X2 = O.bath_diamond(a, 30.0, 0.011, seed=0)
    assert np.array_equal(X, X2)
    assert not np.array_equal(X, O.bath_diamond(a, 30.0, 0.011, seed=1))
    assert np.all(np.linalg.norm(X, axis=1) <= 30.0 + 1e-12)
    # every site is a diamond lattice site: 4x/a integer and parity rule
    q = X * 4.0 / a
    assert np.allclose(q, np.round(q), atol=1e-9)
    qi = np.round(q).astype(int) % 4
    fcc = np.all(qi % 2 == 0, axis=1) & ((qi // 2).sum(axis=1) % 2 == 0)
    shifted = np.all(qi % 2 == 1, axis=1) & (((qi - 1) // 2).sum(axis=1) % 2 == 0)
    assert np.all(fcc | shifted)
    # NV sites excluded; occupancy ~ p
    assert not np.any(np.all(np.abs(X) < 1e-9, axis=1))
    nsites = 8 * (4.0 / 3.0) * np.pi * 30.0 ** 3 / a ** 3
    assert abs(len(X) - 0.011 * nsites) < 5 * np.sqrt(0.011 * nsites)


def test_precision_and_arguments_rejected_without_gpu_calls():
    with pytest.raises(ValueError):
        O._precision_code("fp16")


@pytest.mark.parametrize("k,which", [(4, 0), (4, 2), (4, 1), (4, 5), (3, 0), (2, 0), (5, 0), (5, 2)])
def test_slot_form_reproduces_the_operator(k, which):
    """The flip-mask form the mixed Lanczos kernels build (per-element
    coefficient of every mask) reassembles the oracle's explicit Liouvillian;
    a k-spin spin-1/2 Liouvillian has 1 + 2k one-spin masks (diagonal, a ket
    or bra flip per spin) and k(k-1) two-spin masks (ket-ket and bra-bra pair
    flips): 9 + 12 at k = 4 (n = 256 kernel), 11 + 20 at k = 5 and 13 + 30 at
    k = 6 (n = 1024 / 4096 kernel, at most 16 + 48)."""
    X = O.bath_diamond(3.567, 25.0, 0.02, seed=3)
    C = O.Clusters(X, 7.0, k)
    cl = C.get(k - 1)
    coords = X[list(cl[len(cl) // 2])]
    par = O.cce_params(B=B111)
    T, _, _ = O.cluster_terms(coords, par, which)
    masks, coef, n64 = O.slot_form(T)
    n = T.n
    assert (n64, len(masks) - n64) == (1 + 2 * k, k * (k - 1))
    assert masks[0] == 0
    fields = [sum(((int(mk) >> (2 * q)) & 3) != 0 for q in range(k)) for mk in masks]
    assert all(f <= 1 for f in fields[:n64]) and all(f == 2 for f in fields[n64:])
    A = np.zeros((n, n), complex)
    rows = np.arange(n)
    for s in range(len(masks)):
        A[rows, rows ^ masks[s]] += coef[s]
    nv = P.NVCenter((0, 0, 0), B111)
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    key = {0: "lvA", 1: "lvB", 2: "lvBdag"}.get(which)
    ref = P.full(prob[key], n) if key else -P.full(prob["lvA"], n).conj().T
    assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()


def test_tile_plan_product_equals_flat_product():
    """The large path's tiled product (6 passes for a 10-spin Liouvillian,
    n = 2^20): the host evaluation through the kernels' tile-to-index mapping
    (checked to be a bijection in every pass) equals the flat ELL product;
    small states keep the flat product."""
    rng = np.random.default_rng(0)
    X = np.array([p for p in rng.uniform(-15, 15, (400, 3)) if np.linalg.norm(p) <= 15.0][:10])
    par = O.cce_params(B=B111)
    T, _, _ = O.cluster_terms(X, par, 0)
    assert T.n == 1 << 20
    v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
    npass, wt = O.tile_plan(T, v)
    assert npass == 6
    _, wf = O.tile_plan(T, v, flat=True)
    assert np.abs(wt - wf).max() <= 1e-13 * np.abs(wf).max()
    T9, _, _ = O.cluster_terms(X[:9], par, 0)
    assert O.tile_plan(T9) == 0


def test_two_axis_product_equals_flat_product():
    """The large path's two-axis product of a split Liouvillian (lvA = H0# -
    H1b: every term acts on ket digits only or on bra digits only): the host
    evaluation through the kernels' tile mapping (a bijection in both passes)
    equals the flat ELL product for 8 and 10 spins (n = 2^16, 2^20);
    -lvB^dagger splits too, a Liouvillian with dephasing terms does not."""
    rng = np.random.default_rng(3)
    X = np.array([p for p in rng.uniform(-15, 15, (400, 3)) if np.linalg.norm(p) <= 15.0][:10])
    par = O.cce_params(B=B111)
    for k, which in ((8, 0), (10, 0), (8, 2)):
        T, _, _ = O.cluster_terms(X[:k], par, which)
        v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
        npass, wa = O.tile_plan(T, v, axis=True)
        assert npass == 2
        _, wf = O.tile_plan(T, v, flat=True)
        assert np.abs(wa - wf).max() <= 1e-12 * np.abs(wf).max(), (k, which)
    Td, _, _ = O.cluster_terms(X[:8], O.cce_params(B=B111, dephasing_rate=1e3), 0)
    with pytest.raises(O.OopsError):
        O.tile_plan(Td, np.zeros(Td.n, complex), axis=True)


@pytest.mark.parametrize("dims", [[3, 2, 3], [3, 3, 3, 3], [9, 9], [2, 3, 4]])
def test_padded_local_dimensions_reproduce_the_operator(dims):
    """Spins whose local dimensions are not one power of two (spin-1, its
    Liouville space, mixed kinds) are embedded in the next power of two; the
    host product through the padded tables, read back in the caller's index
    order, equals the explicit Kronecker matrix (hamvec_func3.cpp:307-389
    takes any local dimension)."""
    from mixed_dims import random_table
    T, H = random_table(dims, seed=len(dims))
    v = np.random.default_rng(1).standard_normal(T.n) + 1j * np.random.default_rng(2).standard_normal(T.n)
    npass, w = O.tile_plan(T, v, flat=True)
    assert npass == 0
    assert np.abs(w - H @ v).max() <= 1e-12 * np.abs(H @ v).max()


def test_coordinate_files_round_trip(tmp_path):
    """Coordinate files in the reference's layout (SpinSource.cpp:20-79):
    RoyCoord.xyz reads back 548 13C spins with the coordinates the survey's
    counts are computed from, and writing the reference's exported file
    (dat_example/output/c.xyz, 100 spins) reproduces it byte for byte."""
    X, iso = O.read_xyz(os.path.join(ROOT, "tests", "golden", "RoyCoord.xyz"))
    assert X.shape == (548, 3) and set(iso) == {"13C"}
    rows = [ln.split() for ln in open(os.path.join(ROOT, "tests", "golden", "RoyCoord.xyz")).read().split("\n")[1:]
            if ln.strip()]
    assert np.array_equal(X, np.array([[float(v) for v in r[1:]] for r in rows]))
    src = os.path.join(ROOT, "tests", "golden", "c_export.xyz")
    Y, iso2 = O.read_xyz(src)
    assert Y.shape == (100, 3) and set(iso2) == {"E"}
    out = tmp_path / "c.xyz"
    O.write_xyz(str(out), Y, iso2)
    assert open(out, "rb").read() == open(src, "rb").read()
    with pytest.raises(O.OopsError):
        O.read_xyz(src, max_n=10)


def test_cce_result_export(tmp_path):
    """The result file carries the reference's variables (cce.cpp:190-230):
    per order the L_C(t) and reduced L~_C(t) matrices (nt x clusters), the
    per-order and cumulative products and the time list."""
    X, _ = O.read_xyz(os.path.join(ROOT, "tests", "golden", "RoyCoord.xyz"))
    C = O.Clusters(X, 6.0, 3)
    nt = 7
    rng = np.random.default_rng(0)
    L = 1.0 - 0.01 * rng.random((sum(C.counts()), nt))
    t = np.linspace(0.0, 1e-3, nt)
    path = tmp_path / "cce.npz"
    C.export_npz(str(path), L, t)
    z = np.load(path, allow_pickle=False)
    assert [z[f"CCE{i}"].shape for i in range(3)] == [(nt, c) for c in C.counts()]
    assert z["final_result"].shape == (nt, 3) and z["time_list"].shape == (nt, 1)
    r = C.reduce(L.reshape(-1), nt)
    assert np.array_equal(z["final_result"], r["final"])
    assert np.array_equal(z["CCE1"], L[C.counts()[0]:C.counts()[0] + C.counts()[1]].T)


def test_union_of_baths_enumerates_each_bath_alone():
    """Clusters.union (one enumeration over several baths, no connection
    between them; cDepthFirstPathTracing on the block-diagonal graph): every
    order holds exactly the clusters of each bath enumerated alone, offset by
    the bath's first spin, bath 0's first and in its own order."""
    Xs = [O.bath_diamond(3.567, 25.0, 0.02, seed=s) for s in range(3)]
    CU, first = O.Clusters.union(Xs, 5.0, 4)
    Cs = [O.Clusters(X, 5.0, 4) for X in Xs]
    assert list(first) == [0, len(Xs[0]), len(Xs[0]) + len(Xs[1])]
    for o in range(4):
        got = {tuple(r) for r in CU.get(o).tolist()}
        want = {tuple(int(i) + int(f) for i in r) for f, C in zip(first, Cs) for r in C.get(o).tolist()}
        assert got == want and CU.count(o) == sum(C.count(o) for C in Cs)
        assert np.array_equal(CU.get(o)[:Cs[0].count(o)], Cs[0].get(o))
