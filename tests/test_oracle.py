"""Pins of the oracle (CPU restatement of the reference path) against
independent answers: the reference's own DLARAN.f, scipy/mpmath matrix
exponentials, explicit Kronecker products, closed forms and brute force."""
import os

import numpy as np
import pytest

from oracle import physics as P

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# --------------------------------------------------------------- DLARAN
def test_dlaran_matches_reference_sequence():
    g = load("dlaran_seq.npz")
    seed = [0, 0, 0, 5]
    got = np.array([P.dlaran(seed) for _ in range(len(g["seq"]))])
    assert np.array_equal(got, g["seq"])
    assert seed == list(g["final_seed"])


def test_dlaran_against_reference_build_if_present():
    import ctypes
    ref = os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref", "libdlaran_ref.so")
    if not os.path.exists(ref):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    L = ctypes.CDLL(ref)
    L.dlaran_.restype = ctypes.c_double
    s_ref = (ctypes.c_int * 4)(1234, 17, 4000, 77)  # odd last limb
    seed = [1234, 17, 4000, 77]
    for _ in range(500):
        assert P.dlaran(seed) == L.dlaran_(s_ref)
    assert seed == list(s_ref)


# --------------------------------------------------------------- ZGPADM
def test_zgpadm_against_scipy_and_mpmath():
    g = load("expm_cases.npz")
    off = 0
    for n, t in zip(g["sizes"], g["t"]):
        A = g["mats"][off:off + n * n].reshape(n, n, order="F")
        E = g["outs"][off:off + n * n].reshape(n, n, order="F")
        off += n * n
        got, ns = P.zgpadm(A, t)
        err = np.linalg.norm(got - E) / max(np.linalg.norm(E), 1e-300)
        assert err < 1e-12, (n, t, err)


def test_zgpadm_pauli_rotation_closed_form():
    # exp(-i theta n.sigma / 2) = cos(theta/2) I - i sin(theta/2) n.sigma
    sig = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]], complex),
           np.array([[1, 0], [0, -1]], complex)]
    for theta, nv in [(0.3, (0, 0, 1)), (2.9, (1, 1, 1)), (17.0, (0.2, -0.5, 0.8))]:
        nvec = np.array(nv, float) / np.linalg.norm(nv)
        ns_ = sum(c * s for c, s in zip(nvec, sig))
        got, _ = P.zgpadm(-0.5j * ns_, theta)
        want = np.cos(theta / 2) * np.eye(2) - 1j * np.sin(theta / 2) * ns_
        assert np.abs(got - want).max() < 1e-14


def test_zgpadm_inverse_and_unitarity():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((9, 9)) + 1j * rng.standard_normal((9, 9))
    H = A + A.conj().T
    U, _ = P.zgpadm(-1j * H, 0.8)
    Ui, _ = P.zgpadm(1j * H, 0.8)
    assert np.abs(U @ Ui - np.eye(9)).max() < 1e-12
    assert np.abs(U.conj().T @ U - np.eye(9)).max() < 1e-12


# --------------------------------------------------------------- matvec / expv
def _case(g, c):
    pre = f"c{c}_"
    return P.TermTable(len(g[pre + "spin_dim"]), g[pre + "coeff"], g[pre + "nbody"], g[pre + "pos_off"],
                       g[pre + "pos"], g[pre + "dims"], g[pre + "mat_off"], g[pre + "mats"],
                       g[pre + "spin_dim"], int(np.prod(g[pre + "spin_dim"])))


def test_hamvec_against_explicit_kron():
    g = load("expv_cases.npz")
    for c in range(int(g["ncases"])):
        tt = _case(g, c)
        got = P.hamvec(tt, g[f"c{c}_x"])
        want = g[f"c{c}_Hx"]
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_expv_against_scipy_expm():
    g = load("expv_cases.npz")
    for c in range(int(g["ncases"])):
        tt = _case(g, c)
        n = tt.n
        m = min(30, n - 1)
        if n <= 8:
            continue  # m would have to be tiny: the reference requires m < n
        W, st, rc = P.expv(tt, g[f"c{c}_v"], g[f"c{c}_t"], m=m, tol=1e-12)
        assert rc == 0
        want = g[f"c{c}_w"]
        err = np.abs(W - want).max() / np.abs(want).max()
        assert err < 1e-10, (c, err)
        assert st.nmult > 0 and st.nstep > 0


def test_expv_norm_preserved_for_hermitian():
    g = load("expv_cases.npz")
    tt = _case(g, 3)  # 5 spins, Hermitian
    W, _, rc = P.expv(tt, g["c3_v"], np.linspace(0, 1e-3, 6), m=20)
    assert rc == 0
    assert np.abs(np.linalg.norm(W, axis=0) - 1.0).max() < 1e-11


def test_expv_rejects_m_not_below_n():
    g = load("expv_cases.npz")
    tt = _case(g, 0)  # n = 4
    _, _, rc = P.expv(tt, g["c0_v"], [0.0, 1e-5], m=4)
    assert rc == -3


# --------------------------------------------------------------- physics
B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def test_liouvillian_structure_against_dense_definition():
    nv = P.NVCenter((0, 0, 0), B111)
    a = 3.567
    spins = [P.Spin(np.array([a * 1.0, a * 1.0, 0.0])), P.Spin(np.array([a * 1.25, a * 0.75, -a * 0.25]))]
    prob = P.cluster_problem(spins, B111, nv)
    H0 = P.full(prob["h0"])
    H1 = P.full(prob["h1"])
    L = P.full(prob["lvA"])
    # lvA vec(rho) = vec(rho H0 - H1 rho) in the per-spin block ordering: check on a product operator
    rng = np.random.default_rng(0)
    r1 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    r2 = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
    vec = np.kron(r1.reshape(-1, order="F"), r2.reshape(-1, order="F"))
    rho = np.kron(r1, r2)
    out = L @ vec
    want_mat = rho @ H0 - H1 @ rho
    # map the global matrix to the per-spin block vector ordering
    want = np.zeros(16, complex)
    for i0 in range(2):
        for j0 in range(2):
            for i1 in range(2):
                for j1 in range(2):
                    want[(i0 + 2 * j0) * 4 + (i1 + 2 * j1)] = want_mat[i0 * 2 + i1, j0 * 2 + j1]
    assert np.abs(out - want).max() < 1e-9 * np.abs(want).max()


def test_hahn_cce1_closed_form():
    """One spin-1/2 with conditional fields w0 (m_s=0) and w1 (m_s=1):
    L = 1 - 2 |n0 x n1|^2 sin^2(w0 t/4) sin^2(w1 t/4)."""
    nv = P.NVCenter((0, 0, 0), B111)
    spin = P.Spin(np.array([3.0, 2.0, -4.0]))
    prob = P.cluster_problem([spin], B111, nv)
    tl = np.linspace(0, 2e-4, 41)
    Ld = P.hahn_dense(prob, tl)
    # conditional field vectors from the term tables
    def field(h):
        f = np.zeros(3)
        for kp in h:
            sx, sy, sz = P.spin_ops(2)
            for a, s in enumerate((sx, sy, sz)):
                if np.allclose(kp.mats[0], s):
                    f[a] += kp.coeff
        return f
    f0, f1 = field(prob["h0"]), field(prob["h1"])
    w0, w1 = np.linalg.norm(f0), np.linalg.norm(f1)
    n0, n1 = f0 / w0, f1 / w1
    cr = np.linalg.norm(np.cross(n0, n1)) ** 2
    want = 1 - 2 * cr * np.sin(w0 * tl / 4) ** 2 * np.sin(w1 * tl / 4) ** 2
    assert np.abs(Ld - want).max() < 1e-12


def test_hahn_krylov_oracle_matches_dense_reference_path():
    """The oracle's two-sweep Krylov L(t) equals the reference's production
    dense piecewise propagator (QuantumEvolutionAlgorithm.cpp:60-83)."""
    nv = P.NVCenter((0, 0, 0), B111)
    a = 3.567
    coords = [np.array([a * x for x in v]) for v in ((1, 1, 0), (1.25, 0.75, -0.25), (1.5, 1.0, 0.5))]
    prob = P.cluster_problem([P.Spin(c) for c in coords], B111, nv)
    tl = np.linspace(0, 1e-4, 11)
    tA = P.flatten(prob["lvA"], prob["ldims"])
    tB = P.flatten(prob["lvBdag"], prob["ldims"])
    L, _, rc = P.hahn_batch([tA], [tB], [prob["rho"]], [prob["rho"]], [prob["D"]], tl / 2, m=30, nthreads=1)
    assert rc == 0
    Ld = P.hahn_dense(prob, tl)
    assert np.abs(L[0] - Ld).max() < 1e-10


# --------------------------------------------------------------- enumeration / reduction
def test_enumeration_matches_brute_force():
    rng = np.random.default_rng(5)
    for trial in range(4):
        X = rng.uniform(0, 9.0, size=(11, 3))
        C = P.connection_matrix(X, 4.5)
        cl, _ = P.enumerate_clusters(C, 4)
        bf = P.brute_force_clusters(C, 4)
        for o in range(4):
            assert cl[o] == bf[o]


def _roy():
    path = os.path.join(GOLD, "RoyCoord.xyz")
    lines = open(path).read().split("\n")
    n = int(lines[0])
    return np.array([[float(v) for v in l.split()[1:4]] for l in lines[1:n + 1]])


def test_enumeration_counts_roycoord():
    X = _roy()
    for cut, want in [(6.0, [548, 504, 672, 1107]), (8.0, [548, 1179, 3530, 12662])]:
        cl, _ = P.enumerate_clusters(P.connection_matrix(X, cut), 4)
        assert [len(c) for c in cl] == want


def test_reduction_mobius_identity():
    """CCE with L_C = prod_{S subset C} f_S  (f arbitrary positive) gives
    tilde L_C = f_C, so the final product is prod_C f_C."""
    rng = np.random.default_rng(11)
    X = rng.uniform(0, 8.0, size=(9, 3))
    C = P.connection_matrix(X, 4.0)
    cl, par = P.enumerate_clusters(C, 3)
    nt = 4
    f = [rng.uniform(0.5, 1.5, size=(len(c), nt)) for c in cl]
    index = [{c: i for i, c in enumerate(cl[o])} for o in range(3)]
    L = []
    for o in range(3):
        Lo = np.ones((len(cl[o]), nt))
        for i, c in enumerate(cl[o]):
            for oo in range(o + 1):
                for j, s in enumerate(cl[oo]):
                    if set(s) <= set(c):
                        Lo[i] *= f[oo][j]
        L.append(Lo)
    tilde, final, each = P.cce_reduce(L, par)
    for o in range(3):
        assert np.allclose(tilde[o], f[o], rtol=1e-12)
    want = np.prod(np.concatenate([fo for fo in f]), axis=0)
    assert np.allclose(final[:, -1], want, rtol=1e-12)
