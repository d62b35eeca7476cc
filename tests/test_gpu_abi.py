"""GPU: the link-compatible symbols of the reference's expv library, called
through ctypes exactly as the reference's C++/Fortran would call them
(krylov_expv.h, main_mkl.h, main_cache.h, expokit.h, hamvec_func3.cpp:399,
hamvec_cuda3.cu:661-665), against the golden fixtures (explicit Kronecker
products, scipy expm).  Tolerances: products 1e-12, Krylov results 1e-10
relative to max |w|, the Pade exponential 1e-12 (Frobenius, relative) for
every size, m <= 42 in one workgroup and larger m on rocBLAS / rocSOLVER."""
import ctypes
import os

import numpy as np
import pytest

import oops_amd as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
S = ctypes.c_size_t
I = ctypes.c_int
D = ctypes.c_double


def _g():
    return np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dense_H(g, c):
    """H from the case's terms by explicit Kronecker products (spin 0 slowest)."""
    pre = f"c{c}_"
    dims = g[pre + "spin_dim"]
    n = int(np.prod(dims))
    H = np.zeros((n, n), complex)
    po, pos, mo, mats, nb, coeff = (g[pre + k] for k in ("pos_off", "pos", "mat_off", "mats", "nbody", "coeff"))
    for t in range(len(coeff)):
        ops = [np.eye(d, dtype=complex) for d in dims]
        for b in range(nb[t]):
            body = po[t] + b
            d = int(g[pre + "dims"][body])
            ops[pos[body]] = mats[mo[body]:mo[body] + d * d].reshape(d, d, order="F")
        K = np.ones((1, 1), complex)
        for o in ops:
            K = np.kron(K, o)
        H += coeff[t] * K
    return H


def _sz(a):
    return np.ascontiguousarray(a, np.uint64)


@pytest.mark.gpu
def test_krylov_zgexpv_and_zcooexpv_match_scipy():
    L = O.lib()
    g = _g()
    for c in range(int(g["ncases"])):
        H = _dense_H(g, c)
        n = H.shape[0]
        if n <= 8:
            continue
        A = np.asfortranarray(-1j * H)
        v = np.ascontiguousarray(g[f"c{c}_v"], complex)
        ta = np.ascontiguousarray(g[f"c{c}_t"], float)
        want = g[f"c{c}_w"]
        m = min(30, n - 1)
        w = np.zeros((len(ta), n), complex)
        a_flat = np.ascontiguousarray(A.reshape(-1, order="F"))
        rc = L.krylov_zgexpv(I(n), _p(a_flat), _p(v), I(len(ta)), _p(ta), _p(w), I(10), I(m), D(1e-12), I(0))
        assert rc == 0
        assert np.abs(w.T - want).max() / np.abs(want).max() < 1e-10, c
        r, cidx = np.nonzero(A)
        ia = np.ascontiguousarray(r + 1, np.int32)
        ja = np.ascontiguousarray(cidx + 1, np.int32)
        av = np.ascontiguousarray(A[r, cidx])
        w2 = np.zeros((len(ta), n), complex)
        rc = L.krylov_zcooexpv(I(n), I(len(av)), _p(ia), _p(ja), _p(av), _p(v), I(len(ta)), _p(ta), _p(w2), I(10),
                               I(m), D(1e-12), I(0))
        assert rc == 0
        assert np.abs(w2.T - want).max() / np.abs(want).max() < 1e-10, c


def _table_args(g, c):
    pre = f"c{c}_"
    keep = {k: _sz(g[pre + k]) for k in ("nbody", "pos_off", "pos", "dims", "mat_off", "spin_dim")}
    keep["mats"] = np.ascontiguousarray(g[pre + "mats"], complex)
    keep["coeff"] = np.ascontiguousarray(g[pre + "coeff"], float)
    keep["n"] = int(np.prod(g[pre + "spin_dim"]))
    return keep


@pytest.mark.gpu
def test_main_mkl_and_main_cache_match_scipy():
    L = O.lib()
    g = _g()
    for c in range(int(g["ncases"])):
        k = _table_args(g, c)
        n = k["n"]
        if n <= 8:
            continue
        v = np.ascontiguousarray(g[f"c{c}_v"], complex)
        ta = np.ascontiguousarray(g[f"c{c}_t"], float)
        want = g[f"c{c}_w"]
        common = [ctypes.byref(S(len(k["spin_dim"]))), ctypes.byref(S(len(k["coeff"]))), _p(k["coeff"]),
                  _p(k["nbody"]), _p(k["pos_off"]), _p(k["pos"]), _p(k["dims"]), _p(k["mat_off"]), _p(k["mats"]),
                  ctypes.byref(S(n)), _p(k["spin_dim"]), _p(v), ctypes.byref(S(len(k["pos"]))),
                  ctypes.byref(S(len(k["mats"]))), ctypes.byref(S(10))]
        tail = [ctypes.byref(S(len(ta))), _p(ta), ctypes.byref(S(min(30, n - 1))), ctypes.byref(D(1e-12)),
                ctypes.byref(S(0))]
        w = np.zeros((len(ta), n), complex)
        L.main_mkl_(*common, *tail, _p(w), ctypes.byref(S(w.size)))
        assert np.abs(w.T - want).max() / np.abs(want).max() < 1e-10, c
        w2 = np.zeros((len(ta), n), complex)
        L.main_cache_(*common, ctypes.byref(S(256)), _p(np.array([2 ** 31 - 1, 65535, 65535], np.uint64)), *tail,
                      _p(w2), ctypes.byref(S(w2.size)))
        assert np.abs(w2.T - want).max() / np.abs(want).max() < 1e-10, c


def _hip():
    """The HIP runtime the library itself uses (already loaded with it)."""
    O.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _to_dev(h, arr):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
    assert h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
    return p


@pytest.mark.gpu
def test_hamvec_func3_and_hamvec_cuda3_with_complex_coefficients():
    hip = _hip()
    L = O.lib()
    g = _g()
    for c in range(int(g["ncases"])):
        k = _table_args(g, c)
        n = k["n"]
        x = np.ascontiguousarray(g[f"c{c}_x"], complex)
        want = -1j * g[f"c{c}_Hx"]                      # coefficients -i c_t, as MAIN_CACHE passes them
        cz = np.ascontiguousarray(-1j * k["coeff"], complex)
        w = np.zeros(n, complex)
        med = np.zeros(n, complex)
        args = [ctypes.byref(I(len(k["spin_dim"]))), ctypes.byref(I(len(k["coeff"]))), _p(cz), _p(k["nbody"]),
                _p(k["pos_off"]), _p(k["pos"]), _p(k["dims"]), _p(k["mat_off"])]
        strides = _sz(np.ones(len(k["spin_dim"])))
        L.hamvec_func3_(*args, _p(k["mats"]), ctypes.byref(S(n)), _p(k["spin_dim"]), _p(strides), _p(strides),
                        _p(x), _p(w), _p(med))
        assert np.abs(w - want).max() <= 1e-12 * np.abs(want).max(), c
        h = S(0)
        L.hamvec_cuda3_init_(ctypes.byref(h))
        assert h.value != 0
        dm = _to_dev(hip, k["mats"])
        dv = _to_dev(hip, x)
        dw = _to_dev(hip, np.zeros(n, complex))
        L.hamvec_cuda3_(ctypes.byref(h), *args, ctypes.byref(S(dm.value)), ctypes.byref(S(n)),
                        _p(k["spin_dim"]), _p(strides), _p(strides), ctypes.byref(S(dv.value)),
                        ctypes.byref(S(dw.value)), ctypes.byref(S(0)), ctypes.byref(S(0)), ctypes.byref(S(256)),
                        _p(np.array([2 ** 31 - 1, 65535, 65535], np.uint64)))
        got = np.zeros(n, complex)
        assert hip.hipMemcpy(got.ctypes.data, dw, got.nbytes, 2) == 0  # device to host
        for ptr in (dm, dv, dw):
            hip.hipFree(ptr)
        L.hamvec_cuda3_term_(ctypes.byref(h))
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), c


@pytest.mark.gpu
def test_zgpadm_matches_scipy_fixtures():
    L = O.lib()
    g = np.load(os.path.join(GOLD, "expm_cases.npz"), allow_pickle=False)
    off = 0
    for n, t in zip(g["sizes"], g["t"]):
        n = int(n)
        A = np.ascontiguousarray(g["mats"][off:off + n * n])
        E = g["outs"][off:off + n * n]
        off += n * n
        wsp = np.zeros(4 * n * n + 7, complex)
        ipiv = np.zeros(n, np.int32)
        iexph, ns, iflag = I(0), I(0), I(0)
        L.zgpadm_(ctypes.byref(I(6)), ctypes.byref(I(n)), ctypes.byref(D(float(t))), _p(A), ctypes.byref(I(n)),
                  _p(wsp), ctypes.byref(I(len(wsp))), _p(ipiv), ctypes.byref(iexph), ctypes.byref(ns),
                  ctypes.byref(iflag))
        assert iflag.value == 0
        got = wsp[iexph.value - 1:iexph.value - 1 + n * n]
        assert np.linalg.norm(got - E) / np.linalg.norm(E) < 1e-12, (n, t)


def _zgpadm(A, t):
    n = A.shape[0]
    a = np.asfortranarray(A).reshape(-1, order="F").copy()
    wsp = np.zeros(4 * n * n + 7, complex)
    ipiv = np.zeros(n, np.int32)
    iexph, ns, iflag = I(0), I(0), I(0)
    O.lib().zgpadm_(ctypes.byref(I(6)), ctypes.byref(I(n)), ctypes.byref(D(float(t))), _p(a), ctypes.byref(I(n)),
                    _p(wsp), ctypes.byref(I(len(wsp))), _p(ipiv), ctypes.byref(iexph), ctypes.byref(ns),
                    ctypes.byref(iflag))
    E = wsp[iexph.value - 1:iexph.value - 1 + n * n].reshape(n, n, order="F")
    return E, ns.value, iflag.value


@pytest.mark.gpu
@pytest.mark.parametrize("n,s", [(43, 0.2), (64, 3.0), (256, 40.0), (1024, 8.0)])
def test_zgpadm_dense_sizes_match_scipy(n, s):
    """m above the one-workgroup limit (42): the dense propagators the
    reference's MatExp::pade_exp_mat takes from ZGPADM (MatExp.cpp:29-77,
    d = 4^k up to 4096), on rocBLAS products and a rocSOLVER solve, against
    scipy expm; the squaring count is ZGPADM's (expokit.f:798-800)."""
    import scipy.linalg as sla
    rng = np.random.default_rng(n)
    A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * s / np.sqrt(n)
    t = 0.9
    E, ns, iflag = _zgpadm(A, t)
    want = sla.expm(t * A)
    hnorm = abs(t) * np.abs(A).sum(axis=1).max()
    print(f"zgpadm n={n}: ns {ns}, |E - expm| / |expm| {np.linalg.norm(E - want) / np.linalg.norm(want):.1e}",
          flush=True)
    assert iflag == 0
    assert ns == max(0, int(np.log(hnorm) / np.log(2.0)) + 2)
    assert np.linalg.norm(E - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.gpu
def test_zgpadm_dense_liouville_propagator():
    """The dense propagator of a 5-spin cluster Liouvillian (1024 x 1024, the
    reference's dense-propagator CCE step, QuantumEvolutionAlgorithm.cpp:60-83)
    through zgpadm_ against scipy expm; unitary-in-Liouville-space checks
    ride along (exp(-i lvA dt) applied to rho0 keeps the trace)."""
    import scipy.linalg as sla
    from oracle import physics as P
    X = O.bath_diamond(3.567, 14.0, 0.05, seed=2)
    C = O.Clusters(X, 8.0, 5)
    cl = C.get(4)
    assert len(cl) > 0
    nv = P.NVCenter((0, 0, 0), np.ones(3) * 0.05 / np.sqrt(3.0))
    prob = P.cluster_problem([P.Spin(c) for c in X[list(cl[0])]], np.ones(3) * 0.05 / np.sqrt(3.0), nv)
    n = int(np.prod(prob["ldims"]))
    A = -1j * P.full(prob["lvA"], n)
    dt = 2e-6
    E, ns, iflag = _zgpadm(A, dt)
    want = sla.expm(dt * A)
    err = np.linalg.norm(E - want) / np.linalg.norm(want)
    print(f"Liouville-5 propagator: ns {ns}, rel err {err:.1e}", flush=True)
    assert iflag == 0 and err < 1e-12
    assert abs(np.vdot(prob["rho"], E @ prob["rho"]) - np.vdot(prob["rho"], want @ prob["rho"])) < 1e-12
