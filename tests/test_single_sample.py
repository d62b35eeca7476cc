"""Single-sample CCE (a pure bath state, source/app/single_sample_cce.cpp:35-128):
pins of the oracle's restatement and the library's host builder, on the CPU.

* the bath states come from glibc's rand() after srand(seed): the oracle's
  restatement of the generator is checked against the C library itself;
* the bath mean field against hand-derived values for spins on a line;
* the single-spin Hahn echo and CPMG-N of a pure state against 2x2 rotation
  closed forms with hand-derived fields;
* the library's Hamiltonians and product states against the oracle's."""
import ctypes

import numpy as np
import pytest

import oops_amd as O
from oracle import physics as P

SQ3 = np.sqrt(3.0)
N111 = np.ones(3) / SQ3
B111 = 0.05 * N111
PREF_CC_200 = 5967.5365   # hbar mu0/(4 pi) gamma_C^2 / (2 A)^3, rad/s (CODATA 2018, 10.7084 MHz/T)


@pytest.mark.parametrize("seed", [0, 1, 7, 12345, 2 ** 31 - 1, 2 ** 31 + 5, 2 ** 32 - 1])
def test_glibc_rand_restatement_matches_the_c_library(seed):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    g = P.glibc_rand(seed)
    assert [libc.rand() for _ in range(3000)] == [next(g) for _ in range(3000)]


def test_library_bath_states_match_oracle():
    for seed in (0, 3, 99, 2 ** 31 + 1):
        assert np.array_equal(O.bath_states(seed, 2000), P.bath_states(seed, 2000))
    s = O.bath_states(7, 4000)
    assert 0.45 < s.mean() < 0.55


def test_bath_field_closed_form_on_a_line():
    """Spin 0 at the origin, spin 1 2 A above it: J = PREF diag(1, 1, -2), so
    spin 1 in state 0 (<Sz> = +1/2) gives (0, 0, -PREF) and in state 1
    (0, 0, +PREF); a spin of the cluster itself contributes nothing."""
    bath = [P.Spin(np.zeros(3)), P.Spin(np.array([0.0, 0.0, 2.0])), P.Spin(np.array([0.0, 0.0, 4.0]))]
    f = P.bath_field(bath[0], bath, [0, 0, 1], exclude=[0, 2])
    assert np.abs(f - np.array([0.0, 0.0, -PREF_CC_200])).max() <= 1e-5 * PREF_CC_200
    f = P.bath_field(bath[0], bath, [0, 1, 1], exclude=[0, 2])
    assert np.abs(f - np.array([0.0, 0.0, PREF_CC_200])).max() <= 1e-5 * PREF_CC_200
    # spin 2 at 4 A adds 1/8 of the same (both excluded above)
    f = P.bath_field(bath[0], bath, [0, 0, 0], exclude=[0])
    assert np.abs(f - np.array([0.0, 0.0, -PREF_CC_200 * (1 + 1 / 8)])).max() <= 1e-5 * PREF_CC_200
    assert not np.any(P.bath_field(bath[0], bath, [0, 0, 0], exclude=[0, 1, 2]))


SIG = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]], complex),
       np.array([[1, 0], [0, -1]], complex)]


def _rot(w, t):
    a = np.linalg.norm(w)
    ns = sum(c * s for c, s in zip(w / a, SIG))
    return np.cos(a * t / 2) * np.eye(2) - 1j * np.sin(a * t / 2) * ns


def _pure_closed_form(w0, w1, psi, npulse, t):
    """psi_1 under segments H0, H1, H0, ... and psi_2 under H1, H0, ..., with
    pi pulses at t(2i+1)/(2N) (end intervals t/(2N), inner t/N);
    L = Re <psi_1, psi_2>."""
    segs = [t / 2, t / 2] if npulse == 1 else [t / (2 * npulse)] + [t / npulse] * (npulse - 1) + [t / (2 * npulse)]
    p1, p2 = psi.copy(), psi.copy()
    for j, s in enumerate(segs):
        p1 = _rot(w0 if j % 2 == 0 else w1, s) @ p1
        p2 = _rot(w1 if j % 2 == 0 else w0, s) @ p2
    return float(np.real(np.vdot(p1, p2)))


@pytest.mark.parametrize("npulse", [1, 2, 4])
@pytest.mark.parametrize("bstate", [0, 1])
def test_single_spin_pure_state_against_closed_form(npulse, bstate):
    """A 13C 5 A above the NV electron with one bath spin 2 A above it: the
    conditional fields are worked out by hand (Zeeman -gamma B, the m_s = +1
    hyperfine field PREF_eC (1, 1, -2)/sqrt3, the bath field (0, 0, -/+PREF)),
    and L(t) of the oracle's dense evolution equals the rotation closed form."""
    nv = P.NVCenter((0.0, 0.0, 0.0), B111)
    e = nv.espin.coord
    X = np.array([e + [0.0, 0.0, 5.0], e + [0.0, 0.0, 7.0]])
    states = np.array([bstate, 0], np.int32)
    prob = P.single_sample_problem(X, [0], B111, nv, states)
    g = P.ISOTOPES["13C"][1]
    pref_ec = P.HBAR * P.MU0 / (4 * P.PI) * P.ISOTOPES["E"][1] * g / (5e-10) ** 3
    pref_cc = P.HBAR * P.MU0 / (4 * P.PI) * g * g / (2e-10) ** 3
    bath = np.array([0.0, 0.0, -pref_cc])          # bath spin in state 0: <Sz> = +1/2
    w0 = -g * 0.05 * N111 + bath
    w1 = w0 + pref_ec * np.array([1.0, 1.0, -2.0]) / SQ3
    tl = np.linspace(0.0, 6e-5 * max(1, npulse / 2), 31)
    got = P.single_sample_dense(prob, tl, npulse)
    want = np.array([_pure_closed_form(w0, w1, prob["psi"], npulse, t) for t in tl])
    assert np.abs(got - want).max() < 1e-10
    assert np.ptp(want) > 1e-3


def test_library_single_sample_terms_match_oracle():
    X = O.bath_diamond(3.567, 15.0, 0.011, seed=3)
    states = O.bath_states(11, len(X))
    par = O.cce_params(B=B111)
    nv = P.NVCenter((0.0, 0.0, 0.0), B111)
    C = O.Clusters(X, 8.0, 3)
    for order in range(3):
        for cl in C.get(order)[:3]:
            prob = P.single_sample_problem(X, list(cl), B111, nv, states)
            for which, key in ((0, "h0"), (1, "h1")):
                T, psi = O.single_sample_terms(X, cl, par, states, which)
                oT = P.TermTable(T.spin_dim.size, T.coeff, T.nbody, T.pos_offset, T.pos, T.dims, T.mat_offset,
                                 T.mats, T.spin_dim, T.n)
                M = np.array([P.hamvec(oT, e) for e in np.eye(T.n)]).T
                want = P.full(prob[key], T.n)
                assert np.abs(M - want).max() <= 1e-13 * np.abs(want).max()
                assert np.array_equal(psi, prob["psi"])
