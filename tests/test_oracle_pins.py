"""Closed-form pins of the oracle's coupling coefficients and pulse sequences.

Each check uses numbers derived by hand for a special geometry or from
published constants (CODATA 2018, NMR tables), never the oracle's own general
formula, so that a plausible mistake in the oracle (a dropped unit factor, a
wrong sign, a transposed tensor, a swapped segment) fails one of them.

Reference functions pinned (paths relative to the reference's src/):
  dipole / zeeman / dipole_field   source/misc/misc.cpp:11-57
  isotope constants                source/spin/Spin/SpinData.cpp:4-17
  NV eigenstates                   source/app/DefectCenter.cpp:14-45
  Pulse_Interval / riffle          source/misc/misc.cpp:59-84, include/misc/misc.h:25-40
  CPMG evolution                   source/app/ensemble_cce.cpp:33-55,
                                   source/quantum/QuantumEvolution/QuantumEvolutionAlgorithm.cpp:60-83
"""
import numpy as np
import pytest

from oracle import physics as P

SQ3 = np.sqrt(3.0)
N111 = np.ones(3) / SQ3
# 13C-13C dipolar prefactor hbar (mu0 / 4 pi) gamma_C^2 / r^3 in rad/s, worked
# out by hand from CODATA 2018 hbar = 1.054571817e-34 J s, mu0 / 4 pi = 1e-7
# and gamma_C / 2 pi = 10.7084 MHz/T:
PREF_CC_154 = 13071.424   # r = 1.54 A (the C-C bond): 2080 Hz, the textbook ~2.1 kHz
PREF_CC_200 = 5967.5365   # r = 2 A
# NV electron - 13C at 5 A (gamma_e / 2 pi = -28.02495 GHz/T): -159.08 kHz
PREF_EC_500 = -999528.81
# the constants above come from slightly different gamma / hbar digits than
# the reference's table; they agree to a few parts in 1e6
REL_CONST = 1e-5


def _spin(xyz, iso="13C"):
    return P.Spin(np.asarray(xyz, float), iso)


def test_isotope_constants_against_published_values():
    assert P.ISOTOPES["13C"][1] / (2 * np.pi) == pytest.approx(10.7084e6, rel=1e-5)
    assert P.ISOTOPES["E"][1] / (2 * np.pi) == pytest.approx(-28.02495e9, rel=1e-6)
    assert P.ISOTOPES["NVe"][2] == pytest.approx(2870.0)          # D in MHz
    assert P.ISOTOPES["13C"][0] == 2 and P.ISOTOPES["NVe"][0] == 3
    assert P.HBAR == pytest.approx(1.054571817e-34, rel=1e-7)
    assert P.MU0 / (4 * np.pi) == pytest.approx(1e-7, rel=1e-9)


def test_dipole_pair_along_z_is_diag_1_1_minus2():
    J = P.dipole(_spin((0.0, 0.0, 0.0)), _spin((0.0, 0.0, 1.54))).reshape(3, 3)
    want = PREF_CC_154 * np.diag([1.0, 1.0, -2.0])
    assert np.abs(J - want).max() <= REL_CONST * PREF_CC_154 * 2
    assert J[0, 0] / (2 * np.pi) == pytest.approx(2080.4, rel=1e-4)
    # the tensor does not depend on which spin is first
    J2 = P.dipole(_spin((0.0, 0.0, 1.54)), _spin((0.0, 0.0, 0.0))).reshape(3, 3)
    assert np.array_equal(J, J2)
    # coincident spins give no coupling (DISTANCE_EPSILON)
    assert not np.any(P.dipole(_spin((1.0, 2.0, 3.0)), _spin((1.0, 2.0, 3.0))))


def test_dipole_pair_along_111_has_minus_pref_off_diagonal():
    r = 2.0 * N111
    J = P.dipole(_spin(r), _spin((0.0, 0.0, 0.0))).reshape(3, 3)
    want = PREF_CC_200 * np.array([[0.0, -1.0, -1.0], [-1.0, 0.0, -1.0], [-1.0, -1.0, 0.0]])
    assert np.abs(J - want).max() <= REL_CONST * PREF_CC_200
    # and along x: diag(-2, 1, 1), inverse-cube scaling with distance
    Jx = P.dipole(_spin((4.0, 0.0, 0.0)), _spin((0.0, 0.0, 0.0))).reshape(3, 3)
    assert np.abs(Jx - PREF_CC_200 / 8.0 * np.diag([-2.0, 1.0, 1.0])).max() <= REL_CONST * PREF_CC_200


def test_zeeman_sign_and_quadrupole_slots():
    g = P.ISOTOPES["13C"][1]
    B = np.array([0.1, -0.2, 0.3])
    z = P.zeeman(_spin((0, 0, 0)), B)
    assert np.allclose(z, [-0.1 * g, 0.2 * g, -0.3 * g, 0.0, 0.0, 0.0], rtol=0, atol=1e-9)
    # physical sign: H = -gamma B.S, so for gamma > 0 the lower level has the
    # spin along B, split by gamma |B|
    sx, sy, sz = P.spin_ops(2)
    H = z[0] * sx + z[1] * sy + z[2] * sz
    ev, vec = np.linalg.eigh(H)
    Bn = np.linalg.norm(B)
    assert np.allclose(ev, [-g * Bn / 2, g * Bn / 2], rtol=1e-12)
    low = vec[:, 0]
    s_low = np.real([low.conj() @ o @ low for o in (sx, sy, sz)])
    assert np.allclose(s_low, 0.5 * B / Bn, atol=1e-12)
    # spin-1 slots (e/3, -e/3, q) multiply Sx^2, Sy^2, Sz^2: the NV zero-field
    # splitting D = 2.87 GHz sits in the Sz^2 slot
    znv = P.zeeman(_spin((0, 0, 0), "NVe"), np.zeros(3))
    assert np.array_equal(znv, [0.0, 0.0, 0.0, 0.0, 0.0, 2870.0])


def _nv():
    return P.NVCenter((0.0, 0.0, 0.0), 0.05 * N111)


def test_nv_eigenstates_along_111():
    nv = _nv()
    D = 2 * np.pi * 2.87e9
    geB = 1.760859708e11 * 0.05
    assert np.allclose(nv.evals, [0.0, D - geB, D + geB], rtol=1e-12, atol=1e-3)
    s = [P.spin_vector(3, nv.state(k)) for k in range(3)]
    assert np.abs(s[0]).max() < 1e-12                    # m_s = 0
    assert np.abs(s[1] - N111).max() < 1e-12             # m_s = +1 (gamma_e < 0: the lower of the two)
    assert np.abs(s[2] + N111).max() < 1e-12             # m_s = -1


def test_dipole_field_of_nv_on_13c_along_z():
    nv = _nv()
    e = nv.espin.coord
    c = _spin(e + np.array([0.0, 0.0, 5.0]))
    # J = PREF_EC diag(1, 1, -2) for a spin straight above the electron;
    # m_s = +1 along [111] gives <S> = (1,1,1)/sqrt3, so the field is
    # PREF_EC (1, 1, -2)/sqrt3; m_s = 0 gives none
    f1 = P.dipole_field(c, nv.espin, nv.state(1))
    want = PREF_EC_500 * np.array([1.0, 1.0, -2.0]) / SQ3
    assert np.abs(f1 - want).max() <= REL_CONST * abs(PREF_EC_500)
    f0 = P.dipole_field(c, nv.espin, nv.state(0))
    assert np.abs(f0).max() < 1e-9 * abs(PREF_EC_500)
    # off-axis geometry: the tensor is symmetric, so the field is the same for
    # either reshape orientation; check it against J computed by hand for r along x
    cx = _spin(e + np.array([5.0, 0.0, 0.0]))
    fx = P.dipole_field(cx, nv.espin, nv.state(1))
    assert np.abs(fx - PREF_EC_500 * np.array([-2.0, 1.0, 1.0]) / SQ3).max() <= REL_CONST * abs(PREF_EC_500)


def _cce1_fields():
    """Conditional fields of a 13C 5 A above the NV electron, by hand."""
    g = P.ISOTOPES["13C"][1]
    B = 0.05
    pref = P.HBAR * P.MU0 / (4 * P.PI) * P.ISOTOPES["E"][1] * g / (5e-10) ** 3   # pinned constants
    w0 = -g * B * N111
    w1 = -g * B * N111 + pref * np.array([1.0, 1.0, -2.0]) / SQ3
    return w0, w1


def test_hahn_cce1_with_fields_derived_by_hand():
    """The whole single-spin chain (coupling, Zeeman, mean field, Liouvillian,
    density state, observable) against L = 1 - 2|n0 x n1|^2 sin^2(w0 t/4) sin^2(w1 t/4)."""
    nv = _nv()
    spin = _spin(nv.espin.coord + np.array([0.0, 0.0, 5.0]))
    prob = P.cluster_problem([spin], 0.05 * N111, nv)
    tl = np.linspace(0.0, 4e-5, 81)
    Ld = P.hahn_dense(prob, tl)
    w0, w1 = _cce1_fields()
    a0, a1 = np.linalg.norm(w0), np.linalg.norm(w1)
    cr = np.linalg.norm(np.cross(w0 / a0, w1 / a1)) ** 2
    want = 1 - 2 * cr * np.sin(a0 * tl / 4) ** 2 * np.sin(a1 * tl / 4) ** 2
    assert cr > 0.1                      # a real echo modulation, not L = 1
    assert np.abs(Ld - want).max() < 1e-11


def test_pulse_interval_and_riffle_hand_values():
    assert P.pulse_interval(1) == [0.5, 0.5]
    assert P.pulse_interval(2) == [0.25, 0.5, 0.25]
    assert np.allclose(P.pulse_interval(4), [1 / 8, 1 / 4, 1 / 4, 1 / 4, 1 / 8], rtol=0, atol=1e-16)
    assert np.allclose(P.pulse_interval(3), [1 / 6, 1 / 3, 1 / 3, 1 / 6], rtol=0, atol=1e-16)
    for n in (1, 2, 4, 8, 16):
        assert sum(P.pulse_interval(n)) == pytest.approx(1.0, abs=1e-15)
    assert P.riffle("A", "B", 1) == ["A", "B"]
    assert P.riffle("A", "B", 2) == ["A", "B", "A"]
    assert P.riffle("A", "B", 3) == ["A", "B", "A", "B"]
    assert P.riffle("A", "B", 4) == ["A", "B", "A", "B", "A"]


SIG = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]], complex),
       np.array([[1, 0], [0, -1]], complex)]


def _su2(w, t):
    """exp(-i (w.S) t) for S = sigma/2: cos(|w|t/2) I - i sin(|w|t/2) n.sigma."""
    a = np.linalg.norm(w)
    if a == 0.0:
        return np.eye(2, dtype=complex)
    ns = sum(c * s for c, s in zip(w / a, SIG))
    return np.cos(a * t / 2) * np.eye(2) - 1j * np.sin(a * t / 2) * ns


def _cpmg_closed_form(w0, w1, npulse, t):
    """CPMG-N coherence of one spin-1/2 with rho0 = I/2 built from 2x2
    rotations only.  With pi pulses at t(2i+1)/(2N), the electron's coherence
    sees H0 and H1 alternately on the two sides; the i-th free interval has
    length t/(2N) at the ends and t/N inside.  rho(t) = X rho0 Y with
    X = prod exp(+i H_other s), Y = prod exp(-i H_this s) in sequence order,
    and L = Re Tr(Y X) / 2 (written without the oracle's segment helpers)."""
    if npulse == 1:
        segs = [t / 2, t / 2]
    else:
        segs = [t / (2 * npulse)] + [t / npulse] * (npulse - 1) + [t / (2 * npulse)]
    X = np.eye(2, dtype=complex)
    Y = np.eye(2, dtype=complex)
    for j, s in enumerate(segs):
        this, other = (w0, w1) if j % 2 == 0 else (w1, w0)
        X = _su2(other, -s) @ X          # exp(+i H_other s) applied after the earlier segments
        Y = Y @ _su2(this, s)
    return float(np.real(np.trace(Y @ X)) / 2.0)


def _single_spin_terms(w):
    sx, sy, sz = P.spin_ops(2)
    return [P.KronProd((2,), (0,), float(c), (m,)) for c, m in zip(w, (sx, sy, sz)) if c != 0.0]


def _oracle_cpmg(w0, w1, npulse, tl):
    h0, h1 = _single_spin_terms(w0), _single_spin_terms(w1)
    lvA = P.full(P.liouvillian(h0, h1), 4)
    lvB = P.full(P.liouvillian(h1, h0), 4)
    rho, D = P.density_vector([P.Spin(np.zeros(3))])
    ops = P.riffle(lvA, lvB, npulse)
    st = P.piecewise_dense(ops, P.pulse_interval(npulse), tl[1] - tl[0], rho, len(tl))
    return np.real(D * (rho.conj() @ st))


@pytest.mark.parametrize("npulse", [1, 2, 4, 8, 16])
def test_cpmg_against_su2_closed_form(npulse):
    w0 = np.array([1.3e5, -4.0e4, 2.2e5])
    w1 = np.array([-7.0e4, 1.9e5, 6.0e4])
    tl = np.linspace(0.0, 6e-5 * max(1.0, npulse / 2.0), 61)
    got = _oracle_cpmg(w0, w1, npulse, tl)
    want = np.array([_cpmg_closed_form(w0, w1, npulse, t) for t in tl])
    assert np.abs(got - want).max() < 1e-12
    assert np.ptp(want) > 0.01           # the sequence really modulates L
    # a segment mistake (equal segments) would be caught at these fields
    if npulse >= 2:
        eq = np.array([_cpmg_closed_form_equal(w0, w1, npulse, t) for t in tl])
        assert np.abs(eq - want).max() > 1e-3


def _cpmg_closed_form_equal(w0, w1, npulse, t):
    segs = [t / (npulse + 1)] * (npulse + 1)
    X = np.eye(2, dtype=complex)
    Y = np.eye(2, dtype=complex)
    for j, s in enumerate(segs):
        this, other = (w0, w1) if j % 2 == 0 else (w1, w0)
        X = _su2(other, -s) @ X
        Y = Y @ _su2(this, s)
    return float(np.real(np.trace(Y @ X)) / 2.0)


@pytest.mark.parametrize("npulse", [1, 2, 4, 8, 16])
def test_cpmg_refocuses_parallel_fields(npulse):
    """w0 || w1: the two conditional Hamiltonians commute and CPMG-N spends
    half the time under each on either side, so L = 1 at every time."""
    w0 = np.array([1.0e5, 2.0e5, -5.0e4])
    tl = np.linspace(0.0, 2e-4, 41)
    got = _oracle_cpmg(w0, 0.37 * w0, npulse, tl)
    assert np.abs(got - 1.0).max() < 1e-12


@pytest.mark.parametrize("npulse", [1, 2, 8])
def test_cpmg_cluster_problem_matches_closed_form(npulse):
    """The NV + one 13C problem the CCE pipeline builds, under CPMG-N, against
    the SU(2) closed form with the hand-derived fields."""
    nv = _nv()
    spin = _spin(nv.espin.coord + np.array([0.0, 0.0, 5.0]))
    prob = P.cluster_problem([spin], 0.05 * N111, nv)
    tl = np.linspace(0.0, 3e-5, 31)
    A = P.full(prob["lvA"], 4)
    Bm = P.full(prob["lvB"], 4)
    st = P.piecewise_dense(P.riffle(A, Bm, npulse), P.pulse_interval(npulse), tl[1] - tl[0], prob["rho"], len(tl))
    got = np.real(prob["D"] * (prob["rho"].conj() @ st))
    w0, w1 = _cce1_fields()
    want = np.array([_cpmg_closed_form(w0, w1, npulse, t) for t in tl])
    assert np.abs(got - want).max() < 1e-10


def test_product_dlaran_fill_matches_reference_sequence():
    """The library's DLARAN (host code, no GPU) against the committed draws of
    the reference's DLARAN.f from seed (0, 0, 0, 5)."""
    import os
    import oops_amd as O
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "dlaran_seq.npz"), allow_pickle=False)
    seed = [0, 0, 0, 5]
    seq = O.dlaran_fill(seed, len(g["seq"]))
    assert np.array_equal(seq, g["seq"])
    assert seed == list(g["final_seed"])
    # split draws continue the same stream
    s1 = [0, 0, 0, 5]
    a = O.dlaran_fill(s1, 1000)
    b = O.dlaran_fill(s1, len(g["seq"]) - 1000)
    assert np.array_equal(np.concatenate([a, b]), g["seq"])


@pytest.mark.parametrize("mult", [2, 3])
def test_spin_operators_obey_the_spin_algebra(mult):
    """[Sx, Sy] = i Sz (cyclic), S^2 = s(s+1), Sz = diag(s, ..., -s)."""
    sx, sy, sz = P.spin_ops(mult)
    s = (mult - 1) / 2.0
    for a, b, c in ((sx, sy, sz), (sy, sz, sx), (sz, sx, sy)):
        assert np.abs(a @ b - b @ a - 1j * c).max() < 1e-15
    assert np.abs(sx @ sx + sy @ sy + sz @ sz - s * (s + 1) * np.eye(mult)).max() < 1e-15
    assert np.allclose(np.diag(sz), np.arange(s, -s - 1, -1))
