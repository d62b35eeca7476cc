"""ORACLE — TEST INFRASTRUCTURE ONLY.

numpy restatement of the reference's operator construction, cluster
enumeration and CCE reduction.  Imported only by tests/, by
``__graft_entry__.smoke()`` and by bench.py's ``cpu_baseline`` leg, as the
checker.  Nothing in the product path imports this module.

Reference anchors (paths relative to the reference's ``src/``):
  spin operators            source/spin/Spin/Spin.cpp:28-95
  isotope constants         source/spin/Spin/SpinData.cpp:4-17
  dipole / zeeman / field   source/misc/misc.cpp:11-57
  term forms and order      source/spin/SpinInteraction/SpinInteractionComponent.cpp:184-304,
                            source/spin/SpinInteraction/SpinInteraction.cpp:29-65 (``!= 0`` filter)
  Hamiltonian assembly      source/app/ensemble_cce.cpp:57-71
  Liouvillian lvA / lvB     source/app/ensemble_cce.cpp:41-42, 85-91
  SHARP / FLAT / Expand     source/kron/KronProd.cpp:9-28, 84-107, 169-180
  operator minus            source/quantum/QuantumOperator/QuantumOperator.cpp:33-49
  density vector            source/spin/SpinInteraction/SpinState.cpp:9-23, include/quantum/MixedState.h:44,
                            source/kron/KronProd.cpp:68-82
  observable                source/app/ensemble_cce.cpp:104-113
  NV eigenstates            source/app/DefectCenter.cpp:14-45
  CPMG segments / riffle    source/misc/misc.cpp:59-84, include/misc/misc.h:25-40
  enumeration               source/spin/SpinCluster/SpinClusterAlgorithm.cpp:28-148,
                            source/spin/SpinCluster/SpinClusterIndex.cpp:55-64,
                            source/spin/SpinCollection/SpinCollection.cpp:31-57
  sub-clusters              source/spin/SpinCluster/SpinCluster.cpp:82-98
  reduction / product       source/app/cce.cpp:151-189
  single-sample CCE         source/app/single_sample_cce.cpp:35-128 (bath states, mean
                            field, Hilbert sweeps, observable), bath field with an
                            exclusion mask SpinInteraction.cpp:159-173,
                            SpinInteractionComponent.cpp:433-451
"""
from __future__ import annotations

import ctypes
import itertools
import os
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# constants (Armadillo ``datum`` values are version dependent; these are the
# values this project fixes, shared with the product code)
HBAR = 1.0545718e-34
MU0 = 1.25663706143592e-06
PI = 3.14159265358979323846
DISTANCE_EPSILON = 1e-10
DIAMOND_LATTICE_CONST_NV = 3.57  # DefectCenter.cpp:16

# isotope -> (multiplicity, gamma [rad/s/T], omegaQ, eta)   SpinData.cpp:4-17
ISOTOPES = {
    "13C": (2, 6.728284e7, 0.0, 0.0),
    "14C": (3, 1.9337792e7, 2.96, 0.0),   # the reference's key for 14N data
    "15C": (2, -2.71261804e7, 0.0, 0.0),  # the reference's key for 15N data
    "E": (2, -1.760859708e11, 0.0, 0.0),
    "NVe": (3, -1.760859708e11, 2.87e3, 0.0),
}


def spin_ops(mult: int):
    """(sx, sy, sz) for multiplicity 2 or 3 (Spin.cpp:28-95)."""
    if mult == 2:
        sx = 0.5 * np.array([[0, 1], [1, 0]], dtype=complex)
        sy = 0.5 * np.array([[0, -1j], [1j, 0]], dtype=complex)
        sz = 0.5 * np.array([[1, 0], [0, -1]], dtype=complex)
    elif mult == 3:
        sx = np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]], dtype=complex) / np.sqrt(2.0)
        sy = np.array([[0, -1j, 0], [1j, 0, -1j], [0, 1j, 0]], dtype=complex) / np.sqrt(2.0)
        sz = np.array([[1, 0, 0], [0, 0, 0], [0, 0, -1]], dtype=complex)
    else:
        raise ValueError("only spin-1/2 and spin-1")
    return sx, sy, sz


@dataclass
class Spin:
    coord: np.ndarray
    isotope: str = "13C"

    @property
    def mult(self):
        return ISOTOPES[self.isotope][0]

    @property
    def gamma(self):
        return ISOTOPES[self.isotope][1]

    @property
    def omegaQ(self):
        return ISOTOPES[self.isotope][2]

    @property
    def eta(self):
        return ISOTOPES[self.isotope][3]


def dipole(s1: Spin, s2: Spin) -> np.ndarray:
    r = np.asarray(s1.coord, float) - np.asarray(s2.coord, float)
    d = float(np.sqrt(np.dot(r, r)))
    if d <= DISTANCE_EPSILON:
        return np.zeros(9)
    n = r / d
    d0 = d * 1e-10
    pref = HBAR * MU0 / (4.0 * PI) * (s1.gamma * s2.gamma) / (d0 * d0 * d0)
    nx, ny, nz = n
    return pref * np.array([1.0 - 3.0 * nx * nx, -3.0 * nx * ny, -3.0 * nx * nz,
                            -3.0 * ny * nx, 1.0 - 3.0 * ny * ny, -3.0 * ny * nz,
                            -3.0 * nz * nx, -3.0 * nz * ny, 1.0 - 3.0 * nz * nz])


def zeeman(s: Spin, B) -> np.ndarray:
    g = s.gamma
    return np.array([-g * B[0], -g * B[1], -g * B[2], s.eta / 3.0, -s.eta / 3.0, s.omegaQ])


def spin_vector(mult: int, state: np.ndarray) -> np.ndarray:
    sx, sy, sz = spin_ops(mult)
    st = np.asarray(state, complex)
    return np.array([np.real(st.conj() @ op @ st) for op in (sx, sy, sz)])


def dipole_field(s: Spin, src: Spin, src_state: np.ndarray) -> np.ndarray:
    dm = dipole(s, src).reshape(3, 3, order="F")
    return dm @ spin_vector(src.mult, src_state)


class NVCenter:
    """NV electron at nitrogen + (3.57/4)(1,1,1) (DefectCenter.cpp:14-31)."""

    def __init__(self, position=(0.0, 0.0, 0.0), B=(0.0, 0.0, 0.0)):
        pos = np.asarray(position, float)
        d4 = DIAMOND_LATTICE_CONST_NV / 4.0
        self.espin = Spin(pos + np.array([d4, d4, d4]), "NVe")
        self.B = np.asarray(B, float)
        sx, sy, sz = spin_ops(3)
        omegaQ = self.espin.omegaQ * 2.0 * PI * 1e6
        g = self.espin.gamma
        s111 = (sx + sy + sz) / np.sqrt(3.0)
        H = omegaQ * s111 @ s111 + g * (self.B[0] * sx + self.B[1] * sy + self.B[2] * sz)
        self.evals, self.evecs = np.linalg.eigh(H)

    def state(self, idx: int) -> np.ndarray:
        return self.evecs[:, idx]


# ---------------------------------------------------------------------------
# KronProd / SumKronProd
@dataclass
class KronProd:
    dims: tuple
    indices: tuple
    coeff: float
    mats: tuple

    def full(self) -> np.ndarray:
        ops = [np.eye(d, dtype=complex) for d in self.dims]
        for i, m in zip(self.indices, self.mats):
            ops[i] = m
        out = np.ones((1, 1), dtype=complex)
        for o in ops:
            out = np.kron(out, o)
        return self.coeff * out

    def vectorize(self) -> np.ndarray:
        mats = [np.eye(d, dtype=complex) for d in self.dims]
        for i, m in zip(self.indices, self.mats):
            mats[i] = m
        out = np.ones(1, dtype=complex)
        for m in mats:
            out = np.kron(out, m.reshape(-1, order="F"))
        return self.coeff * out


def sharp(m):
    return np.kron(m.T, np.eye(m.shape[0]))


def flat(m):
    return np.kron(np.eye(m.shape[0]), m)


def expand(terms, func):
    return [KronProd(tuple(d * d for d in kp.dims), kp.indices, kp.coeff,
                     tuple(func(np.asarray(m)) for m in kp.mats)) for kp in terms]


def scaled(terms, f):
    return [KronProd(kp.dims, kp.indices, kp.coeff * f, kp.mats) for kp in terms]


def full(terms, dim=None) -> np.ndarray:
    if not terms:
        return np.zeros((dim, dim), dtype=complex)
    out = terms[0].full()
    for kp in terms[1:]:
        out = out + kp.full()
    return out


# ---------------------------------------------------------------------------
# term tables in the reference's order
def _filtered(dims, idx, coeffs, mats):
    out = []
    for c, m in zip(coeffs, mats):
        if c != 0.0:
            out.append(KronProd(tuple(dims), tuple(idx), float(c), tuple(m)))
    return out


def dipolar_terms(spins):
    dims = [s.mult for s in spins]
    out = []
    for i in range(len(spins)):
        for j in range(i + 1, len(spins)):
            a, b = spin_ops(spins[i].mult), spin_ops(spins[j].mult)
            mats = [(a[p], b[q]) for p in range(3) for q in range(3)]
            out += _filtered(dims, (i, j), dipole(spins[i], spins[j]), mats)
    return out


def single_forms(s):
    sx, sy, sz = spin_ops(s.mult)
    return [(sx,), (sy,), (sz,), (sx @ sx,), (sy @ sy,), (sz @ sz,)]


def zeeman_terms(spins, B):
    dims = [s.mult for s in spins]
    out = []
    for i, s in enumerate(spins):
        out += _filtered(dims, (i,), zeeman(s, B), single_forms(s))
    return out


def field_terms(spins, center: Spin, state):
    dims = [s.mult for s in spins]
    out = []
    for i, s in enumerate(spins):
        f = dipole_field(s, center, state)
        out += _filtered(dims, (i,), [f[0], f[1], f[2], 0.0, 0.0, 0.0], single_forms(s))
    return out


def hamiltonian(spins, B, center: Spin, state):
    """create_spin_hamiltonian (ensemble_cce.cpp:57-71): dipolar + Zeeman + field."""
    return dipolar_terms(spins) + zeeman_terms(spins, B) + field_terms(spins, center, state)


def dephasing_terms(spins, rate_hz, axis):
    """create_incoherent_operator (ensemble_cce.cpp:73-84; SpinInteractionComponent.cpp:285-304)."""
    rate = 2.0 * PI * rate_hz
    ax = np.asarray(axis, float)
    ax = ax / np.linalg.norm(ax)
    nx, ny, nz = ax
    sig = np.array([[nz, nx - 1j * ny], [nx + 1j * ny, -nz]], dtype=complex)
    dm = 1j * (np.kron(sig.T, sig) - np.eye(4))
    dims = [s.mult ** 2 for s in spins]
    out = []
    for i in range(len(spins)):
        out += _filtered(dims, (i,), [rate], [(dm,)])
    return out


def liouvillian(h_sharp, h_flat, dephase=()):
    """Liouvillian(h_sharp, SHARP) - Liouvillian(h_flat, FLAT) + dephase."""
    return expand(h_sharp, sharp) + scaled(expand(h_flat, flat), -1.0) + list(dephase)


def adjoint_negated(terms):
    """-(sum c kron A)^dagger = sum (-c) kron A^dagger (bra-side generator)."""
    return [KronProd(kp.dims, kp.indices, -kp.coeff, tuple(np.conj(np.asarray(m).T) for m in kp.mats))
            for kp in terms]


def density_vector(spins, pol=(0.0, 0.0, 0.0)):
    """vec rho0 = vecterize(sum pol.S + I/D) in the per-spin-block ordering."""
    dims = [s.mult for s in spins]
    terms = []
    for i, s in enumerate(spins):
        sx, sy, sz = spin_ops(s.mult)
        terms += _filtered(dims, (i,), [pol[0], pol[1], pol[2], 0.0, 0.0, 0.0], single_forms(s))
    D = int(np.prod(dims))
    terms.append(KronProd(tuple(dims), (), 1.0 / D, ()))
    v = terms[0].vectorize()
    for kp in terms[1:]:
        v = v + kp.vectorize()
    return v, D


def pulse_interval(n):
    t = [0.0] + [(2.0 * i + 1.0) / (2.0 * n) for i in range(n)] + [1.0]
    return [t[i + 1] - t[i] for i in range(n + 1)]


def riffle(a, b, n):
    out = []
    for _ in range((n + 1) // 2):
        out += [a, b]
    if n % 2 == 0:
        out.append(a)
    return out


# ---------------------------------------------------------------------------
# flattened term table (MatExp.cpp:216-262)
@dataclass
class TermTable:
    nspin: int
    coeff: np.ndarray
    nbody: np.ndarray
    pos_off: np.ndarray
    pos: np.ndarray
    dims: np.ndarray
    mat_off: np.ndarray
    mats: np.ndarray
    spin_dim: np.ndarray
    n: int
    _keep: list = field(default_factory=list)


def flatten(terms, spin_dim) -> TermTable:
    spin_dim = np.asarray(spin_dim, np.int64)
    coeff = np.array([kp.coeff for kp in terms], float)
    nbody = np.array([len(kp.indices) for kp in terms], np.int64)
    pos_off = np.zeros(len(terms) + 1, np.int64)
    pos_off[1:] = np.cumsum(nbody)
    pos = np.array([i for kp in terms for i in kp.indices], np.int64)
    mats_l = [np.asarray(m, complex) for kp in terms for m in kp.mats]
    dims = np.array([m.shape[0] for m in mats_l], np.int64)
    mat_off = np.zeros(len(mats_l) + 1, np.int64)
    mat_off[1:] = np.cumsum(dims * dims)
    mats = (np.concatenate([m.reshape(-1, order="F") for m in mats_l])
            if mats_l else np.zeros(1, complex))
    return TermTable(len(spin_dim), coeff, nbody, pos_off, pos if pos.size else np.zeros(1, np.int64),
                     dims if dims.size else np.zeros(1, np.int64), mat_off, mats, spin_dim,
                     int(np.prod(spin_dim)))


# ---------------------------------------------------------------------------
# cluster enumeration
def connection_matrix(coords, cutoff):
    """cSpinCollection::make + getConnectionMatrix: d(i,j) for j>=i, symmetrised, <= cutoff."""
    X = np.asarray(coords, float)
    diff = X[:, None, :] - X[None, :, :]
    d = np.sqrt(np.einsum("ijk,ijk->ij", diff, diff))
    iu = np.triu(np.ones(d.shape, bool))
    d = np.where(iu, d, d.T)
    C = (d <= cutoff).astype(np.int8)
    np.fill_diagonal(C, 0)
    return C


def enumerate_clusters(C, max_order):
    """cDepthFirstPathTracing::generate restated with Python sets.

    Returns (clusters, parents): clusters[k] is the sorted list (size, then
    lexicographic) of connected vertex tuples of size k+1; parents[k][i] is
    the list of positions (in clusters[k-1]) of the parents that generated
    cluster i, in insertion order."""
    n = C.shape[0]
    nbrs = [np.nonzero(C[i])[0].tolist() for i in range(n)]
    clusters = [[(i,) for i in range(n)]]
    parents = [[[] for _ in range(n)]]
    for order in range(1, max_order):
        prev = clusters[-1]
        gen = {}
        for pi, par in enumerate(prev):
            pset = set(par)
            cand = set()
            for v in par:
                cand.update(nbrs[v])
            for c in sorted(cand):
                if c in pset:
                    continue
                key = tuple(sorted(par + (c,)))
                gen.setdefault(key, []).append(pi)
        keys = sorted(gen.keys())
        clusters.append(keys)
        parents.append([gen[k] for k in keys])
    return clusters, parents


def brute_force_clusters(C, max_order):
    """All connected induced vertex subsets of size 1..max_order (tiny N only)."""
    n = C.shape[0]
    out = []
    for k in range(1, max_order + 1):
        lst = []
        for sub in itertools.combinations(range(n), k):
            seen = {sub[0]}
            stack = [sub[0]]
            s = set(sub)
            while stack:
                v = stack.pop()
                for u in sub:
                    if u not in seen and C[v, u]:
                        seen.add(u)
                        stack.append(u)
            if len(seen) == k:
                lst.append(sub)
        out.append(lst)
    return out


def sub_clusters(parents, order, index, memo=None):
    """getSubClusters: all (order, position) reachable through parent links."""
    if memo is None:
        memo = {}
    key = (order, index)
    if key in memo:
        return memo[key]
    res = set()
    if order > 0:
        for p in parents[order][index]:
            res.add((order - 1, p))
            res |= sub_clusters(parents, order - 1, p, memo)
    memo[key] = res
    return res


def cce_reduce(L_list, parents):
    """cce_coherence_reduction + compuate_final_coherence (cce.cpp:151-189).

    L_list[k]: array (nclusters_k, nt).  Returns (tilde list, final, each_order)."""
    memo = {}
    tilde = []
    for order, Lk in enumerate(L_list):
        tk = np.array(Lk, float, copy=True)
        for j in range(Lk.shape[0]):
            r = np.array(Lk[j], float, copy=True)
            for (o, p) in sorted(sub_clusters(parents, order, j, memo)):
                r = r / tilde[o][p]
            tk[j] = r
        tilde.append(tk)
    nt = L_list[0].shape[1]
    final = np.zeros((nt, len(L_list)))
    each = np.zeros((nt, len(L_list)))
    acc = np.ones(nt)
    for order, tk in enumerate(tilde):
        r = np.ones(nt)
        for j in range(tk.shape[0]):
            r = r * tk[j]
        each[:, order] = r
        acc = acc * r
        final[:, order] = acc
    return tilde, final, each


# ---------------------------------------------------------------------------
# ctypes binding of the C++ restatement (oracle/expv_oracle.cpp)
class OrcExpvStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in
                ("nmult", "nexph", "nscale", "nstep", "nreject", "ibrkflag", "mbrkdwn", "iflag")] + \
               [(n, ctypes.c_double) for n in
                ("step_min", "step_max", "x_error", "s_error", "tbrkdwn", "t_now", "hump",
                 "scaled_norm", "anorm")]


class OrcTermTable(ctypes.Structure):
    _fields_ = [("nspin", ctypes.c_int32), ("nterm", ctypes.c_int32),
                ("coeff", ctypes.c_void_p), ("nbody", ctypes.c_void_p), ("pos_off", ctypes.c_void_p),
                ("pos", ctypes.c_void_p), ("dims", ctypes.c_void_p), ("mat_off", ctypes.c_void_p),
                ("mats", ctypes.c_void_p), ("spin_dim", ctypes.c_void_p), ("n", ctypes.c_int64)]


_LIB = None
HERE = os.path.dirname(os.path.abspath(__file__))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liborc_expv.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run oracle/build.py")
        L = ctypes.CDLL(path)
        L.orc_dlaran.restype = ctypes.c_double
        L.orc_dlaran.argtypes = [ctypes.c_void_p]
        L.orc_zgpadm.restype = ctypes.c_int
        L.orc_zgpadm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p,
                                 ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_hamvec.restype = None
        L.orc_expv.restype = ctypes.c_int
        L.orc_hahn_batch.restype = ctypes.c_int
        L.orc_piecewise_dense.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def term_struct(tt: TermTable) -> OrcTermTable:
    s = OrcTermTable()
    s.nspin = tt.nspin
    s.nterm = len(tt.coeff)
    arrs = [np.ascontiguousarray(x) for x in (tt.coeff, tt.nbody, tt.pos_off, tt.pos, tt.dims,
                                               tt.mat_off, tt.mats, tt.spin_dim)]
    tt._keep.append(arrs)
    (s.coeff, s.nbody, s.pos_off, s.pos, s.dims, s.mat_off, s.mats, s.spin_dim) = [a.ctypes.data for a in arrs]
    s.n = tt.n
    return s


def dlaran(iseed):
    seed = (ctypes.c_int32 * 4)(*iseed)
    v = lib().orc_dlaran(ctypes.addressof(seed))
    iseed[:] = list(seed)
    return v


def zgpadm(H, t, ideg=6):
    H = np.asfortranarray(np.asarray(H, complex))
    m = H.shape[0]
    out = np.zeros((m, m), complex, order="F")
    ns = ctypes.c_int(0)
    rc = lib().orc_zgpadm(ideg, m, float(t), _p(H), m, _p(out), ctypes.byref(ns))
    if rc != 0:
        raise RuntimeError(f"orc_zgpadm failed {rc}")
    return out, ns.value


def hamvec(tt: TermTable, v, coeff_complex=None):
    c = (np.asarray(coeff_complex, complex) if coeff_complex is not None
         else tt.coeff.astype(complex))
    v = np.ascontiguousarray(v, complex)
    w = np.zeros(tt.n, complex)
    wm = np.zeros(tt.n, complex)
    lib().orc_hamvec(ctypes.c_int32(tt.nspin), ctypes.c_int32(len(tt.coeff)), _p(c), _p(tt.nbody),
                     _p(tt.pos_off), _p(tt.pos), _p(tt.dims), _p(tt.mat_off), _p(tt.mats),
                     ctypes.c_int64(tt.n), _p(tt.spin_dim), _p(v), _p(w), _p(wm))
    return w


def expv(tt: TermTable, v, ta, m=30, tol=1e-12, klim=10, iseed=None):
    ta = np.ascontiguousarray(ta, float)
    v = np.ascontiguousarray(v, complex)
    out = np.zeros((len(ta), tt.n), complex)
    seed = (ctypes.c_int32 * 4)(*(iseed if iseed is not None else (0, 0, 0, 5)))
    st = OrcExpvStats()
    rc = lib().orc_expv(ctypes.c_int32(tt.nspin), ctypes.c_int32(len(tt.coeff)), _p(tt.coeff),
                        _p(tt.nbody), _p(tt.pos_off), _p(tt.pos), _p(tt.dims), _p(tt.mat_off),
                        _p(tt.mats), ctypes.c_int64(tt.n), _p(tt.spin_dim), _p(v),
                        ctypes.c_int32(klim), ctypes.c_int32(len(ta)), _p(ta), ctypes.c_int32(m),
                        ctypes.c_double(tol), _p(out), ctypes.byref(seed), ctypes.byref(st))
    if iseed is not None:
        iseed[:] = list(seed)
    return out.T.copy(), st, rc


def hahn_batch(opsA, opsB, u0s, y0s, scales, ta, m=30, tol=1e-12, klim=10, nthreads=None):
    nclu = len(opsA)
    A = (OrcTermTable * nclu)(*[term_struct(t) for t in opsA])
    B = (OrcTermTable * nclu)(*[term_struct(t) for t in opsB])
    u0s = [np.ascontiguousarray(u, complex) for u in u0s]
    y0s = [np.ascontiguousarray(y, complex) for y in y0s]
    up = (ctypes.c_void_p * nclu)(*[u.ctypes.data for u in u0s])
    yp = (ctypes.c_void_p * nclu)(*[y.ctypes.data for y in y0s])
    sc = np.ascontiguousarray(scales, float)
    ta = np.ascontiguousarray(ta, float)
    L = np.zeros((nclu, len(ta)))
    stats = (OrcExpvStats * (2 * nclu))()
    nth = nthreads or os.cpu_count() or 1
    rc = lib().orc_hahn_batch(ctypes.c_int32(nclu), A, B, up, yp, _p(sc), ctypes.c_int32(klim),
                              ctypes.c_int32(len(ta)), _p(ta), ctypes.c_int32(m),
                              ctypes.c_double(tol), _p(L), stats, ctypes.c_int32(nth))
    return L, stats, rc


def piecewise_dense(ops, seg, dt, v, nt):
    d = ops[0].shape[0]
    P = np.concatenate([np.asfortranarray(o, complex).reshape(-1, order="F") for o in ops])
    seg = np.ascontiguousarray(seg, float)
    v = np.ascontiguousarray(v, complex)
    out = np.zeros((nt, d), complex)
    rc = lib().orc_piecewise_dense(ctypes.c_int32(d), ctypes.c_int32(len(ops)), _p(P), _p(seg),
                                   ctypes.c_double(dt), _p(v), ctypes.c_int32(nt), _p(out))
    if rc != 0:
        raise RuntimeError("orc_piecewise_dense failed")
    return out.T.copy()


# ---------------------------------------------------------------------------
# per-cluster problem (EnsembleCCE::cluster_evolution, Hahn echo) in oracle form
def cluster_problem(spins, B, nv: NVCenter, state0=0, state1=1, dephasing_rate=0.0,
                    dephasing_axis=(1.0, 1.0, 1.0), pol=(0.0, 0.0, 0.0)):
    h0 = hamiltonian(spins, B, nv.espin, nv.state(state0))
    h1 = hamiltonian(spins, B, nv.espin, nv.state(state1))
    deph = dephasing_terms(spins, dephasing_rate, dephasing_axis)
    lvA = liouvillian(h0, h1, deph)
    lvB = liouvillian(h1, h0, deph)
    rho, D = density_vector(spins, pol)
    ldims = [s.mult ** 2 for s in spins]
    return dict(lvA=lvA, lvB=lvB, lvBdag=adjoint_negated(lvB), rho=rho, D=D, ldims=ldims,
                h0=h0, h1=h1)


def hahn_dense(prob, t_list):
    """Reference production path: PiecewiseFullMatrixVectorEvolution with
    riffle(lvA, lvB, 1) and segments [1/2, 1/2], then Re(D <rho0, state>)."""
    n = int(np.prod(prob["ldims"]))
    A = full(prob["lvA"], n)
    Bm = full(prob["lvB"], n)
    dt = t_list[1] - t_list[0]
    st = piecewise_dense([A, Bm], [0.5, 0.5], dt, prob["rho"], len(t_list))
    return np.real(prob["D"] * (prob["rho"].conj() @ st))


# ---------------------------------------------------------------------------
# single-sample CCE (a pure bath state, Hilbert space)
def glibc_rand(seed):
    """glibc rand() after srand(seed): random(3)'s TYPE_3 additive feedback
    generator (31 words seeded by Park-Miller 16807 mod 2^31-1 with Schrage's
    decomposition, r_i = r_{i-31} + r_{i-3} mod 2^32, 310 values discarded,
    outputs r_i >> 1).  Yields the rand() sequence."""
    def tdiv(a, b):
        q = abs(a) // abs(b)
        return q if (a >= 0) == (b >= 0) else -q
    w = seed & 0xFFFFFFFF
    if w >= 2 ** 31:
        w -= 2 ** 32
    if w == 0:
        w = 1
    r = [w]
    for i in range(1, 31):
        hi = tdiv(r[i - 1], 127773)
        lo = r[i - 1] - hi * 127773
        x = 16807 * lo - 2836 * hi
        if x < 0:
            x += 2147483647
        r.append(x)
    for i in range(31, 34):
        r.append(r[i - 31])
    for i in range(34, 344):
        r.append((r[i - 31] + r[i - 3]) & 0xFFFFFFFF)
    while True:
        v = (r[-31] + r[-3]) & 0xFFFFFFFF
        r.append(v)
        if len(r) > 4096:
            del r[:-64]
        yield v >> 1


def bath_states(seed, nspin):
    """prepare_bath_state (single_sample_cce.cpp:35-46): srand(seed), then
    spin i (spin-list order) in basis state rand() % 2."""
    g = glibc_rand(seed)
    return np.array([next(g) % 2 for _ in range(nspin)], np.int32)


def basis_state(mult, b):
    v = np.zeros(mult, complex)
    v[b] = 1.0
    return v


def bath_field(spin, bath, states, exclude):
    """DipolarField(spin_list, bath, states, exclude_idx): sum over the bath in
    its order of mask_j * dipole_field(spin, bath_j, state_j), mask 0 on the
    excluded (cluster) spins (SpinInteraction.cpp:159-173,
    SpinInteractionComponent.cpp:433-451)."""
    mask = np.ones(len(bath))
    mask[list(exclude)] = 0.0
    f = np.zeros(3)
    for j, src in enumerate(bath):
        f = f + mask[j] * dipole_field(spin, src, basis_state(src.mult, states[j]))
    return f


def single_sample_problem(coords, cluster, B, nv: NVCenter, states, state0=0, state1=1):
    """create_spin_hamiltonian (single_sample_cce.cpp:75-93) for both NV states
    and create_cluster_state (:95-105)."""
    bath = [Spin(np.asarray(x, float)) for x in coords]
    spins = [bath[i] for i in cluster]
    dims = [s.mult for s in spins]

    def ham(state):
        h = dipolar_terms(spins) + zeeman_terms(spins, B) + field_terms(spins, nv.espin, state)
        for i, s in enumerate(spins):
            f = bath_field(s, bath, states, cluster)
            h += _filtered(dims, (i,), [f[0], f[1], f[2], 0.0, 0.0, 0.0], single_forms(s))
        return h
    psi = np.ones(1, complex)
    for i in cluster:
        psi = np.kron(psi, basis_state(2, states[i]))
    return dict(h0=ham(nv.state(state0)), h1=ham(nv.state(state1)), psi=psi, dims=dims)


def single_sample_dense(prob, t_list, npulse):
    """cluster_evolution + calc_observables (single_sample_cce.cpp:48-73,
    120-128) with the dense piecewise propagator
    (QuantumEvolutionAlgorithm.cpp:60-83): psi_1 under riffle(H0, H1, N),
    psi_2 under riffle(H1, H0, N), L = Re cdot(psi_1, psi_2)."""
    n = int(np.prod(prob["dims"]))
    H0 = full(prob["h0"], n)
    H1 = full(prob["h1"], n)
    dt = t_list[1] - t_list[0]
    seg = pulse_interval(npulse)
    s1 = piecewise_dense(riffle(H0, H1, npulse), seg, dt, prob["psi"], len(t_list))
    s2 = piecewise_dense(riffle(H1, H0, npulse), seg, dt, prob["psi"], len(t_list))
    return np.real(np.sum(s1.conj() * s2, axis=0))
