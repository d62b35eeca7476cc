"""Writes the golden fixtures under tests/golden/ (run in the build container).

Every expected value comes from an independent source, never from the code
under test:
  * expm_cases.npz   — scipy.linalg.expm on seeded complex matrices (the
                       ZGPADM pin) and, for 2x2 cases, mpmath at 50 digits;
  * expv_cases.npz   — small seeded spin operators in the reference's term
                       format, v and a time list; w(t) = scipy expm(-iHt) v
                       built from numpy.kron (pins the matvec and the Krylov
                       driver);
  * dlaran_seq.npz   — the first draws of the reference's own DLARAN.f
                       (compiled from its source into oracle/_ref by
                       oracle/build.py) from iseed = (0,0,0,5), cross-checked
                       against the closed-form 48-bit LCG.
"""
import ctypes
import os
import sys

import numpy as np
import scipy.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def expm_cases():
    rng = np.random.default_rng(20261202)
    sizes = [2, 3, 5, 8, 16, 32]
    scales = [1e-3, 0.3, 2.0, 40.0]
    mats, ts, outs, shapes = [], [], [], []
    for n in sizes:
        for s in scales:
            A = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * s / np.sqrt(n)
            t = float(rng.uniform(0.2, 1.5))
            mats.append(A.reshape(-1, order="F"))
            ts.append(t)
            outs.append(sla.expm(t * A).reshape(-1, order="F"))
            shapes.append(n)
    # Hessenberg-like (Arnoldi shaped) cases with -i scaling
    for n in (12, 22, 32):
        H = np.triu(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)), -1) * 3.0
        t = 0.7
        mats.append((-1j * H).reshape(-1, order="F"))
        ts.append(t)
        outs.append(sla.expm(t * -1j * H).reshape(-1, order="F"))
        shapes.append(n)
    # 2x2 at 50 digits
    import mpmath as mp
    mp.mp.dps = 50
    for k in range(4):
        A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
        t = 1.3
        M = mp.matrix([[mp.mpc(complex(A[i, j])) for j in range(2)] for i in range(2)]) * t
        E = mp.expm(M)
        out = np.array([[complex(E[i, j]) for j in range(2)] for i in range(2)])
        mats.append(A.reshape(-1, order="F"))
        ts.append(t)
        outs.append(out.reshape(-1, order="F"))
        shapes.append(2)
    flat_m = np.concatenate(mats)
    flat_o = np.concatenate(outs)
    np.savez_compressed(os.path.join(HERE, "expm_cases.npz"), sizes=np.array(shapes), t=np.array(ts),
                        mats=flat_m, outs=flat_o)


def spin_ops():
    sx = 0.5 * np.array([[0, 1], [1, 0]], complex)
    sy = 0.5 * np.array([[0, -1j], [1j, 0]], complex)
    sz = 0.5 * np.array([[1, 0], [0, -1]], complex)
    return [sx, sy, sz]


def expv_cases():
    """Random spin-1/2 Hamiltonians H = sum_i b_i.S_i + sum_{i<j} S_i.J_ij.S_j
    (k = 2..5 spins) in the flattened term format, plus a Liouville case."""
    rng = np.random.default_rng(7)
    S = spin_ops()
    cases = []
    for k in (2, 3, 4, 5, 6):
        terms = []
        for i in range(k):
            for a in range(3):
                terms.append((float(rng.normal() * 1e3), [i], [S[a]]))
        for i in range(k):
            for j in range(i + 1, k):
                for a in range(3):
                    for b in range(3):
                        terms.append((float(rng.normal() * 3e2), [i, j], [S[a], S[b]]))
        terms.append((float(rng.normal() * 10.0), [], []))  # identity term
        dims = [2] * k
        cases.append((terms, dims))
    # Liouville-space operator: H# - H(b) for a 3-spin H, plus a dephasing-like term
    k = 3
    terms = []
    for i in range(k):
        for a in range(3):
            c = float(rng.normal() * 1e3)
            terms.append((c, [i], [np.kron(S[a].T, np.eye(2))]))
            terms.append((-c * 0.9, [i], [np.kron(np.eye(2), S[a])]))
    for i in range(k):
        for j in range(i + 1, k):
            for a in range(3):
                c = float(rng.normal() * 2e2)
                terms.append((c, [i, j], [np.kron(S[a].T, np.eye(2)), np.kron(S[a].T, np.eye(2))]))
                terms.append((-c, [i, j], [np.kron(np.eye(2), S[a]), np.kron(np.eye(2), S[a])]))
    cases.append((terms, [4] * k))
    out = {}
    for ci, (terms, dims) in enumerate(cases):
        n = int(np.prod(dims))
        coeff = np.array([t[0] for t in terms])
        nbody = np.array([len(t[1]) for t in terms], np.int64)
        pos_off = np.concatenate([[0], np.cumsum(nbody)]).astype(np.int64)
        pos = np.array([p for t in terms for p in t[1]], np.int64)
        mats = [m for t in terms for m in t[2]]
        mdims = np.array([m.shape[0] for m in mats], np.int64)
        mat_off = np.concatenate([[0], np.cumsum(mdims * mdims)]).astype(np.int64)
        matv = np.concatenate([m.reshape(-1, order="F") for m in mats])
        # dense H by explicit Kronecker products (spin 0 slowest)
        H = np.zeros((n, n), complex)
        for c, idx, ms in terms:
            ops = [np.eye(d, dtype=complex) for d in dims]
            for p, m in zip(idx, ms):
                ops[p] = m
            K = np.ones((1, 1), complex)
            for o in ops:
                K = np.kron(K, o)
            H += c * K
        v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        v /= np.linalg.norm(v)
        tl = np.array([0.0, 1e-5, 2.5e-5, 1e-4, 3e-4])
        W = np.stack([sla.expm(-1j * H * t) @ v for t in tl], axis=1)
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        pre = f"c{ci}_"
        out.update({pre + "coeff": coeff, pre + "nbody": nbody, pre + "pos_off": pos_off, pre + "pos": pos,
                    pre + "dims": mdims, pre + "mat_off": mat_off, pre + "mats": matv,
                    pre + "spin_dim": np.array(dims, np.int64), pre + "v": v, pre + "t": tl, pre + "w": W,
                    pre + "x": x, pre + "Hx": H @ x})
    out["ncases"] = np.array(len(cases))
    np.savez_compressed(os.path.join(HERE, "expv_cases.npz"), **out)


def dlaran_seq():
    ref = os.path.join(ROOT, "oracle", "_ref", "libdlaran_ref.so")
    count = 4096
    x = 5
    M = 33952834046453
    closed = np.empty(count)
    for i in range(count):
        x = (x * M) % (1 << 48)
        closed[i] = x / float(1 << 48)
    if os.path.exists(ref):
        L = ctypes.CDLL(ref)
        L.dlaran_.restype = ctypes.c_double
        seed = (ctypes.c_int * 4)(0, 0, 0, 5)
        seq = np.array([L.dlaran_(seed) for _ in range(count)])
        assert np.array_equal(seq, closed), "reference DLARAN disagrees with the closed-form LCG"
        src = "reference DLARAN.f (oracle/_ref)"
    else:
        seq = closed
        src = "closed-form 48-bit LCG"
    np.savez_compressed(os.path.join(HERE, "dlaran_seq.npz"), seq=seq, final_seed=np.array(
        [(x >> 36) & 4095, (x >> 24) & 4095, (x >> 12) & 4095, x & 4095]), source=np.array(src))


if __name__ == "__main__":
    expm_cases()
    expv_cases()
    dlaran_seq()
    print("golden fixtures written to", HERE)
