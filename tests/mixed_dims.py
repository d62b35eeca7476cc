"""A random Hermitian operator on spins of local dimensions that are not one
power of two (spin-1: 3, a spin-1 Liouville space: 9, mixed with spin-1/2),
in the reference's flattened term layout (MatExp.cpp:228-262), and its
explicit matrix from numpy Kronecker products (spin 0 the slowest factor;
explicit=False skips it for sizes a dense matrix does not fit)."""
import numpy as np

import oops_amd as O


def random_table(dims, seed=0, pairs=True, explicit=True):
    rng = np.random.default_rng(seed)
    k = len(dims)
    coeff, nbody, pos_off, pos, bdims, mat_off, mats = [], [], [0], [], [], [0], []

    def herm(d):
        a = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        return a + a.conj().T

    def add(c, bodies):
        coeff.append(c)
        nbody.append(len(bodies))
        for s, m in bodies:
            pos.append(s)
            bdims.append(dims[s])
            mats.extend(m.reshape(-1, order="F"))
            mat_off.append(mat_off[-1] + m.size)
        pos_off.append(len(pos))

    add(0.3, [])
    for s in range(k):
        add(rng.uniform(0.5, 1.5), [(s, herm(dims[s]))])
    if pairs:
        for s in range(k):
            for t in range(s + 1, k):
                a, b = herm(dims[s]), herm(dims[t])
                add(rng.uniform(0.1, 0.5), [(s, a), (t, b)])
    T = O.TermTable(coeff, nbody, pos_off, pos, bdims, mat_off, mats, dims)

    def full():
        n = int(np.prod(dims))
        H = np.zeros((n, n), complex)
        for ti, c in enumerate(coeff):
            ops = [np.eye(d) for d in dims]
            for b in range(pos_off[ti], pos_off[ti + 1]):
                d = bdims[b]
                ops[pos[b]] = np.asarray(mats[mat_off[b]:mat_off[b + 1]]).reshape(d, d, order="F")
            M = np.ones((1, 1))
            for o in ops:
                M = np.kron(M, o)
            H += c * M
        return H
    return T, (full() if explicit else None)
