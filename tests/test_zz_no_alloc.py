"""GPU: oops_hamvec_noalloc, the product entry point that allocates no device
memory (its operator tables are staged in the library's fixed device array),
against oops_hamvec on the same operators: two calls in a row, the second
with a smaller input, through the C ABI with caller-owned device vectors.
The flat product (a 6-spin Hilbert case of the scipy fixtures, then a 5-spin
one) and the tiled product (a 10-spin Liouvillian, n = 2^20, six passes, then
a 4-spin Liouvillian) give the same vectors as the allocating entry point."""
import ctypes
import os

import numpy as np
import pytest

import oops_amd as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _hip():
    O.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def _noalloc(ctx, hip, T, x):
    L = O.lib()
    L.oops_hamvec_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    x = np.ascontiguousarray(x, np.complex128)
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dx), x.nbytes) == 0
    assert hip.hipMalloc(ctypes.byref(dy), x.nbytes) == 0
    try:
        assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) == 0
        rc = L.oops_hamvec_noalloc(ctx.h, ctypes.byref(T.c), dx, dy)
        if rc != 0:
            raise O.OopsError(f"oops_hamvec_noalloc: {rc} {L.oops_last_error(ctx.h)}")
        y = np.zeros_like(x)
        assert hip.hipMemcpy(y.ctypes.data, dy, y.nbytes, 2) == 0
        return y
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)


def _fixture_case(g, c):
    pre = f"c{c}_"
    fields = [g[pre + k] for k in ("coeff", "nbody", "pos_off", "pos", "dims", "mat_off", "mats", "spin_dim")]
    return O.TermTable(*fields), g[pre + "x"]


@pytest.mark.gpu
def test_noalloc_product_equals_hamvec_flat(ctx):
    hip = _hip()
    g = np.load(os.path.join(GOLD, "expv_cases.npz"), allow_pickle=False)
    for c in (4, 3):                      # n = 64, then n = 32
        T, x = _fixture_case(g, c)
        want = O.hamvec(ctx, T, x)
        got = _noalloc(ctx, hip, T, x)
        print(f"no-alloc product n={T.n}: max |diff| {np.abs(got - want).max():.1e}", flush=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
        assert np.abs(want - g[f"c{c}_Hx"]).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.gpu
def test_noalloc_product_equals_hamvec_tiled_then_smaller(ctx):
    hip = _hip()
    rng = np.random.default_rng(0)
    X = np.array([p for p in rng.uniform(-15, 15, (400, 3)) if np.linalg.norm(p) <= 15.0][:10])
    par = O.cce_params(B=B111)
    for k in (10, 4):
        T, _, _ = O.cluster_terms(X[:k], par, 0)
        x = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
        want = O.hamvec(ctx, T, x)
        got = _noalloc(ctx, hip, T, x)
        print(f"no-alloc product n={T.n} ({O.tile_plan(T)} passes): max |diff| {np.abs(got - want).max():.1e}",
              flush=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
