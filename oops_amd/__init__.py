"""oops_amd — MI355X-native CCE Krylov-expv hot path (Python binding of the C ABI).

The numeric work runs in ``liboops_amd.so`` (hand-written HIP kernels for
gfx950 behind ``include/oops_expv.h`` / ``include/oops_cce.h``).  This module
only marshals arguments through ctypes; there is no CPU fallback: if the
library or a GPU is missing the calls raise.

Names mirror the reference's host classes where the reference has them
(``MatExpVector``, ``SumKronProd``, ``cDepthFirstPathTracing``, ``CCE``) so
the tests read like the reference's own call sites.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboops_amd.so")

OOPS_OK = 0
STATS_PER_ITEM = 24  # int32 counters per cluster (oops_hahn_batch_results)


class OopsError(RuntimeError):
    pass


class TermTableC(ctypes.Structure):
    _fields_ = [("nspin", ctypes.c_int32), ("nterm", ctypes.c_int32),
                ("coeff", ctypes.c_void_p), ("nbody", ctypes.c_void_p),
                ("pos_offset", ctypes.c_void_p), ("pos", ctypes.c_void_p),
                ("dims", ctypes.c_void_p), ("mat_offset", ctypes.c_void_p),
                ("mats", ctypes.c_void_p), ("spin_dim", ctypes.c_void_p)]


class HahnItemC(ctypes.Structure):
    _fields_ = [("opA", ctypes.POINTER(TermTableC)), ("opB", ctypes.POINTER(TermTableC)),
                ("u0", ctypes.c_void_p), ("y0", ctypes.c_void_p), ("scale", ctypes.c_double)]


class CCEParamsC(ctypes.Structure):
    _fields_ = [("B", ctypes.c_double * 3), ("nv_pos", ctypes.c_double * 3),
                ("state0", ctypes.c_int32), ("state1", ctypes.c_int32),
                ("dephasing_rate", ctypes.c_double), ("dephasing_axis", ctypes.c_double * 3),
                ("pol", ctypes.c_double * 3)]


_lib = None

EXPORTS = [
    "oops_ctx_create", "oops_ctx_destroy", "oops_last_error", "oops_ctx_info", "oops_ctx_set_launch",
    "oops_hahn_batch_create", "oops_hahn_batch_run", "oops_hahn_batch_results",
    "oops_hahn_batch_timing", "oops_hahn_batch_destroy", "oops_dlaran_fill",
    "oops_bath_diamond", "oops_nv_states", "oops_cce_cluster_terms", "oops_terms_view",
    "oops_terms_sizes", "oops_terms_destroy", "oops_clusters_enumerate", "oops_clusters_count",
    "oops_clusters_get", "oops_clusters_parents", "oops_clusters_destroy", "oops_cce_reduce",
    "oops_cce_hahn_batch_create", "oops_hamvec", "oops_expv", "oops_expv_ex", "oops_echo_batch_create", "oops_echo_batch_launch",
    "oops_echo_batch_query", "oops_echo_batch_wait", "oops_echo_batch_run", "oops_echo_batch_footprint",
    "oops_cce_echo_batch_create", "oops_server_create", "oops_server_submit", "oops_server_poll", "oops_server_state",
    "oops_server_result", "oops_server_pending", "oops_server_destroy", "oops_server_times",
    "oops_hahn_batch_kernels", "oops_echo_batch_capture", "oops_slot_form", "oops_tile_plan", "oops_read_xyz", "oops_write_xyz",
    "oops_hahn_batch_elapsed", "oops_server_failed", "oops_bath_states", "oops_cce_single_sample_terms",
    "oops_cce_single_sample_batch_create", "oops_clusters_enumerate_graph", "oops_cce_op_tables",
    "oops_hamvec_noalloc",
]


def lib():
    """Load liboops_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise OopsError(f"{LIBPATH} missing: run oops_amd/build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.oops_last_error.restype = ctypes.c_char_p
        L.oops_last_error.argtypes = [ctypes.c_void_p]
        L.oops_ctx_create.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.oops_ctx_destroy.argtypes = [ctypes.c_void_p]
        L.oops_hahn_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
        L.oops_hahn_batch_results.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oops_hahn_batch_destroy.argtypes = [ctypes.c_void_p]
        L.oops_bath_diamond.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_uint64,
                                        ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32)]
        L.oops_clusters_enumerate.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                              ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]
        L.oops_clusters_destroy.argtypes = [ctypes.c_void_p]
        L.oops_terms_destroy.argtypes = [ctypes.c_void_p]
        L.oops_server_create.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.oops_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]
        L.oops_server_poll.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.oops_server_state.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]
        L.oops_server_result.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.oops_server_pending.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        L.oops_server_destroy.argtypes = [ctypes.c_void_p]
        L.oops_server_times.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.oops_server_failed.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
        L.oops_hahn_batch_elapsed.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.oops_hamvec.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oops_expv.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
                                ctypes.c_void_p]
        L.oops_expv_ex.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                   ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


KERNEL_NAMES = {0: "dense_pade", 1: "arnoldi_fp64", 2: "arnoldi_mixed", 3: "lanczos_mixed", 4: "dense64_mfma",
                5: "single_state", 6: "lanczos_slot", 7: "lanczos_medium", 8: "lanczos_slot_fp64",
                9: "factored_powers", -1: None}


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _precision_code(p):
    codes = {"fp64": 0, "mixed": 1, "mixed-arnoldi": 2, "mixed-ell": 3, "fp64-lanczos": 4}
    if p not in codes:
        raise ValueError(f"precision must be one of {sorted(codes)}")
    return codes[p]


def _check(rc, ctx=None, what=""):
    if rc != OOPS_OK:
        msg = ""
        if ctx is not None:
            m = lib().oops_last_error(ctx)
            msg = m.decode() if m else ""
        raise OopsError(f"{what} failed with status {rc}: {msg}")


# ---------------------------------------------------------------------------
class TermTable:
    """Owning host copy of an operator in the reference's flattened format
    (MatExp.cpp:228-262) plus its C view."""

    def __init__(self, coeff, nbody, pos_offset, pos, dims, mat_offset, mats, spin_dim):
        self.coeff = np.ascontiguousarray(coeff, np.float64)
        self.nbody = np.ascontiguousarray(nbody, np.int64)
        self.pos_offset = np.ascontiguousarray(pos_offset, np.int64)
        self.pos = np.ascontiguousarray(pos if len(pos) else [0], np.int64)
        self.dims = np.ascontiguousarray(dims if len(dims) else [1], np.int64)
        self.mat_offset = np.ascontiguousarray(mat_offset, np.int64)
        m = np.ascontiguousarray(mats, np.complex128)
        self.mats = m if m.size else np.zeros(1, np.complex128)
        self.spin_dim = np.ascontiguousarray(spin_dim, np.int64)
        self.n = int(np.prod(self.spin_dim))
        self.c = TermTableC(len(self.spin_dim), len(self.coeff), self.coeff.ctypes.data,
                            self.nbody.ctypes.data, self.pos_offset.ctypes.data, self.pos.ctypes.data,
                            self.dims.ctypes.data, self.mat_offset.ctypes.data, self.mats.ctypes.data,
                            self.spin_dim.ctypes.data)

    @classmethod
    def from_fields(cls, t):
        """From any object with the reference's field names (e.g. an oracle table)."""
        return cls(t.coeff, t.nbody, getattr(t, "pos_offset", getattr(t, "pos_off", None)), t.pos, t.dims,
                   getattr(t, "mat_offset", getattr(t, "mat_off", None)), t.mats, t.spin_dim)


class Context:
    """A device context (one HIP stream on one GPU)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().oops_ctx_create(device, ctypes.byref(h))
        if rc != OOPS_OK:
            raise OopsError(f"oops_ctx_create({device}) failed with status {rc} (no HIP device?)")
        self.h = h

    def info(self):
        cu = ctypes.c_int32()
        lds = ctypes.c_int32()
        name = ctypes.create_string_buffer(64)
        _check(lib().oops_ctx_info(self.h, ctypes.byref(cu), ctypes.byref(lds), name, 64), self.h, "ctx_info")
        return {"num_cu": cu.value, "lds_bytes": lds.value, "arch": name.value.decode()}

    def set_launch(self, kernel, wg_per_cu):
        """Workgroups per CU of the persistent batch kernel `kernel` (a name
        of KERNEL_NAMES or its kind) in batches created afterwards; 0 restores
        the default."""
        kind = kernel if isinstance(kernel, int) else {v: k for k, v in KERNEL_NAMES.items()}[kernel]
        _check(lib().oops_ctx_set_launch(self.h, ctypes.c_int32(kind), ctypes.c_int32(wg_per_cu)), self.h,
               "oops_ctx_set_launch")

    def close(self):
        if self.h:
            lib().oops_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def slot_form(H: TermTable, max_slots=4096):
    """Flip-mask form of H (host): (masks, coef, nslot64) with
    (H x)_i = sum_s coef[s, i] x[i ^ masks[s]]; the first nslot64 masks change
    at most one spin's digits.  The layout the mixed Lanczos kernel builds."""
    n64, n32 = ctypes.c_int32(), ctypes.c_int32()
    _check(lib().oops_slot_form(ctypes.byref(H.c), ctypes.c_int32(max_slots), ctypes.byref(n64), ctypes.byref(n32),
                                None, None), None, "oops_slot_form")
    S = n64.value + n32.value
    n = int(np.prod(H.spin_dim))
    masks = np.zeros(S, np.int32)
    coef = np.zeros((S, n), np.complex128)
    _check(lib().oops_slot_form(ctypes.byref(H.c), ctypes.c_int32(max_slots), ctypes.byref(n64), ctypes.byref(n32),
                                _ptr(masks), _ptr(coef)), None, "oops_slot_form")
    return masks, coef, n64.value


def tile_plan(H: TermTable, v=None, flat=False, axis=False):
    """Number of passes of the large path's tiled product for H (0: flat ELL
    product); with v, also H v computed on the host through that plan (or
    through the flat ELL tables with flat=True, or through the two-axis
    product of split Liouvillians with axis=True)."""
    n = ctypes.c_int32()
    mode = 3 if axis else (2 if flat else 1)
    if v is None:
        _check(lib().oops_tile_plan(ctypes.byref(H.c), ctypes.byref(n), ctypes.c_int32(3 if axis else 0), None,
                                    None), None, "oops_tile_plan")
        return n.value
    v = np.ascontiguousarray(v, np.complex128)
    w = np.zeros_like(v)
    _check(lib().oops_tile_plan(ctypes.byref(H.c), ctypes.byref(n), ctypes.c_int32(mode), _ptr(v), _ptr(w)), None,
           "oops_tile_plan")
    return n.value, w


def read_xyz(path, max_n=1 << 20):
    """Spin coordinates from a coordinate file (count, then "isotope x y z"
    lines): (coords [n, 3] in Angstrom, isotope names)."""
    n = ctypes.c_int32()
    coords = np.zeros((max_n, 3))
    iso = ctypes.create_string_buffer(16 * max_n)
    _check(lib().oops_read_xyz(os.fsencode(path), ctypes.c_int32(max_n), _ptr(coords), iso, ctypes.byref(n)), None,
           "oops_read_xyz")
    raw = iso.raw[:16 * n.value]
    names = [raw[16 * i:16 * (i + 1)].split(b"\0", 1)[0].decode() for i in range(n.value)]
    return coords[:n.value].copy(), names


def write_xyz(path, coords, isotopes=None):
    """Write coordinates in the same format (the reference's export layout)."""
    coords = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    n = len(coords)
    iso = None
    if isotopes is not None:
        iso = ctypes.create_string_buffer(16 * n)
        for i, name in enumerate(isotopes):
            b = name.encode()[:15]
            iso[16 * i:16 * i + len(b)] = b
    _check(lib().oops_write_xyz(os.fsencode(path), ctypes.c_int32(n), _ptr(coords), iso), None, "oops_write_xyz")


def hamvec(ctx: Context, H: TermTable, v):
    """w = H v on the device (hamvec_cuda3_, hamvec_cuda3.cu:661-665)."""
    v = np.ascontiguousarray(v, np.complex128)
    if v.shape != (H.n,):
        raise ValueError(f"v must have length {H.n}")
    w = np.empty_like(v)
    _check(lib().oops_hamvec(ctx.h, ctypes.byref(H.c), _ptr(v), _ptr(w)), ctx.h, "oops_hamvec")
    return w


def expv(ctx: Context, H: TermTable, v, ta, m=30, klim=10, tol=1e-12, overlaps=False, precision="fp64",
         steps="reference"):
    """exp(-i H t) v at every time of ta (MatExpVector::runInexplicitGPU,
    MatExp.cpp:367-521).  Returns (w_seq [n x nt], stats) or, with
    overlaps=True, (<v, w(t_k)> [nt], stats).  precision "fp32" stores the
    Krylov basis in complex FP32.  steps "reference": a step ends at every
    output time (Expokit's schedule); "inside": steps run past output times
    and evaluate the ones they contain from their own basis."""
    v = np.ascontiguousarray(v, np.complex128)
    ta = np.ascontiguousarray(ta, np.float64)
    if v.shape != (H.n,):
        raise ValueError(f"v must have length {H.n}")
    out = np.empty(len(ta) if overlaps else (len(ta), H.n), np.complex128)
    st = np.zeros(8)
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be fp64 or fp32")
    if steps not in ("reference", "inside"):
        raise ValueError("steps must be reference or inside")
    flags = (1 if precision == "fp32" else 0) | (2 if steps == "inside" else 0)
    rc = lib().oops_expv_ex(ctx.h, ctypes.byref(H.c), _ptr(v), ctypes.c_int32(len(ta)), _ptr(ta),
                            ctypes.c_int32(klim), ctypes.c_int32(m), ctypes.c_double(tol),
                            ctypes.c_int32(flags), ctypes.c_int32(1 if overlaps else 0),
                            _ptr(out), _ptr(st))
    _check(rc, ctx.h, "oops_expv")
    keys = ["nmult", "nexph", "nscale", "nstep", "nreject", "ibrkflag", "anorm", "kernel_ms"]
    stats = dict(zip(keys, st.tolist()))
    return (out if overlaps else out.T), stats


class RequestC(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("radius", ctypes.c_double), ("p", ctypes.c_double),
                ("B", ctypes.c_double * 3), ("cutoff", ctypes.c_double), ("order", ctypes.c_int32),
                ("npulse", ctypes.c_int32), ("nt", ctypes.c_int32), ("t_end", ctypes.c_double)]


class EchoItemC(ctypes.Structure):
    _fields_ = [("opA", ctypes.POINTER(TermTableC)), ("opB", ctypes.POINTER(TermTableC)),
                ("opC", ctypes.POINTER(TermTableC)), ("opD", ctypes.POINTER(TermTableC)),
                ("u0", ctypes.c_void_p), ("y0", ctypes.c_void_p), ("scale", ctypes.c_double),
                ("npulse", ctypes.c_int32), ("nt", ctypes.c_int32), ("t", ctypes.c_void_p)]


class HahnBatch:
    """Device-resident batch of Hahn-echo clusters (oops_hahn_batch_*)."""

    def __init__(self, ctx: Context, handle, nitems: int):
        self.ctx = ctx
        self.h = handle
        self.nitems = nitems
        self.nt = 0

    @classmethod
    def from_items(cls, ctx: Context, items, m=30, klim=10, tol=1e-12, precision="fp64"):
        """items: list of (TermTable A, TermTable B', u0, y0, scale)."""
        keep = []
        arr = (HahnItemC * len(items))()
        for i, (A, B, u0, y0, sc) in enumerate(items):
            u0 = np.ascontiguousarray(u0, np.complex128)
            y0 = np.ascontiguousarray(y0, np.complex128)
            keep += [A, B, u0, y0]
            arr[i].opA = ctypes.pointer(A.c)
            arr[i].opB = ctypes.pointer(B.c)
            arr[i].u0 = u0.ctypes.data
            arr[i].y0 = y0.ctypes.data
            arr[i].scale = float(sc)
        h = ctypes.c_void_p()
        rc = lib().oops_hahn_batch_create(ctx.h, ctypes.c_int32(len(items)), arr, ctypes.c_int32(m),
                                          ctypes.c_int32(klim), ctypes.c_double(tol),
                                          ctypes.c_int32(_precision_code(precision)), ctypes.byref(h))
        _check(rc, ctx.h, "oops_hahn_batch_create")
        return cls(ctx, h, len(items))

    def run(self, tau):
        tau = np.ascontiguousarray(tau, np.float64)
        _check(lib().oops_hahn_batch_run(self.h, ctypes.c_int32(len(tau)), _ptr(tau)), self.ctx.h,
               "oops_hahn_batch_run")
        self.nt = len(tau)

    def results(self):
        L = np.zeros((self.nitems, self.nt))
        st = np.zeros((self.nitems, STATS_PER_ITEM), np.int32)
        _check(lib().oops_hahn_batch_results(self.h, _ptr(L), _ptr(st)), self.ctx.h, "oops_hahn_batch_results")
        return L, st

    def timing(self):
        nb = ctypes.c_int32()
        n = np.zeros(32, np.int32)
        cnt = np.zeros(32, np.int32)
        ms = np.zeros(32)
        _check(lib().oops_hahn_batch_timing(self.h, ctypes.c_int32(32), _ptr(n), _ptr(cnt), _ptr(ms),
                                            ctypes.byref(nb)), self.ctx.h, "oops_hahn_batch_timing")
        k = nb.value
        kind = np.full(32, -1, np.int32)
        _check(lib().oops_hahn_batch_kernels(self.h, ctypes.c_int32(32), _ptr(kind)), self.ctx.h,
               "oops_hahn_batch_kernels")
        return [{"n": int(n[i]), "count": int(cnt[i]), "ms": float(ms[i]), "kernel": KERNEL_NAMES.get(int(kind[i]))}
                for i in range(k)]

    def elapsed(self):
        """Device time of the last run in ms: one interval around every bucket
        (buckets overlap on their own streams, so per-bucket times do not add)."""
        v = ctypes.c_double()
        _check(lib().oops_hahn_batch_elapsed(self.h, ctypes.byref(v)), self.ctx.h, "oops_hahn_batch_elapsed")
        return v.value

    def close(self):
        if self.h:
            lib().oops_hahn_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EchoBatch(HahnBatch):
    """Device batch of pulse-sequence items with their own time grids
    (oops_echo_batch_*): Hahn echo (npulse 1, times are sweep times t/2) or
    CPMG-N (times are t).  Runs asynchronously on its own HIP stream."""

    @classmethod
    def from_items(cls, ctx: Context, items, m=30, klim=10, tol=1e-12, precision="mixed"):
        """items: dicts with opA, opB, [opC, opD] (TermTable), u0, y0, scale, npulse, t."""
        keep = []
        arr = (EchoItemC * len(items))()
        nts = []
        for i, it in enumerate(items):
            u0 = np.ascontiguousarray(it["u0"], np.complex128)
            y0 = np.ascontiguousarray(it["y0"], np.complex128)
            t = np.ascontiguousarray(it["t"], np.float64)
            keep += [u0, y0, t, it["opA"], it["opB"]]
            arr[i].opA = ctypes.pointer(it["opA"].c)
            arr[i].opB = ctypes.pointer(it["opB"].c)
            if it.get("opC") is not None:
                keep += [it["opC"], it["opD"]]
                arr[i].opC = ctypes.pointer(it["opC"].c)
                arr[i].opD = ctypes.pointer(it["opD"].c)
            arr[i].u0 = u0.ctypes.data
            arr[i].y0 = y0.ctypes.data
            arr[i].scale = float(it["scale"])
            arr[i].npulse = int(it.get("npulse", 1))
            arr[i].nt = len(t)
            arr[i].t = t.ctypes.data
            nts.append(len(t))
        h = ctypes.c_void_p()
        rc = lib().oops_echo_batch_create(ctx.h, ctypes.c_int32(len(items)), arr, ctypes.c_int32(m),
                                          ctypes.c_int32(klim), ctypes.c_double(tol),
                                          ctypes.c_int32(_precision_code(precision)), ctypes.byref(h))
        _check(rc, ctx.h, "oops_echo_batch_create")
        b = cls(ctx, h, len(items))
        b.nts = nts
        return b

    def capture(self):
        """Record the batch's launches in a HIP graph; later runs replay it."""
        _check(lib().oops_echo_batch_capture(self.h), self.ctx.h, "oops_echo_batch_capture")

    def launch(self):
        _check(lib().oops_echo_batch_launch(self.h), self.ctx.h, "oops_echo_batch_launch")

    def done(self):
        d = ctypes.c_int32()
        _check(lib().oops_echo_batch_query(self.h, ctypes.byref(d)), self.ctx.h, "oops_echo_batch_query")
        return bool(d.value)

    def wait(self):
        _check(lib().oops_echo_batch_wait(self.h), self.ctx.h, "oops_echo_batch_wait")

    def run(self):
        _check(lib().oops_echo_batch_run(self.h), self.ctx.h, "oops_echo_batch_run")

    def footprint(self):
        v = ctypes.c_int64()
        _check(lib().oops_echo_batch_footprint(self.h, ctypes.byref(v)), self.ctx.h, "oops_echo_batch_footprint")
        return v.value

    def results(self):
        """(list of per-item L arrays, stats)"""
        L = np.zeros(sum(self.nts))
        st = np.zeros((self.nitems, STATS_PER_ITEM), np.int32)
        _check(lib().oops_hahn_batch_results(self.h, _ptr(L), _ptr(st)), self.ctx.h, "oops_hahn_batch_results")
        out, off = [], 0
        for k in self.nts:
            out.append(L[off:off + k].copy())
            off += k
        return out, st


class Server:
    """Serving loop (oops_server_*): continuous batching of CCE requests on
    one device.  A request is a dict with seed, radius, p, B (3), cutoff,
    order, npulse, nt, t_end; its result is the CCE product L(t)."""

    def __init__(self, ctx: Context, hbm_budget=8 << 30, max_inflight=8, m=30, tol=1e-3, precision="mixed",
                 threads=16):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().oops_server_create(ctx.h, int(hbm_budget), int(max_inflight), int(m), float(tol),
                                        _precision_code(precision), int(threads), ctypes.byref(h)),
               ctx.h, "oops_server_create")
        self.h = h
        self.nt = {}

    def submit(self, req):
        r = RequestC()
        r.seed = int(req["seed"])
        r.radius = float(req["radius"])
        r.p = float(req.get("p", 0.011))
        for c in range(3):
            r.B[c] = float(req["B"][c])
        r.cutoff = float(req.get("cutoff", 8.0))
        r.order = int(req["order"])
        r.npulse = int(req.get("npulse", 1))
        r.nt = int(req["nt"])
        r.t_end = float(req["t_end"])
        i = ctypes.c_int64()
        _check(lib().oops_server_submit(self.h, ctypes.byref(r), ctypes.byref(i)), self.ctx.h, "oops_server_submit")
        self.nt[i.value] = r.nt
        return i.value

    def poll(self):
        c = ctypes.c_int32()
        _check(lib().oops_server_poll(self.h, ctypes.byref(c)), self.ctx.h, "oops_server_poll")
        return c.value

    def state(self, rid):
        s = ctypes.c_int32()
        _check(lib().oops_server_state(self.h, ctypes.c_int64(rid), ctypes.byref(s)), self.ctx.h, "oops_server_state")
        return s.value

    def pending(self):
        p = ctypes.c_int32()
        _check(lib().oops_server_pending(self.h, ctypes.byref(p)), self.ctx.h, "oops_server_pending")
        return p.value

    def result(self, rid):
        L = np.zeros(self.nt[rid])
        lat = ctypes.c_double()
        ncl = ctypes.c_int64()
        fp = ctypes.c_int64()
        _check(lib().oops_server_result(self.h, ctypes.c_int64(rid), _ptr(L), ctypes.byref(lat), ctypes.byref(ncl),
                                        ctypes.byref(fp)), self.ctx.h, "oops_server_result")
        tb, tl = ctypes.c_double(), ctypes.c_double()
        _check(lib().oops_server_times(self.h, ctypes.c_int64(rid), ctypes.byref(tb), ctypes.byref(tl)),
               self.ctx.h, "oops_server_times")
        return {"L": L, "latency_s": lat.value, "clusters": ncl.value, "footprint": fp.value,
                "built_s": tb.value, "launched_s": tl.value}

    def failed(self, rid):
        """Clusters of a finished request whose integration failed."""
        v = ctypes.c_int64()
        _check(lib().oops_server_failed(self.h, ctypes.c_int64(rid), ctypes.byref(v)), self.ctx.h,
               "oops_server_failed")
        return v.value

    def drain(self, sleep=0.002):
        import time
        while self.pending():
            self.poll()
            time.sleep(sleep)

    def close(self):
        if self.h:
            lib().oops_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# CCE host pipeline
def cce_params(B=(0.0, 0.0, 0.0), nv_pos=(0.0, 0.0, 0.0), state0=0, state1=1, dephasing_rate=0.0,
               dephasing_axis=(1.0, 1.0, 1.0), pol=(0.0, 0.0, 0.0)) -> CCEParamsC:
    p = CCEParamsC()
    p.B[:] = list(B)
    p.nv_pos[:] = list(nv_pos)
    p.state0 = state0
    p.state1 = state1
    p.dephasing_rate = dephasing_rate
    p.dephasing_axis[:] = list(dephasing_axis)
    p.pol[:] = list(pol)
    return p


def bath_diamond(a=3.567, radius=30.0, p=0.011, seed=0, shape="sphere"):
    """Seeded diamond-lattice 13C bath (Angstrom), NV sites excluded."""
    sh = 0 if shape == "sphere" else 1
    n = ctypes.c_int32()
    _check(lib().oops_bath_diamond(a, radius, p, seed, sh, ctypes.c_void_p(0), 0, ctypes.byref(n)), None,
           "oops_bath_diamond")
    out = np.zeros((max(n.value, 1), 3))
    _check(lib().oops_bath_diamond(a, radius, p, seed, sh, _ptr(out), n.value, ctypes.byref(n)), None,
           "oops_bath_diamond")
    return out[:n.value].copy()


def nv_states(pos, B):
    pos = np.ascontiguousarray(pos, float)
    B = np.ascontiguousarray(B, float)
    ev = np.zeros(3)
    vec = np.zeros(9, np.complex128)
    _check(lib().oops_nv_states(_ptr(pos), _ptr(B), _ptr(ev), _ptr(vec)), None, "oops_nv_states")
    return ev, vec.reshape(3, 3, order="F")


def _terms_from_handle(h):
    """Copy an oops_terms handle into a TermTable and destroy the handle."""
    try:
        v = TermTableC()
        _check(lib().oops_terms_view(h, ctypes.byref(v)), None, "oops_terms_view")
        nterm = v.nterm
        npos = ctypes.c_int64()
        nmat = ctypes.c_int64()
        lib().oops_terms_sizes(h, None, ctypes.byref(npos), ctypes.byref(nmat))

        def arr(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(dt)), shape=(count,)).copy()

        coeff = arr(v.coeff, nterm, ctypes.c_double)
        nbody = arr(v.nbody, nterm, ctypes.c_int64)
        pos_off = arr(v.pos_offset, nterm + 1, ctypes.c_int64)
        pos = arr(v.pos, npos.value, ctypes.c_int64)
        dims = arr(v.dims, npos.value, ctypes.c_int64)
        mat_off = arr(v.mat_offset, npos.value + 1, ctypes.c_int64)
        mats = arr(v.mats, 2 * nmat.value, ctypes.c_double).view(np.complex128)
        spin_dim = arr(v.spin_dim, v.nspin, ctypes.c_int64)
        return TermTable(coeff, nbody, pos_off, pos, dims, mat_off, mats, spin_dim)
    finally:
        lib().oops_terms_destroy(h)


def cluster_terms(coords, params: CCEParamsC, which=0):
    """Term table of one cluster: which 0 lvA, 1 lvB, 2 -lvB^dagger, 3 H0, 4 H1.
    Returns (TermTable copy, rho0, D)."""
    X = np.ascontiguousarray(coords, float).reshape(-1, 3)
    k = X.shape[0]
    n = 4 ** k
    rho = np.zeros(n, np.complex128)
    D = ctypes.c_double()
    h = ctypes.c_void_p()
    _check(lib().oops_cce_cluster_terms(_ptr(X), ctypes.c_int32(k), ctypes.byref(params), ctypes.c_int32(which),
                                        ctypes.byref(h), _ptr(rho), ctypes.byref(D)), None, "oops_cce_cluster_terms")
    tt = _terms_from_handle(h)
    if which >= 3:
        rho = rho[: 2 ** k]
    return tt, rho, D.value


class Clusters:
    """cDepthFirstPathTracing result: clusters per order in the reference's set order."""

    def __init__(self, coords, cutoff, max_order):
        X = np.ascontiguousarray(coords, float).reshape(-1, 3)
        self.coords = X
        self.max_order = max_order
        h = ctypes.c_void_p()
        _check(lib().oops_clusters_enumerate(_ptr(X), X.shape[0], float(cutoff), max_order, ctypes.byref(h)),
               None, "oops_clusters_enumerate")
        self.h = h

    @classmethod
    def union(cls, baths, cutoff, max_order):
        """The clusters of several baths in one enumeration: their spins
        concatenated (bath b's spin i is spin first[b] + i), no connection
        between baths (cDepthFirstPathTracing on the block-diagonal connection
        graph, SpinClusterAlgorithm.h:68-82), so one batch covers every bath.
        Within each order the clusters of bath 0 come first, in their own
        set order.  Returns (Clusters, first)."""
        baths = [np.ascontiguousarray(X, float).reshape(-1, 3) for X in baths]
        first = np.cumsum([0] + [len(X) for X in baths])[:-1]
        offsets, neighbors = [0], []
        for f, X in zip(first, baths):
            d = X[:, None, :] - X[None, :, :]
            # the distance as oops_clusters_enumerate forms it
            dist = np.sqrt((d[..., 0] * d[..., 0] + d[..., 2] * d[..., 2]) + d[..., 1] * d[..., 1])
            close = dist <= cutoff
            np.fill_diagonal(close, False)
            for i in range(len(X)):
                nb = np.nonzero(close[i])[0] + f
                neighbors.extend(nb.tolist())
                offsets.append(len(neighbors))
        coords = np.concatenate(baths) if baths else np.zeros((0, 3))
        off = np.asarray(offsets, np.int64)
        nbr = np.asarray(neighbors if neighbors else [0], np.int32)
        h = ctypes.c_void_p()
        _check(lib().oops_clusters_enumerate_graph(ctypes.c_int32(len(coords)), _ptr(off), _ptr(nbr),
                                                   ctypes.c_int32(max_order), ctypes.byref(h)),
               None, "oops_clusters_enumerate_graph")
        obj = cls.__new__(cls)
        obj.coords = np.ascontiguousarray(coords)
        obj.max_order = max_order
        obj.h = h
        return obj, first

    def count(self, order):
        c = ctypes.c_int64()
        _check(lib().oops_clusters_count(self.h, ctypes.c_int32(order), ctypes.byref(c)), None, "count")
        return c.value

    def counts(self):
        return [self.count(o) for o in range(self.max_order)]

    def get(self, order):
        c = self.count(order)
        out = np.zeros((c, order + 1), np.int32)
        if c:
            _check(lib().oops_clusters_get(self.h, ctypes.c_int32(order), _ptr(out)), None, "get")
        return out

    def parents(self, order):
        c = self.count(order)
        off = np.zeros(c + 1, np.int64)
        lib().oops_clusters_parents(self.h, ctypes.c_int32(order), _ptr(off), ctypes.c_void_p(0))
        lst = np.zeros(max(int(off[-1]), 1), np.int32)
        lib().oops_clusters_parents(self.h, ctypes.c_int32(order), _ptr(off), _ptr(lst))
        return [lst[off[i]:off[i + 1]].tolist() for i in range(c)]

    def export_npz(self, path, L, time_list):
        """Write a CCE result in the variables of the reference's result file
        (CCE::export_mat_file, cce.cpp:190-230) as an .npz instead of a MATLAB
        file: CCE{i} and CCE{i}_tilder (nt x clusters of order i+1: L_C(t) and
        the reduced L~_C(t)), final_result_each_order, final_result (nt x
        max order) and time_list.  L: every cluster's L_C(t), cluster-major."""
        t = np.ascontiguousarray(time_list, np.float64)
        nt = len(t)
        L = np.ascontiguousarray(L, np.float64).reshape(-1, nt)
        r = self.reduce(L.reshape(-1), nt)
        til = r["tilde"].reshape(-1, nt)
        out, base = {}, 0
        for i, c in enumerate(self.counts()):
            out[f"CCE{i}"] = L[base:base + c].T.copy()
            out[f"CCE{i}_tilder"] = til[base:base + c].T.copy()
            base += c
        out["final_result_each_order"] = r["each"]
        out["final_result"] = r["final"]
        out["time_list"] = t.reshape(nt, 1)
        np.savez(path, **out)
        return out

    def reduce(self, L, nt):
        """CCE reduction: returns dict(tilde, final, each, log_abs, sign)."""
        L = np.ascontiguousarray(L, float)
        K = self.max_order
        til = np.zeros_like(L)
        fin = np.zeros((nt, K))
        each = np.zeros((nt, K))
        la = np.zeros((nt, K))
        sg = np.zeros((nt, K), np.int32)
        _check(lib().oops_cce_reduce(self.h, ctypes.c_int32(nt), _ptr(L), _ptr(til), _ptr(fin), _ptr(each),
                                     _ptr(la), _ptr(sg)), None, "oops_cce_reduce")
        return {"tilde": til, "final": fin, "each": each, "log_abs": la, "sign": sg}

    def hahn_batch(self, ctx: Context, params: CCEParamsC, m=30, klim=10, tol=1e-12, precision="fp64", select=None):
        """Device batch of the Hahn-echo sweeps of every cluster, or of the
        ascending global cluster indices in `select`."""
        h = ctypes.c_void_p()
        sel = None if select is None else np.ascontiguousarray(select, np.int64)
        nsel = 0 if sel is None else len(sel)
        rc = lib().oops_cce_hahn_batch_create(ctx.h, _ptr(self.coords), self.h, ctypes.byref(params),
                                              ctypes.c_int64(nsel), _ptr(sel), ctypes.c_int32(m),
                                              ctypes.c_int32(klim), ctypes.c_double(tol),
                                              ctypes.c_int32(_precision_code(precision)), ctypes.byref(h))
        _check(rc, ctx.h, "oops_cce_hahn_batch_create")
        return HahnBatch(ctx, h, sum(self.counts()) if sel is None else nsel)

    def echo_batch(self, ctx: Context, params: CCEParamsC, npulse, t, m=30, klim=10, tol=1e-3, precision="mixed"):
        """Device batch of every cluster under a Hahn echo (npulse 1; t holds
        the sweep times t/2) or CPMG-N (t holds the times t), one time grid for
        all clusters (oops_cce_echo_batch_create)."""
        t = np.ascontiguousarray(t, np.float64)
        h = ctypes.c_void_p()
        rc = lib().oops_cce_echo_batch_create(ctx.h, _ptr(self.coords), self.h, ctypes.byref(params),
                                              ctypes.c_int32(npulse), ctypes.c_int32(len(t)), _ptr(t),
                                              ctypes.c_int32(m), ctypes.c_int32(klim), ctypes.c_double(tol),
                                              ctypes.c_int32(_precision_code(precision)), ctypes.byref(h))
        _check(rc, ctx.h, "oops_cce_echo_batch_create")
        nclu = sum(self.counts())
        b = EchoBatch(ctx, h, nclu)
        b.nts = [len(t)] * nclu
        return b

    def single_sample_batch(self, ctx: Context, params: CCEParamsC, states, npulse, t, m=30, klim=10, tol=1e-12,
                            precision="fp64"):
        """Device batch of every cluster under the single-sample evolution of
        a pure bath state (oops_cce_single_sample_batch_create): psi_1 under
        riffle(H0, H1, N), psi_2 under riffle(H1, H0, N), L = Re <psi_1, psi_2>
        at the times t (uniform grids from 0 as the reference assumes)."""
        t = np.ascontiguousarray(t, np.float64)
        st = np.ascontiguousarray(states, np.int32)
        h = ctypes.c_void_p()
        rc = lib().oops_cce_single_sample_batch_create(ctx.h, _ptr(self.coords), self.h, ctypes.byref(params),
                                                       _ptr(st), ctypes.c_int32(npulse), ctypes.c_int32(len(t)),
                                                       _ptr(t), ctypes.c_int32(m), ctypes.c_int32(klim),
                                                       ctypes.c_double(tol), ctypes.c_int32(_precision_code(precision)),
                                                       ctypes.byref(h))
        _check(rc, ctx.h, "oops_cce_single_sample_batch_create")
        nclu = sum(self.counts())
        b = EchoBatch(ctx, h, nclu)
        b.nts = [len(t)] * nclu
        return b

    def close(self):
        if self.h:
            lib().oops_clusters_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def dlaran_fill(iseed, count):
    seed = (ctypes.c_int32 * 4)(*iseed)
    out = np.zeros(count)
    _check(lib().oops_dlaran_fill(seed, ctypes.c_int64(count), _ptr(out)), None, "oops_dlaran_fill")
    iseed[:] = list(seed)
    return out


def bath_states(seed, nspin):
    """Basis state (0: m = +1/2, 1: m = -1/2) of every bath spin for a
    single-sample bath: glibc rand() % 2 after srand(seed)."""
    out = np.zeros(nspin, np.int32)
    _check(lib().oops_bath_states(ctypes.c_uint32(seed), ctypes.c_int32(nspin), _ptr(out)), None, "oops_bath_states")
    return out


def single_sample_terms(coords, cluster, params: CCEParamsC, states, which=0):
    """(TermTable of H0 (which 0) or H1 (which 1) with the bath mean field,
    the cluster's product state psi) of one cluster of a single-sample bath."""
    X = np.ascontiguousarray(coords, np.float64)
    cl = np.ascontiguousarray(cluster, np.int32)
    st = np.ascontiguousarray(states, np.int32)
    psi = np.zeros(1 << len(cl), np.complex128)
    h = ctypes.c_void_p()
    _check(lib().oops_cce_single_sample_terms(_ptr(X), ctypes.c_int32(len(X)), _ptr(cl), ctypes.c_int32(len(cl)),
                                              ctypes.byref(params), _ptr(st), ctypes.c_int32(which), ctypes.byref(h),
                                              _ptr(psi)), None, "oops_cce_single_sample_terms")
    return _terms_from_handle(h), psi


def op_tables(ctx: Context, coords, params: CCEParamsC, which=0, device=True):
    """Compiled tables of one cluster's generator from the device term builder
    (device=True) or the host compiler: dict of counts, locs, val, off, eslot, masks."""
    X = np.ascontiguousarray(coords, np.float64).reshape(-1, 3)
    k = len(X)
    max_loc, max_ent = 64, 1 << 14
    counts = np.zeros(6, np.int32)
    locs = np.zeros(6 * max_loc, np.int32)
    val = np.zeros(2 * max_ent)
    off = np.zeros(max_ent, np.int32)
    eslot = np.zeros(max_ent, np.uint8)
    masks = np.zeros(32, np.int32)
    _check(lib().oops_cce_op_tables(ctx.h, _ptr(X), ctypes.c_int32(k), ctypes.byref(params), ctypes.c_int32(which),
                                    ctypes.c_int32(1 if device else 0), ctypes.c_int32(max_loc), ctypes.c_int32(max_ent),
                                    _ptr(counts), _ptr(locs), _ptr(val), _ptr(off), _ptr(eslot), _ptr(masks)),
           ctx.h, "oops_cce_op_tables")
    nl, ne = int(counts[0]), int(counts[1])
    return {"counts": counts.copy(), "locs": locs[:6 * nl].reshape(nl, 6).copy(),
            "val": val[:2 * ne].view(np.complex128).copy(), "off": off[:ne].copy(), "eslot": eslot[:ne].copy(),
            "masks": masks.copy()}
