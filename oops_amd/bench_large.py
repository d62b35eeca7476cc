"""Single large-cluster Krylov expv on one MI355X (the HBM-bound workload).

Two states: a 14-spin 13C cluster in Hilbert space (n = 16384, the
conditional Hamiltonian with the NV in m_s = 0, i.e. dipolar + Zeeman) and a
10-spin cluster in Liouville space (n = 1 048 576, lvA).  Positions uniform
in a 1.5 nm sphere from a seed, full dipolar coupling (no cutoff), B = 0.05 T
along [111], v a seeded normalised random complex vector, m = 40, 1000 steps
of dt = 1 us (1001 output times), FP64 and FP32-basis variants.

Achieved bandwidth = algorithmic bytes / device time of the Krylov steps,
with the per-cycle bytes of a Krylov cycle of size m, n b (m^2 + 7m + 4)
(b = 16 B for the FP64 basis, 8 B for the FP32 basis), against MI355X's
8 TB/s.

    python -m oops_amd.bench_large [--nt 1001] [--only hilbert|liouville] [--steps reference|inside]
"""
import argparse
import json
import sys
import time

import numpy as np

PEAK_HBM_GBS = 8000.0


def sphere(k, radius, seed):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < k:
        p = rng.uniform(-radius, radius, 3)
        if np.linalg.norm(p) <= radius:
            out.append(p)
    return np.array(out)


def run_case(ctx, label, k, which, nt, m, tol, precision, seed=0, steps="reference"):
    import oops_amd as O
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    X = sphere(k, 15.0, seed)
    T, _, _ = O.cluster_terms(X, par, which)
    rng = np.random.default_rng(seed + 1)
    v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
    v /= np.linalg.norm(v)
    tl = np.arange(nt) * 1e-6
    t0 = time.perf_counter()
    ov, st = O.expv(ctx, T, v, tl, m=m, tol=tol, overlaps=True, precision=precision, steps=steps)
    wall = time.perf_counter() - t0
    b = 8 if precision == "fp32" else 16
    cycle_bytes = T.n * b * (m * m + 7 * m + 4)
    alg = st["nstep"] * cycle_bytes
    gbs = alg / (st["kernel_ms"] * 1e-3) / 1e9
    out = {"case": label, "n": T.n, "spins": k, "terms": int(len(T.coeff)), "precision": precision, "steps": steps, "m": m,
           "tol": tol, "nt": nt, "wall_s": wall, "device_ms": st["kernel_ms"], "krylov_steps": int(st["nstep"]),
           "products": int(st["nmult"]), "rejections": int(st["nreject"]), "anorm": st["anorm"],
           "algorithmic_bytes": alg, "achieved_GBs": gbs, "frac_of_8TBs": gbs / PEAK_HBM_GBS,
           "overlap_last": [float(ov[-1].real), float(ov[-1].imag)]}
    print(json.dumps(out), flush=True)
    return out, ov


def main(argv=None):
    import oops_amd as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--nt", type=int, default=1001)
    ap.add_argument("--m", type=int, default=40)
    ap.add_argument("--only", default=None)
    ap.add_argument("--steps", default="reference", choices=["reference", "inside"],
                    help="reference: a step ends at every output time; inside: output times inside steps")
    a = ap.parse_args(argv)
    ctx = O.Context(0)
    res = []
    for label, k, which in (("hilbert", 14, 3), ("liouville", 10, 0)):
        if a.only and a.only != label:
            continue
        r64, ov64 = run_case(ctx, label, k, which, a.nt, a.m, 1e-12, "fp64", steps=a.steps)
        r32, ov32 = run_case(ctx, label, k, which, a.nt, a.m, 1e-12, "fp32", steps=a.steps)
        r32["max_abs_overlap_diff_vs_fp64"] = float(np.abs(ov32 - ov64).max())
        res += [r64, r32]
        print(f"# {label}: fp32 vs fp64 max |<v,w>| difference {r32['max_abs_overlap_diff_vs_fp64']:.2e}",
              file=sys.stderr, flush=True)
    print(json.dumps({"large_single_cluster": res}), flush=True)


if __name__ == "__main__":
    main()
