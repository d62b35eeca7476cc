"""A few Krylov steps of the two single-state workloads (a 10-spin Liouville
state, n = 2^20, and a 14-spin Hilbert state, n = 16 384; 1.5 nm sphere,
B = 0.05 T along [111], m = 40, tol = 1e-12, outputs at 0-3 us), short
enough to run under a kernel tracer and read off the per-kernel costs of an
Arnoldi cycle:

    rocprofv3 --kernel-trace --stats -f csv -d DIR -o run -- python3 -m oops_amd.trace_large
"""
import sys

import numpy as np


def main():
    import oops_amd as O
    from oops_amd.bench_large import sphere
    ctx = O.Context(0)
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    for k, which, label in ((10, 0, "liouville-10"), (14, 3, "hilbert-14")):
        T, _, _ = O.cluster_terms(sphere(k, 15.0, 0), par, which)
        rng = np.random.default_rng(1)
        v = rng.standard_normal(T.n) + 1j * rng.standard_normal(T.n)
        v /= np.linalg.norm(v)
        ov, st = O.expv(ctx, T, v, np.arange(4) * 1e-6, m=40, tol=1e-12, overlaps=True)
        print(f"{label}: n={T.n} Krylov steps {st['nstep']} products {st['nmult']} device {st['kernel_ms']:.1f} ms",
              flush=True)


if __name__ == "__main__":
    main()
    sys.exit(0)
