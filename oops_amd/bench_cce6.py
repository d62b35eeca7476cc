"""Mixed-precision CCE-6 Hahn echo in Liouville space (dims 4 .. 4096).

Workload: diamond 13C bath r = 4 nm, p = 0.011, cutoff 5 A, clusters of 1-6
spins, B = 0.05 T along [111], 401 points on 0-5 ms, m = 30, tol 1e-3.
Mixed precision: n <= 256 on the batched kernels (flip-mask Lanczos for
n = 256), n = 1024 and 4096 on the batched medium kernel (1024 threads per
work unit, flip-mask product).  The baths of seeds 0 .. S-1 are enumerated
together (Clusters.union: their spins concatenated, no connection between
baths), so one batch holds every cluster and each bucket is one launch whose
work units fill the device (with one batch per bath, the launches share
the process's four hardware queues and run four at a time).

Reported: cluster expv/s over all seeds (device time of the concurrent
launches), per-bucket kernels and times, the CCE-6 coherence of seed 0, and
the error of every L_c(t) of seed 0 (every cluster, every time point)
against FP64 (m = 30, tol = 1e-12: the batched Arnoldi kernel for n <= 256,
the single-state path for n = 1024 / 4096).

    python -m oops_amd.bench_cce6 [--seeds 8] [--no-fp64]
"""
import argparse
import json
import sys
import threading
import time

import numpy as np


def _alive(label, t0, done):
    while not done.wait(20.0):
        print(f"# {label}: {time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)


def main(argv=None):
    import oops_amd as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=1)
    ap.add_argument("--no-fp64", dest="no_fp64", action="store_true")
    a = ap.parse_args(argv)
    ctx = O.Context(0)
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    tl = np.linspace(0.0, 5e-3, 401)
    t0 = time.perf_counter()
    Xs = [O.bath_diamond(3.567, 40.0, 0.011, seed=seed) for seed in range(a.seeds)]
    X0 = Xs[0]
    C0 = O.Clusters(X0, 5.0, 6)
    CU, first = O.Clusters.union(Xs, 5.0, 6)
    counts0 = C0.counts()
    for o in range(6):   # bath 0's clusters lead every order of the union, in their own order
        assert np.array_equal(CU.get(o)[:counts0[o]], C0.get(o))
    batch = CU.echo_batch(ctx, par, 1, tl / 2.0, m=30, tol=1e-3, precision="mixed")
    setup = time.perf_counter() - t0
    cu = CU.counts()
    nclu = int(sum(cu))
    print(f"# {a.seeds} baths, {nclu} clusters, setup {setup:.1f}s", file=sys.stderr, flush=True)
    done = threading.Event()
    t1 = time.perf_counter()
    threading.Thread(target=_alive, args=("mixed batch", t1, done), daemon=True).start()
    batch.run()
    done.set()
    run_s = time.perf_counter() - t1
    print(f"# mixed: {nclu} clusters in {run_s:.2f}s", file=sys.stderr, flush=True)
    Ls, st = batch.results()
    LU = np.stack(Ls)
    base = np.concatenate([[0], np.cumsum(cu)])
    L = np.concatenate([LU[base[o]:base[o] + counts0[o]] for o in range(6)])   # bath 0's clusters
    final = C0.reduce(L.reshape(-1), len(tl))["final"]
    out = {"workload": "CCE-6 Hahn echo, r=4nm p=0.011, cutoff 5A, 401 points on 0-5ms, mixed m=30 tol=1e-3",
           "seeds": a.seeds, "clusters": nclu, "clusters_per_order": [int(c) for c in cu],
           "clusters_per_order_seed0": [int(c) for c in counts0],
           "setup_s": setup, "run_s": run_s, "cluster_expv_per_s": nclu / run_s,
           "buckets": batch.timing(),
           "failed_clusters": int(((st[:, 8] & 0xFF) != 0).sum()),
           "L_cce6_seed0_at_2_5ms": float(final[200, -1])}
    if not a.no_fp64:
        # FP64 on every cluster of seed 0, on the full window
        t2 = time.perf_counter()
        done = threading.Event()
        threading.Thread(target=_alive, args=("fp64 reference", t2, done), daemon=True).start()
        ref = C0.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64")
        ref.run(tl / 2.0)
        done.set()
        L64, st64 = ref.results()
        dL = np.abs(L - L64)
        order = np.repeat(np.arange(6), C0.counts())
        out["fp64_seconds"] = time.perf_counter() - t2
        out["fp64_failed_clusters"] = int(((st64[:, 8] & 0xFF) != 0).sum())
        out["max_abs_dL_every_cluster_and_time"] = float(dL.max())
        out["max_abs_dL_per_order"] = [float(dL[order == o].max()) if np.any(order == o) else 0.0 for o in range(6)]
        print(f"# max |dL| over every L(t): {dL.max():.2e}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
    sys.exit(0)
