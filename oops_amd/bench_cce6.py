"""Mixed-precision CCE-6 Hahn echo in Liouville space (dims 4 .. 4096).

Workload: diamond 13C bath r = 4 nm, p = 0.011, seed 0, cutoff 5 A, clusters of
1-6 spins, B = 0.05 T along [111], 401 points on 0-5 ms, m = 30.  Clusters up
to n = 256 run on the batched kernels (mixed precision), n = 1024 and 4096 on
the single-state path (FP64, several items concurrently).  Reports per-bucket
device time, the CCE-6 coherence and, on the n <= 256 clusters, the largest
|L_mixed - L_FP64|.

    python -m oops_amd.bench_cce6
"""
import json
import sys
import threading
import time

import numpy as np


def main():
    import oops_amd as O
    ctx = O.Context(0)
    X = O.bath_diamond(3.567, 40.0, 0.011, seed=0)
    C = O.Clusters(X, 5.0, 6)
    counts = C.counts()
    par = O.cce_params(B=np.ones(3) * 0.05 / np.sqrt(3.0))
    tau = np.linspace(0.0, 5e-3, 401) / 2
    t0 = time.perf_counter()
    b = C.hahn_batch(ctx, par, m=30, tol=1e-3, precision="mixed")
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    done = threading.Event()

    def progress():   # the batch run is one blocking call; report that it is alive
        while not done.wait(30.0):
            print(f"# running {time.perf_counter() - t1:.0f}s", file=sys.stderr, flush=True)

    th = threading.Thread(target=progress, daemon=True)
    th.start()
    b.run(tau)
    done.set()
    run_s = time.perf_counter() - t1
    print(f"# batch of {int(sum(counts))} clusters in {run_s:.1f}s", file=sys.stderr, flush=True)
    L, st = b.results()
    timing = b.timing()
    small = int(sum(counts[:4]))
    ref = C.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64", select=np.arange(small, dtype=np.int64))
    ref.run(tau)
    L64, _ = ref.results()
    final = C.reduce(L.reshape(-1), len(tau))["final"]
    out = {"workload": "CCE-6 Hahn echo, r=4nm p=0.011 seed 0, cutoff 5A, 401 points on 0-5ms, mixed m=30 tol=1e-3",
           "clusters_per_order": [int(c) for c in counts], "setup_s": setup, "run_s": run_s,
           "cluster_expv_per_s": int(sum(counts)) / run_s, "buckets": timing,
           "failed_clusters": int(((st[:, 8] & 0xFF) != 0).sum()),
           "max_abs_dL_n_le_256": float(np.abs(L[:small] - L64).max()),
           "L_cce6_at_2_5ms": float(final[200, -1]) if final.ndim == 2 else None}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
    sys.exit(0)
