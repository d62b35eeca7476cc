"""Serving benchmark: Poisson arrivals of CCE decoherence requests.

Request mix (the serving configuration): bath seed, radius uniform in
[rmin, rmax] A (p = 0.011, cutoff 8 A), B uniform in [0.01, 0.5] T along
[111], a CPMG-N sequence with N in {1, 2, 4, 8, 16}, CCE order 2-4 and
101-401 time points on [0, t_end].  CPMG-N with N >= 2 is drawn with CCE
order <= --cpmg-max-order (default 4: the 4-spin clusters of CPMG-N requests
run on the factored propagator powers; with OOPS_FACTORED=0 they would run on
the Krylov path with one restart per output time, minutes per request at
r = 5 nm).

Requests are submitted at their arrival times to an oops_amd.Server (mixed
precision, continuous batching); latency = completion - arrival.

    python -m oops_amd.serve_bench --rate 50 --requests 40
"""
import argparse
import json
import sys
import time

import numpy as np


def make_requests(n, seed, rmin, rmax, t_end, cpmg_max_order, max_order):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        npulse = int(rng.choice([1, 2, 4, 8, 16]))
        hi = max_order if npulse == 1 else min(max_order, cpmg_max_order)
        order = int(rng.integers(2, hi + 1))
        field = float(rng.uniform(0.01, 0.5))
        out.append({"seed": int(rng.integers(0, 2 ** 31)), "radius": float(rng.uniform(rmin, rmax)), "p": 0.011,
                    "B": [field / np.sqrt(3.0)] * 3, "cutoff": 8.0, "order": order, "npulse": npulse,
                    "nt": int(rng.integers(101, 402)), "t_end": t_end})
    return out


def run(ctx, requests, rate, seed, budget, inflight, tol, m, threads, log=None):
    import oops_amd as O
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(requests)))
    srv = O.Server(ctx, hbm_budget=budget, max_inflight=inflight, m=m, tol=tol, precision="mixed", threads=threads)
    t0 = time.perf_counter()
    ids = []
    nxt = 0
    done = 0
    last_report = t0
    while done < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            ids.append(srv.submit(requests[nxt]))
            nxt += 1
        done += srv.poll()
        if log and time.perf_counter() - last_report > 5.0:
            last_report = time.perf_counter()
            print(f"# t={now:.1f}s submitted {nxt} done {done}", file=log, flush=True)
        if nxt < len(requests) and done < nxt:
            time.sleep(0.0005)
        elif nxt < len(requests):
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - now)))
        else:
            time.sleep(0.0005)
    wall = time.perf_counter() - t0
    res = [srv.result(i) for i in ids]
    srv.close()
    lat = np.array([r["latency_s"] for r in res])
    clusters = int(sum(r["clusters"] for r in res))
    return {
        "requests": len(requests), "rate_per_s": rate, "wall_s": wall,
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()), "max_latency_s": float(lat.max()),
        "clusters": clusters, "clusters_per_s": clusters / wall,
        "max_footprint_bytes": int(max(r["footprint"] for r in res)),
        "mean_build_s": float(np.mean([r["built_s"] for r in res])),
        "mean_device_s": float(np.mean([r["latency_s"] - r["launched_s"] for r in res])),
        "mix": {"orders": [q["order"] for q in requests], "npulse": [q["npulse"] for q in requests],
                "nt": [q["nt"] for q in requests]},
    }, res


def main(argv=None):
    import oops_amd as O
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=50.0)
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rmin", type=float, default=30.0)
    ap.add_argument("--rmax", type=float, default=50.0)
    ap.add_argument("--t-end", dest="t_end", type=float, default=2e-3)
    ap.add_argument("--max-order", dest="max_order", type=int, default=4)
    ap.add_argument("--cpmg-max-order", dest="cpmg_max_order", type=int, default=4)
    ap.add_argument("--budget-gb", dest="budget_gb", type=float, default=32.0)
    ap.add_argument("--inflight", type=int, default=8)
    ap.add_argument("--tol", type=float, default=1e-3)
    ap.add_argument("--m", type=int, default=30)
    ap.add_argument("--threads", type=int, default=16)
    a = ap.parse_args(argv)
    ctx = O.Context(0)
    reqs = make_requests(a.requests, a.seed, a.rmin, a.rmax, a.t_end, a.cpmg_max_order, a.max_order)
    summary, _ = run(ctx, reqs, a.rate, a.seed, int(a.budget_gb * (1 << 30)), a.inflight, a.tol, a.m, a.threads,
                     log=sys.stderr)
    print(json.dumps(summary), flush=True)
    return summary


if __name__ == "__main__":
    main()
