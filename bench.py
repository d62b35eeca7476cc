"""Benchmark: cluster expv/s of the batched CCE-4 Hahn echo on one MI355X,
mixed precision against FP64.

Workload (the batched CCE-4 configuration): a diamond-lattice 13C bath
(a = 3.567 A, p = 0.011, sphere r = 5 nm, seed 0, NV sites excluded), every
connected cluster of 1-4 spins with pair cutoff 8 A, B = 0.05 T along [111],
Hahn echo sampled at 401 points on 0-5 ms.  One step = the two Krylov sweeps
of EVERY cluster of that enumeration in the mixed-precision kernel (complex
FP32 couplings and basis, FP64 orthogonalisation and Hessenberg exponential;
Krylov dimension and tolerance from oops_amd/tuning.json, written by
oops_amd/autotune.py), reduced on the device to each cluster's L_c(t).
Inputs (term tables, initial vectors) are uploaded before the timed region.

Beside it, the FP64 path (m = 30, tol = 1e-12, the reference's settings) runs
on a stratified sample of the same clusters (every k-th cluster of every
order): its throughput gives the mixed/FP64 ratio and its L_c(t) the error
max |L_mixed - L_fp64| over every sampled cluster and time point.  The same
sample also runs in FP64 with the mixed path's algorithm, Krylov dimension,
tolerance and schedule (precision "fp64-lanczos"), which isolates the gain of
the FP32 parts alone.

The JSON line carries the dominant kernel's roofline (the algorithmic flops
of its executed products and Lanczos steps over its launch time, measured
with HIP events on its stream, against the published FP64/FP32 vector peaks
blended by the flops of each precision; the executed-flop rate beside it), a
CPU baseline (the oracle's restatement of the reference CPU path, FP64, timed
on a bounded sample and extrapolated; see the `sample` field) and a serving
segment (50 req/s, saturation, 70 % of saturation).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X vector FP64, published
PEAK_FP32_TFLOPS = 157.3  # MI355X vector FP32, published (packed FP32; this build does not emit packed ops)


def matvec_flops(n, k):
    """Algorithmic flops of one merged product (one s x s operator per spin,
    one s^2 x s^2 operator per pair, s = 4 Liouville): 8 n (k s + P s^2),
    P = k(k-1)/2.  Returned as (one-spin part, pair part)."""
    s = 4
    P = k * (k - 1) // 2
    return 8.0 * n * k * s, 8.0 * n * P * s * s


def executed_matvec_flops(n, k):
    """Flops the flip-mask product executes: per element one complex
    multiply-add per flip mask, 1 + 2k one-spin masks and k(k-1) two-spin
    masks (21 at k = 4, against the 112 entries of the merged operators).
    Returned as (one-spin part, pair part)."""
    return 8.0 * n * (1 + 2 * k), 8.0 * n * k * (k - 1)


def lanczos_step_flops(n, nstep, m):
    """Per Krylov step: m Lanczos iterations, each one fused <v_j, p>, |p|^2
    reduction and the three-term update (4 complex dot/axpy passes, 32 n),
    plus the (m+1)-term combination of the update (8 n (m+1))."""
    return nstep * (32.0 * n * m + 8.0 * n * (m + 1))


def krylov_flops(n, k, nmult, nstep, m):
    """Algorithmic flops of the Arnoldi path (all FP64): the products + one
    classical Gram-Schmidt pass per Arnoldi step, 8 n m (m+1) per Krylov cycle."""
    one, pair = matvec_flops(n, k)
    return nmult * (one + pair) + nstep * 8.0 * n * m * (m + 1), 0.0


def lanczos_flops(n, k, nmult, nstep, m, pair_fp32=True, executed=False):
    """(FP64 flops, FP32 flops) of the flip-mask Lanczos kernel: the one-spin
    part of every product and the Lanczos recurrence in FP64, the pair part in
    FP32 (mixed) or FP64 (the FP64 variant)."""
    one, pair = (executed_matvec_flops if executed else matvec_flops)(n, k)
    f64 = nmult * one + lanczos_step_flops(n, nstep, m)
    f32 = nmult * pair
    return (f64, f32) if pair_fp32 else (f64 + f32, 0.0)


def blended_peak(f64, f32):
    """Peak of a mix of FP64 and FP32 vector flops: the time the mix takes at
    the two published peaks, inverted."""
    t = f64 / PEAK_FP64_TFLOPS + f32 / PEAK_FP32_TFLOPS
    return (f64 + f32) / t if t > 0 else PEAK_FP64_TFLOPS


# device function of each kernel kind (oops_hahn_batch_kernels), for the
# rocprofv3 summaries
DEVICE_NAMES = {"lanczos_slot": "cce_hahn_lanczos_slot_kernel<{m}, false>",
                "lanczos_slot_fp64": "cce_hahn_lanczos_slot_kernel<{m}, true>", "lanczos_mixed": "cce_hahn_lanczos_kernel<{m}>",
                "arnoldi_fp64": "cce_hahn_krylov_kernel<{m}, false>", "arnoldi_mixed": "cce_hahn_krylov_kernel<{m}, true>",
                "dense64_mfma": "cce_dense64_kernel", "dense_pade": "cce_hahn_dense_kernel"}


def build_workload(args):
    import oops_amd as O
    X = O.bath_diamond(3.567, args.radius, 0.011, seed=args.seed, shape="sphere")
    C = O.Clusters(X, args.cutoff, args.order)
    par = O.cce_params(B=np.ones(3) * args.field / np.sqrt(3.0))
    tlist = np.linspace(0.0, args.t_end, args.nt)
    return X, C, par, tlist


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(X, C, par, tlist, args, counts):
    """Oracle (CPU restatement of the reference path, FP64, tol 1e-12) on a
    bounded sample: `threads` clusters of every order (evenly spaced through
    the order) over the first quarter of the output points, one cluster per
    thread on every host core this process may use (at most 16, the host's
    share of one GPU).  Krylov steps are uniform in sweep time on the uniform
    grid, so the per-cluster time is scaled by the ratio of sweep times and
    weighted by the cluster mix."""
    import oops_amd as O
    from oracle import physics as P
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, min(16, nproc))
    nsamp = (len(tlist) - 1) // 4 + 1          # 101 of 401 points
    tau = tlist[:nsamp] / 2.0
    scale_t = (tlist[-1] / 2.0) / tau[-1]
    per = threads if args.cpu_per is None else args.cpu_per
    total_cpu = 0.0
    detail = []
    for order in range(args.order):
        cl = C.get(order)
        idx = np.unique(np.linspace(0, len(cl) - 1, min(per, len(cl))).astype(int))
        A, B, R, D = [], [], [], []
        for r in cl[idx]:
            a, rho, d = O.cluster_terms(X[list(r)], par, 0)
            b, _, _ = O.cluster_terms(X[list(r)], par, 2)
            A.append(P.TermTable(a.spin_dim.size, a.coeff, a.nbody, a.pos_offset, a.pos, a.dims, a.mat_offset,
                                 a.mats, a.spin_dim, a.n))
            B.append(P.TermTable(b.spin_dim.size, b.coeff, b.nbody, b.pos_offset, b.pos, b.dims, b.mat_offset,
                                 b.mats, b.spin_dim, b.n))
            R.append(rho)
            D.append(d)
        t0 = time.perf_counter()
        c0 = time.process_time()
        _, _, rc = P.hahn_batch(A, B, R, R, D, tau, m=args.m, tol=args.tol, nthreads=threads)
        cpu = time.process_time() - c0     # CPU seconds of every thread of the process
        wall = time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle failed")
        core_s = cpu / len(idx) * scale_t  # core-seconds per cluster over the whole window
        total_cpu += core_s * counts[order] / threads
        detail.append(f"order {order + 1}: {len(idx)} clusters, {cpu:.2f} core-s in {wall:.2f} s")
    value = sum(counts) / total_cpu
    return {
        "value": value,
        "unit": "cluster expv/s",
        "cores": threads,
        "kind": "port",
        "nproc": nproc,
        "cpu_model": cpu_model(),
        "sample": (f"oracle C++ restatement (g++ -O3, FP64, m={args.m}, tol={args.tol:g}, one cluster per thread on "
                   f"{threads} threads) on {per} clusters per order over the first {nsamp} of {len(tlist)} output "
                   f"points (tau <= {tau[-1]:.3g} s), scaled x{scale_t:.2f} in sweep time and weighted by the "
                   f"cluster mix {counts}; " + "; ".join(detail)),
    }


def stratified(counts, frac):
    """Every round(1/frac)-th cluster of every order (global indices)."""
    step = max(1, int(round(1.0 / frac)))
    sel, base = [], 0
    for c in counts:
        sel += list(range(base, base + c, step))
        base += c
    return np.array(sel, np.int64)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when no summary covers it."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_cce4_r5nm_pmc.md"))):
        txt = open(path).read()
        m = re.search(r"^kernel:\s*(.+?)\s*$", txt, re.M)
        t = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+-]+)", txt)
        if m and t and m.group(1) == kernel:
            return float(t.group(1))
    return None


def tuned_mixed():
    """The autotuner's result: (m, tol, product form, launch shapes, source)."""
    path = os.path.join(ROOT, "oops_amd", "tuning.json")
    if os.path.exists(path):
        with open(path) as f:
            t = json.load(f)
        return (int(t["mixed"]["m"]), float(t["mixed"]["tol"]), t["mixed"].get("precision", "mixed"),
                t.get("launch", {}), "oops_amd/tuning.json")
    return 30, 1e-4, "mixed", {}, "default"


def roofline(counts, st, timing, m, kernel):
    """Roofline of one Krylov kernel over all its launches (the buckets that
    ran it): algorithmic flops of the executed products and Krylov steps per
    launch over the average launch time (HIP events on the launch stream).
    The kernels are VALU kernels; their flops are part FP64 and part FP32
    (pair terms of the mixed flip-mask kernel), so the peak is the published
    FP64 / FP32 vector peaks blended by the flops of each precision."""
    base, f64, f32, e64, e32, ms, nl = 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0
    by_n = {b["n"]: b for b in timing}
    for o, cnt in enumerate(counts):
        n = 4 ** (o + 1)
        if n in by_n and by_n[n]["kernel"] == kernel and cnt > 0:
            sl = st[base:base + cnt]
            nmult = sl[:, 1].sum() + sl[:, 5].sum()
            nstep = sl[:, 0].sum() + sl[:, 4].sum()
            if kernel.startswith("lanczos_slot"):
                pair32 = kernel == "lanczos_slot"
                a64, a32 = lanczos_flops(n, o + 1, nmult, nstep, m, pair32)
                x64, x32 = lanczos_flops(n, o + 1, nmult, nstep, m, pair32, executed=True)
            else:
                a64, a32 = krylov_flops(n, o + 1, nmult, nstep, m)
                x64, x32 = a64, a32
            f64 += a64
            f32 += a32
            e64 += x64
            e32 += x32
            ms += by_n[n]["ms"]
            nl += 1
        base += cnt
    flops = f64 + f32
    sec = ms * 1e-3
    achieved = flops / sec / 1e12 if ms > 0 else 0.0
    peak = blended_peak(f64, f32)
    return {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None, "launches": nl,
            "avg_launch_ms": ms / max(nl, 1), "algorithmic_flops_per_launch": flops / max(nl, 1),
            "fp32_share_of_flops": f32 / flops if flops else 0.0,
            "peak_note": (f"published vector peaks FP64 {PEAK_FP64_TFLOPS} / FP32 {PEAK_FP32_TFLOPS} TFLOP/s "
                          "blended by the algorithmic flops of each precision"),
            "executed_tflops": (e64 + e32) / sec / 1e12 if ms > 0 else 0.0,
            "executed_frac": ((e64 + e32) / sec / 1e12) / blended_peak(e64, e32) if ms > 0 else 0.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--radius", type=float, default=50.0)
    ap.add_argument("--cutoff", type=float, default=8.0)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--field", type=float, default=0.05)
    ap.add_argument("--t-end", dest="t_end", type=float, default=5e-3)
    ap.add_argument("--nt", type=int, default=401)
    ap.add_argument("--m", type=int, default=30, help="Krylov dimension of the FP64 path")
    ap.add_argument("--tol", type=float, default=1e-12, help="tolerance of the FP64 path")
    ap.add_argument("--mixed-m", dest="mixed_m", type=int, default=None)
    ap.add_argument("--mixed-tol", dest="mixed_tol", type=float, default=None)
    ap.add_argument("--fp64-frac", dest="fp64_frac", type=float, default=0.25,
                    help="fraction of every order's clusters run on the FP64 path")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", dest="no_fp64", action="store_true",
                    help="skip the FP64 comparison (profiling runs)")
    ap.add_argument("--serve-requests", dest="serve_requests", type=int, default=20,
                    help="requests of the serving segment (Poisson arrivals at 50/s); 0 skips it")
    ap.add_argument("--cpu-per", dest="cpu_per", type=int, default=None)
    ap.add_argument("--serve-cpmg-order", dest="serve_cpmg_order", type=int, default=4,
                    help="highest CCE order of the serving mix's CPMG-N requests")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build targets one GPU; --gpus must be 1 "
                                   "(clusters are independent: across GPUs each rank would take a "
                                   "cost-balanced slice of every bucket and one all-reduce of the "
                                   "nt-point log L(t) would combine them)"}))
        sys.exit(2)

    import oops_amd as O
    mm, mtol, mprec, launch, msrc = tuned_mixed()
    mm = args.mixed_m or mm
    mtol = args.mixed_tol or mtol
    ctx = O.Context(0)
    for kname, wg in launch.items():
        ctx.set_launch(kname, int(wg))
    X, C, par, tlist = build_workload(args)
    counts = C.counts()
    nclu = sum(counts)
    tau = tlist / 2.0
    t0 = time.perf_counter()
    batch = C.hahn_batch(ctx, par, m=mm, tol=mtol, precision=mprec)
    setup_mixed_s = time.perf_counter() - t0
    sel = stratified(counts, args.fp64_frac)
    batch64 = None if args.no_fp64 else C.hahn_batch(ctx, par, m=args.m, tol=args.tol, precision="fp64",
                                                        select=sel)
    setup_s = time.perf_counter() - t0
    print(f"# bath {len(X)} spins, clusters per order {counts} ({nclu}), fp64 sample {len(sel)}, "
          f"setup {setup_s:.2f}s", file=sys.stderr, flush=True)
    # prime code objects on a small batch (not a step)
    Cs = O.Clusters(X[:40], args.cutoff, 2)
    Cs.hahn_batch(ctx, par, m=mm, tol=mtol, precision=mprec).run(tau[:3])
    if batch64 is not None:
        Cs.hahn_batch(ctx, par, m=args.m, tol=args.tol).run(tau[:3])
    for w in range(args.warmup):
        batch.run(tau)
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for s in range(args.steps):
        batch.run(tau)             # synchronises the device before returning
        L, st = batch.results()    # every cluster's L(t) and counters back on the host
        print(f"# step {s + 1}/{args.steps} done {time.perf_counter() - t0:.2f}s", file=sys.stderr, flush=True)
    elapsed = time.perf_counter() - t0
    timing = batch.timing()
    value = nclu * args.steps / elapsed

    # dominant kernel: the kind with the most device time
    per_kind = {}
    for bkt in timing:
        per_kind[bkt["kernel"]] = per_kind.get(bkt["kernel"], 0.0) + bkt["ms"]
    dom = max(per_kind, key=per_kind.get)
    roof = roofline(counts, st, timing, mm, dom)
    roof["kernel"] = DEVICE_NAMES.get(dom, dom).format(m=mm)
    roof["traffic"] = measured_traffic(roof["kernel"])
    fp64 = None
    if batch64 is not None:
        # FP64 path on the stratified sample (same clusters, same times)
        t1 = time.perf_counter()
        batch64.run(tau)
        el64 = time.perf_counter() - t1
        L64, st64 = batch64.results()
        rate64 = len(sel) / el64
        print(f"# fp64 sample {len(sel)} clusters in {el64:.2f}s", file=sys.stderr, flush=True)
        dL = np.abs(L[sel] - L64)
        order_of = np.repeat(np.arange(len(counts)), counts)[sel]
        err_per_order = [float(dL[order_of == o].max()) if np.any(order_of == o) else 0.0
                         for o in range(len(counts))]
        sample_counts = [int(np.sum(order_of == o)) for o in range(len(counts))]
        roof64 = roofline(sample_counts, st64, batch64.timing(), args.m, "arnoldi_fp64")
        fp64 = {"m": args.m, "tol": args.tol, "sample_clusters": int(len(sel)), "value": rate64, "seconds": el64,
                "buckets": batch64.timing(), "failed_clusters": int(((st64[:, 8] & 0xFF) != 0).sum()),
                "roofline_frac": roof64["frac"], "roofline_achieved": roof64["achieved"],
                "mixed_over_fp64": value / rate64, "max_abs_dL": float(dL.max()),
                "max_abs_dL_per_order": err_per_order}
        # the same sample in FP64 with the mixed path's algorithm, Krylov
        # dimension, tolerance and schedule: what the FP32 parts alone gain
        bl = C.hahn_batch(ctx, par, m=mm, tol=mtol, precision="fp64-lanczos", select=sel)
        t2 = time.perf_counter()
        bl.run(tau)
        Ll, stl = bl.results()
        ell = time.perf_counter() - t2
        mixed_sample = C.hahn_batch(ctx, par, m=mm, tol=mtol, precision=mprec, select=sel)
        t3 = time.perf_counter()
        mixed_sample.run(tau)
        Lms, _ = mixed_sample.results()
        elm = time.perf_counter() - t3
        fp64["same_algorithm"] = {
            "precision": "fp64-lanczos", "m": mm, "tol": mtol, "value": len(sel) / ell, "seconds": ell,
            "mixed_seconds_same_sample": elm, "mixed_over_fp64_same_algorithm": ell / elm,
            "buckets": bl.timing(), "failed_clusters": int(((stl[:, 8] & 0xFF) != 0).sum()),
            "max_abs_dL_vs_fp64_reference_schedule": float(np.abs(Ll - L64).max()),
            "max_abs_dL_mixed_vs_fp64_same_algorithm": float(np.abs(Lms - Ll).max())}
        print(f"# fp64 same algorithm: {ell:.2f}s, mixed on the sample {elm:.2f}s", file=sys.stderr, flush=True)
        bl.close()
        mixed_sample.close()
    out = {
        "metric": "cluster expv/sec at CCE-4 (mixed vs FP64, |\u0394L|\u22641e-6); p99 request latency",
        "value": value,
        "unit": "cluster expv/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "batched CCE-4 Hahn echo, diamond 13C bath r=5nm p=0.011 seed 0, cutoff 8A, "
                               "B=0.05T||[111], 401 points on 0-5ms",
                   "clusters": nclu, "clusters_per_order": counts, "spins": int(len(X)),
                   "setup_s": setup_mixed_s,
                   "setup": "operator tables of every cluster built and uploaded (device term builder), untimed",
                   "mixed": {"m": mm, "tol": mtol, "precision": mprec, "launch": launch, "source": msrc, "buckets": timing,
                             "failed_clusters": int(((st[:, 8] & 0xFF) != 0).sum())},
                   "fp64": fp64,
                   "mixed_over_fp64": fp64["mixed_over_fp64"] if fp64 else None,
                   "max_abs_dL": fp64["max_abs_dL"] if fp64 else None},
        "roofline": roof,
    }
    if args.serve_requests > 0:
        # serving segment: the serving request mix (bath r 3-5 nm, B 0.01-0.5 T,
        # CPMG-N N in {1,2,4,8,16}, CCE 2-4, 101-401 points) with Poisson
        # arrivals at 50 req/s; then the same mix submitted at once (the
        # saturation throughput, requests/s of device capacity) and at 70 % of
        # that rate, where latency is service time rather than queueing
        from oops_amd import serve_bench
        reqs = serve_bench.make_requests(args.serve_requests, 0, 30.0, 50.0, 2e-3, args.serve_cpmg_order, 4)
        summary, _ = serve_bench.run(ctx, reqs, 50.0, 0, 32 << 30, 8, mtol, mm, 16, log=sys.stderr)
        summary.pop("mix", None)
        out["serving"] = summary
        print(f"# serving p50 {summary['p50_latency_s']:.2f}s p99 {summary['p99_latency_s']:.2f}s",
              file=sys.stderr, flush=True)
        sat, _ = serve_bench.run(ctx, reqs, 1e9, 0, 32 << 30, 8, mtol, mm, 16, log=sys.stderr)
        cap = sat["requests"] / sat["wall_s"]
        nsub = max(4, args.serve_requests // 2)
        sub, _ = serve_bench.run(ctx, reqs[:nsub], 0.7 * cap, 1, 32 << 30, 8, mtol, mm, 16, log=sys.stderr)
        out["serving"]["saturation_req_per_s"] = cap
        out["serving"]["saturation_clusters_per_s"] = sat["clusters_per_s"]
        out["serving"]["below_saturation"] = {k: sub[k] for k in ("requests", "rate_per_s", "p50_latency_s",
                                                                 "p99_latency_s", "mean_device_s", "clusters_per_s")}
        print(f"# serving capacity {cap:.2f} req/s; at {0.7 * cap:.2f} req/s p50 {sub['p50_latency_s']:.2f}s "
              f"p99 {sub['p99_latency_s']:.2f}s", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(X, C, par, tlist, args, counts)
        out["config"]["mixed_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
