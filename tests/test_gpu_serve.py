"""GPU: the serving loop (oops_server_*).

A request's CCE product L(t) is the same when it runs alone and when it runs
in a mix with other requests admitted while it is on the device (each cluster
is computed independently; tolerance 1e-12), and equals the batch API plus the
host reduction on the same bath."""
import time

import numpy as np
import pytest

import oops_amd as O

B111 = np.ones(3) * 0.05 / np.sqrt(3.0)


def _req(seed, radius, order, npulse, nt, field=0.05):
    return {"seed": seed, "radius": radius, "p": 0.011, "B": list(np.ones(3) * field / np.sqrt(3.0)),
            "cutoff": 8.0, "order": order, "npulse": npulse, "nt": nt, "t_end": 1e-3}


@pytest.mark.gpu
def test_request_alone_equals_request_in_the_mix(ctx):
    target = _req(7, 22.0, 3, 2, 21)
    s = O.Server(ctx, hbm_budget=1 << 30, max_inflight=4, tol=1e-4, threads=8)
    rid = s.submit(target)
    s.drain()
    alone = s.result(rid)["L"]
    s.close()
    s = O.Server(ctx, hbm_budget=1 << 30, max_inflight=4, tol=1e-4, threads=8)
    others = [_req(1, 20.0, 2, 1, 31), _req(2, 25.0, 3, 4, 17, 0.2), _req(3, 18.0, 3, 1, 41)]
    ids = [s.submit(others[0]), s.submit(others[1])]
    time.sleep(0.05)
    s.poll()
    mid = s.submit(target)  # arrives while the others are being built or run
    ids.append(s.submit(others[2]))
    s.drain()
    mixed = s.result(mid)["L"]
    for i in ids:
        r = s.result(i)
        assert r["clusters"] > 0 and np.all(np.isfinite(r["L"])) and abs(r["L"][0] - 1.0) < 1e-9
    s.close()
    print("max |alone - mix| =", np.abs(alone - mixed).max(), flush=True)
    assert np.abs(alone - mixed).max() <= 1e-12


@pytest.mark.gpu
def test_served_hahn_equals_batch_and_reduction(ctx):
    req = _req(4, 24.0, 3, 1, 26)
    s = O.Server(ctx, hbm_budget=1 << 30, max_inflight=2, tol=1e-4, threads=8)
    rid = s.submit(req)
    s.drain()
    L = s.result(rid)["L"]
    s.close()
    X = O.bath_diamond(3.567, 24.0, 0.011, seed=4)
    C = O.Clusters(X, 8.0, 3)
    nt = 26
    tl = (1e-3 * np.arange(nt)) / (nt - 1)          # the server's grid, same rounding
    b = C.hahn_batch(ctx, O.cce_params(B=B111), m=30, tol=1e-4, precision="mixed")
    b.run(tl * 0.5)
    Lc, _ = b.results()
    want = C.reduce(Lc.reshape(-1), len(tl))["final"][:, -1]
    assert np.abs(L - want).max() <= 1e-12


@pytest.mark.gpu
def test_failed_clusters_fail_the_request(ctx, monkeypatch):
    """A step limit forced from the host (OOPS_MAX_STEPS) makes the 4-spin
    clusters' integrations fail: the batch reports their failure flags with
    NaN in place of unwritten outputs, and the server fails the request
    instead of reducing garbage; a request without failures still completes."""
    X = O.bath_diamond(3.567, 15.0, 0.011, seed=3)
    C = O.Clusters(X, 8.0, 4)
    assert C.count(3) > 0
    monkeypatch.setenv("OOPS_MAX_STEPS", "2")
    b = C.hahn_batch(ctx, O.cce_params(B=B111), m=30, tol=1e-4, precision="mixed")
    b.run(np.linspace(0.0, 5e-4, 11))
    L, st = b.results()
    four = slice(sum(C.counts()[:3]), None)
    bad = (st[:, 8] & 0xFF) != 0
    print(f"forced failures: {int(bad.sum())} of {len(bad)} clusters", flush=True)
    assert bad[four].all() and not bad[:sum(C.counts()[:3])].any()
    assert np.isnan(L[bad]).any()
    s = O.Server(ctx, hbm_budget=1 << 30, max_inflight=2, tol=1e-4, threads=8)
    rid = s.submit({"seed": 3, "radius": 15.0, "p": 0.011, "B": list(B111), "cutoff": 8.0, "order": 4,
                    "npulse": 1, "nt": 11, "t_end": 1e-3})
    ok = s.submit(_req(4, 15.0, 3, 1, 11))
    s.drain()
    assert s.state(rid) == -5 and s.failed(rid) == C.count(3)
    with pytest.raises(O.OopsError):
        s.result(rid)
    assert s.state(ok) == 4 and s.failed(ok) == 0 and np.all(np.isfinite(s.result(ok)["L"]))
    s.close()


@pytest.mark.gpu
def test_serving_mix_at_size_equals_batches_and_fp64(ctx):
    """Four requests drawn from the serving mix at its sizes (r 3-5 nm,
    CCE 2-4, CPMG-N up to N = 16, 101-401 points on 0-2 ms, B 0.01-0.5 T
    along [111]; a CCE-4 CPMG-8 request on the factored propagator powers)
    served together: every served L(t) equals the batch API plus the host CCE
    reduction on the same bath (1e-12), and the CCE-4 Hahn request's clusters
    agree with FP64 (m = 30, tol = 1e-12) to 1e-6 at every point."""
    t_end, m, tol = 2e-3, 30, 3e-3
    reqs = [{"seed": 11, "radius": 30.0, "order": 4, "npulse": 1, "nt": 401, "field": 0.05},
            {"seed": 12, "radius": 40.0, "order": 3, "npulse": 16, "nt": 201, "field": 0.3},
            {"seed": 13, "radius": 50.0, "order": 2, "npulse": 8, "nt": 101, "field": 0.01},
            {"seed": 14, "radius": 35.0, "order": 4, "npulse": 8, "nt": 301, "field": 0.2}]
    s = O.Server(ctx, hbm_budget=32 << 30, max_inflight=4, m=m, tol=tol, threads=16)
    ids = []
    for r in reqs:
        q = {"seed": r["seed"], "radius": r["radius"], "p": 0.011, "B": list(np.ones(3) * r["field"] / np.sqrt(3.0)),
             "cutoff": 8.0, "order": r["order"], "npulse": r["npulse"], "nt": r["nt"], "t_end": t_end}
        ids.append(s.submit(q))
    s.drain()
    served = [s.result(i) for i in ids]
    s.close()
    for r, got in zip(reqs, served):
        X = O.bath_diamond(3.567, r["radius"], 0.011, seed=r["seed"])
        C = O.Clusters(X, 8.0, r["order"])
        par = O.cce_params(B=np.ones(3) * r["field"] / np.sqrt(3.0))
        t = t_end * np.arange(r["nt"]) / (r["nt"] - 1)
        if r["npulse"] == 1:
            t = t * 0.5
        b = C.echo_batch(ctx, par, r["npulse"], t, m=m, tol=tol, precision="mixed")
        b.run()
        Ls, st = b.results()
        assert ((st[:, 8] & 0xFF) == 0).all()
        Lc = np.stack(Ls)
        want = C.reduce(Lc.reshape(-1), r["nt"])["final"][:, -1]
        print(f"request r={r['radius']} CCE-{r['order']} N={r['npulse']} nt={r['nt']}: {got['clusters']} clusters, "
              f"L(t_end) {got['L'][-1]:.6f}, max |served - batch| {np.abs(got['L'] - want).max():.1e}", flush=True)
        assert got["clusters"] == sum(C.counts())
        assert np.abs(got["L"] - want).max() <= 1e-12
        if r["order"] == 4 and r["npulse"] == 1:
            ref = C.hahn_batch(ctx, par, m=30, tol=1e-12, precision="fp64")
            ref.run(t)
            L64, st64 = ref.results()
            assert ((st64[:, 8] & 0xFF) == 0).all()
            d = np.abs(Lc - L64).max()
            print(f"  CCE-4 request clusters vs FP64: max |dL| {d:.2e}", flush=True)
            assert d <= 1e-6
