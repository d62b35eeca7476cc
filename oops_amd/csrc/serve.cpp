// Serving loop: continuous batching of CCE decoherence requests on one GPU.
//
// Pipeline per request: queued -> built on a host worker (bath, cluster
// enumeration, every cluster's operators) -> admitted when an in-flight slot
// and enough of the device-memory budget are free (upload + asynchronous
// launch of its size-bucketed kernels on its own stream) -> completed on a
// later poll (results, CCE product, memory back to the budget).  HIP calls
// happen only on the thread that calls oops_server_poll.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/oops_cce.h"
#include "../../include/oops_expv.h"
#include "../../include/oops_serve.h"
#include "host_common.h"
#include "oops_kernels.h"
#include "term_builder.h"

using namespace oops;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

enum State : int32_t { kQueued = 0, kBuilding = 1, kBuilt = 2, kRunning = 3, kDone = 4 };

struct Req {
    int64_t id = 0;
    oops_request r{};
    double t_submit = 0.0;
    int32_t state = kQueued;
    int32_t err = 0;
    std::vector<double> coords;
    oops_clusters* C = nullptr;
    CceItems items;
    // device term builder path: clusters by size, builder constants, rho0 and times
    bool device_tables = false;
    std::vector<BuiltGroup> groups;
    int64_t nitems = 0;
    TermBuilderArgs consts{};
    std::vector<std::vector<double>> rho;
    std::vector<double> t;
    int64_t est_bytes = 0;
    oops_hahn_batch* batch = nullptr;
    int64_t footprint = 0;
    int64_t nclusters = 0;
    int64_t failed = 0;   // clusters whose integration failed (flags in their stats)
    std::vector<double> L;
    double latency = 0.0;
    double t_built = 0.0, t_launched = 0.0;  // since submit
    ~Req() {
        if (batch) oops_hahn_batch_destroy(batch);
        if (C) oops_clusters_destroy(C);
    }
};

// Device bytes a request will hold, matching what batch_create allocates:
// the merged ELL tables (an upper bound: one d x d block per spin, one
// d^2 x d^2 per pair), vectors and results, plus per state dimension
//   n = 64   the dense-propagator workspace (one workgroup per CU),
//   n = 256  the flip-mask coefficient tables and the Hahn-sweep state store
//            of the slot kernel (two workgroups per CU),
//   n = 1024 / 4096 (mixed) the FP32 bases and flip-mask tables of the medium
//            kernel's workgroups and its state store (both sweeps of every Hahn
//            item at every output time, left out above 32 GB as batch_create does),
//   n > 256 on the single-state FP64 path: up to eight items in flight, each
//            with its basis, work vectors and stored ket states.
int64_t workspace_bytes(int64_t n, int64_t items, int64_t hahn, int64_t nt_hahn, int64_t nt_max, int num_cu, int m,
                        int precision);

int64_t estimate_bytes(const CceItems& it, int num_cu, int m, int precision) {
    struct Dim { int64_t items = 0, hahn = 0, nt_hahn = 0, nt_max = 0; };
    std::map<int64_t, Dim> dims;
    int64_t b = 0;
    for (const EchoIn& e : it.items) {
        int64_t n = 1;
        for (int s = 0; s < e.opA->nspin; ++s) n *= e.opA->spin_dim[s];
        Dim& d = dims[n];
        ++d.items;
        d.nt_max = std::max<int64_t>(d.nt_max, e.nt);
        if (e.npulse == 1) {
            ++d.hahn;
            d.nt_hahn = std::max<int64_t>(d.nt_hahn, e.nt);
        }
        const oops_term_table* ops[4] = {e.opA, e.opB, e.opC, e.opD};
        for (const oops_term_table* o : ops) {
            if (!o) continue;
            const int64_t k = o->nspin, dl = o->spin_dim[0];
            const int64_t ent = k * dl * 4 + k * (k - 1) / 2 * dl * dl * 16;
            b += 64 + (k + k * (k - 1) / 2) * 16 + ent * 20 + kMedMaskStride * 4 + kSlotMaskStride * 4;
        }
        b += 32 * n + 8 * (int64_t)e.nt + 4 * kStatsPerItem + 64;
    }
    for (const auto& kv : dims)
        b += workspace_bytes(kv.first, kv.second.items, kv.second.hahn, kv.second.nt_hahn, kv.second.nt_max, num_cu,
                             m, precision);
    return b + 4096;
}

// Device bytes of the per-bucket workspaces of `items` items of dimension n
// (hahn of them Hahn echoes), as batch_create allocates them.
int64_t workspace_bytes(int64_t n, int64_t items, int64_t hahn, int64_t nt_hahn, int64_t nt_max, int num_cu, int m,
                        int precision) {
    int64_t b = 0;
    const int64_t c16 = 16;   // bytes of one complex FP64
    if (n == 64) {
        b += 8 * dense64_workspace_doubles((int)std::min<int64_t>(items, num_cu));
    } else if (n > m && n <= 256 && (precision == 1 || precision == 4)) {
        const int64_t grid = std::min<int64_t>(items, (precision == 4 ? 1 : 2) * (int64_t)num_cu);
        b += c16 * slot_table_double2(precision == 4) * 4 * grid;
        b += c16 * 256 * nt_hahn * std::min<int64_t>(hahn, grid);
    } else if ((n == 1024 || n == 4096) && precision == 1) {
        const int64_t grid = std::min<int64_t>(2 * items, num_cu);
        b += 8 * medium_workspace_float2((int)grid, m, (int)n);
        b += c16 * 4 * ((int64_t)kMedSlot64 * n + (int64_t)kMedSlot32 * n / 2) * grid;
        const int64_t store = c16 * 2 * n * nt_hahn * hahn;
        if (store <= (int64_t(32) << 30)) b += store;
    } else if (n > 256) {
        const int64_t inflight = std::min<int64_t>(items, 8);
        b += inflight * c16 * n * (2 * (m + 1) + 8 + nt_max);
    }
    return b;
}

// Device bytes of a batch built by the device term builder: the fixed-stride
// tables of every generator (batch_create_built), the shared rho0 and DLARAN
// vectors, the per-item sequences and results, and the workspaces.
int64_t estimate_bytes_built(const std::vector<BuiltGroup>& groups, int npulse, int64_t nt, int num_cu, int m,
                             int precision) {
    int64_t b = 4096;
    const int nops = npulse > 1 ? 4 : 2;
    for (const BuiltGroup& g : groups) {
        const int64_t k = g.k, ni = (int64_t)g.items.size(), n = int64_t(1) << (2 * k), P = k * (k - 1) / 2;
        if (!ni) continue;
        if (factored_applies((int)k, npulse, 0.0)) {   // spins, per-item arrays and L only
            b += ni * (4 * k + 4 * 4 + 8 + 8 + 4 * kStatsPerItem + 8 * nt) + 8 * nt;
            continue;
        }
        const int64_t ent = 16 * k + 128 * P;
        b += ni * nops * (64 + (k + P) * 16 + ent * 21 + (kSlotMaskStride + kMedMaskStride + kTBSummary) * 4);
        b += ni * (4 * 4 + 8 + 8 + 4 + 4 + 8 + 8 + 4 * kStatsPerItem + 8 * nt) + 48 * n + 8 * nt;
        b += workspace_bytes(n, ni, npulse == 1 ? ni : 0, npulse == 1 ? nt : 0, nt, num_cu, m, precision);
    }
    return b;
}

}  // namespace

struct oops_server {
    oops_ctx* ctx = nullptr;
    int64_t budget = 0, used = 0;
    int32_t max_inflight = 1, m = 30, precision = 1, inner_threads = 1;
    double tol = 1e-3;
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    std::vector<std::thread> workers;
    std::map<int64_t, std::unique_ptr<Req>> reqs;
    std::deque<int64_t> to_build, built;
    std::vector<int64_t> running;
    int64_t next_id = 1;

    void worker() {
        for (;;) {
            Req* q = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !to_build.empty(); });
                if (stop) return;
                q = reqs[to_build.front()].get();
                to_build.pop_front();
                q->state = kBuilding;
            }
            int rc = build(*q);
            std::lock_guard<std::mutex> lk(mu);
            if (rc) {
                q->err = rc;
                q->state = rc;
            } else {
                q->state = kBuilt;
                q->t_built = now_s() - q->t_submit;
                built.push_back(q->id);
            }
        }
    }

    int build(Req& q) {
        const oops_request& r = q.r;
        if (r.order < 1 || r.order > 6 || r.npulse < 1 || r.nt < 2 || !(r.t_end > 0.0)) return OOPS_ERR_ARG;
        const double a = 3.567;
        int32_t n = 0;
        int rc = oops_bath_diamond(a, r.radius, r.p, r.seed, 0, nullptr, 0, &n);
        if (rc && rc != OOPS_ERR_CAPACITY) return rc;
        q.coords.resize(3 * (size_t)std::max(n, 1));
        rc = oops_bath_diamond(a, r.radius, r.p, r.seed, 0, q.coords.data(), n, &n);
        if (rc) return rc;
        rc = oops_clusters_enumerate(q.coords.data(), n, r.cutoff, r.order, &q.C);
        if (rc) return rc;
        oops_cce_params P;
        std::memset(&P, 0, sizeof(P));
        for (int c = 0; c < 3; ++c) P.B[c] = r.B[c];
        P.state0 = 0;
        P.state1 = 1;
        P.dephasing_axis[2] = 1.0;
        std::vector<double> t(r.nt);
        for (int k = 0; k < r.nt; ++k) t[k] = r.t_end * k / (r.nt - 1) * (r.npulse == 1 ? 0.5 : 1.0);
        if (cce_builder_inputs(q.C, &P, q.groups, q.nitems, q.consts, q.rho)) {
            // the operator tables are built on the device when the request is admitted
            q.device_tables = true;
            q.t = t;
            q.nclusters = q.nitems;
            q.est_bytes = estimate_bytes_built(q.groups, r.npulse, r.nt, ctx->num_cu, m, precision);
            return OOPS_OK;
        }
        rc = cce_build_items(q.coords.data(), q.C, &P, r.npulse, r.nt, t.data(), inner_threads, q.items);
        if (rc) return rc;
        q.nclusters = (int64_t)q.items.items.size();
        q.est_bytes = estimate_bytes(q.items, ctx->num_cu, m, precision);
        return OOPS_OK;
    }

    // Admit built requests in arrival order while a slot and the budget allow.
    int admit(int32_t& launched) {
        launched = 0;
        for (;;) {
            Req* q = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                if (built.empty() || (int32_t)running.size() >= max_inflight) return OOPS_OK;
                q = reqs[built.front()].get();
                if (q->est_bytes > budget) {
                    built.pop_front();
                    q->err = q->state = OOPS_ERR_CAPACITY;
                    continue;
                }
                if (!running.empty() && used + q->est_bytes > budget) return OOPS_OK;
                built.pop_front();
            }
            int rc = q->device_tables
                         ? batch_create_built(ctx, q->coords.data(), (int32_t)(q->coords.size() / 3), q->groups,
                                              q->nitems, q->consts, q->rho, q->r.npulse, true, q->r.nt, q->t.data(), m,
                                              10, tol, precision, &q->batch)
                         : batch_create(ctx, (int32_t)q->items.items.size(), q->items.items.data(), true, m, 10, tol,
                                        precision, &q->batch);
            if (!rc) rc = oops_echo_batch_footprint(q->batch, &q->footprint);
            if (!rc) rc = oops_echo_batch_launch(q->batch);
            q->items.items.clear();
            q->items.hold.reset();  // host copies of the operators are no longer needed
            std::lock_guard<std::mutex> lk(mu);
            if (rc) {
                q->err = q->state = rc;
                if (q->batch) { oops_hahn_batch_destroy(q->batch); q->batch = nullptr; }
                continue;
            }
            q->state = kRunning;
            q->t_launched = now_s() - q->t_submit;
            used += q->footprint;
            running.push_back(q->id);
            ++launched;
        }
    }

    int complete(int32_t& done) {
        done = 0;
        std::vector<int64_t> ids;
        {
            std::lock_guard<std::mutex> lk(mu);
            ids = running;
        }
        for (int64_t id : ids) {
            Req* q = nullptr;
            {
                std::lock_guard<std::mutex> lk(mu);
                q = reqs[id].get();
            }
            int32_t fin = 0;
            int rc = oops_echo_batch_query(q->batch, &fin);
            if (!rc && !fin) continue;
            if (!rc) rc = oops_echo_batch_wait(q->batch);
            const int nt = q->r.nt, K = q->r.order;
            std::vector<double> Lall((size_t)q->nclusters * nt), fin_prod((size_t)nt * K);
            std::vector<int32_t> st((size_t)q->nclusters * kStatsPerItem);
            if (!rc) rc = oops_hahn_batch_results(q->batch, Lall.data(), st.data());
            int64_t failed = 0;
            for (int64_t c = 0; c < q->nclusters; ++c)
                if ((st[(size_t)c * kStatsPerItem + 8] & 0xFF) != 0) ++failed;
            if (!rc && failed)
                rc = set_err(ctx, OOPS_ERR_NUMERIC, "request " + std::to_string(id) + ": " + std::to_string(failed) +
                                                        " of " + std::to_string(q->nclusters) +
                                                        " clusters failed (step limit or step-size collapse)");
            if (!rc) rc = oops_cce_reduce(q->C, nt, Lall.data(), nullptr, fin_prod.data(), nullptr, nullptr, nullptr);
            if (!rc) {
                q->L.resize(nt);
                for (int k = 0; k < nt; ++k) q->L[k] = fin_prod[(size_t)k * K + (K - 1)];
            }
            oops_hahn_batch_destroy(q->batch);
            q->batch = nullptr;
            std::lock_guard<std::mutex> lk(mu);
            used -= q->footprint;
            running.erase(std::find(running.begin(), running.end(), id));
            q->latency = now_s() - q->t_submit;
            q->failed = failed;
            q->state = rc ? rc : kDone;
            q->err = rc;
            ++done;
        }
        return OOPS_OK;
    }
};

extern "C" {

int oops_server_create(oops_ctx* ctx, int64_t hbm_budget, int32_t max_inflight, int32_t m, double tol,
                       int32_t precision, int32_t threads, oops_server** out) {
    if (!ctx || !out || hbm_budget <= 0 || max_inflight < 1 || threads < 1) return OOPS_ERR_ARG;
    std::unique_ptr<oops_server> s(new oops_server());
    s->ctx = ctx;
    s->budget = hbm_budget;
    s->max_inflight = max_inflight;
    s->m = m;
    s->tol = tol;
    s->precision = precision;
    const int nworkers = std::max(1, std::min(4, threads / 4));
    s->inner_threads = std::max(1, threads / nworkers);
    oops_server* raw = s.get();
    for (int w = 0; w < nworkers; ++w) s->workers.emplace_back([raw] { raw->worker(); });
    *out = s.release();
    return OOPS_OK;
}

int oops_server_submit(oops_server* s, const oops_request* r, int64_t* id) {
    if (!s || !r) return OOPS_ERR_ARG;
    std::unique_ptr<Req> q(new Req());
    q->r = *r;
    q->t_submit = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    q->id = s->next_id++;
    if (id) *id = q->id;
    s->to_build.push_back(q->id);
    s->reqs[q->id] = std::move(q);
    s->cv.notify_one();
    return OOPS_OK;
}

int oops_server_poll(oops_server* s, int32_t* completed) {
    if (!s) return OOPS_ERR_ARG;
    if (hipSetDevice(s->ctx->device) != hipSuccess) return set_err(s->ctx, OOPS_ERR_HIP, "hipSetDevice");
    int32_t done = 0, launched = 0;
    int rc = s->complete(done);
    if (!rc) rc = s->admit(launched);
    if (completed) *completed = done;
    return rc;
}

int oops_server_state(oops_server* s, int64_t id, int32_t* state) {
    if (!s || !state) return OOPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto f = s->reqs.find(id);
    if (f == s->reqs.end()) return OOPS_ERR_ARG;
    *state = f->second->state;
    return OOPS_OK;
}

int oops_server_result(oops_server* s, int64_t id, double* L, double* latency_s, int64_t* nclusters,
                       int64_t* footprint) {
    if (!s) return OOPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto f = s->reqs.find(id);
    if (f == s->reqs.end()) return OOPS_ERR_ARG;
    const Req& q = *f->second;
    if (q.state != kDone) return q.state < 0 ? q.state : OOPS_ERR_ARG;
    if (L) std::memcpy(L, q.L.data(), sizeof(double) * q.L.size());
    if (latency_s) *latency_s = q.latency;
    if (nclusters) *nclusters = q.nclusters;
    if (footprint) *footprint = q.footprint;
    return OOPS_OK;
}

int oops_server_times(oops_server* s, int64_t id, double* built_s, double* launched_s) {
    if (!s) return OOPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto f = s->reqs.find(id);
    if (f == s->reqs.end()) return OOPS_ERR_ARG;
    if (built_s) *built_s = f->second->t_built;
    if (launched_s) *launched_s = f->second->t_launched;
    return OOPS_OK;
}

int oops_server_failed(oops_server* s, int64_t id, int64_t* failed) {
    if (!s || !failed) return OOPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto f = s->reqs.find(id);
    if (f == s->reqs.end()) return OOPS_ERR_ARG;
    *failed = f->second->failed;
    return OOPS_OK;
}

int oops_server_pending(oops_server* s, int32_t* pending) {
    if (!s || !pending) return OOPS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t p = 0;
    for (auto& kv : s->reqs)
        if (kv.second->state >= 0 && kv.second->state != kDone) ++p;
    *pending = p;
    return OOPS_OK;
}

int oops_server_destroy(oops_server* s) {
    if (!s) return OOPS_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv.notify_all();
    for (auto& w : s->workers) w.join();
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (int64_t id : s->running) {
        Req* q = s->reqs[id].get();
        if (q->batch) oops_echo_batch_wait(q->batch);
    }
    delete s;
    return OOPS_OK;
}

}  // extern "C"
