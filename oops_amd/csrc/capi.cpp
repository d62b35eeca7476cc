// Host runtime behind include/oops_expv.h: operator compilation, device
// buffers, bucketed launches and timing.  The product path has no CPU
// fallback: every numeric entry point runs on the GPU and returns an error
// status when no HIP device is available.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oops_expv.h"
#include "oops_kernels.h"
#include "host_common.h"
#include "term_builder.h"

using oops::DevOpHeader;
using oops::DevOpLocal;

namespace oops {

int set_err(oops_ctx* ctx, int code, const std::string& msg) {
    static std::mutex mu;   // host worker threads (large items, serving builders) share a context
    if (ctx) {
        std::lock_guard<std::mutex> lock(mu);
        ctx->err = msg;
    }
    return code;
}

int hip_check(oops_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess) return OOPS_OK;
    return set_err(ctx, OOPS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Operator compiler: merge the reference's term list into one dense local
// matrix per spin and per spin pair, then store each as an ELL row table.
// max_locals: capacity of the kernel's local-operator table (LDS).
int compile_operator(const oops_term_table* T, CompiledOp& out, std::string& err, int max_locals,
                     const std::complex<double>* ccoeff, int max_row) {
    if (!T || T->nspin <= 0 || T->nterm < 0) { err = "empty term table"; return OOPS_ERR_ARG; }
    const int ns = T->nspin;
    const int64_t d = T->spin_dim[0];
    for (int s = 0; s < ns; ++s)
        if (T->spin_dim[s] != d) { err = "all spins of one operator must share the local dimension"; return OOPS_ERR_UNSUPPORTED; }
    int ld = 0;
    while ((1LL << ld) < d) ++ld;
    if ((1LL << ld) != d || d < 2) { err = "local dimension must be a power of two"; return OOPS_ERR_UNSUPPORTED; }
    int64_t n = 1;
    for (int s = 0; s < ns; ++s) n *= d;
    if (n > (1LL << 30)) { err = "state too large"; return OOPS_ERR_UNSUPPORTED; }
    typedef std::complex<double> cplx;
    const cplx* mats = reinterpret_cast<const cplx*>(T->mats);
    cplx c0(0.0, 0.0);
    std::vector<std::vector<cplx>> single(ns);
    std::map<std::pair<int, int>, std::vector<cplx>> pair;
    for (int t = 0; t < T->nterm; ++t) {
        const cplx c = ccoeff ? ccoeff[t] : cplx(T->coeff[t], 0.0);
        const int64_t nb = T->nbody[t];
        const int64_t p0 = T->pos_offset[t];
        if (nb == 0) {
            c0 += c;
        } else if (nb == 1) {
            const int s = (int)T->pos[p0];
            if (s < 0 || s >= ns || T->dims[p0] != d) { err = "bad 1-body term"; return OOPS_ERR_ARG; }
            auto& M = single[s];
            if (M.empty()) M.assign(d * d, cplx(0.0));
            const cplx* A = mats + T->mat_offset[p0];
            for (int64_t k = 0; k < d * d; ++k) M[k] += c * A[k];
        } else if (nb == 2) {
            int s = (int)T->pos[p0], u = (int)T->pos[p0 + 1];
            if (s < 0 || s >= ns || u < 0 || u >= ns || s == u || T->dims[p0] != d || T->dims[p0 + 1] != d) {
                err = "bad 2-body term";
                return OOPS_ERR_ARG;
            }
            const cplx* A = mats + T->mat_offset[p0];
            const cplx* B = mats + T->mat_offset[p0 + 1];
            if (s > u) { std::swap(s, u); std::swap(A, B); }
            auto& M = pair[{s, u}];
            const int64_t dd = d * d;
            if (M.empty()) M.assign(dd * dd, cplx(0.0));
            // kron(A, B) column-major: row (ra*d + rb), col (ca*d + cb)
            for (int64_t ca = 0; ca < d; ++ca)
                for (int64_t cb = 0; cb < d; ++cb)
                    for (int64_t ra = 0; ra < d; ++ra)
                        for (int64_t rb = 0; rb < d; ++rb)
                            M[(ra * d + rb) + dd * (ca * d + cb)] += c * A[ra + d * ca] * B[rb + d * cb];
        } else {
            err = "terms with more than two bodies are not supported on the GPU path";
            return OOPS_ERR_UNSUPPORTED;
        }
    }
    out = CompiledOp();
    out.n = n;
    out.nspin = ns;
    out.ld_bits = ld;
    out.c0 = c0;
    auto stride = [&](int s) { return (int64_t)1 << (ld * (ns - 1 - s)); };
    // flip-mask slots (slot 0 = mask 0, the diagonal)
    std::map<int64_t, int> slot64_of, slot32_of;
    out.mask64.push_back(0);
    slot64_of[0] = 0;
    auto slot_of = [&](int64_t mask) -> uint8_t {
        int fields = 0;
        for (int q = 0; q < ns; ++q)
            if ((mask >> (ld * (ns - 1 - q))) & (d - 1)) ++fields;
        auto& tab = fields <= 1 ? slot64_of : slot32_of;
        auto& lst = fields <= 1 ? out.mask64 : out.mask32;
        auto f = tab.find(mask);
        int id;
        if (f != tab.end()) {
            id = f->second;
        } else {
            id = (int)lst.size();
            tab[mask] = id;
            lst.push_back((int32_t)mask);
        }
        if (id > 127) return 0;  // counted below; the slot path is then unavailable
        return (uint8_t)(fields <= 1 ? id : (id | 128));
    };
    auto add_local = [&](int nb, int s, int u, const std::vector<cplx>& M, int64_t rows) {
        int R = 0;
        for (int64_t r = 0; r < rows; ++r) {
            int cnt = 0;
            for (int64_t c = 0; c < rows; ++c) if (M[r + rows * c] != cplx(0.0)) ++cnt;
            R = std::max(R, cnt);
        }
        if (R == 0) return;
        R = R <= 4 ? 4 : (R <= 8 ? 8 : (R <= 16 ? 16 : (R + 3) / 4 * 4));  // kernel row widths
        DevOpLocal lo;
        lo.nb = (int16_t)nb; lo.s = (int16_t)s; lo.t = (int16_t)u; lo.R = (int16_t)R;
        lo.ent = (int32_t)out.val.size();
        lo.rows = (int32_t)rows;
        for (int64_t r = 0; r < rows; ++r) {
            int cnt = 0;
            for (int64_t c = 0; c < rows; ++c) {
                const cplx v = M[r + rows * c];
                if (v == cplx(0.0)) continue;
                int64_t off, mask;
                if (nb == 1) {
                    off = c * stride(s);
                    mask = (r ^ c) * stride(s);
                } else {
                    off = (c / d) * stride(s) + (c % d) * stride(u);
                    mask = ((r / d) ^ (c / d)) * stride(s) + ((r % d) ^ (c % d)) * stride(u);
                }
                out.val.push_back(v);
                out.off.push_back((int32_t)off);
                out.eslot.push_back(slot_of(mask));
                ++cnt;
            }
            for (; cnt < R; ++cnt) { out.val.push_back(cplx(0.0)); out.off.push_back(0); out.eslot.push_back(0); }
        }
        out.locs.push_back(lo);
        out.max_row = std::max(out.max_row, R);
    };
    for (int s = 0; s < ns; ++s)
        if (!single[s].empty()) add_local(1, s, s, single[s], d);
    out.ent2_first = (int32_t)out.val.size();
    for (auto& kv : pair) add_local(2, kv.first.first, kv.first.second, kv.second, d * d);
    out.nslot64 = (int)out.mask64.size();
    out.nslot32 = (int)out.mask32.size();
    if (out.nslot64 > 128 || out.nslot32 > 128) out.nslot64 = -1;
    // Hermitian check of every merged local operator (sufficient for H = H^dagger)
    auto is_herm = [](const std::vector<cplx>& M, int64_t rows) {
        double mx = 0.0;
        for (const auto& z : M) mx = std::max(mx, std::abs(z));
        for (int64_t r = 0; r < rows; ++r)
            for (int64_t c = 0; c <= r; ++c)
                if (std::abs(M[r + rows * c] - std::conj(M[c + rows * r])) > 1e-13 * mx) return false;
        return true;
    };
    out.hermitian = c0.imag() == 0.0;
    for (int s = 0; s < ns && out.hermitian; ++s)
        if (!single[s].empty() && !is_herm(single[s], d)) out.hermitian = false;
    for (auto& kv : pair)
        if (out.hermitian && !is_herm(kv.second, d * d)) out.hermitian = false;
    if ((int)out.locs.size() > max_locals) { err = "too many local operators"; return OOPS_ERR_CAPACITY; }
    if (out.max_row > max_row) {
        err = "local operator rows wider than " + std::to_string(max_row) + " entries";
        return OOPS_ERR_UNSUPPORTED;
    }
    return OOPS_OK;
}

// LAPACK DLARAN: x <- x * 33952834046453 mod 2^48, value x / 2^48.
void dlaran_fill(int32_t* iseed, int64_t count, double* out) {
    const uint64_t mult = 33952834046453ull;
    const uint64_t mask48 = (1ull << 48) - 1;
    uint64_t x = ((uint64_t)iseed[0] << 36) | ((uint64_t)iseed[1] << 24) | ((uint64_t)iseed[2] << 12) | (uint64_t)iseed[3];
    for (int64_t i = 0; i < count; ++i) {
        // (x * mult) mod 2^48 with 64-bit wrap-around arithmetic
        x = (x * mult) & mask48;
        out[i] = std::ldexp((double)x, -48);
    }
    iseed[0] = (int32_t)((x >> 36) & 4095);
    iseed[1] = (int32_t)((x >> 24) & 4095);
    iseed[2] = (int32_t)((x >> 12) & 4095);
    iseed[3] = (int32_t)(x & 4095);
}

template <typename T>
int dev_upload(oops_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
    b.release();
    if (v.empty()) return OOPS_OK;
    int rc = hip_check(ctx, hipMalloc(&b.p, v.size() * sizeof(T)), "hipMalloc");
    if (rc) return rc;
    b.bytes = v.size() * sizeof(T);
    return hip_check(ctx, hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// L(t) results of a bucket, filled with NaN (all bits set): a cluster whose
// sweeps fail leaves its outputs unwritten, and NaN cannot pass for a result.
int alloc_results(oops_ctx* ctx, DevBuf& b, int64_t count) {
    b.release();
    const size_t bytes = sizeof(double) * (size_t)count;
    int rc = hip_check(ctx, hipMalloc(&b.p, bytes), "hipMalloc L");
    if (rc) return rc;
    b.bytes = bytes;
    return hip_check(ctx, hipMemset(b.p, 0xFF, bytes), "hipMemset L");
}

// Step limit of every Krylov integration (OOPS_MAX_STEPS overrides it, so
// that a caller can force the failure path from the host).
int32_t max_steps_limit() {
    if (const char* e = std::getenv("OOPS_MAX_STEPS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v > 0 && v < 50000000) return (int32_t)v;
    }
    return 50000000;
}

}  // namespace oops

using namespace oops;

struct oops_hahn_batch {
    oops_ctx* ctx = nullptr;
    int32_t nitems = 0, m = 30, klim = 10, precision = 0;
    bool lanczos64 = false;   // precision 4: FP64 slot Lanczos for the buckets that fit it, else the FP64 paths
    double tol = 1e-12;
    struct Bucket {
        int n = 0;
        std::vector<int32_t> items;  // global item indices
        DevBuf hdr, loc, val, off, opA, opB, opC, opD, voff, u0, y0, scale, rnd, L, stats, queue;
        DevBuf npulse, nt, tau_off, L_off, tau;  // per-item sequences (echo batches)
        int64_t L_len = 0;                       // doubles of L (echo batches)
        int grid = 0;
        bool hermitian = true;  // every generator Hermitian: mixed precision may use Lanczos
        bool cpmg = false;      // some item has N >= 2 pulses or two chains
        bool two_chain = false; // some item runs two forward chains (single-sample form)
        bool uniform = true;    // every item's times are k * dt from 0 (dense propagator powers apply)
        DevBuf work;            // dense64 workspace
        int dense_grid = 0;
        DevBuf eslot, smask, slotws;  // slot form of the mixed Lanczos kernel
        bool medium = false;          // n = 1024 / 4096 on the batched medium kernel
        bool medslot = true;          // medium: every operator fits kMedSlot64 / kMedSlot32 (flip-mask form)
        int ms64 = 1, ms32 = 0;       // medium flip-mask slots per generator (bucket maxima)
        DevBuf medmask, medtab;       // medium flip-mask masks and per-workgroup coefficient tables
        DevBuf medws;                 // its FP32 Lanczos bases (one per workgroup)
        int medium_grid = 0;
        bool slots = true;            // every operator fits kSlot64 / kSlot32
        int slot_grid = 0;
        DevBuf sweepws;               // slot kernel: sweep-A states at every output time, per workgroup
        int sweep_nt = 0;             // output times per workgroup that sweepws holds
        // states above 256 elements: every item on the single-state path, in turn
        bool large = false;
        bool device_tables = false;   // hdr, loc, val, off, eslot, smask, medmask written by the term builder
        // factored propagator powers (unpolarised k-spin clusters, uniform grid):
        // the kernel builds H0 / H1 from the coordinates, no operator tables
        bool factored = false;
        DevBuf fcoords, fspins;
        FactoredArgs fconst;          // physics constants, k and N (pointers filled at launch)
        int kernel = -1;  // kind that ran last: see oops_hahn_batch_kernels
        std::vector<CompiledOp> hA, hB;
        std::vector<std::vector<double>> hu0, hy0;
        std::vector<double> hscale, hL, htau;
        std::vector<int32_t> hstats, hnt;
        std::vector<int64_t> htau_off;
        double ms = 0.0;
    };
    std::vector<Bucket> buckets;
    DevBuf tau;
    int32_t nt = 0;
    // echo batches: per-item times fixed at creation, own stream, asynchronous launch
    bool per_item = false;
    std::vector<int64_t> item_L_off;  // offset of item i's L in the concatenated result
    std::vector<int32_t> item_nt;
    int64_t footprint = 0;            // device bytes held by the batch
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;
    bool launched = false;
    hipGraph_t graph = nullptr;        // captured launch sequence (oops_echo_batch_capture)
    hipGraphExec_t graph_exec = nullptr;
    std::vector<hipStream_t> bstreams; // oops_hahn_batch_run: one stream per bucket (buckets overlap)
    double total_ms = 0.0;             // device interval of the last run, first launch to last completion
    ~oops_hahn_batch() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        for (auto& q : bstreams) if (q) (void)hipStreamDestroy(q);
    }
};



extern "C" {

int oops_ctx_create(int32_t device, oops_ctx** out) {
    if (!out) return OOPS_ERR_ARG;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return OOPS_ERR_HIP;
    if (device < 0 || device >= count) return OOPS_ERR_ARG;
    std::unique_ptr<oops_ctx> c(new oops_ctx());
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return OOPS_ERR_HIP;
    if (hipStreamCreate(&c->stream) != hipSuccess) return OOPS_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return OOPS_ERR_HIP;
    c->num_cu = prop.multiProcessorCount;
    c->lds_bytes = (int)prop.sharedMemPerBlock;
    c->name = prop.gcnArchName;
    if (oops::noalloc_array_address(device, &c->noalloc) != OOPS_OK) return OOPS_ERR_HIP;
    *out = c.release();
    return OOPS_OK;
}

int oops_ctx_destroy(oops_ctx* ctx) {
    if (!ctx) return OOPS_OK;
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return OOPS_OK;
}

const char* oops_last_error(const oops_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int oops_ctx_set_launch(oops_ctx* ctx, int32_t kind, int32_t wg_per_cu) {
    if (!ctx) return OOPS_ERR_ARG;
    if (kind < 0 || kind > 9 || kind == 5 || wg_per_cu < 0 || wg_per_cu > 16)
        return set_err(ctx, OOPS_ERR_ARG, "kind must be a batch kernel kind (0-4, 6-9), wg_per_cu in [0, 16]");
    ctx->wg_per_cu[kind] = wg_per_cu;
    return OOPS_OK;
}

int oops_ctx_info(const oops_ctx* ctx, int32_t* num_cu, int32_t* lds_bytes, char* name, int32_t name_len) {
    if (!ctx) return OOPS_ERR_ARG;
    if (num_cu) *num_cu = ctx->num_cu;
    if (lds_bytes) *lds_bytes = ctx->lds_bytes;
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s", ctx->name.c_str());
    }
    return OOPS_OK;
}

int oops_dlaran_fill(int32_t* iseed, int64_t count, double* out) {
    if (!iseed || count < 0 || (count > 0 && !out)) return OOPS_ERR_ARG;
    dlaran_fill(iseed, count, out);
    return OOPS_OK;
}

}  // extern "C"

static int sweep_workspace(oops_ctx* ctx, oops_hahn_batch::Bucket& B, int ntmax);

// Workgroups of a persistent batch kernel: units of work (items or sweeps)
// pull from the bucket's device queue, so any grid gives the same results;
// the default per-CU count is what the kernel's LDS and registers let one CU
// hold, oops_ctx_set_launch overrides it per kernel kind.
static int persistent_grid(const oops_ctx* ctx, int kind, int def_per_cu, int64_t units) {
    const int per_cu = ctx->wg_per_cu[kind] > 0 ? ctx->wg_per_cu[kind] : def_per_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(units, (int64_t)ctx->num_cu * per_cu));
}

// Host side of one bucket before upload: packed operator tables (left empty
// when the device term builder wrote them to the bucket directly) and the
// per-item sequences.
struct BucketHost {
    std::vector<DevOpHeader> hdr;
    std::vector<DevOpLocal> loc;
    std::vector<double> val;
    std::vector<int32_t> off;
    std::vector<int32_t> opA, opB;
    std::vector<int64_t> voff;
    std::vector<double> u0, y0, scale;
    std::vector<int32_t> opC, opD, npulse, nt;
    std::vector<int64_t> tau_off, L_off;
    std::vector<double> tau;
    std::vector<uint8_t> eslot;
    std::vector<int32_t> smask;
    std::vector<int32_t> mmask;   // kMedMaskStride masks per operator (medium flip-mask form)
};

static int finalize_batch(oops_ctx* ctx, std::unique_ptr<oops_hahn_batch>& b, std::vector<BucketHost>& hosts,
                          bool per_item, int32_t m, int32_t precision, bool ell_only, bool lanczos64,
                          oops_hahn_batch** out);

int oops::batch_create(oops_ctx* ctx, int32_t nitems, const EchoIn* items, bool per_item, int32_t m, int32_t klim,
                       double tol, int32_t precision, oops_hahn_batch** out) {
    if (!ctx || !out || nitems < 0 || (nitems > 0 && !items)) return set_err(ctx, OOPS_ERR_ARG, "bad arguments");
    *out = nullptr;
    if (precision < 0 || precision > 4)
        return set_err(ctx, OOPS_ERR_ARG,
                       "precision must be 0 (FP64), 1 (mixed), 2 (mixed, Arnoldi), 3 (mixed, ELL product) or 4 "
                       "(FP64 Lanczos in the flip-mask form)");
    const bool ell_only = precision == 3;   // mixed Lanczos without the slot form
    if (ell_only) precision = 1;
    const bool lanczos64 = precision == 4;  // FP64 with the mixed path's Lanczos algorithm and schedule
    if (lanczos64) precision = 0;
    if (m != 20 && m != 30) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "Krylov dimension must be 20 or 30");
    if (klim < 1 || klim > 16) return set_err(ctx, OOPS_ERR_ARG, "klim must be in [1, 16]");
    std::unique_ptr<oops_hahn_batch> b(new oops_hahn_batch());
    b->ctx = ctx;
    b->nitems = nitems;
    b->m = m;
    b->klim = klim;
    b->tol = tol;
    b->precision = precision;
    b->lanczos64 = lanczos64;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    // compile all operators, bucket by n
    std::map<int64_t, int> bucket_of;
    typedef BucketHost Host;
    b->per_item = per_item;
    std::vector<Host> hosts;
    const int emax = krylov_entry_capacity(m);
    int64_t L_total = 0;
    for (int32_t it = 0; it < nitems; ++it) {
        const EchoIn& I = items[it];
        const bool two = I.chains == 2;   // single-sample form: two forward chains
        const bool cpmg = I.npulse > 1 || two;
        if (I.npulse < 1 || I.npulse > 64) return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": npulse must be in [1, 64]");
        if (I.chains != 1 && I.chains != 2) return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": chains must be 1 or 2");
        if (per_item) {
            if (I.nt < 1 || !I.t) return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": no output times");
            for (int k = 0; k < I.nt; ++k)
                if (!(I.t[k] >= (k ? I.t[k - 1] : 0.0))) return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": times must be ascending and >= 0");
        }
        if (cpmg && (!I.opC || !I.opD)) return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": CPMG-N needs lvB and -lvA^dagger");
        CompiledOp A, B, Cc, Dd;
        std::string err;
        int rc = compile_operator(I.opA, A, err);
        if (!rc) rc = compile_operator(I.opB, B, err);
        if (!rc && cpmg) rc = compile_operator(I.opC, Cc, err);
        if (!rc && cpmg) rc = compile_operator(I.opD, Dd, err);
        if (rc) return set_err(ctx, rc, "item " + std::to_string(it) + ": " + err);
        if (A.n != B.n || (cpmg && (Cc.n != A.n || Dd.n != A.n)))
            return set_err(ctx, OOPS_ERR_ARG, "item " + std::to_string(it) + ": sweep dimensions differ");
        bool uniform_item = false;
        if (per_item && I.nt >= 2) {
            const double dt = I.t[1] - I.t[0];
            uniform_item = I.t[0] == 0.0 && dt > 0.0;
            for (int k = 2; k < I.nt && uniform_item; ++k)
                uniform_item = std::fabs(I.t[k] - k * dt) <= 1e-12 * I.t[I.nt - 1];
        }
        if (cpmg && !(A.n <= m || (A.n == 64 && uniform_item) ||
                      ((precision == 1 || (lanczos64 && A.n <= 256)) && A.hermitian && B.hermitian &&
                       Cc.hermitian && Dd.hermitian)))
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "item " + std::to_string(it) +
                           ": CPMG-N runs on the dense paths (n <= m, or n = 64 on a uniform grid) or the mixed "
                           "Lanczos path (Hermitian generators)");
        const int64_t n = A.n;
        const bool dense = n <= m;
        if (dense && n > kDenseNMAX)
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "item " + std::to_string(it) + ": n <= m needs n <= 16");
        // n = 1024 / 4096 with mixed precision and Hermitian generators: the
        // batched medium kernel; other states above 256 elements: the
        // single-state path, item by item
        auto fits_medium = [&](const CompiledOp& o) {
            return o.hermitian && (int)o.val.size() <= kMediumMaxEnt && (int)o.locs.size() <= kMaxLocals;
        };
        const bool medium = !dense && !two && (n == 1024 || n == 4096) && precision == 1 && fits_medium(A) &&
                            fits_medium(B) && (!cpmg || (fits_medium(Cc) && fits_medium(Dd)));
        const bool large = !dense && n > 256 && !medium;
        if (large && cpmg)
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "item " + std::to_string(it) + ": CPMG-N needs n <= 256");
        if (large && (A.locs.size() > 1024 || B.locs.size() > 1024))
            return set_err(ctx, OOPS_ERR_CAPACITY, "item " + std::to_string(it) + ": too many local operators");
        const int cap = dense ? kDenseMaxEnt : (medium ? kMediumMaxEnt : emax);
        if (!large && ((int)A.val.size() > cap || (int)B.val.size() > cap ||
                       (cpmg && ((int)Cc.val.size() > cap || (int)Dd.val.size() > cap))))
            return set_err(ctx, OOPS_ERR_CAPACITY, "item " + std::to_string(it) + ": operator table exceeds LDS capacity");
        const int64_t key = 2 * n + (medium ? 1 : 0);
        auto f = bucket_of.find(key);
        int bk;
        if (f == bucket_of.end()) {
            bk = (int)b->buckets.size();
            bucket_of[key] = bk;
            b->buckets.emplace_back();
            b->buckets.back().n = (int)n;
            b->buckets.back().medium = medium;
            hosts.emplace_back();
        } else {
            bk = f->second;
        }
        Host& H = hosts[bk];
        b->buckets[bk].items.push_back(it);
        if (large) {
            auto& BB = b->buckets[bk];
            BB.large = true;
            BB.hA.push_back(std::move(A));
            BB.hB.push_back(std::move(B));
            BB.hu0.emplace_back(I.u0, I.u0 + 2 * n);
            BB.hy0.emplace_back(I.y0, I.y0 + 2 * n);
            BB.hscale.push_back(I.scale);
            if (per_item) {
                BB.hnt.push_back(I.nt);
                BB.htau_off.push_back((int64_t)BB.htau.size());
                BB.htau.insert(BB.htau.end(), I.t, I.t + I.nt);
                b->item_L_off.push_back(L_total);
                b->item_nt.push_back(I.nt);
                L_total += I.nt;
            }
            continue;
        }
        b->buckets[bk].hermitian = b->buckets[bk].hermitian && A.hermitian && B.hermitian &&
                                   (!cpmg || (Cc.hermitian && Dd.hermitian));
        b->buckets[bk].cpmg = b->buckets[bk].cpmg || cpmg;
        b->buckets[bk].two_chain = b->buckets[bk].two_chain || two;
        for (int s = 0; s < 4; ++s) {
            if (s >= 2 && !cpmg) {
                (s == 2 ? H.opC : H.opD).push_back(-1);
                continue;
            }
            const CompiledOp& C = s == 0 ? A : (s == 1 ? B : (s == 2 ? Cc : Dd));
            DevOpHeader h;
            std::memset(&h, 0, sizeof(h));
            h.n = (int32_t)C.n;
            h.nspin = C.nspin;
            h.ld_bits = C.ld_bits;
            h.nloc = (int32_t)C.locs.size();
            h.nent = (int32_t)C.val.size();
            h.loc_off = (int32_t)H.loc.size();
            h.ent_off = (int32_t)(H.val.size() / 2);
            h.max_row = C.max_row;
            h.c0re = C.c0.real();
            h.c0im = C.c0.imag();
            h.ent2_first = C.ent2_first;
            (s == 0 ? H.opA : (s == 1 ? H.opB : (s == 2 ? H.opC : H.opD))).push_back((int32_t)H.hdr.size());
            H.hdr.push_back(h);
            H.loc.insert(H.loc.end(), C.locs.begin(), C.locs.end());
            for (auto& v : C.val) { H.val.push_back(v.real()); H.val.push_back(v.imag()); }
            H.off.insert(H.off.end(), C.off.begin(), C.off.end());
            H.eslot.insert(H.eslot.end(), C.eslot.begin(), C.eslot.end());
            if (C.nslot64 < 0 || C.nslot64 > kSlot64 || C.nslot32 > kSlot32) b->buckets[bk].slots = false;
            if (C.nslot64 < 1 || C.nslot64 > kMedSlot64 || C.nslot32 > kMedSlot32) {
                b->buckets[bk].medslot = false;
            } else {
                b->buckets[bk].ms64 = std::max(b->buckets[bk].ms64, C.nslot64);
                b->buckets[bk].ms32 = std::max(b->buckets[bk].ms32, C.nslot32);
            }
            for (int q = 0; q < kMedMaskStride; ++q) {
                int32_t mk = 0;
                if (q < kMedSlot64 && C.nslot64 > 0 && q < C.nslot64) mk = C.mask64[q];
                if (q >= kMedSlot64 && q - kMedSlot64 < C.nslot32) mk = C.mask32[q - kMedSlot64];
                H.mmask.push_back(mk);
            }
            for (int q = 0; q < kSlotMaskStride; ++q) {
                int32_t mk = 0;
                if (q < kSlot64 && q < (int)C.mask64.size()) mk = C.mask64[q];
                if (q >= kSlot64 && q - kSlot64 < kSlot32 && q - kSlot64 < (int)C.mask32.size()) mk = C.mask32[q - kSlot64];
                H.smask.push_back(mk);
            }
        }
        H.voff.push_back((int64_t)(H.u0.size() / 2));
        for (int64_t i = 0; i < 2 * n; ++i) {
            H.u0.push_back(I.u0[i]);
            H.y0.push_back(I.y0[i]);
        }
        H.scale.push_back(I.scale);
        H.npulse.push_back(I.npulse | (two ? kTwoChainFlag : 0));
        if (per_item) {
            if (I.nt >= 2) b->buckets[bk].uniform = b->buckets[bk].uniform && uniform_item;
            H.nt.push_back(I.nt);
            H.tau_off.push_back((int64_t)H.tau.size());
            H.tau.insert(H.tau.end(), I.t, I.t + I.nt);
            H.L_off.push_back(b->buckets[bk].L_len);
            b->buckets[bk].L_len += I.nt;
            b->item_L_off.push_back(L_total);
            b->item_nt.push_back(I.nt);
            L_total += I.nt;
        }
    }
    return finalize_batch(ctx, b, hosts, per_item, m, precision, ell_only, lanczos64, out);
}

// Bucket checks, uploads and workspaces of a batch whose buckets and host
// tables are assembled (the host compiler's path or the device term builder's).
static int finalize_batch(oops_ctx* ctx, std::unique_ptr<oops_hahn_batch>& b, std::vector<BucketHost>& hosts,
                          bool per_item, int32_t m, int32_t precision, bool ell_only, bool lanczos64,
                          oops_hahn_batch** out) {
    typedef BucketHost Host;
    for (auto& B : b->buckets)
        if (B.cpmg && !B.factored &&
            !(B.n <= m || (B.n == 64 && B.uniform) || ((precision == 1 || lanczos64) && B.hermitian)))
            return set_err(ctx, OOPS_ERR_UNSUPPORTED,
                           "CPMG-N items of dimension " + std::to_string(B.n) +
                           " need one uniform time grid from 0 for the whole bucket, or mixed precision");
    for (size_t bk = 0; bk < b->buckets.size(); ++bk) {
        auto& B = b->buckets[bk];
        Host& H = hosts[bk];
        if (B.large) continue;
        if (B.factored) {
            int rc = dev_upload(ctx, B.queue, std::vector<int32_t>(4, 0));
            rc = rc ? rc : dev_upload(ctx, B.stats, std::vector<int32_t>(B.items.size() * kStatsPerItem, 0));
            rc = rc ? rc : dev_upload(ctx, B.nt, H.nt);
            rc = rc ? rc : dev_upload(ctx, B.tau_off, H.tau_off);
            rc = rc ? rc : dev_upload(ctx, B.L_off, H.L_off);
            rc = rc ? rc : dev_upload(ctx, B.tau, H.tau);
            if (!rc) rc = alloc_results(ctx, B.L, std::max<int64_t>(B.L_len, 1));
            if (rc) return rc;
            B.grid = persistent_grid(ctx, 9, 2, (int64_t)B.items.size());
            continue;
        }
        std::vector<double> rnd(2 * B.n + 16);
        int32_t seed[4] = {0, 0, 0, 5};
        dlaran_fill(seed, (int64_t)rnd.size(), rnd.data());
        int rc = 0;
        if (!B.device_tables) {
            rc = rc ? rc : dev_upload(ctx, B.hdr, H.hdr);
            rc = rc ? rc : dev_upload(ctx, B.loc, H.loc);
            rc = rc ? rc : dev_upload(ctx, B.val, H.val);
            rc = rc ? rc : dev_upload(ctx, B.off, H.off);
        }
        rc = rc ? rc : dev_upload(ctx, B.opA, H.opA);
        rc = rc ? rc : dev_upload(ctx, B.opB, H.opB);
        rc = rc ? rc : dev_upload(ctx, B.voff, H.voff);
        rc = rc ? rc : dev_upload(ctx, B.u0, H.u0);
        rc = rc ? rc : dev_upload(ctx, B.y0, H.y0);
        rc = rc ? rc : dev_upload(ctx, B.scale, H.scale);
        rc = rc ? rc : dev_upload(ctx, B.rnd, rnd);
        rc = rc ? rc : dev_upload(ctx, B.queue, std::vector<int32_t>(4, 0));
        rc = rc ? rc : dev_upload(ctx, B.stats, std::vector<int32_t>(B.items.size() * kStatsPerItem, 0));
        if (B.cpmg) {
            rc = rc ? rc : dev_upload(ctx, B.opC, H.opC);
            rc = rc ? rc : dev_upload(ctx, B.opD, H.opD);
        }
        rc = rc ? rc : dev_upload(ctx, B.npulse, H.npulse);
        if (per_item) {
            rc = rc ? rc : dev_upload(ctx, B.nt, H.nt);
            rc = rc ? rc : dev_upload(ctx, B.tau_off, H.tau_off);
            rc = rc ? rc : dev_upload(ctx, B.L_off, H.L_off);
            rc = rc ? rc : dev_upload(ctx, B.tau, H.tau);
            if (!rc) rc = alloc_results(ctx, B.L, std::max<int64_t>(B.L_len, 1));
        }
        if (rc) return rc;
        const bool dense = B.n <= m;
        // the kind bucket_launch reports for B.grid: dense Pade, FP64 or
        // mixed Arnoldi, mixed ELL Lanczos
        const int gkind = dense ? 0 : (precision == 1 && B.hermitian ? 3 : (precision == 0 ? 1 : 2));
        B.grid = persistent_grid(ctx, gkind, dense ? 4 : 2, (int64_t)B.items.size());
        if (B.medium) {
            // one workgroup per CU; Hahn items run as two units (one per sweep)
            B.medium_grid = persistent_grid(ctx, 7, 1, 2 * (int64_t)B.items.size());
            const size_t bytes = sizeof(float) * 2 * (size_t)medium_workspace_float2(B.medium_grid, m, B.n);
            rc = hip_check(ctx, hipMalloc(&B.medws.p, bytes), "hipMalloc medium workspace");
            if (rc) return rc;
            B.medws.bytes = bytes;
            // The flip-mask form streams a work unit's per-element coefficient
            // tables (1.8 MB for a 6-spin Liouvillian) from HBM every product;
            // about 110 units running at once saturate it (CCE-6, n = 4096:
            // 70 units 2.1 s, 232 units 4.5 s, 1736 units 37 s).  The merged
            // ELL form reads its tables from LDS (1.5x slower per unit when
            // alone, 3.3 s, but 26.6 s for 1736 units), so buckets with more
            // units than 2/3 of the CUs take it.
            const int64_t units = 2 * (int64_t)B.items.size();
            B.medslot = B.medslot && !ell_only && 3 * units <= 2 * (int64_t)ctx->num_cu;
            if (B.medslot) {
                // flip-mask tables: per workgroup four generators of ms64 FP64 + ms32 FP32 slots
                const int64_t gstride = (int64_t)B.ms64 * B.n + ((int64_t)B.ms32 * B.n + 1) / 2;
                const size_t tb = sizeof(double) * 2 * 4 * (size_t)gstride * (size_t)B.medium_grid;
                if (!B.device_tables) {
                    rc = dev_upload(ctx, B.eslot, H.eslot);
                    rc = rc ? rc : dev_upload(ctx, B.medmask, H.mmask);
                }
                rc = rc ? rc : hip_check(ctx, hipMalloc(&B.medtab.p, tb), "hipMalloc medium tables");
                if (rc) return rc;
                B.medtab.bytes = tb;
            }
        }
        B.slots = B.slots && !B.medium && !ell_only && (precision == 1 || lanczos64) && B.hermitian && !dense &&
                  B.n <= 256 && (m == 20 || m == 30);
        if (B.slots) {
            // mixed: two workgroups per CU, FP64 variant: one; 2 (Hahn) or 4
            // (CPMG-N) coefficient tables each
            B.slot_grid = persistent_grid(ctx, lanczos64 ? 8 : 6, lanczos64 ? 1 : 2, (int64_t)B.items.size());
            if (!B.device_tables) {
                rc = rc ? rc : dev_upload(ctx, B.eslot, H.eslot);
                rc = rc ? rc : dev_upload(ctx, B.smask, H.smask);
            }
            const size_t bytes = sizeof(double) * 2 * (size_t)slot_table_double2(lanczos64) * 4 * (size_t)B.slot_grid;
            rc = rc ? rc : hip_check(ctx, hipMalloc(&B.slotws.p, bytes), "hipMalloc slot tables");
            if (!rc) B.slotws.bytes = bytes;
            if (rc) return rc;
        }
        if (B.two_chain && B.n > m && !B.slots)
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "two-chain (single-sample) items of dimension " +
                                                          std::to_string(B.n) + " need the dense path (n <= m) or "
                                                          "the flip-mask Lanczos kernel (precision 1 or 4, Hermitian)");
        if (lanczos64 && B.cpmg && !B.slots && !(B.n <= m || (B.n == 64 && B.uniform)))
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "FP64 Lanczos CPMG-N needs the flip-mask form (n = " +
                                                          std::to_string(B.n) + ")");
        if (per_item && (B.slots || B.medium)) {
            int ntmax = 0;
            for (size_t i = 0; i < H.nt.size(); ++i)
                if (H.npulse[i] == 1) ntmax = std::max(ntmax, (int)H.nt[i]);
            rc = sweep_workspace(ctx, B, ntmax);
            if (rc) return rc;
        }
        if (B.n == 64) {
            // one workgroup per CU (the kernel holds 152 KB of LDS)
            B.dense_grid = persistent_grid(ctx, 4, 1, (int64_t)B.items.size());
            const size_t bytes = sizeof(double) * (size_t)dense64_workspace_doubles(B.dense_grid);
            rc = hip_check(ctx, hipMalloc(&B.work.p, bytes), "hipMalloc dense workspace");
            if (rc) return rc;
            B.work.bytes = bytes;
        }
    }
    for (auto& B : b->buckets)
        for (DevBuf* d : {&B.hdr, &B.loc, &B.val, &B.off, &B.opA, &B.opB, &B.opC, &B.opD, &B.voff, &B.u0, &B.y0,
                          &B.scale, &B.rnd, &B.L, &B.stats, &B.queue, &B.npulse, &B.nt, &B.tau_off, &B.L_off, &B.tau,
                          &B.work, &B.eslot, &B.smask, &B.slotws, &B.sweepws, &B.medws, &B.medmask, &B.medtab,
                          &B.fcoords, &B.fspins})
            b->footprint += (int64_t)d->bytes;
    if (per_item) {
        if (hipStreamCreate(&b->stream) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipStreamCreate");
        b->ev.resize(2 * b->buckets.size());
        for (auto& e : b->ev)
            if (hipEventCreate(&e) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipEventCreate");
    }
    *out = b.release();
    return OOPS_OK;
}

// The factored propagator powers take the 4-spin clusters of an unpolarised
// CPMG-N batch (N >= 2) without dephasing on a uniform grid: the dense
// D x D propagators are exact and far cheaper than restarting the Krylov
// sweeps at every output time.  OOPS_FACTORED=0 keeps them on the Lanczos
// path; OOPS_FACTORED=all takes the Hahn echo (N = 1) too.
bool oops::factored_applies(int k, int npulse, double deph_rate) {
    const char* e = std::getenv("OOPS_FACTORED");
    const std::string mode = e ? e : "";
    if (mode == "0") return false;
    if (k != 4 || deph_rate != 0.0) return false;
    return npulse >= 2 || mode == "all";
}

int oops::batch_create_built(oops_ctx* ctx, const double* coords, int32_t nspin, const std::vector<BuiltGroup>& groups,
                             int64_t nitems, const TermBuilderArgs& consts,
                             const std::vector<std::vector<double>>& rho, int32_t npulse, bool per_item, int32_t nt,
                             const double* t, int32_t m, int32_t klim, double tol, int32_t precision,
                             oops_hahn_batch** out) {
    if (!ctx || !out || !coords || nitems < 0) return set_err(ctx, OOPS_ERR_ARG, "bad arguments");
    *out = nullptr;
    if (precision < 0 || precision > 4) return set_err(ctx, OOPS_ERR_ARG, "precision must be 0..4");
    const bool ell_only = precision == 3;
    if (ell_only) precision = 1;
    const bool lanczos64 = precision == 4;
    if (lanczos64) precision = 0;
    if (m != 20 && m != 30) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "Krylov dimension must be 20 or 30");
    if (klim < 1 || klim > 16) return set_err(ctx, OOPS_ERR_ARG, "klim must be in [1, 16]");
    if (npulse < 1 || npulse > 64) return set_err(ctx, OOPS_ERR_ARG, "npulse must be in [1, 64]");
    if (per_item) {
        if (nt < 1 || !t) return set_err(ctx, OOPS_ERR_ARG, "no output times");
        for (int q = 0; q < nt; ++q)
            if (!(t[q] >= (q ? t[q - 1] : 0.0))) return set_err(ctx, OOPS_ERR_ARG, "times must be ascending and >= 0");
    }
    const bool cpmg = npulse > 1;
    const int nops = cpmg ? 4 : 2;
    std::unique_ptr<oops_hahn_batch> b(new oops_hahn_batch());
    b->ctx = ctx;
    b->nitems = (int32_t)nitems;
    b->m = m;
    b->klim = klim;
    b->tol = tol;
    b->precision = precision;
    b->lanczos64 = lanczos64;
    b->per_item = per_item;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    bool uniform = false;
    if (per_item && nt >= 2) {
        const double dt = t[1] - t[0];
        uniform = t[0] == 0.0 && dt > 0.0;
        for (int q = 2; q < nt && uniform; ++q) uniform = std::fabs(t[q] - q * dt) <= 1e-12 * t[nt - 1];
    }
    DevBuf dcoords;
    int rc = dev_upload(ctx, dcoords, std::vector<double>(coords, coords + 3 * (size_t)nspin));
    if (rc) return rc;
    std::vector<BucketHost> hosts;
    if (per_item) {
        b->item_L_off.assign((size_t)nitems, 0);
        b->item_nt.assign((size_t)nitems, nt);
    }
    std::vector<int64_t> L_len_of_item_order;
    for (const BuiltGroup& g : groups) {
        const int k = g.k;
        const int64_t ni = (int64_t)g.items.size();
        if (ni == 0) continue;
        if (k < 1 || k > 4 || (int64_t)g.spins.size() != ni * k)
            return set_err(ctx, OOPS_ERR_ARG, "device term builder: clusters of 1-4 spins");
        const int n = 1 << (2 * k);
        if (per_item && uniform && factored_applies(k, npulse, consts.deph_rate)) {
            // factored propagator powers: H0 / H1 built in the kernel, no tables
            b->buckets.emplace_back();
            auto& B = b->buckets.back();
            hosts.emplace_back();
            BucketHost& H = hosts.back();
            B.n = n;
            B.factored = true;
            B.cpmg = cpmg;
            B.uniform = true;
            B.slots = false;
            B.medslot = false;
            B.items.assign(g.items.begin(), g.items.end());
            rc = dev_upload(ctx, B.fspins, g.spins);
            rc = rc ? rc : dev_upload(ctx, B.fcoords, std::vector<double>(coords, coords + 3 * (size_t)nspin));
            if (rc) return rc;
            std::memset(&B.fconst, 0, sizeof(B.fconst));
            for (int c = 0; c < 3; ++c) {
                B.fconst.zee[c] = consts.zee[c];
                B.fconst.center[c] = consts.center[c];
                B.fconst.sv0[c] = consts.sv0[c];
                B.fconst.sv1[c] = consts.sv1[c];
            }
            B.fconst.gamma_c = consts.gamma_c;
            B.fconst.gamma_e = consts.gamma_e;
            B.fconst.k = k;
            B.fconst.npulse = npulse;
            for (int64_t i = 0; i < ni; ++i) {
                H.npulse.push_back(npulse);
                H.nt.push_back(nt);
                H.tau_off.push_back(0);
                H.L_off.push_back(B.L_len);
                B.L_len += nt;
            }
            H.tau.assign(t, t + nt);
            continue;
        }
        const int P = k * (k - 1) / 2;
        const int loc_stride = k + P;
        const int ent_stride = 16 * k + 128 * P;
        b->buckets.emplace_back();
        auto& B = b->buckets.back();
        hosts.emplace_back();
        BucketHost& H = hosts.back();
        B.n = n;
        B.device_tables = true;
        B.cpmg = cpmg;
        B.uniform = uniform;
        B.items.assign(g.items.begin(), g.items.end()) ;
        const int64_t nop = ni * nops;
        DevBuf dcl, summary;
        rc = dev_upload(ctx, dcl, g.spins);
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.hdr.p, sizeof(DevOpHeader) * nop), "hipMalloc hdr");
        if (!rc) B.hdr.bytes = sizeof(DevOpHeader) * nop;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.loc.p, sizeof(DevOpLocal) * nop * loc_stride), "hipMalloc loc");
        if (!rc) B.loc.bytes = sizeof(DevOpLocal) * nop * loc_stride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.val.p, sizeof(double2) * nop * ent_stride), "hipMalloc val");
        if (!rc) B.val.bytes = sizeof(double2) * nop * ent_stride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.off.p, sizeof(int32_t) * nop * ent_stride), "hipMalloc off");
        if (!rc) B.off.bytes = sizeof(int32_t) * nop * ent_stride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.eslot.p, nop * ent_stride), "hipMalloc eslot");
        if (!rc) B.eslot.bytes = nop * ent_stride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.smask.p, sizeof(int32_t) * nop * kSlotMaskStride), "hipMalloc smask");
        if (!rc) B.smask.bytes = sizeof(int32_t) * nop * kSlotMaskStride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&B.medmask.p, sizeof(int32_t) * nop * kMedMaskStride), "hipMalloc mmask");
        if (!rc) B.medmask.bytes = sizeof(int32_t) * nop * kMedMaskStride;
        rc = rc ? rc : hip_check(ctx, hipMalloc(&summary.p, sizeof(int32_t) * nop * kTBSummary), "hipMalloc summary");
        if (rc) return rc;
        TermBuilderArgs a = consts;
        a.coords = (const double*)dcoords.p;
        a.clusters = (const int32_t*)dcl.p;
        a.nitems = (int32_t)ni;
        a.k = k;
        a.nops = nops;
        a.hdr = (DevOpHeader*)B.hdr.p;
        a.loc = (DevOpLocal*)B.loc.p;
        a.val = (double2*)B.val.p;
        a.off = (int32_t*)B.off.p;
        a.eslot = (uint8_t*)B.eslot.p;
        a.smask = (int32_t*)B.smask.p;
        a.mmask = (int32_t*)B.medmask.p;
        a.summary = (int32_t*)summary.p;
        a.loc_stride = loc_stride;
        a.ent_stride = ent_stride;
        const int lrc = launch_term_builder(a, (int)nop, ctx->stream);
        if (lrc) return set_err(ctx, OOPS_ERR_HIP, "term builder launch failed: " + std::to_string(lrc));
        std::vector<int32_t> sm((size_t)nop * kTBSummary);
        rc = hip_check(ctx, hipMemcpyAsync(sm.data(), summary.p, sizeof(int32_t) * sm.size(), hipMemcpyDeviceToHost,
                                           ctx->stream), "D2H builder summary");
        rc = rc ? rc : hip_check(ctx, hipStreamSynchronize(ctx->stream), "term builder");
        if (rc) return rc;
        // bucket properties from every operator, as the host compiler decides them
        const bool dense = n <= m;
        const int cap = dense ? kDenseMaxEnt : krylov_entry_capacity(m);
        for (int64_t o = 0; o < nop; ++o) {
            const int32_t* q = &sm[(size_t)o * kTBSummary];
            if (q[6]) return set_err(ctx, OOPS_ERR_CAPACITY, "device term builder: table overflow");
            if (q[1] > cap) return set_err(ctx, OOPS_ERR_CAPACITY, "operator table exceeds LDS capacity");
            if (q[2] > 16) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "local operator rows wider than 16 entries");
            B.hermitian = B.hermitian && q[5] != 0;
            if (q[3] < 0 || q[3] > kSlot64 || q[4] > kSlot32) B.slots = false;
            if (q[3] < 1 || q[3] > kMedSlot64 || q[4] > kMedSlot32) B.medslot = false;
        }
        if (cpmg && !(n <= m || (n == 64 && uniform) || ((precision == 1 || lanczos64) && B.hermitian)))
            return set_err(ctx, OOPS_ERR_UNSUPPORTED, "CPMG-N items of dimension " + std::to_string(n) +
                                                          " need the dense paths or Lanczos (Hermitian generators)");
        for (int64_t i = 0; i < ni; ++i) {
            H.opA.push_back((int32_t)(i * nops));
            H.opB.push_back((int32_t)(i * nops + 1));
            H.opC.push_back(cpmg ? (int32_t)(i * nops + 2) : -1);
            H.opD.push_back(cpmg ? (int32_t)(i * nops + 3) : -1);
            H.voff.push_back(0);   // every item starts from the same rho0
            H.scale.push_back((double)(1 << k));
            H.npulse.push_back(npulse);
            if (per_item) {
                H.nt.push_back(nt);
                H.tau_off.push_back(0);
                H.L_off.push_back(B.L_len);
                B.L_len += nt;
            }
        }
        if (per_item) H.tau.assign(t, t + nt);
        const std::vector<double>& r0 = rho.at(k);
        if ((int)r0.size() != 2 * n) return set_err(ctx, OOPS_ERR_ARG, "rho0 of the wrong size");
        H.u0 = r0;
        H.y0 = r0;
    }
    if (per_item) {
        // items are numbered by cluster; every item has nt outputs
        for (int64_t i = 0; i < nitems; ++i) b->item_L_off[i] = i * nt;
    }
    return finalize_batch(ctx, b, hosts, per_item, m, precision, ell_only, lanczos64, out);
}

// Hahn-sweep state stores, kept when already large enough.  Slot kernel: the
// sweep-A state (256 complex FP64) at ntmax output times per workgroup.
// Medium kernel: both sweeps' states of every item at ntmax times (items x 2 x
// n x ntmax complex FP64).  Above kStoreMax bytes, or when the allocation
// fails, the bucket runs without a store (both sweeps of an item in one
// workgroup, a step cut at every output time); results are the same to the
// tolerance.
constexpr size_t kStoreMax = size_t(32) << 30;
static int sweep_workspace(oops_ctx* ctx, oops_hahn_batch::Bucket& B, int ntmax) {
    (void)ctx;
    if (ntmax <= 0) return OOPS_OK;
    if (const char* e = std::getenv("OOPS_SWEEP_STORE"))   // "0": never use a store
        if (e[0] == '0') {
            B.sweepws.release();
            B.sweep_nt = 0;
            return OOPS_OK;
        }
    size_t bytes = 0;
    if (B.slots && B.slot_grid > 0) bytes = sizeof(double) * 2 * 256 * (size_t)ntmax * (size_t)B.slot_grid;
    else if (B.medium) bytes = sizeof(double) * 2 * 2 * (size_t)B.n * (size_t)ntmax * B.items.size();
    else return OOPS_OK;
    if (ntmax <= B.sweep_nt && B.sweepws.p) return OOPS_OK;
    B.sweepws.release();
    B.sweep_nt = 0;
    if (bytes > kStoreMax) return OOPS_OK;
    if (hipMalloc(&B.sweepws.p, bytes) != hipSuccess) {
        B.sweepws.p = nullptr;
        (void)hipGetLastError();
        return OOPS_OK;
    }
    B.sweepws.bytes = bytes;
    B.sweep_nt = ntmax;
    return OOPS_OK;
}

// Kernel arguments of bucket B; per-item sequences when the batch has them.
static KrylovArgs bucket_args(const oops_hahn_batch* b, const oops_hahn_batch::Bucket& B, int32_t nt, const double* tau) {
    KrylovArgs a;
    std::memset(&a, 0, sizeof(a));
    a.hdr = (const DevOpHeader*)B.hdr.p;
    a.loc = (const DevOpLocal*)B.loc.p;
    a.val = (const double2*)B.val.p;
    a.off = (const int32_t*)B.off.p;
    a.item_opA = (const int32_t*)B.opA.p;
    a.item_opB = (const int32_t*)B.opB.p;
    a.item_opC = (const int32_t*)B.opC.p;
    a.item_opD = (const int32_t*)B.opD.p;
    a.item_npulse = (const int32_t*)B.npulse.p;
    a.vec_off = (const int64_t*)B.voff.p;
    a.u0 = (const double2*)B.u0.p;
    a.y0 = (const double2*)B.y0.p;
    a.scale = (const double*)B.scale.p;
    a.rnd = (const double*)B.rnd.p;
    a.L = (double*)B.L.p;
    a.stats = (int32_t*)B.stats.p;
    a.queue = (int32_t*)B.queue.p;
    a.eslot = (const uint8_t*)B.eslot.p;
    a.smask = (const int32_t*)B.smask.p;
    a.slotws = (double2*)B.slotws.p;
    a.sweepws = (double2*)B.sweepws.p;
    a.ws_nt = B.sweep_nt;
    if (B.medium && B.medtab.p) {
        a.medmask = (const int32_t*)B.medmask.p;
        a.medtab = (double2*)B.medtab.p;
        a.ms64 = B.ms64;
        a.ms32 = B.ms32;
    }
    a.nitems = (int32_t)B.items.size();
    a.n = B.n;
    a.klim = b->klim;
    a.max_steps = max_steps_limit();
    a.tol = b->tol;
    if (b->per_item) {
        a.item_nt = (const int32_t*)B.nt.p;
        a.item_tau_off = (const int64_t*)B.tau_off.p;
        a.item_L_off = (const int64_t*)B.L_off.p;
        a.tau = (const double*)B.tau.p;
        a.nt = 0;
    } else {
        a.tau = tau;
        a.nt = nt;
    }
    return a;
}

// Buckets of large states: each item's two sweeps on the single-state path
// (synchronous); L and stats kept on the host.
static int bucket_run_large(oops_hahn_batch* b, oops_hahn_batch::Bucket& B, int32_t nt, const double* tau) {
    const size_t ni = B.items.size();
    B.kernel = 5;
    int64_t total = 0;
    for (size_t i = 0; i < ni; ++i) total += b->per_item ? B.hnt[i] : nt;
    B.hL.assign((size_t)total, std::numeric_limits<double>::quiet_NaN());
    B.hstats.assign(ni * kStatsPerItem, 0);
    std::vector<int64_t> off(ni + 1, 0);
    for (size_t i = 0; i < ni; ++i) off[i + 1] = off[i] + (b->per_item ? B.hnt[i] : nt);
    // items run concurrently on host threads, each on its own stream (the
    // per-step host synchronisation of one item overlaps the kernels of the
    // others); device memory per item ~ n (m + 3 + nt) complex
    const double per_item = 16.0 * (double)B.n * (b->m + 3 + (double)(off[ni] / std::max<size_t>(ni, 1)));
    const int nthreads = (int)std::max(1.0, std::min({8.0, (double)ni, 16e9 / per_item}));
    std::atomic<size_t> next{0};
    std::atomic<int> first_rc{0};
    auto worker = [&]() {
        if (hipSetDevice(b->ctx->device) != hipSuccess) { first_rc = OOPS_ERR_HIP; return; }
        for (size_t i; (i = next++) < ni && first_rc.load() == 0;) {
            const int nti = b->per_item ? B.hnt[i] : nt;
            const double* ti = b->per_item ? B.htau.data() + B.htau_off[i] : tau;
            const int rc = large_hahn_item(b->ctx, B.hA[i], B.hB[i], B.hu0[i].data(), B.hy0[i].data(), B.hscale[i],
                                           nti, ti, b->m, b->klim, b->tol, B.hL.data() + off[i],
                                           B.hstats.data() + i * kStatsPerItem);
            if (rc) {
                int zero = 0;
                first_rc.compare_exchange_strong(zero, rc);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    return first_rc.load();
}

static int bucket_launch(oops_hahn_batch* b, oops_hahn_batch::Bucket& B, const KrylovArgs& a, hipStream_t st,
                         bool uniform) {
    const bool dense64 = B.n == 64 && uniform && B.work.p && !B.two_chain;
    const bool lanczos = b->precision == 1 && B.hermitian && B.n > b->m;
    int lrc;
    if (B.factored) {
        B.kernel = 9;
        FactoredArgs f = B.fconst;
        f.coords = (const double*)B.fcoords.p;
        f.clusters = (const int32_t*)B.fspins.p;
        f.nitems = (int32_t)B.items.size();
        f.item_nt = a.item_nt;
        f.item_tau_off = a.item_tau_off;
        f.item_L_off = a.item_L_off;
        f.tau = a.tau;
        f.L = a.L;
        f.stats = a.stats;
        f.queue = a.queue;
        launch_zero_counter(a.queue, st);
        lrc = launch_factored(f, B.grid, st);
    } else if (dense64) {
        B.kernel = 4;
        launch_zero_counter(a.queue, st);
        lrc = launch_dense64(a, (double*)B.work.p, B.dense_grid, st);
    } else if (B.medium && B.medws.p) {
        B.kernel = 7;
        launch_zero_counter(a.queue, st);
        lrc = launch_hahn_medium(a, (float2*)B.medws.p, b->m, B.medium_grid, st);
    } else if (b->lanczos64 && B.slots && B.slotws.p && B.hermitian && B.n > b->m) {
        B.kernel = 8;
        launch_zero_counter(a.queue, st);
        lrc = launch_hahn_lanczos_slot(a, b->m, B.slot_grid, st, true);
    } else if (lanczos && B.slots && B.slotws.p) {
        B.kernel = 6;
        launch_zero_counter(a.queue, st);
        lrc = launch_hahn_lanczos_slot(a, b->m, B.slot_grid, st);
    } else if (lanczos) {
        B.kernel = 3;
        launch_zero_counter(a.queue, st);
        lrc = launch_hahn_lanczos(a, b->m, B.grid, st);
    } else {
        B.kernel = B.n <= b->m ? 0 : (b->precision == 0 ? 1 : 2);
        lrc = launch_hahn_bucket(a, b->m, b->precision == 0 ? 0 : 1, B.grid, st);
    }
    if (lrc != 0)
        return set_err(b->ctx, lrc < 0 ? OOPS_ERR_UNSUPPORTED : OOPS_ERR_HIP,
                       "launch failed for n=" + std::to_string(B.n) + " code " + std::to_string(lrc));
    return OOPS_OK;
}

extern "C" {

int oops_hahn_batch_create(oops_ctx* ctx, int32_t nitems, const oops_hahn_item* items, int32_t m, int32_t klim,
                           double tol, int32_t precision, oops_hahn_batch** out) {
    std::vector<EchoIn> in(std::max(nitems, 0));
    for (int32_t i = 0; i < nitems; ++i)
        in[i] = EchoIn{items[i].opA, items[i].opB, nullptr, nullptr, items[i].u0, items[i].y0, items[i].scale, 1, 0,
                       nullptr};
    return batch_create(ctx, nitems, nitems > 0 && items ? in.data() : nullptr, false, m, klim, tol, precision, out);
}

int oops_echo_batch_create(oops_ctx* ctx, int32_t nitems, const oops_echo_item* items, int32_t m, int32_t klim,
                           double tol, int32_t precision, oops_hahn_batch** out) {
    std::vector<EchoIn> in(std::max(nitems, 0));
    for (int32_t i = 0; i < nitems; ++i)
        in[i] = EchoIn{items[i].opA, items[i].opB, items[i].opC, items[i].opD, items[i].u0, items[i].y0,
                       items[i].scale, items[i].npulse, items[i].nt, items[i].t};
    return batch_create(ctx, nitems, nitems > 0 && items ? in.data() : nullptr, true, m, klim, tol, precision, out);
}

int oops_echo_batch_capture(oops_hahn_batch* b) {
    if (!b || !b->per_item) return OOPS_ERR_ARG;
    oops_ctx* ctx = b->ctx;
    for (auto& B : b->buckets)
        if (B.large) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "batches with states above 256 elements run on the host loop");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    if (b->graph_exec) return OOPS_OK;
    int rc = hip_check(ctx, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (rc) return rc;
    int lrc = OOPS_OK;
    for (auto& B : b->buckets) {
        const KrylovArgs a = bucket_args(b, B, 0, nullptr);
        lrc = bucket_launch(b, B, a, b->stream, B.uniform);
        if (lrc) break;
    }
    hipGraph_t g = nullptr;
    rc = hip_check(ctx, hipStreamEndCapture(b->stream, &g), "hipStreamEndCapture");
    if (lrc) { if (g) (void)hipGraphDestroy(g); return lrc; }
    if (rc) return rc;
    b->graph = g;
    return hip_check(ctx, hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0), "hipGraphInstantiate");
}

int oops_echo_batch_launch(oops_hahn_batch* b) {
    if (!b || !b->per_item) return OOPS_ERR_ARG;
    oops_ctx* ctx = b->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    if (b->graph_exec) {
        (void)hipEventRecord(b->ev[0], b->stream);
        int rc = hip_check(ctx, hipGraphLaunch(b->graph_exec, b->stream), "hipGraphLaunch");
        (void)hipEventRecord(b->ev[1], b->stream);
        b->launched = rc == OOPS_OK;
        return rc;
    }
    for (size_t bk = 0; bk < b->buckets.size(); ++bk) {
        auto& B = b->buckets[bk];
        if (B.large) {
            (void)hipEventRecord(b->ev[2 * bk], b->stream);
            int rc = bucket_run_large(b, B, 0, nullptr);
            (void)hipEventRecord(b->ev[2 * bk + 1], b->stream);
            if (rc) return rc;
            continue;
        }
        const KrylovArgs a = bucket_args(b, B, 0, nullptr);
        (void)hipEventRecord(b->ev[2 * bk], b->stream);
        int rc = bucket_launch(b, B, a, b->stream, B.uniform);
        (void)hipEventRecord(b->ev[2 * bk + 1], b->stream);
        if (rc) return rc;
    }
    b->launched = true;
    return OOPS_OK;
}

int oops_echo_batch_query(oops_hahn_batch* b, int32_t* done) {
    if (!b || !done || !b->per_item) return OOPS_ERR_ARG;
    if (!b->launched) { *done = 0; return OOPS_OK; }
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) { *done = 1; return OOPS_OK; }
    if (e == hipErrorNotReady) { *done = 0; return OOPS_OK; }
    return hip_check(b->ctx, e, "hipStreamQuery");
}

int oops_echo_batch_wait(oops_hahn_batch* b) {
    if (!b || !b->per_item || !b->launched) return OOPS_ERR_ARG;
    int rc = hip_check(b->ctx, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    if (rc) return rc;
    if (b->graph_exec) {  // one interval for the whole replayed graph
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev[0], b->ev[1]);
        for (auto& B : b->buckets) B.ms = ms / std::max<size_t>(1, b->buckets.size());
        b->total_ms = ms;
        return OOPS_OK;
    }
    b->total_ms = 0.0;
    for (size_t bk = 0; bk < b->buckets.size(); ++bk) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, b->ev[2 * bk], b->ev[2 * bk + 1]);
        b->buckets[bk].ms = ms;
        b->total_ms += ms;   // one stream: the buckets run one after another
    }
    return OOPS_OK;
}

int oops_echo_batch_run(oops_hahn_batch* b) {
    int rc = oops_echo_batch_launch(b);
    return rc ? rc : oops_echo_batch_wait(b);
}

int oops_echo_batch_footprint(const oops_hahn_batch* b, int64_t* bytes) {
    if (!b || !bytes) return OOPS_ERR_ARG;
    *bytes = b->footprint;
    return OOPS_OK;
}

int oops_hahn_batch_run(oops_hahn_batch* b, int32_t nt, const double* tau) {
    if (!b || nt <= 0 || !tau) return OOPS_ERR_ARG;
    oops_ctx* ctx = b->ctx;
    if (b->per_item) return set_err(ctx, OOPS_ERR_ARG, "echo batches carry their times: use oops_echo_batch_run");
    for (int k = 1; k < nt; ++k)
        if (!(tau[k] >= tau[k - 1])) return set_err(ctx, OOPS_ERR_ARG, "sweep times must be ascending");
    if (tau[0] < 0.0) return set_err(ctx, OOPS_ERR_ARG, "sweep times must be >= 0");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    int rc = dev_upload(ctx, b->tau, std::vector<double>(tau, tau + nt));
    if (rc) return rc;
    b->nt = nt;
    // uniform grid from 0: the dense propagator powers apply to n = 64 buckets
    bool uniform = nt >= 2 && tau[0] == 0.0 && tau[1] > 0.0;
    for (int k = 2; k < nt && uniform; ++k) uniform = std::fabs(tau[k] - k * tau[1]) <= 1e-12 * tau[nt - 1];
    std::vector<hipEvent_t> ev(2 * b->buckets.size() + 2);
    for (auto& e : ev) (void)hipEventCreate(&e);
    hipEvent_t ev_all0 = ev[ev.size() - 2], ev_all1 = ev[ev.size() - 1];
    // every bucket on its own stream, so that buckets that cannot fill the
    // device (few large clusters) overlap; allocations first, then the
    // launches, then the host-driven buckets of large states
    if (b->bstreams.size() != b->buckets.size()) {
        for (auto& q : b->bstreams) if (q) (void)hipStreamDestroy(q);
        b->bstreams.assign(b->buckets.size(), nullptr);
        for (auto& q : b->bstreams)
            if (!rc) rc = hip_check(ctx, hipStreamCreate(&q), "hipStreamCreate");
    }
    for (size_t bk = 0; bk < b->buckets.size() && !rc; ++bk) {
        auto& B = b->buckets[bk];
        if (B.large) continue;
        rc = alloc_results(ctx, B.L, (int64_t)B.items.size() * nt);
        if (rc) break;
        rc = sweep_workspace(ctx, B, nt);
    }
    // one interval around the whole run: every bucket stream starts after
    // ev_all0 on the context stream, which waits for every bucket before ev_all1
    if (!rc) rc = hip_check(ctx, hipEventRecord(ev_all0, ctx->stream), "hipEventRecord");
    for (size_t bk = 0; bk < b->buckets.size() && !rc; ++bk) {
        auto& B = b->buckets[bk];
        if (B.large) continue;
        hipStream_t st = b->bstreams[bk];
        rc = hip_check(ctx, hipStreamWaitEvent(st, ev_all0, 0), "hipStreamWaitEvent");
        if (rc) break;
        const KrylovArgs a = bucket_args(b, B, nt, (const double*)b->tau.p);
        (void)hipEventRecord(ev[2 * bk], st);
        rc = bucket_launch(b, B, a, st, uniform);
        (void)hipEventRecord(ev[2 * bk + 1], st);
    }
    for (size_t bk = 0; bk < b->buckets.size() && !rc; ++bk) {
        auto& B = b->buckets[bk];
        if (!B.large) continue;
        (void)hipEventRecord(ev[2 * bk], ctx->stream);
        rc = bucket_run_large(b, B, nt, tau);
        (void)hipEventRecord(ev[2 * bk + 1], ctx->stream);
    }
    for (size_t bk = 0; bk < b->buckets.size() && !rc; ++bk)
        if (!b->buckets[bk].large)
            rc = hip_check(ctx, hipStreamWaitEvent(ctx->stream, ev[2 * bk + 1], 0), "hipStreamWaitEvent");
    if (!rc) rc = hip_check(ctx, hipEventRecord(ev_all1, ctx->stream), "hipEventRecord");
    for (auto& q : b->bstreams) {
        if (!q) continue;
        const hipError_t e = hipStreamSynchronize(q);
        if (!rc) rc = hip_check(ctx, e, "hipStreamSynchronize bucket");
    }
    hipError_t se = hipStreamSynchronize(ctx->stream);
    if (!rc) rc = hip_check(ctx, se, "hipStreamSynchronize");
    if (!rc) {
        for (size_t bk = 0; bk < b->buckets.size(); ++bk) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[2 * bk], ev[2 * bk + 1]);
            b->buckets[bk].ms = ms;
        }
        float all = 0.f;
        (void)hipEventElapsedTime(&all, ev_all0, ev_all1);
        b->total_ms = all;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

int oops_hahn_batch_results(oops_hahn_batch* b, double* L, int32_t* stats) {
    if (!b) return OOPS_ERR_ARG;
    oops_ctx* ctx = b->ctx;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    for (auto& B : b->buckets) {
        const size_t ni = B.items.size();
        if (B.large) {
            int64_t off = 0;
            for (size_t i = 0; i < ni; ++i) {
                const int32_t it = B.items[i];
                const int nti = b->per_item ? B.hnt[i] : b->nt;
                if (L) std::memcpy(L + (b->per_item ? b->item_L_off[it] : (int64_t)it * b->nt), B.hL.data() + off,
                                   sizeof(double) * nti);
                if (stats) std::memcpy(stats + (size_t)it * kStatsPerItem, B.hstats.data() + i * kStatsPerItem,
                                       sizeof(int32_t) * kStatsPerItem);
                off += nti;
            }
            continue;
        }
        if (L && b->per_item) {
            std::vector<double> h(B.L_len);
            int rc = hip_check(ctx, hipMemcpy(h.data(), B.L.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy L");
            if (rc) return rc;
            int64_t off = 0;
            for (size_t i = 0; i < ni; ++i) {
                const int32_t it = B.items[i];
                std::memcpy(L + b->item_L_off[it], h.data() + off, sizeof(double) * b->item_nt[it]);
                off += b->item_nt[it];
            }
        } else if (L) {
            if (!B.L.p) return set_err(ctx, OOPS_ERR_ARG, "batch has not been run");
            std::vector<double> h(ni * b->nt);
            int rc = hip_check(ctx, hipMemcpy(h.data(), B.L.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy L");
            if (rc) return rc;
            for (size_t i = 0; i < ni; ++i)
                std::memcpy(L + (size_t)B.items[i] * b->nt, h.data() + i * b->nt, sizeof(double) * b->nt);
        }
        if (stats) {
            std::vector<int32_t> h(ni * kStatsPerItem);
            int rc = hip_check(ctx, hipMemcpy(h.data(), B.stats.p, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost), "hipMemcpy stats");
            if (rc) return rc;
            for (size_t i = 0; i < ni; ++i)
                std::memcpy(stats + (size_t)B.items[i] * kStatsPerItem, h.data() + i * kStatsPerItem,
                            sizeof(int32_t) * kStatsPerItem);
        }
    }
    return OOPS_OK;
}

int oops_hahn_batch_timing(const oops_hahn_batch* b, int32_t max_buckets, int32_t* n, int32_t* count, double* ms,
                           int32_t* nbuckets) {
    if (!b) return OOPS_ERR_ARG;
    if (nbuckets) *nbuckets = (int32_t)b->buckets.size();
    for (int k = 0; k < (int)b->buckets.size() && k < max_buckets; ++k) {
        if (n) n[k] = b->buckets[k].n;
        if (count) count[k] = (int32_t)b->buckets[k].items.size();
        if (ms) ms[k] = b->buckets[k].ms;
    }
    return OOPS_OK;
}

int oops_hahn_batch_elapsed(const oops_hahn_batch* b, double* ms) {
    if (!b || !ms) return OOPS_ERR_ARG;
    *ms = b->total_ms;
    return OOPS_OK;
}

int oops_hahn_batch_kernels(const oops_hahn_batch* b, int32_t max_buckets, int32_t* kind) {
    if (!b || !kind) return OOPS_ERR_ARG;
    for (int k = 0; k < (int)b->buckets.size() && k < max_buckets; ++k) kind[k] = b->buckets[k].kernel;
    return OOPS_OK;
}

int oops_slot_form(const oops_term_table* T, int32_t max_slots, int32_t* nslot64, int32_t* nslot32, int32_t* masks,
                   double* coef) {
    if (!T || !nslot64 || !nslot32) return OOPS_ERR_ARG;
    CompiledOp C;
    std::string err;
    int rc = compile_operator(T, C, err, 1 << 20);
    if (rc) return rc;
    if (C.nslot64 < 0) return OOPS_ERR_CAPACITY;
    *nslot64 = C.nslot64;
    *nslot32 = C.nslot32;
    const int S = C.nslot64 + C.nslot32;
    if (!masks || !coef) return OOPS_OK;   // counts only
    if (S > max_slots) return OOPS_ERR_CAPACITY;
    for (int s = 0; s < C.nslot64; ++s) masks[s] = C.mask64[s];
    for (int s = 0; s < C.nslot32; ++s) masks[C.nslot64 + s] = C.mask32[s];
    const int64_t n = C.n;
    std::vector<std::complex<double>> tab((size_t)S * n, 0.0);
    const int ld = C.ld_bits;
    const int64_t lmask = (1LL << ld) - 1;
    const int top = C.nspin - 1;
    // the same accumulation as the device's per-element table build
    for (int64_t i = 0; i < n; ++i) {
        tab[i] += C.c0;
        for (const DevOpLocal& lo : C.locs) {
            int64_t row = (i >> (ld * (top - lo.s))) & lmask;
            if (lo.nb == 2) row = (row << ld) | ((i >> (ld * (top - lo.t))) & lmask);
            const int64_t e0 = lo.ent + row * lo.R;
            for (int e = 0; e < lo.R; ++e) {
                const uint8_t sl = C.eslot[e0 + e];
                const int slot = (sl & 128) ? C.nslot64 + (sl & 127) : sl;
                tab[(size_t)slot * n + i] += C.val[e0 + e];
            }
        }
    }
    std::memcpy(coef, tab.data(), sizeof(double) * 2 * tab.size());
    return OOPS_OK;
}

int oops_hahn_batch_destroy(oops_hahn_batch* b) {
    if (!b) return OOPS_OK;
    if (b->ctx) (void)hipSetDevice(b->ctx->device);
    delete b;
    return OOPS_OK;
}

}  // extern "C"
