// Host driver of the large-state Krylov path (kernels in expv_large.hip).
//
// Restates the control flow of the reference's multi-output Krylov driver
// (src/expv/source/hamvec_zgexpv_w_mkl_profile.f:4-538; its GPU twin
// src/expv/source/zgexpv_cache.f:4-690) with every vector on the device:
// norm estimate (Lanczos k + power iteration, :210-265), first step
// (:329-334), m Arnoldi steps + one extra product, augmented Hessenberg,
// Pade(6) on the device, local error estimate, rejection (gamma 0.9,
// delta 1.2), update, 2-significant-digit step rounding, output at every
// requested time.  Only scalars (breakdown flag, norms, two entries of the
// small exponential) come back per Krylov step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/oops_expv.h"
#include "host_common.h"
#include "oops_large.h"

using namespace oops;

namespace {

typedef std::complex<double> cplx;

double round_step(double t) {
    if (!(t > 0.0) || !std::isfinite(t)) return std::numeric_limits<double>::infinity();
    const double sqr1 = std::sqrt(0.1);
    const double x = std::log10(t) - sqr1;
    const int k = (int)(x >= 0.0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)) - 1;
    double p = 1.0;
    for (int i = 0; i < std::abs(k); ++i) p *= 10.0;
    if (k < 0) p = 1.0 / p;
    return std::trunc(t / p + 0.55) * p;
}

struct LargeRun {
    oops_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    CompiledOp op;
    DevBuf hdr, loc, val, off;
    LargeOpArgs args{};
    int64_t n = 0;
    int m = 0, ldh = 0, grid = 0, nblk = 0;
    DevBuf V, p0, p1, w, v0, part, hess, coef, scal, flags, F, status, wpart;
    int rc = 0;
};

template <typename T>
int upload(oops_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
    b.release();
    if (v.empty()) return OOPS_OK;
    int rc = hip_check(ctx, hipMalloc(&b.p, v.size() * sizeof(T)), "hipMalloc");
    if (rc) return rc;
    b.bytes = v.size() * sizeof(T);
    return hip_check(ctx, hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int alloc(oops_ctx* ctx, DevBuf& b, size_t bytes) {
    b.release();
    int rc = hip_check(ctx, hipMalloc(&b.p, bytes), "hipMalloc");
    if (!rc) b.bytes = bytes;
    return rc;
}

int setup(oops_ctx* ctx, const oops_term_table* T, int m, LargeRun& R) {
    std::string err;
    int rc = compile_operator(T, R.op, err, kLargeMaxLocals);
    if (rc) return set_err(ctx, rc, err);
    R.ctx = ctx;
    R.n = R.op.n;
    R.m = m;
    R.ldh = m + 2;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    R.grid = large_grid(R.n, ctx->num_cu);
    R.nblk = std::min(R.grid, 512);
    DevOpHeader h;
    std::memset(&h, 0, sizeof(h));
    h.n = (int32_t)R.op.n;
    h.nspin = R.op.nspin;
    h.ld_bits = R.op.ld_bits;
    h.nloc = (int32_t)R.op.locs.size();
    h.nent = (int32_t)R.op.val.size();
    h.max_row = R.op.max_row;
    h.c0re = R.op.c0.real();
    h.c0im = R.op.c0.imag();
    h.ent2_first = R.op.ent2_first;
    std::vector<double> vals;
    for (auto& z : R.op.val) { vals.push_back(z.real()); vals.push_back(z.imag()); }
    if (vals.empty()) { vals.push_back(0.0); vals.push_back(0.0); }
    std::vector<int32_t> offs(R.op.off);
    if (offs.empty()) offs.push_back(0);
    std::vector<DevOpLocal> locs(R.op.locs);
    if (locs.empty()) locs.push_back(DevOpLocal{});
    rc = upload(ctx, R.hdr, std::vector<DevOpHeader>{h});
    rc = rc ? rc : upload(ctx, R.loc, locs);
    rc = rc ? rc : upload(ctx, R.val, vals);
    rc = rc ? rc : upload(ctx, R.off, offs);
    if (rc) return rc;
    R.args.hdr = (const DevOpHeader*)R.hdr.p;
    R.args.loc = (const DevOpLocal*)R.loc.p;
    R.args.val = (const double2*)R.val.p;
    R.args.off = (const int32_t*)R.off.p;
    const size_t vb = sizeof(double2) * (size_t)R.n;
    rc = rc ? rc : alloc(ctx, R.V, vb * (m + 1));
    rc = rc ? rc : alloc(ctx, R.p0, vb);
    rc = rc ? rc : alloc(ctx, R.p1, vb);
    rc = rc ? rc : alloc(ctx, R.w, vb);
    rc = rc ? rc : alloc(ctx, R.v0, vb);
    rc = rc ? rc : alloc(ctx, R.part, sizeof(double) * 2 * (size_t)R.nblk * (m + 4));
    // one spare column: the 2x2-tiled products of the Pade kernel read one
    // element past an odd-sized matrix (into discarded accumulators)
    rc = rc ? rc : alloc(ctx, R.hess, sizeof(double2) * R.ldh * (R.ldh + 1));
    rc = rc ? rc : alloc(ctx, R.coef, sizeof(double2) * (m + 2));
    rc = rc ? rc : alloc(ctx, R.scal, sizeof(double) * 8);
    rc = rc ? rc : alloc(ctx, R.flags, sizeof(int) * 4);
    rc = rc ? rc : alloc(ctx, R.F, sizeof(double2) * (m + 2));
    rc = rc ? rc : alloc(ctx, R.status, sizeof(int) * 4);
    rc = rc ? rc : alloc(ctx, R.wpart, sizeof(double) * R.grid);
    return rc;
}

// complex dot <a, b> (conj a) on the device, returned to the host
int dev_dot(LargeRun& R, const double2* a, const double2* b, cplx& out) {
    launch_large_dots(a, 0, b, R.n, 0, 0, (double*)R.part.p, R.nblk, R.s);
    std::vector<double> h(2 * R.nblk);
    int rc = hip_check(R.ctx, hipMemcpyAsync(h.data(), R.part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, R.s),
                       "D2H dot");
    rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync dot");
    double re = 0.0, im = 0.0;
    for (int k = 0; k < R.nblk; ++k) { re += h[2 * k]; im += h[2 * k + 1]; }
    out = cplx(re, im);
    return rc;
}

int dev_norm(LargeRun& R, const double2* a, double& out) {
    launch_large_dots(a, 0, a, R.n, -1, 1, (double*)R.part.p, R.nblk, R.s);
    std::vector<double> h(2 * R.nblk);
    int rc = hip_check(R.ctx, hipMemcpyAsync(h.data(), R.part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, R.s),
                       "D2H norm");
    rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync norm");
    double s2 = 0.0;
    for (int k = 0; k < R.nblk; ++k) s2 += h[2 * k];
    out = std::sqrt(s2);
    return rc;
}

// Lanczos + power iteration norm estimate on +H (reference :210-265)
int norm_estimate(LargeRun& R, int klim, double tol, double& anorm) {
    const int64_t n = R.n;
    int k = klim;
    if (k > n) k = (int)n;
    std::vector<double> rnd(2 * n + k);
    int32_t seed[4] = {0, 0, 0, 5};
    dlaran_fill(seed, (int64_t)rnd.size(), rnd.data());
    double2* q = (double2*)R.p0.p;
    double2* qm = (double2*)R.p1.p;
    double2* wv = (double2*)R.w.p;
    int rc = hip_check(R.ctx, hipMemcpyAsync(q, rnd.data(), sizeof(double2) * n, hipMemcpyHostToDevice, R.s), "H2D q");
    rc = rc ? rc : hip_check(R.ctx, hipMemsetAsync(qm, 0, sizeof(double2) * n, R.s), "memset");
    if (rc) return rc;
    double bj = 0.0;
    rc = dev_norm(R, q, bj);
    if (rc) return rc;
    launch_large_axpby(make_double2(1.0 / bj, 0.0), q, make_double2(0.0, 0.0), q, n, R.grid, R.s);
    std::vector<double> tk((size_t)k * k, 0.0);
    double bjm = 0.0;
    for (int j = 1; j <= k; ++j) {
        launch_large_matvec(R.args, q, nullptr, wv, nullptr, n, 0, R.grid, R.s);
        cplx aj;
        rc = dev_dot(R, q, wv, aj);
        if (rc) return rc;
        tk[(size_t)(j - 1) * k + (j - 1)] = aj.real();
        launch_large_axpby(make_double2(-aj.real(), -aj.imag()), q, make_double2(1.0, 0.0), wv, n, R.grid, R.s);
        launch_large_axpby(make_double2(-bjm, 0.0), qm, make_double2(1.0, 0.0), wv, n, R.grid, R.s);
        rc = dev_norm(R, wv, bj);
        if (rc) return rc;
        if (j == k) break;
        tk[(size_t)(j - 1) * k + j] = bj;
        tk[(size_t)j * k + (j - 1)] = bj;
        if (bj < tol) break;
        std::swap(q, qm);  // qm <- q
        launch_large_axpby(make_double2(1.0 / bj, 0.0), wv, make_double2(0.0, 0.0), q, n, R.grid, R.s);
        bjm = bj;
    }
    std::vector<double> qr(k), wr(k);
    for (int j = 0; j < k; ++j) qr[j] = rnd[2 * n + j];
    double cjm = 0.0, cj = 0.0;
    for (int it = 0; it < 10000; ++it) {
        for (int i = 0; i < k; ++i) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += tk[(size_t)j * k + i] * qr[j];
            wr[i] = s;
        }
        cj = 0.0;
        for (int i = 0; i < k; ++i) cj += qr[i] * wr[i];
        if (std::fabs(1.0 - cjm / cj) < tol) break;
        cjm = cj;
        double nr = 0.0;
        for (int i = 0; i < k; ++i) nr += wr[i] * wr[i];
        cj = std::sqrt(nr);
        for (int i = 0; i < k; ++i) qr[i] = wr[i] / cj;
    }
    anorm = std::fabs(cj) + std::fabs(bj * qr[k - 1]);
    return OOPS_OK;
}

struct Counters {
    int64_t nmult = 0, nexph = 0, nscale = 0, nstep = 0, nreject = 0, ibrk = 0;
    double anorm = 0.0, kernel_ms = 0.0;
};

// The Expokit integration.  out_mode 0: w_seq (n x nt) to the host;
// 1: overlaps <v0, w(t_k)> to the host (2 doubles per time).
int run_expv(LargeRun& R, const double* v_host, int nt, const double* ta, int klim, double tol, int out_mode,
             double* out, Counters& C) {
    const int m = R.m;
    const int64_t n = R.n;
    const double delta = 1.2, gamma = 0.9, break_tol = 1.0e-7;
    const int64_t max_steps = 50000000;
    int rc = norm_estimate(R, klim, tol, C.anorm);
    if (rc) return rc;
    const double anorm = C.anorm;
    rc = hip_check(R.ctx, hipMemcpyAsync(R.w.p, v_host, sizeof(double2) * n, hipMemcpyHostToDevice, R.s), "H2D v");
    rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(R.v0.p, v_host, sizeof(double2) * n, hipMemcpyHostToDevice, R.s),
                             "H2D v0");
    if (rc) return rc;
    double beta = 0.0;
    rc = dev_norm(R, (const double2*)R.w.p, beta);
    if (rc) return rc;
    const double eps = 2.220446049250313e-16;
    if (tol <= eps) tol = std::sqrt(eps);
    double xm = 1.0 / (double)m;
    double t_new;
    {
        const double p2 = tol * std::pow((m + 1) / 2.72, m + 1) * std::sqrt(2.0 * 3.14 * (m + 1));
        t_new = round_step((1.0 / anorm) * std::pow(p2 / (4.0 * beta * anorm), xm));
    }
    double t_now = 0.0;
    double2* V = (double2*)R.V.p;
    double2* hess = (double2*)R.hess.p;
    double* scal = (double*)R.scal.p;
    hipEvent_t e0, e1;
    hipEventCreate(&e0);
    hipEventCreate(&e1);
    float ms_acc = 0.0f;
    for (int tt = 0; tt < nt; ++tt) {
        const double t_out = ta[tt];
        while (t_now < t_out) {
            if (C.nstep >= max_steps) return set_err(R.ctx, OOPS_ERR_NUMERIC, "step limit reached");
            ++C.nstep;
            double t_step = std::min(t_out - t_now, t_new);
            int k1 = 2, mb = m;
            hipEventRecord(e0, R.s);
            (void)hipMemsetAsync(hess, 0, sizeof(double2) * R.ldh * (R.ldh + 1), R.s);
            (void)hipMemsetAsync(R.flags.p, 0, sizeof(int) * 4, R.s);
            const double inv_beta = 1.0 / beta;
            (void)hipMemcpyAsync(scal, &inv_beta, sizeof(double), hipMemcpyHostToDevice, R.s);
            double2* bufs[2] = {(double2*)R.p0.p, (double2*)R.p1.p};
            const double2* x = (const double2*)R.w.p;
            for (int j = 0; j < m; ++j) {
                double2* p = bufs[j & 1];
                launch_large_matvec(R.args, x, scal, p, V + (size_t)j * n, n, 1, R.grid, R.s);
                launch_large_dots(V, n, p, n, j, 0, (double*)R.part.p, R.nblk, R.s);
                launch_large_finish((double*)R.part.p, R.nblk, j + 1, j, 0, hess, R.ldh, (double2*)R.coef.p, scal,
                                    (int*)R.flags.p, break_tol, R.s);
                launch_large_update(V, n, p, n, j, (double2*)R.coef.p, R.grid, R.s);
                launch_large_dots(V, n, p, n, j, 0, (double*)R.part.p, R.nblk, R.s);
                launch_large_finish((double*)R.part.p, R.nblk, j + 1, j, 1, hess, R.ldh, (double2*)R.coef.p, scal,
                                    (int*)R.flags.p, break_tol, R.s);
                launch_large_update(V, n, p, n, j, (double2*)R.coef.p, R.grid, R.s);
                launch_large_dots(p, 0, p, n, -1, 1, (double*)R.part.p, R.nblk, R.s);
                launch_large_finish((double*)R.part.p, R.nblk, 1, j, 2, hess, R.ldh, nullptr, scal, (int*)R.flags.p,
                                    break_tol, R.s);
                x = p;
            }
            // extra product for the error estimate: V[m] = p / h(m, m-1), p' = -iH V[m]
            double2* pe = bufs[m & 1];
            launch_large_matvec(R.args, x, scal, pe, V + (size_t)m * n, n, 1, R.grid, R.s);
            launch_large_dots(pe, 0, pe, n, -1, 1, (double*)R.part.p, R.nblk, R.s);
            launch_large_finish((double*)R.part.p, R.nblk, 1, 0, 3, hess, R.ldh, nullptr, scal, (int*)R.flags.p,
                                break_tol, R.s);
            int fl[4];
            double sc[4];
            rc = hip_check(R.ctx, hipMemcpyAsync(fl, R.flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, R.s), "D2H");
            rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(sc, scal, sizeof(double) * 4, hipMemcpyDeviceToHost, R.s), "D2H");
            rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync arnoldi");
            if (rc) return rc;
            double avnorm = sc[2];
            if (fl[0]) {
                // happy breakdown at column mb: the products after it ran on a
                // zero vector and are not counted
                k1 = 0;
                mb = fl[1];
                t_step = t_out - t_now;
                C.ibrk = 1;
                C.nmult += mb;
            } else {
                C.nmult += m + 1;
            }
            const double2 one = make_double2(1.0, 0.0);
            (void)hipMemcpyAsync(hess + (m + 1) + (size_t)R.ldh * m, &one, sizeof(double2), hipMemcpyHostToDevice, R.s);
            double err_loc = 0.0;
            int mx;
            std::vector<double2> Fh(m + 2);
            for (;;) {
                ++C.nexph;
                mx = mb + k1;
                if (launch_large_pade(hess, R.ldh, mx, t_step, (double2*)R.F.p, (int*)R.status.p, R.s) != 0)
                    return set_err(R.ctx, OOPS_ERR_HIP, "Pade kernel launch");
                int st = 0;
                rc = hip_check(R.ctx, hipMemcpyAsync(Fh.data(), R.F.p, sizeof(double2) * mx, hipMemcpyDeviceToHost, R.s), "D2H F");
                rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(&st, R.status.p, sizeof(int), hipMemcpyDeviceToHost, R.s), "D2H st");
                rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync pade");
                if (rc) return rc;
                if (st < 0) return set_err(R.ctx, OOPS_ERR_NUMERIC, "small exponential failed");
                C.nscale += st;
                if (k1 == 0) {
                    err_loc = tol;
                } else {
                    const double p1 = std::abs(cplx(Fh[m].x, Fh[m].y)) * beta;
                    const double p2 = std::abs(cplx(Fh[m + 1].x, Fh[m + 1].y)) * beta * avnorm;
                    if (p1 > 10.0 * p2) { err_loc = p2; xm = 1.0 / (double)m; }
                    else if (p1 > p2) { err_loc = (p1 * p2) / (p1 - p2); xm = 1.0 / (double)m; }
                    else { err_loc = p1; xm = 1.0 / (double)(m - 1); }
                }
                if (k1 != 0 && err_loc > delta * t_step * tol) {
                    t_step = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
                    ++C.nreject;
                    if (!std::isfinite(t_step)) return set_err(R.ctx, OOPS_ERR_NUMERIC, "step size collapsed");
                    continue;
                }
                break;
            }
            mx = mb + std::max(0, k1 - 1);
            launch_large_combine(V, n, mx, (double2*)R.F.p, beta, (double2*)R.w.p, n, (double*)R.wpart.p, R.grid, R.s);
            hipEventRecord(e1, R.s);
            std::vector<double> wp(R.grid);
            rc = hip_check(R.ctx, hipMemcpyAsync(wp.data(), R.wpart.p, sizeof(double) * R.grid, hipMemcpyDeviceToHost, R.s),
                           "D2H beta");
            rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync combine");
            if (rc) return rc;
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            ms_acc += ms;
            double s2 = 0.0;
            for (double v : wp) s2 += v;
            beta = std::sqrt(s2);
            t_new = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
            t_now += t_step;
        }
        if (out) {
            if (out_mode == 0) {
                rc = hip_check(R.ctx, hipMemcpyAsync(out + 2 * (size_t)n * tt, R.w.p, sizeof(double2) * n,
                                                     hipMemcpyDeviceToHost, R.s), "D2H w");
                rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync w");
            } else {
                cplx ov;
                rc = dev_dot(R, (const double2*)R.v0.p, (const double2*)R.w.p, ov);
                out[2 * tt] = ov.real();
                out[2 * tt + 1] = ov.imag();
            }
            if (rc) return rc;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    C.kernel_ms = ms_acc;
    return OOPS_OK;
}

}  // namespace

extern "C" {

int oops_hamvec(oops_ctx* ctx, const oops_term_table* H, const double* v, double* w) {
    if (!ctx || !H || !v || !w) return OOPS_ERR_ARG;
    LargeRun R;
    int rc = setup(ctx, H, 1, R);
    if (rc) return rc;
    R.s = nullptr;
    rc = hip_check(ctx, hipMemcpy(R.p0.p, v, sizeof(double2) * R.n, hipMemcpyHostToDevice), "H2D v");
    if (rc) return rc;
    launch_large_matvec(R.args, (const double2*)R.p0.p, nullptr, (double2*)R.p1.p, nullptr, R.n, 0, R.grid, nullptr);
    rc = hip_check(ctx, hipGetLastError(), "matvec launch");
    rc = rc ? rc : hip_check(ctx, hipMemcpy(w, R.p1.p, sizeof(double2) * R.n, hipMemcpyDeviceToHost), "D2H w");
    return rc;
}

int oops_expv(oops_ctx* ctx, const oops_term_table* H, const double* v, int32_t nt, const double* ta, int32_t klim,
              int32_t m, double tol, int32_t out_mode, double* out, double* stats) {
    if (!ctx || !H || !v || nt <= 0 || !ta) return set_err(ctx, OOPS_ERR_ARG, "bad arguments");
    if (m < 2 || m > kLargeMaxM) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "m must be in [2, 40] on the large path");
    if (klim < 1 || klim > 64) return set_err(ctx, OOPS_ERR_ARG, "klim must be in [1, 64]");
    for (int k = 1; k < nt; ++k)
        if (!(ta[k] >= ta[k - 1])) return set_err(ctx, OOPS_ERR_ARG, "output times must be ascending");
    LargeRun R;
    int rc = setup(ctx, H, m, R);
    if (rc) return rc;
    if (m >= R.n) return set_err(ctx, OOPS_ERR_ARG, "the Krylov dimension must be smaller than the state dimension");
    hipStreamCreate(&R.s);
    Counters C;
    rc = run_expv(R, v, nt, ta, klim, tol, out_mode, out, C);
    hipStreamDestroy(R.s);
    if (stats) {
        stats[0] = (double)C.nmult;
        stats[1] = (double)C.nexph;
        stats[2] = (double)C.nscale;
        stats[3] = (double)C.nstep;
        stats[4] = (double)C.nreject;
        stats[5] = (double)C.ibrk;
        stats[6] = C.anorm;
        stats[7] = C.kernel_ms;
    }
    return rc;
}

// Fortran-ABI twin of the reference's MAIN_CACHE (all arguments by reference,
// size_t integers as MatExpVector::runInexplicitGPU passes them,
// src/source/math/MatExp.cpp:487-511).  maxThreadsPerBlock / maxGridSize are
// accepted and ignored (launch shapes are chosen here).  Runs on device 0.
void main_cache_(size_t* nspin, size_t* nTerm, double* coeff_lst, size_t* nbody_lst, size_t* pos_i_idx,
                 size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
                 size_t* nspin_dim, double* v, size_t* pos_idx, size_t* mat_idx, size_t* k, size_t* maxThreadsPerBlock,
                 size_t* maxGridSize, size_t* tn, double* ta, size_t* m, double* tol, size_t* itrace, double* w_seq,
                 size_t* w_seq_len) {
    (void)pos_idx; (void)mat_idx; (void)maxThreadsPerBlock; (void)maxGridSize; (void)itrace; (void)ham_dim;
    (void)w_seq_len;
    oops_term_table T;
    T.nspin = (int32_t)*nspin;
    T.nterm = (int32_t)*nTerm;
    T.coeff = coeff_lst;
    T.nbody = reinterpret_cast<const int64_t*>(nbody_lst);
    T.pos_offset = reinterpret_cast<const int64_t*>(pos_i_idx);
    T.pos = reinterpret_cast<const int64_t*>(pos_i_lst);
    T.dims = reinterpret_cast<const int64_t*>(dim_i_lst);
    T.mat_offset = reinterpret_cast<const int64_t*>(mat_i_idx);
    T.mats = mat_i_lst;
    T.spin_dim = reinterpret_cast<const int64_t*>(nspin_dim);
    oops_ctx* ctx = nullptr;
    if (oops_ctx_create(0, &ctx) != OOPS_OK) return;
    oops_expv(ctx, &T, v, (int32_t)*tn, ta, (int32_t)*k, (int32_t)*m, *tol, 0, w_seq, nullptr);
    oops_ctx_destroy(ctx);
}

}  // extern "C"
