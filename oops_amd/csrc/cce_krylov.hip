// Batched, size-bucketed Krylov expv for CCE clusters on gfx950.
//
// One workgroup (NT = 256 threads, 4 waves) owns one cluster at a time and
// pulls clusters from a device work counter until the bucket is drained.
// Thread i owns element i of the Liouville state (n <= 256), so the whole
// Krylov basis V[0..M] of that element lives in registers (FP64, or FP32 for
// v1..vm on the mixed path); the matvec input
// is staged once per product in LDS.  Both Hahn-echo sweeps of a cluster are
// integrated in the same workgroup, interleaved output time by output time,
// and reduced on the spot to
//     L(t_k) = Re( scale * < expv(B', y0, tau_k), expv(A, u0, tau_k) > ),
// so only nt doubles per cluster ever reach HBM.
//
// The integrator follows the reference's multi-output Expokit driver
// (src/expv/source/hamvec_zgexpv_w_mkl_profile.f:4-538 and its GPU twin
// src/expv/source/zgexpv_cache.f:4-690): Lanczos(k) + power-iteration norm
// estimate (:210-265), Expokit's first step (:329-334), m Arnoldi steps plus
// one extra product (:357-393), the 2-corrected (m+2) augmentation (:401),
// Pade(6) (expokit.f:762-934), the local error estimate and step rejection
// with gamma = 0.9 / delta = 1.2 (:431-472), the update (:476-494), and the
// 2-significant-digit step rounding.  Orthogonalisation is classical
// Gram-Schmidt with one full re-orthogonalisation pass (two block reductions
// per step instead of j), which equals the reference's modified Gram-Schmidt
// to rounding.  Clusters with n <= m (where the reference stops, :279) go to
// the dense kernel below instead.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "devcommon.h"
#include "oops_kernels.h"

namespace oops {

constexpr int kNT = 256;

template <int M>
struct KrylovLds {
    static constexpr int MH = M + 2;
    static constexpr int MAT = MH * MH;
    static constexpr int NMAX = 256;
    static constexpr int EMAX = kKrylovMaxEnt;  // ELL entries per operator
};

struct IntegState {
    cd w;
    double beta, t_now, t_new, anorm, rndoff, xm;
    int nstep, nmult, nrej, nexph, nscale, brk, fail, n;
#ifdef OOPS_PHASE_TIMERS
    unsigned long long tmv, tred, texp, tupd;
#endif
};

#ifdef OOPS_PHASE_TIMERS
#define OOPS_TIC(v) const unsigned long long v = clock64()
#define OOPS_TOC(g, field, v) (g).field += clock64() - (v)
#else
#define OOPS_TIC(v)
#define OOPS_TOC(g, field, v)
#endif

// One operator in LDS (ELL tables of the merged local operators).
struct OpLds {
    DevOpHeader hdr;
    DevOpLocal locs[kMaxLocals];
    alignas(16) cd vals[kKrylovMaxEnt];
    alignas(16) int32_t offs[kKrylovMaxEnt];
};

// LDS of one workgroup (< 80 KB so two workgroups share a CU): the operator
// tables and the product input are only needed during the Arnoldi phase and
// the Pade workspace only during the small exponential, so they share one
// region; the operator is re-read from L2 at the start of every Krylov step.
template <int M>
struct KrylovSmem {
    cd hess[KrylovLds<M>::MAT];
    cd F[KrylovLds<M>::MH];
    int piv[128];
    double red[3 * (kNT / kWave) * 64];
    double scal[16];
    double tdiag[16];
    double toff[16];
    int ints[8];
    union Region {
        struct Arn {
            OpLds op;
            cd xs[KrylovLds<M>::NMAX];
            float2 xs32[KrylovLds<M>::NMAX];
        } a;
        cd pw[3][KrylovLds<M>::MAT];
    } u;
};

// Global operator tables of a bucket.
struct OpSrc {
    const DevOpHeader* hdr;
    const DevOpLocal* loc;
    const double2* val;
    const int32_t* off;
};

// Expokit's step rounding to two significant digits (+inf stays +inf).
__device__ __forceinline__ double round_step2(double t) {
    if (!(t > 0.0) || isinf(t)) return __builtin_huge_val();
    const double sqr1 = 0.31622776601683794;  // sqrt(0.1)
    const double x = log10(t) - sqr1;
    const int k = (int)(x >= 0.0 ? floor(x + 0.5) : -floor(-x + 0.5)) - 1;
    double p = 1.0;
    const int a = k < 0 ? -k : k;
    for (int i = 0; i < a; ++i) p *= 10.0;
    if (k < 0) p = 1.0 / p;
    return trunc(t / p + 0.55) * p;
}

// Stage x_i for the product (FP64, plus an FP32 copy on the mixed path).
template <bool MIXED>
__device__ __forceinline__ void stage_x(cd* xs, float2* xs32, int i, int n, cd x) {
    if (i < n) {
        xs[i] = x;
        if (MIXED) xs32[i] = make_float2((float)x.re, (float)x.im);
    }
}

// y_i = H x (x staged in LDS by the caller), zero for i >= n.
template <bool MIXED>
__device__ __forceinline__ cd matvec_elem(const OpLds& o, const cd* xs, const float2* xs32, int i) {
    if (i >= o.hdr.n) return mkcd(0.0, 0.0);
    if (MIXED) return op_apply_elem_mixed(i, xs, xs32, o.hdr, o.locs, o.vals, o.offs);
    return op_apply_elem(i, xs, o.hdr, o.locs, o.vals, o.offs);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) v = fmax(v, __shfl_xor(v, mask, kWave));
    return v;
}

__device__ __forceinline__ cd readlane_cd(cd x, int c) {
    return mkcd(lane_val(x.re, c), lane_val(x.im, c));
}

// ----------------------------------------------------------------------------
// Norm estimate: k Lanczos steps on +H from the DLARAN start vector, then
// power iteration on the k x k tridiagonal (reference :210-265).
template <int M, bool MIXED>
__device__ double lanczos_anorm(KrylovSmem<M>& S, BlockRed<kNT>& br, const OpLds& o, const double* rnd, int klim,
                                double tol) {
    const int tid = threadIdx.x;
    const int n = o.hdr.n;
    int k = klim;
    if (k > n) k = n;
    double* tdiag = S.tdiag;
    double* toff = S.toff;
    for (int q2 = tid; q2 < 16; q2 += kNT) { tdiag[q2] = 0.0; toff[q2] = 0.0; }
    cd q = tid < n ? mkcd(rnd[2 * tid], rnd[2 * tid + 1]) : mkcd(0.0, 0.0);
    cd qm = mkcd(0.0, 0.0);
    {
        const auto t0 = block_sum1<1, kNT>(br, [&](int) { return cabs2(q); });
        q = cscale(q, 1.0 / sqrt(t0.get(0)));
    }
    double bjm = 0.0, bj = 0.0;
    for (int j = 1; j <= k; ++j) {
        stage_x<MIXED>(S.u.a.xs, S.u.a.xs32, tid, n, q);
        __syncthreads();
        cd w = matvec_elem<MIXED>(o, S.u.a.xs, S.u.a.xs32, tid);
        const auto ta = block_sum1<2, kNT>(br, [&](int c) {
            return c == 0 ? fma(q.re, w.re, q.im * w.im) : fma(q.re, w.im, -q.im * w.re);
        });
        const cd aj = mkcd(ta.get(0), ta.get(1));
        w = cfms(aj, q, w);
        w = mkcd(w.re - bjm * qm.re, w.im - bjm * qm.im);
        const auto tb = block_sum1<1, kNT>(br, [&](int) { return cabs2(w); });
        bj = sqrt(tb.get(0));
        if (tid == 0) tdiag[j - 1] = aj.re;
        if (j == k) break;
        if (tid == 0) toff[j - 1] = bj;
        if (bj < tol) break;
        qm = q;
        q = cscale(w, 1.0 / bj);
        bjm = bj;
    }
    __syncthreads();
    if (tid == 0) {
        // power iteration on T (k x k symmetric tridiagonal; entries not set stay 0)
        double qr[16], wr[16];
        for (int j = 0; j < k; ++j) qr[j] = rnd[2 * n + j];
        double cjm = 0.0, cj = 0.0;
        for (int it = 0; it < 10000; ++it) {
            for (int i = 0; i < k; ++i) {
                double s = tdiag[i] * qr[i];
                if (i > 0) s += toff[i - 1] * qr[i - 1];
                if (i + 1 < k) s += toff[i] * qr[i + 1];
                wr[i] = s;
            }
            cj = 0.0;
            for (int i = 0; i < k; ++i) cj += qr[i] * wr[i];
            if (fabs(1.0 - cjm / cj) < tol) break;
            cjm = cj;
            double nr = 0.0;
            for (int i = 0; i < k; ++i) nr += wr[i] * wr[i];
            cj = sqrt(nr);
            for (int i = 0; i < k; ++i) qr[i] = wr[i] / cj;
        }
        S.scal[0] = fabs(cj) + fabs(bj * qr[k - 1]);
    }
    __syncthreads();
    const double an = S.scal[0];
    __syncthreads();
    return an;
}

// ----------------------------------------------------------------------------
// Register-resident Krylov basis of one state element.  FP64 path: all m+1
// vectors in FP64.  Mixed path: v0 = w/beta in FP64 and v1..vm stored in
// FP32 (half the registers); every product with them accumulates in FP64.
template <int M, bool B32>
struct Basis;

template <int M>
struct Basis<M, false> {
    cd v[M + 1];
    __device__ __forceinline__ cd get(int l) const { return v[l]; }
    __device__ __forceinline__ void set(int l, cd x) { v[l] = x; }
};

template <int M>
struct Basis<M, true> {
    cd v0;
    float2 v[M];
    __device__ __forceinline__ cd get(int l) const {
        if (l == 0) return v0;
        const float2 f = v[l > 0 ? l - 1 : 0];
        return mkcd((double)f.x, (double)f.y);
    }
    __device__ __forceinline__ void set(int l, cd x) {
        if (l == 0) v0 = x;
        else v[l > 0 ? l - 1 : 0] = make_float2((float)x.re, (float)x.im);
    }
};

// ----------------------------------------------------------------------------
// One Arnoldi step J (compile-time, so the basis stays in registers and the
// reduction width is fixed): p = -iH V[J], two classical Gram-Schmidt passes
// (one block reduction each), column J of the Hessenberg matrix,
// V[J+1] = p / h(J+1,J).
template <int M, bool MIXED>
struct ArnCtx {
    KrylovSmem<M>* S;
    BlockRed<kNT>* br;
    const OpLds* o;
    IntegState* g;
    double t_out;
    double t_step;
    int k1;
    int mb;
    bool brk;
};

template <int M, bool MIXED, int J>
__device__ __forceinline__ void arnoldi_step(ArnCtx<M, MIXED>& c, Basis<M, MIXED>& V) {
    if (c.brk) return;
    constexpr int MH = M + 2;
    constexpr int NV = 2 * (J + 1) + 1;
    constexpr int P = pow2ceil(NV);
    const int tid = threadIdx.x;
    const int n = c.o->hdr.n;
    OOPS_TIC(tm0);
    stage_x<MIXED>(c.S->u.a.xs, c.S->u.a.xs32, tid, n, V.get(J));
    __syncthreads();
    cd p = cmulmi(matvec_elem<MIXED>(*c.o, c.S->u.a.xs, c.S->u.a.xs32, tid));
    ++c.g->nmult;
    OOPS_TOC(*c.g, tmv, tm0);
    OOPS_TIC(tr0);
    // pass 1: h = V^H p
    const auto t1 = block_sum1<P, kNT>(*c.br, [&](int k) -> double {
        if (k >= 2 * (J + 1)) return 0.0;
        const cd b = V.get(k >> 1);
        return (k & 1) ? fma(b.re, p.im, -b.im * p.re) : fma(b.re, p.re, b.im * p.im);
    });
    cd hcol0 = mkcd(0.0, 0.0);
#pragma unroll
    for (int l = 0; l <= J; ++l) {
        const cd h = mkcd(t1.get(2 * l), t1.get(2 * l + 1));
        if (tid == 0) c.S->hess[l + MH * J] = h;
        p = cfms(h, V.get(l), p);
    }
    (void)hcol0;
    // pass 2 (re-orthogonalisation) and |p|^2 after pass 1
    const auto t2 = block_sum1<P, kNT>(*c.br, [&](int k) -> double {
        if (k < 2 * (J + 1)) {
            const cd b = V.get(k >> 1);
            return (k & 1) ? fma(b.re, p.im, -b.im * p.re) : fma(b.re, p.re, b.im * p.im);
        }
        if (k == 2 * (J + 1)) return tid < n ? cabs2(p) : 0.0;
        return 0.0;
    });
    double corr = 0.0;
#pragma unroll
    for (int l = 0; l <= J; ++l) {
        const cd h2 = mkcd(t2.get(2 * l), t2.get(2 * l + 1));
        p = cfms(h2, V.get(l), p);
        corr += cabs2(h2);
        if (tid == 0) c.S->hess[l + MH * J] = cadd(c.S->hess[l + MH * J], h2);
    }
    const double hj1j = sqrt(fmax(t2.get(2 * (J + 1)) - corr, 0.0));
    OOPS_TOC(*c.g, tred, tr0);
    if (hj1j <= 1.0e-7) {
        c.k1 = 0;
        c.g->brk = 1;
        c.mb = J + 1;
        c.t_step = c.t_out - c.g->t_now;
        c.brk = true;
    } else {
        if (tid == 0) c.S->hess[(J + 1) + MH * J] = mkcd(hj1j, 0.0);
        V.set(J + 1, cscale(p, 1.0 / hj1j));
    }
}

template <int M, bool MIXED, int... Js>
__device__ __forceinline__ void arnoldi_all(ArnCtx<M, MIXED>& c, Basis<M, MIXED>& V, std::integer_sequence<int, Js...>) {
    (arnoldi_step<M, MIXED, Js>(c, V), ...);
}

// ----------------------------------------------------------------------------
// Krylov steps with Expokit's step control; advances g to t_out.
template <int M, bool MIXED>
__device__ __forceinline__ void advance(KrylovSmem<M>& S, BlockRed<kNT>& br, IntegState& g, Basis<M, MIXED>& V,
                                        double t_out, int max_steps, const OpSrc& src, int op, double tol) {
    const OpLds& o = S.u.a.op;
    constexpr int MH = M + 2;
    const int tid = threadIdx.x;
    const double delta = 1.2, gamma = 0.9;
    const int n = g.n;
    while (g.t_now < t_out) {
        if (g.nstep >= max_steps) { g.fail |= 1; break; }
        ++g.nstep;
        for (int idx = tid; idx < MH * MH; idx += kNT) S.hess[idx] = mkcd(0.0, 0.0);
        __syncthreads();  // the Pade workspace of the previous step is dead
        op_load_lds<kNT, MIXED>(src, op, S.u.a.op);
        V.set(0, cscale(g.w, 1.0 / g.beta));
        ArnCtx<M, MIXED> c{&S, &br, &o, &g, t_out, fmin(t_out - g.t_now, g.t_new), 2, M, false};
        __syncthreads();
        arnoldi_all<M, MIXED>(c, V, std::make_integer_sequence<int, M>{});
        const bool brk = c.brk;
        const int k1 = c.k1;
        const int mb = c.mb;
        double t_step = c.t_step;
        double avnorm = 0.0;
        if (!brk) {
            stage_x<MIXED>(S.u.a.xs, S.u.a.xs32, tid, n, V.get(M));
            __syncthreads();
            const cd p = cmulmi(matvec_elem<MIXED>(o, S.u.a.xs, S.u.a.xs32, tid));
            ++g.nmult;
            const auto tn = block_sum1<1, kNT>(br, [&](int) { return cabs2(p); });
            avnorm = sqrt(tn.get(0));
        }
        if (tid == 0) S.hess[(M + 1) + MH * M] = mkcd(1.0, 0.0);
        __syncthreads();
        // ---- small exponential, error estimate, rejection
        double err_loc = 0.0;
        int mx = 0;
        OOPS_TIC(te0);
        for (;;) {
            ++g.nexph;
            mx = mb + k1;
            const int ns = pade6_first_column3<kNT>(S.hess, MH, mx, t_step, S.u.pw[0], S.u.pw[1], S.u.pw[2], S.F,
                                                   S.piv, S.scal + 8);
            if (ns < 0) { g.fail |= 2; return; }
            g.nscale += ns;
            if (k1 == 0) {
                err_loc = tol;
            } else {
                const double p1 = cabs(S.F[M]) * g.beta;
                const double p2 = cabs(S.F[M + 1]) * g.beta * avnorm;
                if (p1 > 10.0 * p2) { err_loc = p2; g.xm = 1.0 / (double)M; }
                else if (p1 > p2) { err_loc = (p1 * p2) / (p1 - p2); g.xm = 1.0 / (double)M; }
                else { err_loc = p1; g.xm = 1.0 / (double)(M - 1); }
            }
            if (k1 != 0 && err_loc > delta * t_step * tol) {
                t_step = gamma * t_step * pow(t_step * tol / err_loc, g.xm);
                t_step = round_step2(t_step);
                ++g.nrej;
                __syncthreads();
                if (!(t_step < 1e300) || g.nrej > max_steps) { g.fail |= 4; return; }
                continue;
            }
            break;
        }
        OOPS_TOC(g, texp, te0);
        OOPS_TIC(tu0);
        // ---- w = beta * V(:, 0:mx) * F(0:mx, 0)
        mx = mb + (k1 - 1 > 0 ? k1 - 1 : 0);
        cd acc = mkcd(0.0, 0.0);
#pragma unroll
        for (int j = 0; j <= M; ++j)
            if (j < mx) acc = cfma(V.get(j), S.F[j], acc);
        g.w = cscale(acc, g.beta);
        {
            const auto tb = block_sum1<1, kNT>(br, [&](int) { return tid < n ? cabs2(g.w) : 0.0; });
            g.beta = sqrt(tb.get(0));
        }
        g.t_new = gamma * t_step * pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        g.t_now += t_step;
        OOPS_TOC(g, tupd, tu0);
    }
}

template <int M, bool MIXED>
__device__ void init_integ(KrylovSmem<M>& S, BlockRed<kNT>& br, IntegState& g, cd w0, const double* rnd, int klim,
                           double tol, const OpSrc& src, int op, int m) {
    const int tid = threadIdx.x;
    __syncthreads();
    op_load_lds<kNT, MIXED>(src, op, S.u.a.op);
    __syncthreads();
    const OpLds& o = S.u.a.op;
    g.n = o.hdr.n;
    g.w = tid < o.hdr.n ? w0 : mkcd(0.0, 0.0);
    g.anorm = lanczos_anorm<M, MIXED>(S, br, o, rnd, klim, tol);
    const auto tb = block_sum1<1, kNT>(br, [&](int) { return cabs2(g.w); });
    g.beta = sqrt(tb.get(0));
    g.t_now = 0.0;
    g.xm = 1.0 / (double)m;
    g.rndoff = 2.220446049250313e-16 * g.anorm;
    const double p2 = tol * pow((m + 1) / 2.72, (double)(m + 1)) * sqrt(2.0 * 3.14 * (m + 1));
    g.t_new = (1.0 / g.anorm) * pow(p2 / (4.0 * g.beta * g.anorm), g.xm);
    g.t_new = round_step2(g.t_new);
    g.nstep = g.nmult = g.nrej = g.nexph = g.nscale = g.brk = g.fail = 0;
#ifdef OOPS_PHASE_TIMERS
    g.tmv = g.tred = g.texp = g.tupd = 0;
#endif
}

template <int NT, bool MIXED>
__device__ __forceinline__ void op_load_lds(const OpSrc& a, int op, OpLds& o) {
    o.hdr = a.hdr[op];
    const DevOpHeader h = a.hdr[op];
    for (int k = threadIdx.x; k < h.nloc; k += NT) o.locs[k] = a.loc[h.loc_off + k];
    float2* v32 = reinterpret_cast<float2*>(o.vals + h.ent2_first);
    for (int k = threadIdx.x; k < h.nent; k += NT) {
        const double2 v = a.val[h.ent_off + k];
        if (MIXED && k >= h.ent2_first) v32[k - h.ent2_first] = make_float2((float)v.x, (float)v.y);
        else o.vals[k] = mkcd(v.x, v.y);
        o.offs[k] = a.off[h.ent_off + k];
    }
}

template <int M, bool MIXED>
__global__ __launch_bounds__(kNT) void cce_hahn_krylov_kernel(KrylovArgs a) {
    __shared__ KrylovSmem<M> S;
    const int tid = threadIdx.x;
    BlockRed<kNT> br(S.red);
    const double tol = a.tol;
    const double* rnd = a.rnd;
    const int klim = a.klim;
    const int max_steps = a.max_steps;
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int64_t voff = a.vec_off[item];
        const OpSrc src{a.hdr, a.loc, a.val, a.off};
        const int ops[2] = {a.item_opA[item], a.item_opB[item]};
        Basis<M, MIXED> V;
        IntegState gA, gB;
        const int n = a.n;
        const cd u0 = tid < n ? mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y) : mkcd(0.0, 0.0);
        const cd y0 = tid < n ? mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y) : mkcd(0.0, 0.0);
        __syncthreads();
        for (int s = 0; s < 2; ++s) {
            IntegState g;
            init_integ<M, MIXED>(S, br, g, s ? y0 : u0, rnd, klim, tol, src, s ? ops[1] : ops[0], M);
            if (s) gB = g; else gA = g;
        }
        const double scale = a.scale[item];
        for (int k = 0; k < a.nt; ++k) {
            const double t_out = a.tau[k];
            // one call site, so the unrolled Arnoldi code is inlined once
            for (int s = 0; s < 2; ++s) {
                IntegState g = s ? gB : gA;
                advance<M, MIXED>(S, br, g, V, t_out, max_steps, src, s ? ops[1] : ops[0], tol);
                if (s) gB = g; else gA = g;
            }
            const auto tl = block_sum1<1, kNT>(br, [&](int) {
                return fma(gB.w.re, gA.w.re, gB.w.im * gA.w.im);
            });
            if (tid == 0) a.L[(int64_t)item * a.nt + k] = scale * tl.get(0);
            if ((gA.fail | gB.fail) != 0) break;
        }
        if (tid == 0) {
            int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
            st[0] = gA.nstep; st[1] = gA.nmult; st[2] = gA.nrej; st[3] = gA.nexph;
            st[4] = gB.nstep; st[5] = gB.nmult; st[6] = gB.nrej; st[7] = gB.nexph;
            st[8] = gA.fail | (gB.fail << 4) | (gA.brk << 8) | (gB.brk << 9);
            st[9] = gA.nscale + gB.nscale;
            reinterpret_cast<double*>(st + 10)[0] = gA.anorm;
            reinterpret_cast<double*>(st + 12)[0] = gB.anorm;
            st[14] = 0; st[15] = 0;
#ifdef OOPS_PHASE_TIMERS
            st[16] = (int)((gA.tmv + gB.tmv) >> 10);
            st[17] = (int)((gA.tred + gB.tred) >> 10);
            st[18] = (int)((gA.texp + gB.texp) >> 10);
            st[19] = (int)((gA.tupd + gB.tupd) >> 10);
#endif
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// Dense path for n <= m (the reference's Krylov driver stops there):
// w(tau_k) = exp(-i H (tau_k - tau_{k-1})) w(tau_{k-1}) with the exponential
// from Pade(6), reusing it while the output spacing repeats.  One 64-thread
// workgroup per cluster, two integrators, same observable as above.
constexpr int kDenseNT = 64;

struct DenseSmem {
    cd H[2][kDenseNMAX * kDenseNMAX];   // dense generators of sweep A and B
    cd U[2][kDenseNMAX * kDenseNMAX];   // current propagators
    cd work[4][kDenseNMAX * kDenseNMAX];
    cd xs[kDenseNMAX];
    cd ws[2][kDenseNMAX];
    DevOpLocal locs[kMaxLocals];
    cd vals[kDenseMaxEnt];
    int32_t offs[kDenseMaxEnt];
    double red[80];
    double dprev[2];
    int ints[8];
    DevOpHeader hdr;
};

__global__ __launch_bounds__(kDenseNT) void cce_hahn_dense_kernel(KrylovArgs a) {
    __shared__ DenseSmem S;
    const int tid = threadIdx.x;
    const int n = a.n;
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int ops[2] = {a.item_opA[item], a.item_opB[item]};
        const int64_t voff = a.vec_off[item];
        int fail = 0;
        // dense -i*H for both sweeps: column c = -i H e_c
        for (int s = 0; s < 2; ++s) {
            op_load<kDenseNT>(a.hdr, a.loc, a.val, a.off, ops[s], S.hdr, S.locs, S.vals, S.offs);
            __syncthreads();
            const DevOpHeader h = S.hdr;
            for (int c = 0; c < n; ++c) {
                if (tid < n) S.xs[tid] = mkcd(tid == c ? 1.0 : 0.0, 0.0);
                __syncthreads();
                if (tid < n) S.H[s][tid + n * c] = cmulmi(op_apply_elem(tid, S.xs, h, S.locs, S.vals, S.offs));
                __syncthreads();
            }
            if (tid < n) S.ws[s][tid] = s == 0 ? mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y)
                                                : mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y);
            if (tid == 0) S.dprev[s] = -1.0;
            __syncthreads();
        }
        const double scale = a.scale[item];
        double tprev = 0.0;
        for (int k = 0; k < a.nt; ++k) {
            const double d = a.tau[k] - tprev;
            tprev = a.tau[k];
            if (d > 0.0) {
                for (int s = 0; s < 2; ++s) {
                    if (d != S.dprev[s]) {
                        int ns = 0;
                        const int which = lds_pade6<kDenseNT>(S.H[s], n, n, d, S.work[0], S.work[1], S.work[2],
                                                              S.work[3], S.red, S.ints, ns);
                        if (which < 0) { fail = 2; break; }
                        for (int idx = tid; idx < n * n; idx += kDenseNT) S.U[s][idx] = S.work[which][idx];
                        if (tid == 0) S.dprev[s] = d;
                        __syncthreads();
                    }
                    cd acc = mkcd(0.0, 0.0);
                    if (tid < n)
                        for (int c = 0; c < n; ++c) acc = cfma(S.U[s][tid + n * c], S.ws[s][c], acc);
                    __syncthreads();
                    if (tid < n) S.ws[s][tid] = acc;
                    __syncthreads();
                }
            }
            if (fail) break;
            const cd ya = tid < n ? S.ws[1][tid] : mkcd(0.0, 0.0);
            const cd ua = tid < n ? S.ws[0][tid] : mkcd(0.0, 0.0);
            double r[2] = {fma(ya.re, ua.re, ya.im * ua.im), fma(ya.re, ua.im, -ya.im * ua.re)};
            block_sum<2, kDenseNT>(r, S.red, S.red + 64);
            if (tid == 0) a.L[(int64_t)item * a.nt + k] = scale * S.red[64];
            __syncthreads();
        }
        if (tid == 0) {
            int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
            for (int q = 0; q < kStatsPerItem; ++q) st[q] = 0;
            st[8] = fail;
        }
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
__global__ void zero_counter_kernel(int32_t* p) { p[threadIdx.x] = 0; }

int launch_hahn_bucket(const KrylovArgs& a, int m, int precision, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    if (a.n <= m) {
        if (a.n > kDenseNMAX) return -10;
        hipLaunchKernelGGL(cce_hahn_dense_kernel, dim3(grid), dim3(kDenseNT), 0, stream, a);
        return (int)hipGetLastError();
    }
    if (a.n > 256) return -11;
    const bool mixed = precision == 1;
    if (m == 20 && !mixed) hipLaunchKernelGGL((cce_hahn_krylov_kernel<20, false>), dim3(grid), dim3(kNT), 0, stream, a);
    else if (m == 30 && !mixed) hipLaunchKernelGGL((cce_hahn_krylov_kernel<30, false>), dim3(grid), dim3(kNT), 0, stream, a);
    else if (m == 20 && mixed) hipLaunchKernelGGL((cce_hahn_krylov_kernel<20, true>), dim3(grid), dim3(kNT), 0, stream, a);
    else if (m == 30 && mixed) hipLaunchKernelGGL((cce_hahn_krylov_kernel<30, true>), dim3(grid), dim3(kNT), 0, stream, a);
    else return -12;
    return (int)hipGetLastError();
}

int krylov_entry_capacity(int m) { return (m == 20 || m == 30) ? kKrylovMaxEnt : 0; }

}  // namespace oops
