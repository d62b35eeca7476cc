// Batched, size-bucketed Krylov expv for CCE clusters on gfx950.
//
// One workgroup (NT = 256 threads, 4 waves) owns one cluster at a time and
// pulls clusters from a device work counter until the bucket is drained.
// Thread i owns element i of the Liouville state (n <= 256), so the whole
// Krylov basis V[0..M] of that element lives in registers; the matvec input
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
    // region 2 union: Pade workspace (4 matrices) | operator copy + x + reduction
    static constexpr int PADE_BYTES = 4 * MAT * 16;
    static constexpr int XBYTES = NMAX * 16;
    static constexpr int RED_BYTES = (kNT / kWave) * 64 * 8 + 64 * 8;
    static constexpr int LOC_BYTES = kMaxLocals * 16;
    static constexpr int ENT_AVAIL = PADE_BYTES - XBYTES - RED_BYTES - LOC_BYTES - 64;
    static constexpr int EMAX = ENT_AVAIL / 20;  // 16 B value + 4 B offset
    static constexpr int R2 = PADE_BYTES;
};

struct IntegState {
    cd w;
    double beta, t_now, t_new, anorm, rndoff, xm;
    int nstep, nmult, nrej, nexph, nscale, brk, fail;
};

template <int M>
struct KrylovSmem {
    cd hess[KrylovLds<M>::MAT];
    alignas(16) unsigned char r2[KrylovLds<M>::R2];
    double scal[16];
    double tdiag[16];
    double toff[16];
    int ints[8];
    DevOpHeader hdr;
};

// Pointers into the region-2 union while the operator is resident.
template <int M>
struct OpView {
    DevOpLocal* locs;
    cd* vals;
    int32_t* offs;
    cd* xs;
    double* red_ws;
    double* red_out;
    __device__ explicit OpView(unsigned char* r2) {
        unsigned char* p = r2;
        xs = reinterpret_cast<cd*>(p);
        p += KrylovLds<M>::XBYTES;
        red_ws = reinterpret_cast<double*>(p);
        p += (kNT / kWave) * 64 * 8;
        red_out = reinterpret_cast<double*>(p);
        p += 64 * 8;
        locs = reinterpret_cast<DevOpLocal*>(p);
        p += KrylovLds<M>::LOC_BYTES;
        vals = reinterpret_cast<cd*>(p);
        p += (size_t)KrylovLds<M>::EMAX * 16;
        offs = reinterpret_cast<int32_t*>(p);
    }
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

// y_i = H x (x staged in LDS by the caller), zero for i >= n.
template <int M>
__device__ __forceinline__ cd matvec_elem(const OpView<M>& v, const DevOpHeader& h, int i) {
    if (i >= h.n) return mkcd(0.0, 0.0);
    return op_apply_elem(i, v.xs, h, v.locs, v.vals, v.offs);
}

// ----------------------------------------------------------------------------
// Norm estimate: k Lanczos steps on +H from the DLARAN start vector, then
// power iteration on the k x k tridiagonal (reference :210-265).
template <int M>
__device__ double lanczos_anorm(KrylovSmem<M>& S, const OpView<M>& v, const DevOpHeader& h, const double* rnd,
                                int klim, double tol) {
    const int tid = threadIdx.x;
    const int n = h.n;
    int k = klim;
    if (k > n) k = n;
    double* tdiag = S.tdiag;
    double* toff = S.toff;
    for (int q2 = tid; q2 < 16; q2 += kNT) { tdiag[q2] = 0.0; toff[q2] = 0.0; }
    __syncthreads();
    cd q = tid < n ? mkcd(rnd[2 * tid], rnd[2 * tid + 1]) : mkcd(0.0, 0.0);
    cd qm = mkcd(0.0, 0.0);
    {
        double r[1] = {cabs2(q)};
        block_sum<1, kNT>(r, v.red_ws, v.red_out);
        const double bj0 = sqrt(v.red_out[0]);
        q = cscale(q, 1.0 / bj0);
    }
    double bjm = 0.0, bj = 0.0;
    int kk = k;
    for (int j = 1; j <= k; ++j) {
        if (tid < n) v.xs[tid] = q;
        __syncthreads();
        cd w = matvec_elem<M>(v, h, tid);
        double r[2] = {fma(q.re, w.re, q.im * w.im), fma(q.re, w.im, -q.im * w.re)};
        block_sum<2, kNT>(r, v.red_ws, v.red_out);
        const cd aj = mkcd(v.red_out[0], v.red_out[1]);
        w = cfms(aj, q, w);
        w = mkcd(w.re - bjm * qm.re, w.im - bjm * qm.im);
        double r2[1] = {cabs2(w)};
        block_sum<1, kNT>(r2, v.red_ws, v.red_out);
        bj = sqrt(v.red_out[0]);
        if (tid == 0) { tdiag[j - 1] = aj.re; toff[j - 1] = 0.0; }
        if (j == k) break;
        if (tid == 0) toff[j - 1] = bj;
        if (bj < tol) { kk = j; break; }
        qm = q;
        q = cscale(w, 1.0 / bj);
        bjm = bj;
    }
    (void)kk;
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
// One Arnoldi step J (compile-time, so the basis stays in registers and the
// reduction width is fixed): p = -iH V[J], two classical Gram-Schmidt passes,
// column J of the Hessenberg matrix, V[J+1] = p / h(J+1,J).
template <int M>
struct ArnCtx {
    KrylovSmem<M>* S;
    const OpView<M>* v;
    const DevOpHeader* h;
    IntegState* g;
    double t_out;
    double t_step;
    int pad;
    int k1;
    int mb;
    bool brk;
};

template <int M, int J>
__device__ __forceinline__ void arnoldi_step(ArnCtx<M>& c, cd (&V)[M + 1]) {
    if (c.brk) return;
    constexpr int MH = M + 2;
    constexpr int NV = 2 * (J + 1) + 1;
    constexpr int P = pow2ceil(NV);
    const int tid = threadIdx.x;
    const OpView<M>& v = *c.v;
    const int n = c.h->n;
    if (tid < n) v.xs[tid] = V[J];
    __syncthreads();
    cd p = cmulmi(matvec_elem<M>(v, *c.h, tid));
    ++c.g->nmult;
    // pass 1: h = V^H p
    block_sum_gen<P, kNT>(
        [&](int k) -> double {
            if (k >= 2 * (J + 1)) return 0.0;
            const cd& b = V[k >> 1];
            return (k & 1) ? fma(b.re, p.im, -b.im * p.re) : fma(b.re, p.re, b.im * p.im);
        },
        v.red_ws, v.red_out);
    if (tid == 0) {
#pragma unroll
        for (int l = 0; l <= J; ++l) c.S->hess[l + MH * J] = mkcd(v.red_out[2 * l], v.red_out[2 * l + 1]);
    }
#pragma unroll
    for (int l = 0; l <= J; ++l) p = cfms(mkcd(v.red_out[2 * l], v.red_out[2 * l + 1]), V[l], p);
    // pass 2 (re-orthogonalisation) and |p|^2 after pass 1
    block_sum_gen<P, kNT>(
        [&](int k) -> double {
            if (k < 2 * (J + 1)) {
                const cd& b = V[k >> 1];
                return (k & 1) ? fma(b.re, p.im, -b.im * p.re) : fma(b.re, p.re, b.im * p.im);
            }
            if (k == 2 * (J + 1)) return tid < n ? cabs2(p) : 0.0;
            return 0.0;
        },
        v.red_ws, v.red_out);
    double corr = 0.0;
#pragma unroll
    for (int l = 0; l <= J; ++l) {
        const cd h2 = mkcd(v.red_out[2 * l], v.red_out[2 * l + 1]);
        p = cfms(h2, V[l], p);
        corr += cabs2(h2);
    }
    if (tid == 0) {
#pragma unroll
        for (int l = 0; l <= J; ++l)
            c.S->hess[l + MH * J] = cadd(c.S->hess[l + MH * J], mkcd(v.red_out[2 * l], v.red_out[2 * l + 1]));
    }
    const double hj1j = sqrt(fmax(v.red_out[2 * (J + 1)] - corr, 0.0));
    if (hj1j <= 1.0e-7) {
        c.k1 = 0;
        c.g->brk = 1;
        c.mb = J + 1;
        c.t_step = c.t_out - c.g->t_now;
        c.brk = true;
    } else {
        if (tid == 0) c.S->hess[(J + 1) + MH * J] = mkcd(hj1j, 0.0);
        V[J + 1] = cscale(p, 1.0 / hj1j);
    }
    __syncthreads();
}

template <int M, int... Js>
__device__ __forceinline__ void arnoldi_all(ArnCtx<M>& c, cd (&V)[M + 1], std::integer_sequence<int, Js...>) {
    (arnoldi_step<M, Js>(c, V), ...);
}

// ----------------------------------------------------------------------------
// One Arnoldi sweep + Pade + step control; advances g to t_out.
template <int M>
__device__ __forceinline__ void advance(KrylovSmem<M>& S, IntegState& g, cd (&V)[M + 1], double t_out, const KrylovArgs& a, int op,
                        double tol) {
    constexpr int MH = M + 2;
    const int tid = threadIdx.x;
    const double delta = 1.2, gamma = 0.9;
    while (g.t_now < t_out) {
        if (g.nstep >= a.max_steps) { g.fail |= 1; break; }
        ++g.nstep;
        int k1 = 2, mb = M;
        double t_step = fmin(t_out - g.t_now, g.t_new);
        OpView<M> v(S.r2);
        op_load<kNT>(a.hdr, a.loc, a.val, a.off, op, S.hdr, v.locs, v.vals, v.offs);
        for (int idx = tid; idx < MH * MH; idx += kNT) S.hess[idx] = mkcd(0.0, 0.0);
        __syncthreads();
        const DevOpHeader h = S.hdr;
        const int n = h.n;
        V[0] = cscale(g.w, 1.0 / g.beta);
        ArnCtx<M> c{&S, &v, &h, &g, t_out, 0.0, 0, 2, M, false};
        c.t_step = t_step;
        arnoldi_all<M>(c, V, std::make_integer_sequence<int, M>{});
        bool brk = c.brk;
        k1 = c.k1;
        mb = c.mb;
        t_step = c.t_step;
        double avnorm = 0.0;
        if (!brk) {
            if (tid < n) v.xs[tid] = V[M];
            __syncthreads();
            const cd p = cmulmi(matvec_elem<M>(v, h, tid));
            ++g.nmult;
            double r[1] = {cabs2(p)};
            block_sum<1, kNT>(r, v.red_ws, v.red_out);
            avnorm = sqrt(v.red_out[0]);
        }
        if (tid == 0) S.hess[(M + 1) + MH * M] = mkcd(1.0, 0.0);
        __syncthreads();
        // ---- Pade + error estimate + rejection (region 2 now holds the Pade workspace)
        cd* pw = reinterpret_cast<cd*>(S.r2);
        constexpr int MAT = KrylovLds<M>::MAT;
        double err_loc = 0.0;
        cd* F = nullptr;
        int mx = 0;
        for (;;) {
            ++g.nexph;
            mx = mb + k1;
            int ns = 0;
            const int which = lds_pade6<kNT>(S.hess, MH, mx, t_step, pw, pw + MAT, pw + 2 * MAT, pw + 3 * MAT,
                                             S.scal + 8, S.ints, ns);
            if (which < 0) { g.fail |= 2; return; }
            F = pw + which * MAT;
            g.nscale += ns;
            if (k1 == 0) {
                err_loc = tol;
            } else {
                const double p1 = cabs(F[M]) * g.beta;
                const double p2 = cabs(F[M + 1]) * g.beta * avnorm;
                if (p1 > 10.0 * p2) { err_loc = p2; g.xm = 1.0 / (double)M; }
                else if (p1 > p2) { err_loc = (p1 * p2) / (p1 - p2); g.xm = 1.0 / (double)M; }
                else { err_loc = p1; g.xm = 1.0 / (double)(M - 1); }
            }
            if (k1 != 0 && err_loc > delta * t_step * tol) {
                t_step = gamma * t_step * pow(t_step * tol / err_loc, g.xm);
                t_step = round_step2(t_step);
                ++g.nrej;
                __syncthreads();
                if (!(t_step < 1e300) || g.nrej > a.max_steps) { g.fail |= 4; return; }
                continue;
            }
            break;
        }
        // ---- w = beta * V(:, 0:mx) * F(0:mx, 0)
        mx = mb + (k1 - 1 > 0 ? k1 - 1 : 0);
        cd acc = mkcd(0.0, 0.0);
#pragma unroll
        for (int j = 0; j <= M; ++j)
            if (j < mx) acc = cfma(V[j], F[j], acc);
        g.w = cscale(acc, g.beta);
        __syncthreads();  // everyone has read F before region 2 is reused
        {
            OpView<M> v2(S.r2);
            double r[1] = {tid < n ? cabs2(g.w) : 0.0};
            block_sum<1, kNT>(r, v2.red_ws, v2.red_out);
            g.beta = sqrt(v2.red_out[0]);
        }
        g.t_new = gamma * t_step * pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        g.t_now += t_step;
        __syncthreads();
    }
}

template <int M>
__device__ void init_integ(KrylovSmem<M>& S, IntegState& g, cd w0, const KrylovArgs& a, int op, int m) {
    OpView<M> v(S.r2);
    op_load<kNT>(a.hdr, a.loc, a.val, a.off, op, S.hdr, v.locs, v.vals, v.offs);
    __syncthreads();
    const DevOpHeader h = S.hdr;
    const int tid = threadIdx.x;
    g.w = tid < h.n ? w0 : mkcd(0.0, 0.0);
    g.anorm = lanczos_anorm<M>(S, v, h, a.rnd, a.klim, a.tol);
    double r[1] = {cabs2(g.w)};
    block_sum<1, kNT>(r, v.red_ws, v.red_out);
    g.beta = sqrt(v.red_out[0]);
    g.t_now = 0.0;
    g.xm = 1.0 / (double)m;
    g.rndoff = 2.220446049250313e-16 * g.anorm;
    const double p2 = a.tol * pow((m + 1) / 2.72, (double)(m + 1)) * sqrt(2.0 * 3.14 * (m + 1));
    g.t_new = (1.0 / g.anorm) * pow(p2 / (4.0 * g.beta * g.anorm), g.xm);
    g.t_new = round_step2(g.t_new);
    g.nstep = g.nmult = g.nrej = g.nexph = g.nscale = g.brk = g.fail = 0;
    g.nmult = 0;
    __syncthreads();
}

template <int M>
__global__ __launch_bounds__(kNT) void cce_hahn_krylov_kernel(KrylovArgs a) {
    __shared__ KrylovSmem<M> S;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int opA = a.item_opA[item];
        const int opB = a.item_opB[item];
        const int64_t voff = a.vec_off[item];
        cd V[M + 1];
        IntegState gA, gB;
        const int n = a.n;
        const cd u0 = tid < n ? mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y) : mkcd(0.0, 0.0);
        const cd y0 = tid < n ? mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y) : mkcd(0.0, 0.0);
        for (int s = 0; s < 2; ++s) {
            IntegState g;
            init_integ<M>(S, g, s ? y0 : u0, a, s ? opB : opA, M);
            if (s) gB = g; else gA = g;
        }
        const double scale = a.scale[item];
        for (int k = 0; k < a.nt; ++k) {
            const double t_out = a.tau[k];
            // one call site, so the unrolled Arnoldi code is inlined once
            for (int s = 0; s < 2; ++s) {
                IntegState g = s ? gB : gA;
                advance<M>(S, g, V, t_out, a, s ? opB : opA, a.tol);
                if (s) gB = g; else gA = g;
            }
            OpView<M> v(S.r2);
            double r[2] = {fma(gB.w.re, gA.w.re, gB.w.im * gA.w.im), fma(gB.w.re, gA.w.im, -gB.w.im * gA.w.re)};
            block_sum<2, kNT>(r, v.red_ws, v.red_out);
            if (tid == 0) a.L[(int64_t)item * a.nt + k] = scale * v.red_out[0];
            __syncthreads();
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

int launch_hahn_bucket(const KrylovArgs& a, int m, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, a.queue);
    if (a.n <= m) {
        if (a.n > kDenseNMAX) return -10;
        hipLaunchKernelGGL(cce_hahn_dense_kernel, dim3(grid), dim3(kDenseNT), 0, stream, a);
        return (int)hipGetLastError();
    }
    if (a.n > 256) return -11;
    switch (m) {
        case 20: hipLaunchKernelGGL(cce_hahn_krylov_kernel<20>, dim3(grid), dim3(kNT), 0, stream, a); break;
        case 30: hipLaunchKernelGGL(cce_hahn_krylov_kernel<30>, dim3(grid), dim3(kNT), 0, stream, a); break;
        default: return -12;
    }
    return (int)hipGetLastError();
}

int krylov_entry_capacity(int m) {
    switch (m) {
        case 20: return KrylovLds<20>::EMAX;
        case 30: return KrylovLds<30>::EMAX;
        default: return 0;
    }
}

}  // namespace oops
