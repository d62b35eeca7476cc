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

#include "krylov_common.h"

namespace oops {

template <int M>
struct KrylovLds {
    static constexpr int MH = M + 2;
    static constexpr int MAT = MH * MH;
    static constexpr int NMAX = 256;
    static constexpr int EMAX = kKrylovMaxEnt;  // ELL entries per operator
};

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
                t_step = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
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
        g.t_new = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        g.t_now += t_step;
        OOPS_TOC(g, tupd, tu0);
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
            __syncthreads();
            op_load_lds<kNT, MIXED>(src, s ? ops[1] : ops[0], S.u.a.op);
            __syncthreads();
            const NormWork nw{S.u.a.xs, S.u.a.xs32, S.tdiag, S.toff, S.scal};
            init_integ<MIXED>(nw, S.u.a.op, br, g, s ? y0 : u0, rnd, klim, tol, M);
            if (s) gB = g; else gA = g;
        }
        const double scale = a.scale[item];
        const ItemSeq q = item_seq(a, item);
        for (int k = 0; k < q.nt; ++k) {
            const double t_out = q.tau[k];
            // one call site, so the unrolled Arnoldi code is inlined once
            for (int s = 0; s < 2; ++s) {
                IntegState g = s ? gB : gA;
                advance<M, MIXED>(S, br, g, V, t_out, max_steps, src, s ? ops[1] : ops[0], tol);
                if (s) gB = g; else gA = g;
            }
            const auto tl = block_sum1<1, kNT>(br, [&](int) {
                return fma(gB.w.re, gA.w.re, gB.w.im * gA.w.im);
            });
            if (tid == 0) q.L[k] = scale * tl.get(0);
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
// Dense path for n <= m (the reference's Krylov driver stops there).
// Hahn echo: w(tau_k) = exp(-i H (tau_k - tau_{k-1})) w(tau_{k-1}) for both
// sweeps with the exponential from Pade(6), reused while the output spacing
// repeats.  CPMG-N: for every output time the whole sequence
// U_N .. U_0 u0 (QuantumEvolutionAlgorithm.cpp:60-83), L = scale Re<y0, s>.
// One 64-thread workgroup per cluster.
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
        const ItemSeq q = item_seq(a, item);
        const bool cpmg = q.npulse > 1 || q.two_chain;
        // Hahn echo: sweeps under lvA and -lvB^dagger; CPMG-N and the two-chain
        // form: the ket-side generators opA (even segments) and opC (odd)
        const int ops[2] = {a.item_opA[item], cpmg ? a.item_opC[item] : a.item_opB[item]};
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
        for (int k = 0; k < q.nt; ++k) {
            if (cpmg) {
                // s = U_N .. U_0 u0 with U_j = exp(-i L_j f_j t): exponentials of the
                // two generators over t/(2N), applied once (end segments) or twice
                const double t = q.tau[k];
                if (tid < n) S.ws[0][tid] = mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y);
                __syncthreads();
                if (t > 0.0) {
                    for (int g2 = 0; g2 < 2 && !fail; ++g2) {
                        int ns = 0;
                        const int which = lds_pade6<kDenseNT>(S.H[g2], n, n, t * (0.5 / q.npulse), S.work[0],
                                                              S.work[1], S.work[2], S.work[3], S.red, S.ints, ns);
                        if (which < 0) { fail = 2; break; }
                        for (int idx = tid; idx < n * n; idx += kDenseNT) S.U[g2][idx] = S.work[which][idx];
                        __syncthreads();
                    }
                    if (fail) break;
                    for (int j = 0; j <= q.npulse; ++j) {
                        const int reps = (j == 0 || j == q.npulse) ? 1 : 2;
                        for (int rep = 0; rep < reps; ++rep) {
                            cd acc = mkcd(0.0, 0.0);
                            if (tid < n)
                                for (int c = 0; c < n; ++c) acc = cfma(S.U[j & 1][tid + n * c], S.ws[0][c], acc);
                            __syncthreads();
                            if (tid < n) S.ws[0][tid] = acc;
                            __syncthreads();
                        }
                    }
                    if (q.two_chain) {
                        // <y0, bra-chain ket> with the bra chain's generators
                        // opB = opC (even) and opD = opA (odd): apply the
                        // adjoints of its segments N..0 to the ket-side state
                        for (int j = q.npulse; j >= 0; --j) {
                            const int reps = (j == 0 || j == q.npulse) ? 1 : 2;
                            const int u = (j & 1) ^ 1;
                            for (int rep = 0; rep < reps; ++rep) {
                                cd acc = mkcd(0.0, 0.0);
                                if (tid < n)
                                    for (int c = 0; c < n; ++c) acc = cjfma(S.U[u][c + n * tid], S.ws[0][c], acc);
                                __syncthreads();
                                if (tid < n) S.ws[0][tid] = acc;
                                __syncthreads();
                            }
                        }
                    }
                }
                const cd ya = tid < n ? mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y) : mkcd(0.0, 0.0);
                const cd ua = tid < n ? S.ws[0][tid] : mkcd(0.0, 0.0);
                double r[2] = {fma(ya.re, ua.re, ya.im * ua.im), fma(ya.re, ua.im, -ya.im * ua.re)};
                block_sum<2, kDenseNT>(r, S.red, S.red + 64);
                if (tid == 0) q.L[k] = scale * S.red[64];
                __syncthreads();
                continue;
            }
            const double d = q.tau[k] - tprev;
            tprev = q.tau[k];
            if (d > 0.0) {
                for (int s = 0; s < 2; ++s) {
                    // reuse the propagator while the spacing repeats: grids such as
                    // linspace differ from k*dt in the last bits, which changes the
                    // step by ~1e-16 relative, far below any tolerance here
                    if (fabs(d - S.dprev[s]) > 1e-13 * d) {
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
            if (tid == 0) q.L[k] = scale * S.red[64];
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

void launch_zero_counter(int32_t* counter, hipStream_t stream) {
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
}

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
