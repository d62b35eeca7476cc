// Mixed-precision batched Hahn-echo kernel for Hermitian generators:
// Lanczos with FP64 recurrence coefficients and re-orthogonalisation against
// v_j, complex-FP32 couplings and FP32 basis storage, and the Hessenberg
// (here: real symmetric tridiagonal) exponential in FP64 by a Chebyshev
// series instead of Pade(6).
//
// Same work decomposition as cce_krylov.hip (one cluster per 256-thread
// workgroup, thread i owns element i, both sweeps interleaved per output time,
// L_c(t) reduced on the device) and the same step control as the reference's
// Expokit driver (src/expv/source/zgexpv_cache.f:418-661): Krylov dimension m,
// one extra product for the error estimate, the corrected (m+1)-term update,
// the local error estimate from the two augmented components, rejection with
// gamma = 0.9 / delta = 1.2 and 2-significant-digit step rounding.  What
// differs from the FP64 path is how the small problem is solved:
//   * for Hermitian A the Arnoldi matrix of -iA is -iT with T real symmetric
//     tridiagonal, so one step needs one block reduction of <v_j, p> and
//     |p|^2 (the complex projection on v_j is removed in FP64, beta_j^2 =
//     |p|^2 - |<v_j, p>|^2) instead of two of 2(j+1)+1 values;
//   * the small exponential F = exp(tau H_aug) e1 of the (m+2) augmented
//     matrix (zgexpv_cache.f:401; H = -iT, one coupling row -i beta_m e_m^T
//     and the trailing 1) is a Chebyshev series with Bessel coefficients
//     (Tal-Ezer & Kosloff, J. Chem. Phys. 81 (1984) 3967):
//       exp(i tau X) e1 = e^{i tau c} sum_k (2 - [k=0]) i^k J_k(tau r) T_k(S) e1,
//     X = -i H_aug, S = (X - c)/r with [c - r, c + r] a Gershgorin interval
//     of X's real spectrum.  Each term is one product with the tridiagonal
//     augmented matrix, done by one wave with lane j owning row j, so a trial
//     step costs O(K m) with K ~ tau r + 30 instead of a dense O(m^3) Pade.
// The host uses this kernel for mixed precision when every generator of the
// bucket is Hermitian (no dephasing); otherwise the mixed Arnoldi kernel runs.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "krylov_common.h"

namespace oops {

constexpr int kBesMax = 128;  // Bessel coefficients per substep

template <int M>
struct LanczosSmem {
    OpLds op[4];                        // lvA, -lvB^dagger, lvB, -lvA^dagger stay resident
    cd xs[256];
    float2 xs32[256];
    double red[3 * (kNT / kWave) * 64];
    double d[M + 2];                    // diagonal of T
    double e[M + 2];                    // off-diagonal of T (e[j] couples j and j+1)
    cd F[M + 2];                        // Krylov coefficients of the step
    double bes[kBesMax];                // Bessel coefficients J_k(tau r / ns)
    double tdiag[16];
    double toff[16];
    double scal[16];
    int ints[8];
};

template <int M>
struct LzCtx {
    LanczosSmem<M>* S;
    BlockRed<kNT>* br;
    const OpLds* o;
    IntegState* g;
    double t_out;
    double t_step;
    int k1;
    int mb;
    bool brk;
    cd vcur;
    cd vprev;
    double bprev;
};

// Lanczos step J: p = A v_J - beta_{J-1} v_{J-1}; h = <v_J, p> (alpha_J =
// Re h); p -= h v_J; beta_J = |p|; v_{J+1} = p / beta_J (FP32 in the basis).
template <int M, int J>
__device__ __forceinline__ void lanczos_step(LzCtx<M>& c, Basis<M, true>& V) {
    if (c.brk) return;
    const int tid = threadIdx.x;
    const int n = c.o->hdr.n;
    OOPS_TIC(tm0);
    stage_x<true>(c.S->xs, c.S->xs32, tid, n, c.vcur);
    __syncthreads();
    cd p = matvec_elem<true>(*c.o, c.S->xs, c.S->xs32, tid);
    ++c.g->nmult;
    p = mkcd(p.re - c.bprev * c.vprev.re, p.im - c.bprev * c.vprev.im);
    OOPS_TOC(*c.g, tmv, tm0);
    OOPS_TIC(tr0);
    const cd vj = c.vcur;
    // one reduction: <v_J, p> and |p|^2; beta_J^2 = |p|^2 - |<v_J, p>|^2 after
    // removing the complex projection on v_J (FP64)
    const auto r1 = block_sum1<4, kNT>(*c.br, [&](int k) -> double {
        if (k == 0) return fma(vj.re, p.re, vj.im * p.im);
        if (k == 1) return fma(vj.re, p.im, -vj.im * p.re);
        if (k == 2) return cabs2(p);
        return 0.0;
    });
    const cd h = mkcd(r1.get(0), r1.get(1));
    p = cfms(h, vj, p);
    const double alpha = h.re;
    const double beta = sqrt(fmax(r1.get(2) - cabs2(h), 0.0));
    OOPS_TOC(*c.g, tred, tr0);
    if (tid == 0) {
        c.S->d[J] = alpha;
        c.S->e[J] = beta;
    }
    if (beta <= 1.0e-7) {
        c.k1 = 0;
        c.g->brk = 1;
        c.mb = J + 1;
        c.t_step = c.t_out - c.g->t_now;
        c.brk = true;
    } else {
        c.vprev = vj;
        c.vcur = cscale(p, 1.0 / beta);
        V.set(J + 1, c.vcur);
        c.bprev = beta;
    }
}

template <int M, int... Js>
__device__ __forceinline__ void lanczos_all(LzCtx<M>& c, Basis<M, true>& V, std::integer_sequence<int, Js...>) {
    (lanczos_step<M, Js>(c, V), ...);
}

// Coefficients of one trial step tau: F = exp(tau H_aug) e1 for the
// augmented matrix of the Lanczos tridiagonal T (mb x mb) plus, with the
// correction (k1 = 2), rows M (coupling -i beta_m to row M-1) and M+1 (1 from
// row M).  Wave 0 runs the Chebyshev-Bessel series on X = -i H_aug:
//   (X x)_j = -(d_j x_j + e_{j-1} x_{j-1} + e_j x_{j+1})   j < mb
//   (X x)_M = -beta_m x_{M-1},   (X x)_{M+1} = -i x_M.
// Long steps are split into ns equal substeps of tau r / ns <= 40.
// Returns the number of series terms summed.
template <int M>
__device__ int cheb_coeffs(LanczosSmem<M>& S, int mb, int k1, double tau, double beta_m) {
    const int tid = threadIdx.x;
    int terms = 0;
    if (tid < kWave) {
        const int j = tid;
        const int rows = k1 ? M + 2 : mb;
        const double dj = j < mb ? S.d[j] : 0.0;
        const double ej = j + 1 < mb ? S.e[j] : 0.0;                 // coupling to row j+1 inside T
        const double ejm = (j >= 1 && j < mb) ? S.e[j - 1] : 0.0;    // coupling to row j-1 inside T
        // Gershgorin interval of T, then of X = -T (+ 0 for the augmentation)
        double lo = j < mb ? dj - fabs(ejm) - fabs(ej) : 1e300;
        double hi = j < mb ? dj + fabs(ejm) + fabs(ej) : -1e300;
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1) {
            lo = fmin(lo, __shfl_xor(lo, mask, kWave));
            hi = fmax(hi, __shfl_xor(hi, mask, kWave));
        }
        double xlo = -hi, xhi = -lo;
        if (k1) { xlo = fmin(xlo, 0.0); xhi = fmax(xhi, 0.0); }
        const double c = 0.5 * (xlo + xhi);
        const double r = fmax(0.5 * (xhi - xlo), 1e-300);
        const double ir = 1.0 / r;
        // rows of S = (X - c) / r owned by this lane
        const double sdiag = j < mb ? (-dj - c) * ir : (j < rows ? -c * ir : 0.0);
        const double sup = -ej * ir;   // times x_{j+1}
        const double slo = j < mb ? -ejm * ir : 0.0;
        const double sM = (k1 && j == M) ? -beta_m * ir : 0.0;   // row M: times x_{M-1}
        const bool rowM1 = k1 && j == M + 1;                      // row M+1: -i x_M / r
        const double xfull = tau * r;
        const int ns = xfull > 40.0 ? (int)ceil(xfull / 40.0) : 1;
        const double x = xfull / ns;
        // Bessel J_k(x), k < K, by backward recurrence (Miller), normalised
        // with J_0 + 2 sum J_2k = 1; every lane computes the same values.
        int K = (int)ceil(x + 10.0 * cbrt(fmax(x, 1.0)) + 20.0);
        if (K > kBesMax - 2) K = kBesMax - 2;
        if (x == 0.0) {
            if (j == 0) S.bes[0] = 1.0;
            for (int k = 1 + j; k < K; k += kWave) S.bes[k] = 0.0;
        } else {
            const int N = K + 12;
            double jp = 0.0, jc = 1e-280, norm = 0.0;
            const double tx = 2.0 / x;
            for (int k = N; k >= 1; --k) {
                const double jm = (double)k * tx * jc - jp;   // J_{k-1}
                jp = jc;
                jc = jm;
                if (k - 1 < K && j == 0) S.bes[k - 1] = jm;
                if (((k - 1) & 1) == 0) norm += (k - 1 == 0 ? 1.0 : 2.0) * jm;
                if (fabs(jc) > 1e250) {   // rescale the stored tail
                    for (int q = k - 1; q < K && j == 0; ++q) S.bes[q] *= 1e-250;
                    jp *= 1e-250; jc *= 1e-250; norm *= 1e-250;
                }
            }
            const double inorm = 1.0 / norm;
            // wave-synchronous: lane 0 wrote bes; every lane scales its share
            for (int k = j; k < K; k += kWave) S.bes[k] *= inorm;
        }
        // series on the vector, ns times
        cd v = mkcd(j == 0 ? 1.0 : 0.0, 0.0);
        double sn, cs;
        sincos(c * (tau / ns), &sn, &cs);   // e^{i c tau / ns}
        const cd ph = mkcd(cs, sn);
        for (int sub = 0; sub < ns; ++sub) {
            auto apply = [&](cd t) -> cd {
                const double pr = __shfl_up(t.re, 1, kWave), pi = __shfl_up(t.im, 1, kWave);
                const double nr = __shfl_down(t.re, 1, kWave), ni = __shfl_down(t.im, 1, kWave);
                cd y = mkcd(sdiag * t.re + slo * pr + sup * nr, sdiag * t.im + slo * pi + sup * ni);
                if (sM != 0.0) y = mkcd(y.re + sM * pr, y.im + sM * pi);
                if (rowM1) y = mkcd(y.re + pi * ir, y.im - pr * ir);   // -i x_M / r
                return y;
            };
            cd tm = v;            // T_0(S) v
            cd tk = apply(v);     // T_1(S) v
            // acc = J0 T0 + 2 i J1 T1 + sum_k 2 i^k J_k T_k
            cd acc = cscale(tm, S.bes[0]);
            acc = cadd(acc, cscale(cmulmi(cmulmi(cmulmi(tk))), 2.0 * S.bes[1]));  // i = (-i)^3
            cd ik = mkcd(0.0, 1.0);   // i^k
            for (int k = 2; k < K; ++k) {
                const cd y = apply(tk);
                const cd tn = mkcd(2.0 * y.re - tm.re, 2.0 * y.im - tm.im);
                tm = tk;
                tk = tn;
                ik = mkcd(-ik.im, ik.re);   // times i
                acc = cadd(acc, cscale(cmul(ik, tk), 2.0 * S.bes[k]));
            }
            v = cmul(ph, acc);
            if (j >= rows) v = mkcd(0.0, 0.0);
        }
        if (j < rows) S.F[j] = v;
        terms = K * ns;
    }
    __syncthreads();
    return terms;
}

template <int M>
__device__ __forceinline__ void advance_lz(LanczosSmem<M>& S, BlockRed<kNT>& br, IntegState& g, Basis<M, true>& V,
                                           double t_out, int max_steps, const OpLds& o, double tol) {
    const int tid = threadIdx.x;
    const double delta = 1.2, gamma = 0.9;
    while (g.t_now < t_out) {
        if (g.nstep >= max_steps) { g.fail |= 1; break; }
        ++g.nstep;
        const cd v0 = cscale(g.w, 1.0 / g.beta);
        V.set(0, v0);
        LzCtx<M> c{&S, &br, &o, &g, t_out, fmin(t_out - g.t_now, g.t_new), 2, M, false, v0, mkcd(0.0, 0.0), 0.0};
        lanczos_all<M>(c, V, std::make_integer_sequence<int, M>{});
        const bool brk = c.brk;
        const int k1 = c.k1;
        const int mb = c.mb;
        double t_step = c.t_step;
        double avnorm = 0.0;
        if (!brk) {
            stage_x<true>(S.xs, S.xs32, tid, g.n, c.vcur);
            __syncthreads();
            const cd p = matvec_elem<true>(o, S.xs, S.xs32, tid);
            ++g.nmult;
            const auto tn = block_sum1<1, kNT>(br, [&](int) { return cabs2(p); });
            avnorm = sqrt(tn.get(0));
        }
        __syncthreads();
        OOPS_TIC(te0);
        double err_loc = 0.0;
        for (;;) {
            ++g.nexph;
            g.nscale += cheb_coeffs<M>(S, mb, k1, t_step, c.bprev);
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
                if (!(t_step < 1e300) || g.nrej > max_steps) { g.fail |= 4; return; }
                continue;
            }
            break;
        }
        OOPS_TOC(g, texp, te0);
        OOPS_TIC(tu0);
        // w = beta * (sum_{j<mb} F_j v_j + [k1] F_M v_M)
        const int mx = mb + (k1 - 1 > 0 ? k1 - 1 : 0);
        cd acc = mkcd(0.0, 0.0);
#pragma unroll
        for (int j = 0; j <= M; ++j)
            if (j < mx) acc = cfma(V.get(j), S.F[j], acc);
        g.w = cscale(acc, g.beta);
        {
            const auto tb = block_sum1<1, kNT>(br, [&](int) { return tid < g.n ? cabs2(g.w) : 0.0; });
            g.beta = sqrt(tb.get(0));
        }
        g.t_new = gamma * t_step * pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        // land exactly on t_out when the step was clamped to it (segment ends)
        if (t_step >= t_out - g.t_now) g.t_now = t_out;
        else g.t_now += t_step;
        OOPS_TOC(g, tupd, tu0);
    }
}

template <int M>
__global__ __launch_bounds__(kNT) void cce_hahn_lanczos_kernel(KrylovArgs a) {
    __shared__ LanczosSmem<M> S;
    const int tid = threadIdx.x;
    BlockRed<kNT> br(S.red);
    const double tol = a.tol;
    const int max_steps = a.max_steps;
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int64_t voff = a.vec_off[item];
        const ItemSeq q = item_seq(a, item);
        const bool cpmg = q.npulse > 1;
        const OpSrc src{a.hdr, a.loc, a.val, a.off};
        op_load_lds<kNT, true>(src, a.item_opA[item], S.op[0]);
        op_load_lds<kNT, true>(src, a.item_opB[item], S.op[1]);
        if (cpmg) {
            op_load_lds<kNT, true>(src, a.item_opC[item], S.op[2]);
            op_load_lds<kNT, true>(src, a.item_opD[item], S.op[3]);
        }
        __syncthreads();
        Basis<M, true> V;
        IntegState gA, gB;
        const int n = a.n;
        const cd u0 = tid < n ? mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y) : mkcd(0.0, 0.0);
        const cd y0 = tid < n ? mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y) : mkcd(0.0, 0.0);
        const NormWork nw{S.xs, S.xs32, S.tdiag, S.toff, S.scal};
        init_integ<true>(nw, S.op[0], br, gA, u0, a.rnd, a.klim, tol, M);
        init_integ<true>(nw, S.op[1], br, gB, y0, a.rnd, a.klim, tol, M);
        const double beta0A = gA.beta, beta0B = gB.beta;
        const double scale = a.scale[item];
        const int h = cpmg_split(q.npulse);
        for (int k = 0; k < q.nt; ++k) {
            // the active sweep's state is always gA: the two states swap after
            // each sweep (register moves, no dynamically indexed storage)
            for (int s = 0; s < 2; ++s) {
                IntegState& g = gA;
                // Hahn echo: one segment, the sweep continues to tau_k.  CPMG-N:
                // restart from rho0 and run this side's segments over t_k.
                int nseg = 1;
                if (cpmg) {
                    g.w = s ? y0 : u0;
                    g.beta = s ? beta0B : beta0A;
                    g.t_now = 0.0;
                    nseg = s == 0 ? h : q.npulse + 1 - h;
                }
                for (int sg = 0; sg < nseg; ++sg) {
                    int opi = s;
                    double target = q.tau[k];
                    if (cpmg) {
                        const int j = s == 0 ? sg : q.npulse - sg;
                        opi = s == 0 ? ((j & 1) ? 2 : 0) : ((j & 1) ? 1 : 3);
                        target = g.t_now + cpmg_fraction(j, q.npulse) * q.tau[k];
                    }
                    advance_lz<M>(S, br, g, V, target, max_steps, S.op[opi], tol);
                }
                const IntegState tmp = gA;
                gA = gB;
                gB = tmp;
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

int launch_hahn_lanczos(const KrylovArgs& a, int m, int grid, hipStream_t stream) {
    if (a.n > 256) return -11;
    if (m == 20) hipLaunchKernelGGL((cce_hahn_lanczos_kernel<20>), dim3(grid), dim3(kNT), 0, stream, a);
    else if (m == 30) hipLaunchKernelGGL((cce_hahn_lanczos_kernel<30>), dim3(grid), dim3(kNT), 0, stream, a);
    else return -12;
    return (int)hipGetLastError();
}

}  // namespace oops
