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

#include <type_traits>
#include <utility>

#include "cheb.h"
#include "krylov_common.h"

namespace oops {

template <int M>
struct LanczosSmem {
    OpLds op[4];                        // lvA, -lvB^dagger, lvB, -lvA^dagger stay resident
    cd xs[256];
    float2 xs32[256];
    double red[3 * (kNT / kWave) * 64];
    double d[M + 2];                    // diagonal of T
    double e[M + 2];                    // off-diagonal of T (e[j] couples j and j+1)
    cd F[M + 2];                        // Krylov coefficients of the step
    double tdiag[16];
    double toff[16];
    double scal[16];
    ChebTables cheb;                   // Bessel tables of the Chebyshev series
    int ints[8];
#ifdef OOPS_PHASE_TIMERS
    unsigned long long tbes;   // cycles of the Chebyshev set-up (interval, Bessel values)
    unsigned long long tser;   // cycles of the Chebyshev series (wave 0)
    unsigned long long temit;  // cycles of the output-time states inside a step (thread 0)
#endif
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
                t_step = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
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
        g.t_new = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
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

// ---------------------------------------------------------------------------
// Slot form.  With power-of-two local dimensions the local digits of an
// element index are bit fields, so every entry of a merged local operator
// moves element i to i ^ mask, with a mask fixed by the entry's local row and
// column.  Summing all entries of all local operators that share a mask gives
//   (A x)_i = sum_s c_s(i) x[i ^ mask_s],
// 21 terms per element for a 4-spin spin-1/2 Liouvillian instead of 64 ELL
// entries (4 one-body rows of 4, 6 two-body rows of 8).  Each thread holds
// its c_s(i) in registers (FP64 for masks that change at most one spin's
// digits, where the large one-body terms land; FP32 for two-spin flips), so
// one product is 21 LDS reads of x and no table reads.  The per-element
// coefficient tables of a cluster's operators are built once per cluster in
// a per-workgroup global buffer and reloaded into registers when the active
// generator changes.  The Lanczos basis v_1..v_M lives in LDS (thread i owns
// column i), which leaves registers for the coefficients and fits two
// workgroups per CU.

template <bool F64>
struct SlotOp {
    // pair-flip coefficients: complex FP32 (mixed) or FP64 (the FP64 variant)
    typedef typename std::conditional<F64, cd, float2>::type PairC;
    cd c[kSlot64];
    PairC f[kSlot32];
    int mc[kSlot64];   // masks (uniform)
    int mf[kSlot32];
    int n;

    __device__ __forceinline__ void load(const double2* __restrict__ tab, const int32_t* __restrict__ masks, int i,
                                         int n_) {
        n = n_;
        // opaque column index: keeps the 2 x 21 per-slot addresses of every
        // call site from being hoisted (and held) across the kernel's loops
        asm volatile("" : "+v"(i));
#pragma unroll
        for (int s = 0; s < kSlot64; ++s) {
            const double2 v = tab[s * 256 + i];
            c[s] = mkcd(v.x, v.y);
            mc[s] = masks[s];
        }
        if constexpr (F64) {
            const double2* t2 = tab + kSlot64 * 256;
#pragma unroll
            for (int s = 0; s < kSlot32; ++s) {
                const double2 v = t2[s * 256 + i];
                f[s] = mkcd(v.x, v.y);
                mf[s] = masks[kSlot64 + s];
            }
        } else {
            const float2* t32 = reinterpret_cast<const float2*>(tab + kSlot64 * 256);
#pragma unroll
            for (int s = 0; s < kSlot32; ++s) {
                f[s] = t32[s * 256 + i];
                mf[s] = masks[kSlot64 + s];
            }
        }
    }

    // element i of A x with x staged in LDS (FP64 copy xs, FP32 copy xs32).
    // Three rounds of 3 FP64 + 4 pair gathers bound the registers in flight;
    // the empty asm keeps the gather addresses from being hoisted out of the
    // caller's loops (21 more live registers otherwise).
    __device__ __forceinline__ cd apply(const cd* __restrict__ xs, const float2* __restrict__ xs32, int i) const {
        static_assert(kSlot64 == 9 && kSlot32 == 12, "gather rounds assume 9 + 12 slots");
        if (i >= n) return mkcd(0.0, 0.0);
        int ii = i;
        asm volatile("" : "+v"(ii));
        cd acc = mkcd(0.0, 0.0);
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            cd x64[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) x64[t] = xs[ii ^ mc[3 * r + t]];
            if constexpr (F64) {
                cd xp[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) xp[t] = xs[ii ^ mf[4 * r + t]];
#pragma unroll
                for (int t = 0; t < 3; ++t) acc = cfma(c[3 * r + t], x64[t], acc);
#pragma unroll
                for (int t = 0; t < 4; ++t) acc = cfma(f[4 * r + t], xp[t], acc);
            } else {
                float2 x32[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) x32[t] = xs32[ii ^ mf[4 * r + t]];
#pragma unroll
                for (int t = 0; t < 3; ++t) acc = cfma(c[3 * r + t], x64[t], acc);
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float2 cf = f[4 * r + t];
                    ar = fmaf(cf.x, x32[t].x, fmaf(-cf.y, x32[t].y, ar));
                    ai = fmaf(cf.x, x32[t].y, fmaf(cf.y, x32[t].x, ai));
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        return mkcd(acc.re + (double)ar, acc.im + (double)ai);
    }
};

// Coefficient table of operator `op` for this thread's element: FP64 slots
// as double2 [kSlot64][256], then the pair slots as float2 [kSlot32][256]
// (mixed) or double2 [kSlot32][256] (F64).  Every thread writes and reads
// only its own column.
template <bool F64>
__device__ __noinline__ void slot_build(SlotTables a, int op, double2* __restrict__ tab, int i) {
    typedef typename std::conditional<F64, double2, float2>::type PairT;
    const DevOpHeader h = a.hdr[op];
    asm volatile("" : "+v"(i));
    PairT* t2 = reinterpret_cast<PairT*>(tab + kSlot64 * 256);
    for (int s = 0; s < kSlot64; ++s) tab[s * 256 + i] = make_double2(0.0, 0.0);
    for (int s = 0; s < kSlot32; ++s) t2[s * 256 + i] = PairT{0, 0};
    if (i >= h.n) return;
    tab[i] = make_double2(h.c0re, h.c0im);   // slot 0: mask 0
    const int ld = h.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = h.nspin - 1;
    for (int L = 0; L < h.nloc; ++L) {
        const DevOpLocal lo = a.loc[h.loc_off + L];
        int row = (i >> (ld * (top - lo.s))) & lmask;
        if (lo.nb == 2) row = (row << ld) | ((i >> (ld * (top - lo.t))) & lmask);
        const int e0 = h.ent_off + lo.ent + row * lo.R;
        for (int e = 0; e < lo.R; ++e) {
            const double2 v = a.val[e0 + e];
            const int sl = a.eslot[e0 + e];
            // this thread's own column: read-modify-write in program order
            if (sl & 128) {
                PairT t = t2[(sl & 127) * 256 + i];
                if constexpr (F64) {
                    t.x += v.x;
                    t.y += v.y;
                } else {
                    t.x += (float)v.x;
                    t.y += (float)v.y;
                }
                t2[(sl & 127) * 256 + i] = t;
            } else {
                double2 t = tab[sl * 256 + i];
                t.x += v.x;
                t.y += v.y;
                tab[sl * 256 + i] = t;
            }
        }
    }
}

template <int M, bool F64>
struct LzSlotSmem {
    // v_1..v_M (FP32, or FP64 in the F64 variant), column i owned by thread i
    typename std::conditional<F64, double2, float2>::type V[M][kNT];
    cd xs[256];
    float2 xs32[256];
    double red[3 * (kNT / kWave) * 64];
    double d[M + 2];
    double e[M + 2];
    cd F[M + 2];
    double tdiag[16];
    double toff[16];
    double scal[16];
    ChebTables cheb;                   // Bessel tables of the Chebyshev series
    int ints[8];
#ifdef OOPS_PHASE_TIMERS
    unsigned long long tbes;   // cycles of the Chebyshev set-up (interval, Bessel values)
    unsigned long long tser;   // cycles of the Chebyshev series (wave 0)
    unsigned long long temit;  // cycles of the output-time states inside a step (thread 0)
#endif
};

// Output side of a sweep.  mode 0: advance to t_end only (CPMG segments).
// mode 1: store the state at every output time tau[k] in ob (kNT elements per
// time).  mode 2: at every tau[k] take Re <stored state, state> into L[k].
struct SweepOut {
    const double* tau;
    double2* ob;
    double* L;
    double scale;
    int nt;
    int mode;
};

template <int M>
__device__ __forceinline__ void emit_state(BlockRed<kNT>& br, const SweepOut& o, int k, const cd& w) {
    const int tid = threadIdx.x;
    double2* const p = o.ob + (int64_t)k * kNT + tid;
    if (o.mode == 1) {
        *p = make_double2(w.re, w.im);
    } else {
        const double2 u = *p;
        const auto t = block_sum_lanes<1, kNT>(br, [&](int) { return fma(w.re, u.x, w.im * u.y); });
        if (tid == 0) o.L[k] = o.scale * t.get(0);
    }
}

// One sweep of Expokit's integrator (zgexpv_cache.f:440-661) with the Lanczos
// recurrence of a Hermitian generator.  With output times (mode 1 / 2) the
// steps are not cut at the outputs: a step runs to min(t_new, t_end) under the
// usual error control, and every output time inside it is evaluated from the
// same Krylov basis, w(tau_k) = beta V exp((tau_k - t_now) H_aug) e_1, the
// small exponential walking from one output time to the next and on to the
// step end.  The FP64 path keeps the reference's schedule (a step ends at
// every output time); the mixed path only has to meet its tolerance on L(t),
// and sparing the clamped steps halves its Krylov work on output grids finer
// than the natural step.  Mode 2 reads the states mode 1 stored for the same
// output times.
template <int M, bool F64>
__device__ __forceinline__ void advance_slot(LzSlotSmem<M, F64>& S, BlockRed<kNT>& br, IntegState& g,
                                             const SlotOp<F64>& so,
                                             double t_end, const SweepOut& o, int max_steps, double tol) {
    const int tid = threadIdx.x;
    const double delta = 1.2, gamma = 0.9;
    int k = 0;
    if (o.mode)
        for (; k < o.nt && o.tau[k] <= g.t_now; ++k) emit_state<M>(br, o, k, g.w);
    while (g.t_now < t_end) {
        if (g.nstep >= max_steps) { g.fail |= 1; break; }
        ++g.nstep;
        const cd v0 = cscale(g.w, 1.0 / g.beta);
        cd vcur = v0, vprev = mkcd(0.0, 0.0);
        double bprev = 0.0;
        int k1 = 2, mb = M;
        bool brk = false;
        double t_step = fmin(t_end - g.t_now, g.t_new);
#pragma unroll 1
        for (int j = 0; j < M; ++j) {
            OOPS_TIC(tm0);
            stage_x<true>(S.xs, S.xs32, tid, g.n, vcur);
            __syncthreads();
            cd p = so.apply(S.xs, S.xs32, tid);
            ++g.nmult;
            p = mkcd(p.re - bprev * vprev.re, p.im - bprev * vprev.im);
            OOPS_TOC(g, tmv, tm0);
            OOPS_TIC(tr0);
            const cd vj = vcur;
            const auto r1 = block_sum_lanes<3, kNT>(br, [&](int k) -> double {
                if (k == 0) return fma(vj.re, p.re, vj.im * p.im);
                if (k == 1) return fma(vj.re, p.im, -vj.im * p.re);
                if (k == 2) return cabs2(p);
                return 0.0;
            });
            const cd h = mkcd(r1.get(0), r1.get(1));
            p = cfms(h, vj, p);
            const double alpha = h.re;
            const double beta = sqrt(fmax(r1.get(2) - cabs2(h), 0.0));
            OOPS_TOC(g, tred, tr0);
            if (tid == 0) {
                S.d[j] = alpha;
                S.e[j] = beta;
            }
            if (beta <= 1.0e-7) {
                k1 = 0;
                g.brk = 1;
                mb = j + 1;
                t_step = t_end - g.t_now;
                brk = true;
                break;
            }
            vprev = vj;
            vcur = cscale(p, 1.0 / beta);
            if constexpr (F64) S.V[j][tid] = make_double2(vcur.re, vcur.im);
            else S.V[j][tid] = make_float2((float)vcur.re, (float)vcur.im);
            bprev = beta;
        }
        double avnorm = 0.0;
        if (!brk) {
            stage_x<true>(S.xs, S.xs32, tid, g.n, vcur);
            __syncthreads();
            const cd p = so.apply(S.xs, S.xs32, tid);
            ++g.nmult;
            const auto tn = block_sum_lanes<1, kNT>(br, [&](int) { return cabs2(p); });
            avnorm = sqrt(tn.get(0));
        }
        __syncthreads();
        OOPS_TIC(te0);
        // w = beta * (sum_{j<mb} F_j v_j + [k1] F_M v_M)
        const int mx = mb + (k1 - 1 > 0 ? k1 - 1 : 0);
        // four basis vectors per iteration into two accumulators: the LDS
        // reads of an iteration go out together and the two FMA chains
        // interleave
        auto combine = [&]() -> cd {
            cd acc0 = cmul(v0, S.F[0]), acc1 = mkcd(0.0, 0.0);
            int j = 1;
#pragma unroll 1
            for (; j + 3 < mx; j += 4) {
                const auto f0 = S.V[j - 1][tid];
                const auto f1 = S.V[j][tid];
                const auto f2 = S.V[j + 1][tid];
                const auto f3 = S.V[j + 2][tid];
                acc0 = cfma(mkcd((double)f0.x, (double)f0.y), S.F[j], acc0);
                acc1 = cfma(mkcd((double)f1.x, (double)f1.y), S.F[j + 1], acc1);
                acc0 = cfma(mkcd((double)f2.x, (double)f2.y), S.F[j + 2], acc0);
                acc1 = cfma(mkcd((double)f3.x, (double)f3.y), S.F[j + 3], acc1);
            }
#pragma unroll 1
            for (; j < mx; ++j) {
                const auto f = S.V[j - 1][tid];
                acc0 = cfma(mkcd((double)f.x, (double)f.y), S.F[j], acc0);
            }
            return cscale(mkcd(acc0.re + acc1.re, acc0.im + acc1.im), g.beta);
        };
        // Trial step: F walks from e_1 through the output times inside the
        // step (each evaluated and emitted from this basis) to the step end,
        // where Expokit's error estimate decides.  After a rejection the walk
        // restarts for the shorter step; outputs emitted beyond its end are
        // emitted again by the steps that reach them.
        double err_loc = 0.0, t1 = t_end;
        int kend = k;
        for (;;) {
            t1 = t_step >= t_end - g.t_now ? t_end : g.t_now + t_step;
            kend = k;
            if (o.mode)
                while (kend < o.nt && o.tau[kend] < t1) ++kend;
            g.nscale += cheb_walk<M>(S, mb, k1, bprev, g.t_now, t1, o.tau, k, kend - k, g.nexph,
                                     [&](int i) { emit_state<M>(br, o, k + i, combine()); });
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
                if (!(t_step < 1e300) || g.nrej > max_steps) { g.fail |= 4; return; }
                __syncthreads();   // every thread has read F before the next attempt rewrites it
                continue;
            }
            break;
        }
        if (o.mode) k = kend;
        OOPS_TOC(g, texp, te0);
        OOPS_TIC(tu0);
        const cd wnew = combine();
        g.w = wnew;
        {
            const auto tb = block_sum_lanes<1, kNT>(br, [&](int) { return tid < g.n ? cabs2(g.w) : 0.0; });
            g.beta = sqrt(tb.get(0));
        }
        g.t_new = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        g.t_now = t1;
        if (o.mode)
            for (; k < o.nt && o.tau[k] <= g.t_now; ++k) emit_state<M>(br, o, k, g.w);
        OOPS_TOC(g, tupd, tu0);
    }
}

// Coefficient-table stride (double2) of one operator in the slot kernels.
template <bool F64>
constexpr int slot_table_d2() { return F64 ? (kSlot64 + kSlot32) * 256 : kSlotOpD2; }

template <int M, bool F64>
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(F64 ? 1 : 2, F64 ? 1 : 2)))
void cce_hahn_lanczos_slot_kernel(KrylovArgs a) {
    __shared__ LzSlotSmem<M, F64> S;
    constexpr int kTab = slot_table_d2<F64>();
    const int tid = threadIdx.x;
    BlockRed<kNT> br(S.red);
    const double tol = a.tol;
    const int max_steps = a.max_steps;
    double2* const ws = a.slotws + (int64_t)blockIdx.x * 4 * kTab;
#ifdef OOPS_PHASE_TIMERS
    if (tid == 0) S.tbes = S.tser = S.temit = 0;
#endif
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int64_t voff = a.vec_off[item];
        const ItemSeq q = item_seq(a, item);
        const bool cpmg = q.npulse > 1 || q.two_chain;
        const int n = a.n;
        // operator indices (uniform): A, -B^dagger, and for CPMG-N B, -A^dagger
        const int qA = __builtin_amdgcn_readfirstlane(a.item_opA[item]);
        const int qB = __builtin_amdgcn_readfirstlane(a.item_opB[item]);
        const int qC = cpmg ? __builtin_amdgcn_readfirstlane(a.item_opC[item]) : qA;
        const int qD = cpmg ? __builtin_amdgcn_readfirstlane(a.item_opD[item]) : qB;
        const SlotTables tabs = slot_tables(a);
        slot_build<F64>(tabs, qA, ws, tid);
        slot_build<F64>(tabs, qB, ws + kTab, tid);
        if (cpmg) {
            slot_build<F64>(tabs, qC, ws + 2 * kTab, tid);
            slot_build<F64>(tabs, qD, ws + 3 * kTab, tid);
        }
        IntegState gA, gB;
        const cd u0 = tid < n ? mkcd(a.u0[voff + tid].x, a.u0[voff + tid].y) : mkcd(0.0, 0.0);
        const cd y0 = tid < n ? mkcd(a.y0[voff + tid].x, a.y0[voff + tid].y) : mkcd(0.0, 0.0);
        const NormWork nw{S.xs, S.xs32, S.tdiag, S.toff, S.scal};
        const double scale = a.scale[item];
        SlotOp<F64> so;
        so.load(ws, a.smask + (int64_t)qA * kSlotMaskStride, tid, n);
        init_integ_mv<true>(nw, n, [&](int i) { return so.apply(S.xs, S.xs32, i); }, br, gA, u0, a.rnd, a.klim,
                            tol, M);
        so.load(ws + kTab, a.smask + (int64_t)qB * kSlotMaskStride, tid, n);
        init_integ_mv<true>(nw, n, [&](int i) { return so.apply(S.xs, S.xs32, i); }, br, gB, y0, a.rnd, a.klim,
                            tol, M);
        const double beta0A = gA.beta, beta0B = gB.beta;
        const int h = cpmg_split(q.npulse);
        // Hahn echo with a state store: sweep A once over every output time
        // (states stored), then sweep B taking <state_B, state_A> at the same
        // times.  Otherwise (CPMG-N) one pair of sweeps per output time.
        const bool hahn = !cpmg && a.sweepws && q.nt > 0 && q.nt <= a.ws_nt;
        SweepOut o{q.tau, a.sweepws + (int64_t)blockIdx.x * kNT * a.ws_nt, q.L, scale, q.nt, 0};
        const int nouter = hahn ? 1 : q.nt;
        for (int k = 0; k < nouter; ++k) {
            for (int s = 0; s < 2; ++s) {
                IntegState& g = gA;   // active sweep; the states swap after each sweep
                int nseg = 1;
                if (cpmg) {
                    g.w = s ? y0 : u0;
                    g.beta = s ? beta0B : beta0A;
                    g.t_now = 0.0;
                    nseg = q.two_chain ? q.npulse + 1 : (s == 0 ? h : q.npulse + 1 - h);
                }
                if (hahn) {
                    o.mode = 1 + s;
                    if (s == 1 && gB.fail != 0) nseg = 0;   // sweep A failed: no stored states
                }
                for (int sg = 0; sg < nseg; ++sg) {
                    int opi = s;
                    double target = hahn ? q.tau[q.nt - 1] : q.tau[k];
                    if (q.two_chain) {
                        // each side its own chain forward: ket opA / opC, bra opB / opD
                        const int j = sg;
                        opi = s == 0 ? ((j & 1) ? 2 : 0) : ((j & 1) ? 3 : 1);
                        target = g.t_now + cpmg_fraction(j, q.npulse) * q.tau[k];
                    } else if (cpmg) {
                        const int j = s == 0 ? sg : q.npulse - sg;
                        opi = s == 0 ? ((j & 1) ? 2 : 0) : ((j & 1) ? 1 : 3);
                        target = g.t_now + cpmg_fraction(j, q.npulse) * q.tau[k];
                    }
                    const int qo = opi == 0 ? qA : (opi == 1 ? qB : (opi == 2 ? qC : qD));
                    so.load(ws + opi * kTab, a.smask + (int64_t)qo * kSlotMaskStride, tid, n);
                    advance_slot<M, F64>(S, br, g, so, target, o, max_steps, tol);
                }
                const IntegState tmp = gA;
                gA = gB;
                gB = tmp;
            }
            if (!hahn) {
                const auto tl = block_sum_lanes<1, kNT>(br, [&](int) {
                    return fma(gB.w.re, gA.w.re, gB.w.im * gA.w.im);
                });
                if (tid == 0) q.L[k] = scale * tl.get(0);
                if ((gA.fail | gB.fail) != 0) break;
            }
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
            st[20] = (int)(S.tbes >> 10);
            st[21] = (int)(S.tser >> 10);
            st[22] = (int)(S.temit >> 10);
            S.tbes = S.tser = S.temit = 0;
#endif
        }
        __syncthreads();
    }
}

int launch_hahn_lanczos_slot(const KrylovArgs& a, int m, int grid, hipStream_t stream, bool fp64) {
    if (a.n > 256 || !a.slotws || !a.eslot || !a.smask) return -11;
    if (fp64) {
        if (m == 20) hipLaunchKernelGGL((cce_hahn_lanczos_slot_kernel<20, true>), dim3(grid), dim3(kNT), 0, stream, a);
        else if (m == 30) hipLaunchKernelGGL((cce_hahn_lanczos_slot_kernel<30, true>), dim3(grid), dim3(kNT), 0, stream, a);
        else return -12;
    } else {
        if (m == 20) hipLaunchKernelGGL((cce_hahn_lanczos_slot_kernel<20, false>), dim3(grid), dim3(kNT), 0, stream, a);
        else if (m == 30) hipLaunchKernelGGL((cce_hahn_lanczos_slot_kernel<30, false>), dim3(grid), dim3(kNT), 0, stream, a);
        else return -12;
    }
    return (int)hipGetLastError();
}

int64_t slot_table_double2(bool fp64) { return fp64 ? slot_table_d2<true>() : slot_table_d2<false>(); }

int launch_hahn_lanczos(const KrylovArgs& a, int m, int grid, hipStream_t stream) {
    if (a.n > 256) return -11;
    if (m == 20) hipLaunchKernelGGL((cce_hahn_lanczos_kernel<20>), dim3(grid), dim3(kNT), 0, stream, a);
    else if (m == 30) hipLaunchKernelGGL((cce_hahn_lanczos_kernel<30>), dim3(grid), dim3(kNT), 0, stream, a);
    else return -12;
    return (int)hipGetLastError();
}

}  // namespace oops
