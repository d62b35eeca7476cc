// Pieces shared by the batched Krylov kernels (cce_krylov.hip: Arnoldi,
// cce_lanczos.hip: Hermitian Lanczos): integrator state, the operator in LDS,
// Expokit's step rounding, the mixed/FP64 product of one element, the
// reference's Lanczos + power-iteration norm estimate, the register basis.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devcommon.h"
#include "oops_kernels.h"

namespace oops {

constexpr int kNT = 256;

// Compiled operator tables handed to the out-of-line slot-table builders by
// value: a reference to the kernel's argument struct would make the kernel
// keep a copy of the struct in scratch memory and read every field from there.
struct SlotTables {
    const DevOpHeader* hdr;
    const DevOpLocal* loc;
    const double2* val;
    const uint8_t* eslot;
    int ms64, ms32;   // slot counts (medium kernel)
};
__device__ __forceinline__ SlotTables slot_tables(const KrylovArgs& a) {
    return SlotTables{a.hdr, a.loc, a.val, a.eslot, a.ms64, a.ms32};
}

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
    const double x = ool_log10(t) - sqr1;
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

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) v += __shfl_xor(v, mask, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) v = fmax(v, __shfl_xor(v, mask, kWave));
    return v;
}

__device__ __forceinline__ cd readlane_cd(cd x, int c) {
    return mkcd(lane_val(x.re, c), lane_val(x.im, c));
}

// Output times, result row and pulse count of one item.
struct ItemSeq {
    int nt;
    int npulse;
    bool two_chain;   // single-sample form: both sides run their own chain forward
    const double* tau;
    double* L;
};

__device__ __forceinline__ ItemSeq item_seq(const KrylovArgs& a, int item) {
    ItemSeq q;
    q.nt = a.item_nt ? a.item_nt[item] : a.nt;
    const int np = a.item_npulse ? a.item_npulse[item] : 1;
    q.npulse = np & 0xFFFF;
    q.two_chain = (np & kTwoChainFlag) != 0;
    q.tau = a.tau + (a.item_tau_off ? a.item_tau_off[item] : 0);
    q.L = a.L + (a.item_L_off ? a.item_L_off[item] : (int64_t)item * a.nt);
    return q;
}

// CPMG-N (Pulse_Interval / riffle, misc.cpp:59-84): segment j of N+1 has the
// fraction 1/(2N) at both ends and 1/N inside, and generator lvA for even j,
// lvB for odd j.  The ket side applies segments 0..h-1 to rho0, the bra side
// the adjoints of segments N..h, h = (N+2)/2, so that
// <rho0, U_N..U_0 rho0> = <U_h^+..U_N^+ rho0, U_{h-1}..U_0 rho0>.
__device__ __forceinline__ double cpmg_fraction(int j, int npulse) {
    return (j == 0 || j == npulse) ? 0.5 / npulse : 1.0 / npulse;
}
__device__ __forceinline__ int cpmg_split(int npulse) { return (npulse + 2) / 2; }

// LDS scratch of the norm estimate.
struct NormWork {
    cd* xs;
    float2* xs32;
    double* tdiag;  // >= klim entries
    double* toff;
    double* scal;
};

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

// ----------------------------------------------------------------------------
// Norm estimate: k Lanczos steps on +H from the DLARAN start vector, then
// power iteration on the k x k tridiagonal (reference :210-265).
// mv(i) returns element i of the product with the vector staged in S.xs /
// S.xs32.
template <bool MIXED, typename MV>
__device__ double lanczos_anorm_mv(const NormWork& S, BlockRed<kNT>& br, int n, MV&& mv, const double* rnd,
                                   int klim, double tol) {
    const int tid = threadIdx.x;
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
        stage_x<MIXED>(S.xs, S.xs32, tid, n, q);
        __syncthreads();
        cd w = mv(tid);
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
    if (tid < kWave) {
        // power iteration on T (k x k symmetric tridiagonal; entries not set
        // stay 0), one row per lane of wave 0
        const int j = tid;
        const bool row = j < k;
        const double dj = row ? tdiag[j] : 0.0;
        const double lo = (row && j > 0) ? toff[j - 1] : 0.0;
        const double up = (row && j + 1 < k) ? toff[j] : 0.0;
        double qj = row ? rnd[2 * n + j] : 0.0;
        double cjm = 0.0, cj = 0.0;
        for (int it = 0; it < 10000; ++it) {
            const double qm1 = __shfl_up(qj, 1, kWave);
            const double qp1 = __shfl_down(qj, 1, kWave);
            double wj = dj * qj;
            wj += lo * qm1;
            wj += up * qp1;
            cj = wave_sum(qj * wj);
            if (fabs(1.0 - cjm / cj) < tol) break;
            cjm = cj;
            cj = sqrt(wave_sum(wj * wj));
            qj = wj / cj;
        }
        const double qlast = __shfl(qj, k - 1, kWave);
        if (j == 0) S.scal[0] = fabs(cj) + fabs(bj * qlast);
    }
    __syncthreads();
    const double an = S.scal[0];
    __syncthreads();
    return an;
}

template <bool MIXED>
__device__ double lanczos_anorm(const NormWork& S, BlockRed<kNT>& br, const OpLds& o, const double* rnd, int klim,
                                double tol) {
    return lanczos_anorm_mv<MIXED>(S, br, o.hdr.n, [&](int i) { return matvec_elem<MIXED>(o, S.xs, S.xs32, i); },
                                   rnd, klim, tol);
}
// Start one integrator: operator o (already in LDS), start vector w0 (this
// thread's element), norm estimate and Expokit's first step size.
template <bool MIXED, typename MV>
__device__ void init_integ_mv(const NormWork& S, int n, MV&& mv, BlockRed<kNT>& br, IntegState& g, cd w0,
                              const double* rnd, int klim, double tol, int m) {
    const int tid = threadIdx.x;
    g.n = n;
    g.w = tid < n ? w0 : mkcd(0.0, 0.0);
    g.anorm = lanczos_anorm_mv<MIXED>(S, br, n, mv, rnd, klim, tol);
    const auto tb = block_sum1<1, kNT>(br, [&](int) { return cabs2(g.w); });
    g.beta = sqrt(tb.get(0));
    g.t_now = 0.0;
    g.xm = 1.0 / (double)m;
    g.rndoff = 2.220446049250313e-16 * g.anorm;
    const double p2 = tol * ool_pow((m + 1) / 2.72, (double)(m + 1)) * sqrt(2.0 * 3.14 * (m + 1));
    g.t_new = (1.0 / g.anorm) * ool_pow(p2 / (4.0 * g.beta * g.anorm), g.xm);
    g.t_new = round_step2(g.t_new);
    g.nstep = g.nmult = g.nrej = g.nexph = g.nscale = g.brk = g.fail = 0;
#ifdef OOPS_PHASE_TIMERS
    g.tmv = g.tred = g.texp = g.tupd = 0;
#endif
}

template <bool MIXED>
__device__ void init_integ(const NormWork& S, const OpLds& o, BlockRed<kNT>& br, IntegState& g, cd w0,
                           const double* rnd, int klim, double tol, int m) {
    init_integ_mv<MIXED>(S, o.hdr.n, [&](int i) { return matvec_elem<MIXED>(o, S.xs, S.xs32, i); }, br, g, w0, rnd,
                         klim, tol, m);
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

}  // namespace oops
