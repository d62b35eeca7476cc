// Mixed-precision batched echo kernel for states of 1024 and 4096 elements
// (5 and 6 spin-1/2 Liouville spins): one 1024-thread workgroup per cluster,
// thread t owning elements t + 1024 q (q < E = n / 1024).  The product input
// (FP64 and FP32 copies) and the ELL tables of the active generator are held
// whole in LDS (148 KB at n = 4096); the FP32 Lanczos basis v_1..v_M lives in a
// per-workgroup device buffer (each thread reads and writes only its own
// elements).  The Krylov scheme is the one of the n <= 256 Lanczos kernels
// (cce_lanczos.hip): Expokit's step control (zgexpv_cache.f:418-661) with
// Lanczos for Hermitian generators, one fused block reduction per step, the
// Chebyshev exponential of the augmented tridiagonal matrix, and the
// reference's norm estimate (k Lanczos steps from the DLARAN vector, power
// iteration on the tridiagonal, hamvec_zgexpv_w_mkl_profile.f:210-265).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cheb.h"
#include "krylov_common.h"

namespace oops {

namespace {

constexpr int kMT = 1024;
constexpr int kMW = kMT / kWave;

struct MedOpLds {
    DevOpHeader hdr;
    DevOpLocal locs[kMaxLocals];
    alignas(16) cd vals[kMediumMaxEnt];
    alignas(16) int32_t offs[kMediumMaxEnt];
};

template <int M, int E>
struct MedSmem {
    cd xs[kMT * E];
    float2 xs32[kMT * E];
    MedOpLds op;
    double red[2 * kMW * 4];
    double d[M + 2];
    double e[M + 2];
    cd F[M + 2];
    double scal[16];
    ChebTables cheb;
    int ints[8];
    // flip-mask form: the active generator's coefficient tables and masks
    const double2* tab64;
    const float2* tab32;
    int ns64, ns32, slot;
    int mask[kMedMaskStride];
#ifdef OOPS_PHASE_TIMERS
    unsigned long long tbes, tser, temit;
#endif
};

struct MedState {
    double beta, t_now, t_new, anorm, xm;
    int nstep, nmult, nrej, nexph, nscale, brk, fail;
#ifdef OOPS_PHASE_TIMERS
    unsigned long long tmv, tred, texp, tupd;   // cycles: products, reductions, small exponentials, updates
#endif
};

// Generator `op` of the bucket into LDS (2-body values as complex FP32).
__device__ void med_load(const KrylovArgs& a, int op, MedOpLds& o) {
    const DevOpHeader h = a.hdr[op];
    if (threadIdx.x == 0) o.hdr = h;
    for (int k = threadIdx.x; k < h.nloc; k += kMT) o.locs[k] = a.loc[h.loc_off + k];
    float2* v32 = reinterpret_cast<float2*>(o.vals + h.ent2_first);
    for (int k = threadIdx.x; k < h.nent; k += kMT) {
        const double2 v = a.val[h.ent_off + k];
        if (k >= h.ent2_first) v32[k - h.ent2_first] = make_float2((float)v.x, (float)v.y);
        else o.vals[k] = mkcd(v.x, v.y);
        o.offs[k] = a.off[h.ent_off + k];
    }
    __syncthreads();
}

template <int M, int E>
__device__ __forceinline__ void med_stage(MedSmem<M, E>& S, const cd (&x)[E]) {
#pragma unroll
    for (int q = 0; q < E; ++q) {
        const int i = threadIdx.x + kMT * q;
        S.xs[i] = x[q];
        S.xs32[i] = make_float2((float)x[q].re, (float)x[q].im);
    }
}

// Element i = threadIdx.x + 1024 q of A x, x staged in LDS.  Flip-mask form:
// (A x)_i = sum_s c_s(i) x[i ^ mask_s], coefficients streamed from the
// workgroup's device tables (FP64 for one-spin masks, FP32 for two-spin
// masks); otherwise the merged ELL tables in LDS.
template <int M, int E>
__device__ __forceinline__ cd med_product(const MedSmem<M, E>& S, int q) {
    const int i = threadIdx.x + kMT * q;
    if (S.slot) {
        const int n = kMT * E;
        const double2* __restrict__ t64 = S.tab64;
        const float2* __restrict__ t32 = S.tab32;
        cd acc = mkcd(0.0, 0.0);
        const int n64 = S.ns64, n32 = S.ns32;
#pragma unroll 4
        for (int s = 0; s < n64; ++s) {
            const double2 c = t64[(int64_t)s * n + i];
            acc = cfma(mkcd(c.x, c.y), S.xs[i ^ S.mask[s]], acc);
        }
        float ar = 0.f, ai = 0.f;
#pragma unroll 4
        for (int s = 0; s < n32; ++s) {
            const float2 c = t32[(int64_t)s * n + i];
            const float2 x = S.xs32[i ^ S.mask[kMedSlot64 + s]];
            ar = fmaf(c.x, x.x, fmaf(-c.y, x.y, ar));
            ai = fmaf(c.x, x.y, fmaf(c.y, x.x, ai));
        }
        return mkcd(acc.re + (double)ar, acc.im + (double)ai);
    }
    return op_apply_elem_mixed(i, S.xs, S.xs32, S.op.hdr, S.op.locs, S.op.vals, S.op.offs);
}

// The merged ELL product for the E elements i0 + kMT r of this thread, the
// local operators walked once.  kMT r sets the digits of spin 0 only (kMT =
// 2^10, two bits per spin): a local operator without spin 0 has the same row
// for the E elements, so each chunk of its entries is read from LDS once and
// gathered for the E elements (E independent chains); operators on spin 0
// take a row per element.
template <int E>
__device__ __forceinline__ void ell_product_all(const cd* __restrict__ xs, const float2* __restrict__ xs32,
                                                const DevOpHeader& h, const DevOpLocal* __restrict__ locs,
                                                const cd* __restrict__ vals, const int32_t* __restrict__ offs,
                                                cd (&y)[E]) {
    const int i0 = threadIdx.x;
    const cd c0 = mkcd(h.c0re, h.c0im);
    cd acc[E];
    float ar[E], ai[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        acc[r] = cmul(c0, xs[i0 + kMT * r]);
        ar[r] = 0.f;
        ai[r] = 0.f;
    }
    const float2* v32 = reinterpret_cast<const float2*>(vals + h.ent2_first);
    const int ld = h.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = h.nspin - 1;
    for (int L = 0; L < h.nloc; ++L) {
        const DevOpLocal lo = locs[L];
        const bool spin0 = lo.s == 0 || (lo.nb == 2 && lo.t == 0);
        const int shs = ld * (top - lo.s);
        const int sht = ld * (top - lo.t);
        int base[E], e0[E];
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const int i = i0 + kMT * r;
            const int ls = (i >> shs) & lmask;
            int b = i - (ls << shs);
            int row = ls;
            if (lo.nb == 2) {
                const int lt = (i >> sht) & lmask;
                b -= lt << sht;
                row = (ls << ld) | lt;
            }
            base[r] = b;
            e0[r] = lo.ent + row * lo.R;
        }
        if (!spin0) {
            // one row for the E elements: the chunk's entries once, E gathers each
            const int e00 = e0[0];
            if (lo.nb == 2) {
#pragma unroll 1
                for (int e = 0; e < lo.R; e += 4) {
                    const int4 q = *reinterpret_cast<const int4*>(offs + e00 + e);
                    const int o[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const float2 vv = v32[e00 - h.ent2_first + e + u];
#pragma unroll
                        for (int r = 0; r < E; ++r) {
                            const float2 xv = xs32[base[r] + o[u]];
                            ar[r] = fmaf(vv.x, xv.x, fmaf(-vv.y, xv.y, ar[r]));
                            ai[r] = fmaf(vv.x, xv.y, fmaf(vv.y, xv.x, ai[r]));
                        }
                    }
                }
            } else {
#pragma unroll 1
                for (int e = 0; e < lo.R; e += 4) {
                    const int4 q = *reinterpret_cast<const int4*>(offs + e00 + e);
                    const int o[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const cd vv = vals[e00 + e + u];
#pragma unroll
                        for (int r = 0; r < E; ++r) acc[r] = cfma(vv, xs[base[r] + o[u]], acc[r]);
                    }
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < E; ++r) {
                if (lo.nb == 2) {
                    const float2 a32 = ell_row32(xs32, base[r], v32, offs, e0[r], e0[r] - h.ent2_first, lo.R,
                                                 make_float2(ar[r], ai[r]));
                    ar[r] = a32.x;
                    ai[r] = a32.y;
                } else {
                    acc[r] = ell_row(xs, base[r], vals, offs, e0[r], lo.R, acc[r]);
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < E; ++r) y[r] = mkcd(acc[r].re + (double)ar[r], acc[r].im + (double)ai[r]);
}

// All E elements of this thread.  The flip-mask form walks the slots once
// for the E elements (E independent coefficient loads per slot in flight).
template <int M, int E>
__device__ __forceinline__ void med_product_all(const MedSmem<M, E>& S, cd (&y)[E]) {
    if (!S.slot) {
        if constexpr (E > 1) {
            ell_product_all<E>(S.xs, S.xs32, S.op.hdr, S.op.locs, S.op.vals, S.op.offs, y);
        } else {
#pragma unroll
            for (int r = 0; r < E; ++r) y[r] = med_product<M, E>(S, r);
        }
        return;
    }
    const int n = kMT * E;
    const int i0 = threadIdx.x;
    const double2* __restrict__ t64 = S.tab64;
    const float2* __restrict__ t32 = S.tab32;
    const int n64 = S.ns64, n32 = S.ns32;
    cd acc[E];
    float ar[E], ai[E];
#pragma unroll
    for (int r = 0; r < E; ++r) { acc[r] = mkcd(0.0, 0.0); ar[r] = 0.f; ai[r] = 0.f; }
#pragma unroll 4
    for (int s = 0; s < n64; ++s) {
        const int mk = S.mask[s];
        double2 c[E];
#pragma unroll
        for (int r = 0; r < E; ++r) c[r] = t64[(int64_t)s * n + i0 + kMT * r];
#pragma unroll
        for (int r = 0; r < E; ++r) acc[r] = cfma(mkcd(c[r].x, c[r].y), S.xs[(i0 + kMT * r) ^ mk], acc[r]);
    }
#pragma unroll 4
    for (int s = 0; s < n32; ++s) {
        const int mk = S.mask[kMedSlot64 + s];
        float2 c[E];
#pragma unroll
        for (int r = 0; r < E; ++r) c[r] = t32[(int64_t)s * n + i0 + kMT * r];
#pragma unroll
        for (int r = 0; r < E; ++r) {
            const float2 x = S.xs32[(i0 + kMT * r) ^ mk];
            ar[r] = fmaf(c[r].x, x.x, fmaf(-c[r].y, x.y, ar[r]));
            ai[r] = fmaf(c[r].x, x.y, fmaf(c[r].y, x.x, ai[r]));
        }
    }
#pragma unroll
    for (int r = 0; r < E; ++r) y[r] = mkcd(acc[r].re + (double)ar[r], acc[r].im + (double)ai[r]);
}

// Coefficient tables of operator `op` for this thread's elements (FP64 slots
// [ms64][n] then FP32 slots [ms32][n]); every thread writes and reads only its
// own columns.  Slot 0 is mask 0 (the diagonal, with the constant term).
template <int E>
__device__ __noinline__ void med_slot_build(SlotTables a, int op, double2* __restrict__ tab) {
    const DevOpHeader h = a.hdr[op];
    const int n = kMT * E;
    float2* t32 = reinterpret_cast<float2*>(tab + (int64_t)a.ms64 * n);
    const int ld = h.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = h.nspin - 1;
    for (int q = 0; q < E; ++q) {
        int i = threadIdx.x + kMT * q;
        asm volatile("" : "+v"(i));
        for (int s = 0; s < a.ms64; ++s) tab[(int64_t)s * n + i] = make_double2(0.0, 0.0);
        for (int s = 0; s < a.ms32; ++s) t32[(int64_t)s * n + i] = make_float2(0.f, 0.f);
        tab[i] = make_double2(h.c0re, h.c0im);
        for (int L = 0; L < h.nloc; ++L) {
            const DevOpLocal lo = a.loc[h.loc_off + L];
            int row = (i >> (ld * (top - lo.s))) & lmask;
            if (lo.nb == 2) row = (row << ld) | ((i >> (ld * (top - lo.t))) & lmask);
            const int e0 = h.ent_off + lo.ent + row * lo.R;
            for (int e = 0; e < lo.R; ++e) {
                const double2 v = a.val[e0 + e];
                const int sl = a.eslot[e0 + e];
                if (sl & 128) {
                    float2 t = t32[(int64_t)(sl & 127) * n + i];
                    t.x += (float)v.x;
                    t.y += (float)v.y;
                    t32[(int64_t)(sl & 127) * n + i] = t;
                } else {
                    double2 t = tab[(int64_t)sl * n + i];
                    t.x += v.x;
                    t.y += v.y;
                    tab[(int64_t)sl * n + i] = t;
                }
            }
        }
    }
}

// Norm estimate of the generator in LDS: k Lanczos steps on +H from the
// DLARAN start vector, then power iteration on the k x k tridiagonal.
template <int M, int E>
__device__ double med_anorm(MedSmem<M, E>& S, BlockRed<kMT>& br, const double* rnd, int n, int klim, double tol) {
    const int tid = threadIdx.x;
    const int k = klim < n ? klim : n;
    if (tid < 16) { S.scal[tid] = 0.0; }
    cd q[E], qm[E], w[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int i = tid + kMT * r;
        q[r] = mkcd(rnd[2 * i], rnd[2 * i + 1]);
        qm[r] = mkcd(0.0, 0.0);
    }
    {
        const auto t0 = block_sum_lanes<1, kMT, 4>(br, [&](int) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < E; ++r) s += cabs2(q[r]);
            return s;
        });
        const double inv = 1.0 / sqrt(t0.get(0));
#pragma unroll
        for (int r = 0; r < E; ++r) q[r] = cscale(q[r], inv);
    }
    double* tdiag = S.d;   // the tridiagonal of the estimate (before any Krylov step uses d, e)
    double* toff = S.e;
    if (tid < M + 2) { tdiag[tid] = 0.0; toff[tid] = 0.0; }
    double bjm = 0.0, bj = 0.0;
#pragma unroll 1
    for (int j = 1; j <= k; ++j) {
        med_stage<M, E>(S, q);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < E; ++r) w[r] = med_product<M, E>(S, r);
        const auto ta = block_sum_lanes<2, kMT, 4>(br, [&](int c) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < E; ++r)
                s += c == 0 ? fma(q[r].re, w[r].re, q[r].im * w[r].im) : fma(q[r].re, w[r].im, -q[r].im * w[r].re);
            return s;
        });
        const cd aj = mkcd(ta.get(0), ta.get(1));
#pragma unroll
        for (int r = 0; r < E; ++r) {
            w[r] = cfms(aj, q[r], w[r]);
            w[r] = mkcd(w[r].re - bjm * qm[r].re, w[r].im - bjm * qm[r].im);
        }
        const auto tb = block_sum_lanes<1, kMT, 4>(br, [&](int) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < E; ++r) s += cabs2(w[r]);
            return s;
        });
        bj = sqrt(tb.get(0));
        if (tid == 0) {
            tdiag[j - 1] = aj.re;
            if (j < k) toff[j - 1] = bj;
        }
        if (j == k) break;
        if (bj < tol) break;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            qm[r] = q[r];
            q[r] = cscale(w[r], 1.0 / bj);
        }
        bjm = bj;
        __syncthreads();   // the product of the next step overwrites xs
    }
    __syncthreads();
    if (tid < kWave) {
        // power iteration on T (k x k symmetric tridiagonal), one row per lane
        const int j = tid;
        const double dj = j < k ? tdiag[j] : 0.0;
        const double lo = (j > 0 && j < k) ? toff[j - 1] : 0.0;
        const double up = (j + 1 < k) ? toff[j] : 0.0;
        double qj = j < k ? rnd[2 * n + j] : 0.0;
        double cjm = 0.0, cj = 0.0;
        for (int it = 0; it < 10000; ++it) {
            const double qm1 = __shfl_up(qj, 1, kWave);
            const double qp1 = __shfl_down(qj, 1, kWave);
            double wj = dj * qj;
            wj += (j > 0 ? lo * qm1 : 0.0);
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

template <int M, int E>
__device__ void med_init(MedSmem<M, E>& S, BlockRed<kMT>& br, MedState& g, cd (&w)[E], const double* rnd, int n,
                         int klim, double tol) {
    g.anorm = med_anorm<M, E>(S, br, rnd, n, klim, tol);
    const auto tb = block_sum_lanes<1, kMT, 4>(br, [&](int) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < E; ++r) s += cabs2(w[r]);
        return s;
    });
    g.beta = sqrt(tb.get(0));
    g.t_now = 0.0;
    g.xm = 1.0 / (double)M;
    const double p2 = tol * ool_pow((M + 1) / 2.72, (double)(M + 1)) * sqrt(2.0 * 3.14 * (M + 1));
    g.t_new = (1.0 / g.anorm) * ool_pow(p2 / (4.0 * g.beta * g.anorm), g.xm);
    g.t_new = round_step2(g.t_new);
    g.nstep = g.nmult = g.nrej = g.nexph = g.nscale = g.brk = g.fail = 0;
#ifdef OOPS_PHASE_TIMERS
    g.tmv = g.tred = g.texp = g.tupd = 0;
#endif
}

// Stores the state at output time k (ob: n complex FP64 per output time).
template <int E>
__device__ __forceinline__ void med_emit(double2* ob, int n, int k, const cd (&x)[E]) {
#pragma unroll
    for (int r = 0; r < E; ++r)
        ob[(int64_t)k * n + threadIdx.x + kMT * r] = make_double2(x[r].re, x[r].im);
}

// One sweep to t_out.  With an output store (ob != NULL) the steps are not cut
// at the output times otau[0..ont): every output inside a step is evaluated
// from that step's Krylov basis, w(tau_k) = beta V exp((tau_k - t_now) H_aug) e_1,
// and stored; the step itself runs to min(t_new, t_out) under Expokit's error
// control.  Without it the step ends at t_out as in the reference.
template <int M, int E>
__device__ void med_advance(MedSmem<M, E>& S, BlockRed<kMT>& br, MedState& g, cd (&w)[E], float2* __restrict__ Vg,
                            int n, double t_out, int max_steps, double tol, const double* otau, int ont,
                            double2* ob) {
    const int tid = threadIdx.x;
    const double delta = 1.2, gamma = 0.9;
    int k = 0;
    if (ob)
        for (; k < ont && otau[k] <= g.t_now; ++k) med_emit<E>(ob, n, k, w);
    while (g.t_now < t_out) {
        if (g.nstep >= max_steps) { g.fail |= 1; break; }
        ++g.nstep;
        cd v0[E], vcur[E], vprev[E], p[E];
        const double ib = 1.0 / g.beta;
#pragma unroll
        for (int r = 0; r < E; ++r) {
            v0[r] = cscale(w[r], ib);
            vcur[r] = v0[r];
            vprev[r] = mkcd(0.0, 0.0);
        }
        double bprev = 0.0;
        int k1 = 2, mb = M;
        bool brk = false;
        double t_step = fmin(t_out - g.t_now, g.t_new);
#pragma unroll 1
        for (int j = 0; j < M; ++j) {
            OOPS_TIC(tm0);
            med_stage<M, E>(S, vcur);
            __syncthreads();
            med_product_all<M, E>(S, p);
#pragma unroll
            for (int r = 0; r < E; ++r) p[r] = mkcd(p[r].re - bprev * vprev[r].re, p[r].im - bprev * vprev[r].im);
            ++g.nmult;
            OOPS_TOC(g, tmv, tm0);
            OOPS_TIC(tr0);
            const auto r1 = block_sum_lanes<3, kMT, 4>(br, [&](int c) -> double {
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (c == 0) s += fma(vcur[r].re, p[r].re, vcur[r].im * p[r].im);
                    else if (c == 1) s += fma(vcur[r].re, p[r].im, -vcur[r].im * p[r].re);
                    else s += cabs2(p[r]);
                }
                return s;
            });
            const cd h = mkcd(r1.get(0), r1.get(1));
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
                t_step = t_out - g.t_now;
                brk = true;
                break;
            }
            const double inv = 1.0 / beta;
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const cd pr = cfms(h, vcur[r], p[r]);
                vprev[r] = vcur[r];
                vcur[r] = cscale(pr, inv);
                Vg[(int64_t)j * n + tid + kMT * r] = make_float2((float)vcur[r].re, (float)vcur[r].im);
            }
            bprev = beta;
        }
        double avnorm = 0.0;
        if (!brk) {
            med_stage<M, E>(S, vcur);
            __syncthreads();
            med_product_all<M, E>(S, p);
            ++g.nmult;
            const auto tn = block_sum_lanes<1, kMT, 4>(br, [&](int) {
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < E; ++r) s += cabs2(p[r]);
                return s;
            });
            avnorm = sqrt(tn.get(0));
        }
        __syncthreads();
        OOPS_TIC(te0);
        // w = beta * (sum_{j<mb} F_j v_j + [k1] F_M v_M)
        const int mx = mb + (k1 - 1 > 0 ? k1 - 1 : 0);
        // the basis loads of all E elements and of two vectors go out
        // together (each element's sum keeps the order j = 0, 1, ..)
        auto combine = [&](cd (&x)[E]) {
            const cd F0 = S.F[0];
            cd acc[E];
#pragma unroll
            for (int r = 0; r < E; ++r) acc[r] = cmul(v0[r], F0);
            const float2* vb = Vg + tid;
            int j = 1;
#pragma unroll 1
            for (; j + 1 < mx; j += 2) {
                float2 fa[E], fb[E];
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    fa[r] = vb[(int64_t)(j - 1) * n + kMT * r];
                    fb[r] = vb[(int64_t)j * n + kMT * r];
                }
                const cd Fa = S.F[j], Fb = S.F[j + 1];
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    acc[r] = cfma(mkcd((double)fa[r].x, (double)fa[r].y), Fa, acc[r]);
                    acc[r] = cfma(mkcd((double)fb[r].x, (double)fb[r].y), Fb, acc[r]);
                }
            }
            if (j < mx) {
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const float2 f = vb[(int64_t)(j - 1) * n + kMT * r];
                    acc[r] = cfma(mkcd((double)f.x, (double)f.y), S.F[j], acc[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < E; ++r) x[r] = cscale(acc[r], g.beta);
        };
        // Trial step: F walks from e_1 through the output times inside the
        // step (each evaluated from this basis and stored; vcur is free here)
        // to the step end, where Expokit's error estimate decides.  After a
        // rejection the walk restarts for the shorter step; states stored
        // beyond its end are stored again by the steps that reach them.
        double err_loc = 0.0, t1 = t_out;
        int kend = k;
        for (;;) {
            t1 = t_step >= t_out - g.t_now ? t_out : g.t_now + t_step;
            kend = k;
            if (ob)
                while (kend < ont && otau[kend] < t1) ++kend;
            g.nscale += cheb_walk<M>(S, mb, k1, bprev, g.t_now, t1, otau, k, kend - k, g.nexph, [&](int i) {
                combine(vcur);
                med_emit<E>(ob, n, k + i, vcur);
            });
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
        if (ob) k = kend;
        OOPS_TOC(g, texp, te0);
        OOPS_TIC(tu0);
        combine(p);   // the state at the end of the step
#pragma unroll
        for (int r = 0; r < E; ++r) w[r] = p[r];
        {
            const auto tb = block_sum_lanes<1, kMT, 4>(br, [&](int) {
                double s = 0.0;
#pragma unroll
                for (int r = 0; r < E; ++r) s += cabs2(w[r]);
                return s;
            });
            g.beta = sqrt(tb.get(0));
        }
        g.t_new = gamma * t_step * ool_pow(t_step * tol / err_loc, g.xm);
        g.t_new = round_step2(g.t_new);
        g.t_now = t1;
        if (ob)
            for (; k < ont && otau[k] <= g.t_now; ++k) med_emit<E>(ob, n, k, w);
        OOPS_TOC(g, tupd, tu0);
    }
}

// Hahn items with a state store (a.sweepws) run as two work units, one per
// sweep: unit 2i + s integrates sweep s of item i over every output time and
// stores its states; medium_dot_kernel then forms L(t).  CPMG-N items (and
// every item without a store) run both sweeps in one unit, as in the n <= 256
// kernels.
template <int M, int E>
__global__ __launch_bounds__(kMT) void cce_hahn_medium_kernel(KrylovArgs a, float2* medws) {
    __shared__ MedSmem<M, E> S;
    const int tid = threadIdx.x;
    BlockRed<kMT> br(S.red);
    const double tol = a.tol;
    const int max_steps = a.max_steps;
    const int n = kMT * E;
    float2* const Vg = medws + (int64_t)blockIdx.x * M * n;
    const bool split = a.sweepws != nullptr;
    const int nunits = split ? 2 * a.nitems : a.nitems;
    // flip-mask form: four generator tables per workgroup (A, B, and C, D of CPMG-N)
    const bool slot = a.medtab != nullptr;
    const int64_t gstride = (int64_t)a.ms64 * n + ((int64_t)a.ms32 * n + 1) / 2;   // double2 units
    double2* const wtab = slot ? a.medtab + (int64_t)blockIdx.x * 4 * gstride : nullptr;
    // make generator qo (table index opi) the active one; ends with a barrier
    auto activate = [&](int opi, int qo) {
        if (slot) {
            if (tid == 0) {
                S.tab64 = wtab + opi * gstride;
                S.tab32 = reinterpret_cast<const float2*>(wtab + opi * gstride + (int64_t)a.ms64 * n);
                S.ns64 = a.ms64;
                S.ns32 = a.ms32;
                S.slot = 1;
            }
            if (tid < kMedMaskStride) S.mask[tid] = a.medmask[(int64_t)qo * kMedMaskStride + tid];
            __syncthreads();
        } else {
            if (tid == 0) S.slot = 0;
            med_load(a, qo, S.op);
        }
    };
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int unit = S.ints[7];
        __syncthreads();
        if (unit >= nunits) break;
        const int item = split ? unit >> 1 : unit;
        const int side = split ? unit & 1 : 0;
        const int64_t voff = a.vec_off[item];
        const ItemSeq q = item_seq(a, item);
        const bool cpmg = q.npulse > 1;
        const bool hahn = split && !cpmg && q.nt > 0 && q.nt <= a.ws_nt;
        if (split && side == 1 && !hahn) continue;   // unit 0 runs both sweeps of this item
        const int qA = a.item_opA[item], qB = a.item_opB[item];
        const int qC = cpmg ? a.item_opC[item] : qA, qD = cpmg ? a.item_opD[item] : qB;
        if (slot) {   // this unit's generator tables (each thread its own columns)
            const SlotTables tabs = slot_tables(a);
            if (hahn) {
                med_slot_build<E>(tabs, side ? qB : qA, wtab + side * gstride);
            } else {
                med_slot_build<E>(tabs, qA, wtab);
                med_slot_build<E>(tabs, qB, wtab + gstride);
                if (cpmg) {
                    med_slot_build<E>(tabs, qC, wtab + 2 * gstride);
                    med_slot_build<E>(tabs, qD, wtab + 3 * gstride);
                }
            }
        }
        cd wA[E], wB[E];
        MedState gA, gB;
        int loaded;
        double beta0A = 0.0, beta0B = 0.0;
        if (hahn) {
            // this unit's sweep only, in the active slot gA / wA
            const double2* src = side ? a.y0 : a.u0;
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const int i = tid + kMT * r;
                wA[r] = mkcd(src[voff + i].x, src[voff + i].y);
                wB[r] = mkcd(0.0, 0.0);
            }
            loaded = side ? qB : qA;
            activate(side, loaded);
            med_init<M, E>(S, br, gA, wA, a.rnd, n, a.klim, tol);
            gB = gA;
        } else {
#pragma unroll
            for (int r = 0; r < E; ++r) {
                const int i = tid + kMT * r;
                wA[r] = mkcd(a.u0[voff + i].x, a.u0[voff + i].y);
                wB[r] = mkcd(a.y0[voff + i].x, a.y0[voff + i].y);
            }
            activate(0, qA);
            med_init<M, E>(S, br, gA, wA, a.rnd, n, a.klim, tol);
            __syncthreads();
            activate(1, qB);
            med_init<M, E>(S, br, gB, wB, a.rnd, n, a.klim, tol);
            loaded = qB;
            beta0A = gA.beta;
            beta0B = gB.beta;
        }
        const double scale = a.scale[item];
        const int h = cpmg_split(q.npulse);
        double2* const ob = hahn ? a.sweepws + ((int64_t)item * 2 + side) * a.ws_nt * n : nullptr;
        const int nouter = hahn ? 1 : q.nt;
        const int s0 = hahn ? side : 0, s1 = hahn ? side + 1 : 2;
        for (int k = 0; k < nouter; ++k) {
            for (int s = s0; s < s1; ++s) {
                // the active sweep is always gA / wA; the two swap after each
                // sweep (register moves, no dynamically selected storage)
                int nseg = 1;
                if (cpmg) {   // restart from rho0 on this side
                    const double2* src = s ? a.y0 : a.u0;
#pragma unroll
                    for (int r = 0; r < E; ++r) wA[r] = mkcd(src[voff + tid + kMT * r].x, src[voff + tid + kMT * r].y);
                    gA.beta = s ? beta0B : beta0A;
                    gA.t_now = 0.0;
                    nseg = s == 0 ? h : q.npulse + 1 - h;
                }
                for (int sg = 0; sg < nseg; ++sg) {
                    int opi = s;
                    double target = hahn ? q.tau[q.nt - 1] : q.tau[k];
                    if (cpmg) {
                        const int j = s == 0 ? sg : q.npulse - sg;
                        opi = s == 0 ? ((j & 1) ? 2 : 0) : ((j & 1) ? 1 : 3);
                        target = gA.t_now + cpmg_fraction(j, q.npulse) * q.tau[k];
                    }
                    const int qo = opi == 0 ? qA : (opi == 1 ? qB : (opi == 2 ? qC : qD));
                    if (qo != loaded) {
                        __syncthreads();
                        activate(opi, qo);
                        loaded = qo;
                    }
                    med_advance<M, E>(S, br, gA, wA, Vg, n, target, max_steps, tol, q.tau, q.nt, ob);
                }
                if (!hahn) {
                    const MedState tg = gA;
                    gA = gB;
                    gB = tg;
#pragma unroll
                    for (int r = 0; r < E; ++r) {
                        const cd tw = wA[r];
                        wA[r] = wB[r];
                        wB[r] = tw;
                    }
                }
            }
            if (!hahn) {
                const auto tl = block_sum_lanes<1, kMT, 4>(br, [&](int) {
                    double s2 = 0.0;
#pragma unroll
                    for (int r = 0; r < E; ++r) s2 += fma(wB[r].re, wA[r].re, wB[r].im * wA[r].im);
                    return s2;
                });
                if (tid == 0) q.L[k] = scale * tl.get(0);
                if ((gA.fail | gB.fail) != 0) break;
            }
        }
        if (tid == 0) {
            int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
#ifdef OOPS_PHASE_TIMERS
            if (!hahn || side == 0) {   // the unit's own sweep (sweep A of an item)
                st[16] = (int)(gA.tmv >> 10);
                st[17] = (int)(gA.tred >> 10);
                st[18] = (int)(gA.texp >> 10);
                st[19] = (int)(gA.tupd >> 10);
            }
#endif
            if (hahn) {
                // each sweep its own fields; medium_dot_kernel merges 14 / 15 into 8 / 9
                st[4 * side + 0] = gA.nstep; st[4 * side + 1] = gA.nmult;
                st[4 * side + 2] = gA.nrej; st[4 * side + 3] = gA.nexph;
                reinterpret_cast<double*>(st + 10 + 2 * side)[0] = gA.anorm;
                if (side == 0) { st[8] = gA.fail | (gA.brk << 8); st[9] = gA.nscale; }
                else { st[14] = (gA.fail << 4) | (gA.brk << 9); st[15] = gA.nscale; }
            } else {
                st[0] = gA.nstep; st[1] = gA.nmult; st[2] = gA.nrej; st[3] = gA.nexph;
                st[4] = gB.nstep; st[5] = gB.nmult; st[6] = gB.nrej; st[7] = gB.nexph;
                st[8] = gA.fail | (gB.fail << 4) | (gA.brk << 8) | (gB.brk << 9);
                st[9] = gA.nscale + gB.nscale;
                reinterpret_cast<double*>(st + 10)[0] = gA.anorm;
                reinterpret_cast<double*>(st + 12)[0] = gB.anorm;
                st[14] = 0; st[15] = 0;
            }
        }
        __syncthreads();
    }
}

// L(t_k) = scale Re <state_B(t_k), state_A(t_k)> for the items whose sweeps
// ran as separate units (one workgroup per item, one wave per output time),
// and the merge of their per-sweep status fields.
__global__ __launch_bounds__(256) void medium_dot_kernel(KrylovArgs a) {
    const int item = blockIdx.x;
    const ItemSeq q = item_seq(a, item);
    if (q.npulse > 1 || q.nt <= 0 || q.nt > a.ws_nt) return;
    const int n = a.n;
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const double2* A = a.sweepws + (int64_t)item * 2 * a.ws_nt * n;
    const double2* B = A + (int64_t)a.ws_nt * n;
    const double scale = a.scale[item];
    for (int k = wv; k < q.nt; k += 256 / kWave) {
        double s = 0.0;
        for (int i = lane; i < n; i += kWave) {
            const double2 u = A[(int64_t)k * n + i], v = B[(int64_t)k * n + i];
            s = fma(v.x, u.x, fma(v.y, u.y, s));
        }
        s = wave_sum_dpp(s);
        if (lane == 0) q.L[k] = scale * s;
    }
    if (threadIdx.x == 0) {
        int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
        st[8] |= st[14];
        st[9] += st[15];
        st[14] = 0;
        st[15] = 0;
    }
}

}  // namespace

int64_t medium_workspace_float2(int grid, int m, int n) { return (int64_t)grid * m * n; }

int launch_hahn_medium(const KrylovArgs& a, float2* work, int m, int grid, hipStream_t stream) {
    if (!work || (a.n != 1024 && a.n != 4096)) return -11;
    if (a.sweepws && a.ws_nt <= 0) return -13;
    if (a.n == 1024) {
        if (m == 20) hipLaunchKernelGGL((cce_hahn_medium_kernel<20, 1>), dim3(grid), dim3(kMT), 0, stream, a, work);
        else if (m == 30) hipLaunchKernelGGL((cce_hahn_medium_kernel<30, 1>), dim3(grid), dim3(kMT), 0, stream, a, work);
        else return -12;
    } else {
        if (m == 20) hipLaunchKernelGGL((cce_hahn_medium_kernel<20, 4>), dim3(grid), dim3(kMT), 0, stream, a, work);
        else if (m == 30) hipLaunchKernelGGL((cce_hahn_medium_kernel<30, 4>), dim3(grid), dim3(kMT), 0, stream, a, work);
        else return -12;
    }
    if (a.sweepws) hipLaunchKernelGGL(medium_dot_kernel, dim3(a.nitems), dim3(256), 0, stream, a);
    return (int)hipGetLastError();
}

}  // namespace oops
