// Chebyshev-Bessel evaluation of the small exponential of the augmented
// Lanczos tridiagonal matrix, shared by the Lanczos kernels (cce_lanczos.hip,
// cce_medium.hip).  The workgroup's LDS struct provides d, e (the tridiagonal),
// F (the result) and cheb (ChebTables: the Bessel tables).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devcommon.h"

namespace oops {

// Neighbour exchange across the wave (DPP wave shifts, 64-bit as two halves):
// lane j receives lane j-1's value (shr) or lane j+1's (shl); 0 at the ends
// (bound_ctrl: a lane with no source lane reads zero, so the destination
// needs no zeroed old value).
__device__ __forceinline__ double wave_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_shl1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x130, 0xF, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Bessel tables of one batch of small exponentials (the intervals a trial
// step's walk takes through its output times): set up together by
// cheb_setup, one lane per interval, then read by cheb_series.
constexpr int kChebTables = 6;
struct ChebTables {
    double bes[kChebTables][2 * kWave];   // J_k(x_q) before normalisation, k < n[q]
    double x[kChebTables];                // x_q = dt_q r / ns_q
    double inorm[kChebTables];            // 1 / (J_0 + 2 sum J_2k) of the stored values
    double c, r, ir;                      // spectral interval of X: centre, half-width, 1 / half-width
    int n[kChebTables];                   // Miller start of interval q (0: x_q < 1e-12)
    int ns[kChebTables];                  // substeps of interval q
};

// Chebyshev-Bessel evaluation of F = exp(dt H_aug) e1 for the augmented
// matrix of the Lanczos tridiagonal T (mb x mb) plus, with the correction
// (k1 = 2), rows M (coupling -i beta_m to row M-1) and M+1 (1 from row M).
// Wave 0 runs the series on X = -i H_aug:
//   (X x)_j = -(d_j x_j + e_{j-1} x_{j-1} + e_j x_{j+1})   j < mb
//   (X x)_M = -beta_m x_{M-1},   (X x)_{M+1} = -i x_M.
// Lane j owns row j; neighbours come by DPP wave shifts.  Long intervals are
// split into ns equal substeps of dt r / ns <= 64.  J_k(x) for k < 128 lives
// in the lanes (J_k in lane k of `bl`, J_{64+k} in lane k of `bh`) and is
// read with readlane; the series stops at the first k > x with
// |J_k| < 1e-15 (the tail decays faster than geometrically there).
//
// cheb_setup: the spectral interval of X (Gershgorin; T is the same for
// every interval of a step) and the Bessel values of nq <= kChebTables
// intervals dt_of(0..nq), lane q running the Miller recurrence of interval
// q, so a step's walk through its output times costs one recurrence instead
// of one per output.  Wave 0 only; no block barrier (only wave 0 reads the
// tables).
template <int M, typename SM, typename DtOf>
__device__ void cheb_setup(SM& S, int mb, int k1, int nq, DtOf dt_of) {
    static_assert(M + 2 <= kWave, "one lane per row of the augmented matrix");
    const int lane = threadIdx.x;
    if (lane >= kWave) return;
#ifdef OOPS_PHASE_TIMERS
    const unsigned long long tb0 = clock64();
#endif
    ChebTables& T = S.cheb;
    const int j = lane & 31;
    const double dj = j < mb ? S.d[j] : 0.0;
    const double ej = j + 1 < mb ? S.e[j] : 0.0;
    const double ejm = (j >= 1 && j < mb) ? S.e[j - 1] : 0.0;
    // Gershgorin interval of T, then of X = -T (+ 0 for the augmentation)
    double lo = j < mb ? dj - fabs(ejm) - fabs(ej) : 1e300;
    double hi = j < mb ? dj + fabs(ejm) + fabs(ej) : -1e300;
    lo = wave_reduce_dpp(lo, [](double a, double b) { return fmin(a, b); });
    hi = wave_reduce_dpp(hi, [](double a, double b) { return fmax(a, b); });
    double xlo = -hi, xhi = -lo;
    if (k1) { xlo = fmin(xlo, 0.0); xhi = fmax(xhi, 0.0); }
    const double c = 0.5 * (xlo + xhi);
    const double r = fmax(0.5 * (xhi - xlo), 1e-300);
    // interval q of this lane: substeps of dt r / ns <= 64 (the Bessel values
    // below stay within 2e-15 of J_k for x <= 64 with at most 128 recurrence
    // steps)
    const bool mine = lane < nq;
    const double xfull = mine ? dt_of(lane) * r : 0.0;
    const int ns = xfull > 64.0 ? (int)ceil(xfull / 64.0) : 1;
    const double x = xfull / ns;
    int N = 0;
    if (x >= 1e-12) {
        const float xf = (float)fmax(x, 1.0);
        // start index: the J_k the series uses (k up to about x + 3.4 x^(1/3))
        // come out within 2e-15 of J_k for every x <= 64 (checked against
        // scipy's jv on a dense grid of x)
        N = x < 1e-3 ? 20 : (int)ceilf((float)x + 6.0f * __expf(__logf(xf) * (1.0f / 3.0f)) + 20.0f);
        N = min((N + 3) & ~3, 2 * kWave);
    }
    if (lane == 0) {
        T.c = c;
        T.r = r;
        T.ir = 1.0 / r;
    }
    if (mine) {
        T.x[lane] = x;
        T.ns[lane] = ns;
        T.n[lane] = N;
    }
    // Bessel J_k(x) by backward recurrence (Miller) from N, normalised with
    // J_0 + 2 sum J_2k = 1; four steps per iteration, lane q from its own N
    // (iteration s takes k = N - 4s).  From 1e-280 the iterates stay below
    // 1e-15 for every x in [1e-12, 64] with these N (checked on a dense grid
    // of x), so no rescaling is needed.
    const int nit = __builtin_amdgcn_readfirstlane(
        (int)wave_reduce_dpp((double)(N / 4), [](double a, double b) { return fmax(a, b); }));
    double* row = T.bes[mine ? lane : 0];
    const double tx = N ? 2.0 / x : 0.0;
    double jp = 0.0, jc = 1e-280, even = 0.0;
#pragma unroll 1
    for (int it = 0; it < nit; ++it) {
        const int k = N - 4 * it;
        const double j1 = (double)k * tx * jc - jp;          // J_{k-1}
        const double j2 = (double)(k - 1) * tx * j1 - jc;    // J_{k-2}
        const double j3 = (double)(k - 2) * tx * j2 - j1;    // J_{k-3}
        const double j4 = (double)(k - 3) * tx * j3 - j2;    // J_{k-4}
        if (mine && k >= 4) {
            row[k - 1] = j1;
            row[k - 2] = j2;
            row[k - 3] = j3;
            row[k - 4] = j4;
            even += j2 + j4;
            jp = j3;
            jc = j4;
        }
    }
    if (mine && N) T.inorm[lane] = 1.0 / (2.0 * even - jc);   // jc = J_0
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#ifdef OOPS_PHASE_TIMERS
    if (lane == 0) S.tbes += clock64() - tb0;
#endif
}

// F = exp(dt H_aug) e1 (from_f: exp(dt H_aug) F, e.g. to walk a step's output
// times one interval at a time) with table q of the last cheb_setup, whose
// interval q was dt.  Called by every thread; ends with a block barrier.
// Returns the number of series terms.
template <int M, typename SM>
__device__ int cheb_series(SM& S, int mb, int k1, double beta_m, int q, double dt, bool from_f) {
    const int tid = threadIdx.x;
    int terms = 0;
    if (tid < kWave) {
#ifdef OOPS_PHASE_TIMERS
        const unsigned long long ts0 = clock64();
#endif
        const ChebTables& T = S.cheb;
        const int lane = tid;
        const int j = tid & 31;          // row of the augmented matrix
        const bool imh = tid >= 32;      // this lane carries the imaginary part of row j
        const int rows = k1 ? M + 2 : mb;
        const double dj = j < mb ? S.d[j] : 0.0;
        const double ej = j + 1 < mb ? S.e[j] : 0.0;                 // coupling to row j+1 inside T
        const double ejm = (j >= 1 && j < mb) ? S.e[j - 1] : 0.0;    // coupling to row j-1 inside T
        const double c = T.c, ir = T.ir;
        // rows of 2S = 2(X - c) / r owned by this lane (the factor 2 of the
        // Chebyshev recurrence folded in): diagonal, the coupling to x_{j+1}
        // and a complex coupling to x_{j-1} (T inside, -beta_m in row M,
        // -i in row M+1)
        // Row M+1 is carried as y_{M+1} = i x_{M+1} (similarity with
        // diag(1, .., 1, i)): its coupling -i x_M becomes the real 1 x_M, so S
        // is real, T_k(S) maps real vectors to real vectors, and the real and
        // imaginary parts of the iterate run as two independent real
        // recurrences in the two halves of the wave.
        const double sdiag = 2.0 * (j < mb ? (-dj - c) * ir : (j < rows ? -c * ir : 0.0));
        const double sup = -2.0 * ej * ir;
        const double slo = 2.0 * (j < mb ? -ejm * ir
                                         : ((k1 && j == M) ? -beta_m * ir : ((k1 && j == M + 1) ? ir : 0.0)));
        const int ns = T.ns[q];
        const double x = T.x[q];
        const int N = T.n[q];
        double bl = 0.0, bh = 0.0;
        int K = 2;
        if (N == 0) {   // exp(i x s) = 1 + i x s to double precision
            bl = lane == 0 ? 1.0 : (lane == 1 ? 0.5 * x : 0.0);
        } else {
            const double inorm = T.inorm[q];
            bl = lane < N ? T.bes[q][lane] * inorm : 0.0;
            bh = lane + kWave < N ? T.bes[q][lane + kWave] * inorm : 0.0;
            const unsigned long long sl = __ballot((double)lane > x && fabs(bl) < 1e-15);
            const unsigned long long sh = __ballot((double)(lane + kWave) > x && fabs(bh) < 1e-15);
            K = sl ? __ffsll((long long)sl) - 1 : (sh ? kWave + __ffsll((long long)sh) - 1 : N);
            if (K < 2) K = 2;
        }
        K = __builtin_amdgcn_readfirstlane(K);   // uniform (scalar loop control below)
        // 2 J_k for k >= 1 (J_0 as is): the series coefficients
        if (lane != 0) bl *= 2.0;
        bh *= 2.0;
        auto apply2 = [&](double t) -> double {   // 2 S t (one real component per lane)
            return fma(sdiag, t, fma(slo, wave_shr1(t), sup * wave_shl1(t)));
        };
        // 2 S t - tm, the diagonal part first: only two fmas wait for the
        // neighbour exchange
        auto apply2m = [&](double t, double tm) -> double {
            return fma(sup, wave_shl1(t), fma(slo, wave_shr1(t), fma(sdiag, t, -tm)));
        };
        auto other = [&](double t) -> double {    // the value of the other half (lane ^ 32)
            const long long b = __double_as_longlong(t);
            const auto l = __builtin_amdgcn_permlane32_swap((int)b, (int)b, false, false);
            const auto u = __builtin_amdgcn_permlane32_swap((int)(b >> 32), (int)(b >> 32), false, false);
            const double e = __longlong_as_double(((long long)u[0] << 32) | (unsigned)l[0]);
            const double o = __longlong_as_double(((long long)u[1] << 32) | (unsigned)l[1]);
            return imh ? e : o;
        };
        const double sgn = imh ? 1.0 : -1.0;
        double v = lane == 0 ? 1.0 : 0.0;          // e1 (real part of row 0)
        if (from_f) {   // this lane's component of F (row M+1 carried as i x_{M+1})
            const cd f = j < rows ? S.F[j] : mkcd(0.0, 0.0);
            v = (k1 && j == M + 1) ? (imh ? f.re : -f.im) : (imh ? f.im : f.re);
        }
        const cd ph = ool_expi(c * (dt / ns));     // e^{i c dt / ns}
        for (int sub = 0; sub < ns; ++sub) {
            // sum_k c_k i^k T_k v = E + i O with E the even, O the odd terms
            // (signs of i^k folded in), T_{k+1} = 2 S T_k - T_{k-1}
            double tm = v;                        // T_0(S) v
            double tk = 0.5 * apply2(v);          // T_1(S) v
            double E = lane_val(bl, 0) * tm;
            double O = lane_val(bl, 1) * tk;
            auto step = [&](double bk, int R) {   // i^k = (1, i, -1, -i)[R]
                const double tn = apply2m(tk, tm);
                tm = tk;
                tk = tn;
                if (R == 0) E = fma(bk, tk, E);
                else if (R == 1) O = fma(bk, tk, O);
                else if (R == 2) E = fma(-bk, tk, E);
                else O = fma(-bk, tk, O);
            };
            auto tail = [&](double b, int k) {    // one term, i^k by selects (no indexed accumulator)
                const double tn = apply2m(tk, tm);
                tm = tk;
                tk = tn;
                const int R = k & 3;
                const double bk = (R & 2) ? -b : b;
                const double eb = (R & 1) ? 0.0 : bk;
                const double ob = (R & 1) ? bk : 0.0;
                E = fma(eb, tk, E);
                O = fma(ob, tk, O);
            };
            // terms k < 64 take J_k from bl, the others from bh: two loops
            // with no per-term branch on the coefficient's register (the
            // skipped fma of the tail adds an exact zero, so the sum is the
            // same whichever loop takes a term)
            int k = 2;
            const int K1 = min(K, kWave);
            for (; k + 3 < K1; k += 4) {
                step(lane_val(bl, k), 2);
                step(lane_val(bl, k + 1), 3);
                step(lane_val(bl, k + 2), 0);
                step(lane_val(bl, k + 3), 1);
            }
            for (; k < K1; ++k) tail(lane_val(bl, k), k);
            for (; k + 3 < K; k += 4) {   // k = 0 mod 4 from here
                step(lane_val(bh, k - kWave), 0);
                step(lane_val(bh, k - kWave + 1), 1);
                step(lane_val(bh, k - kWave + 2), 2);
                step(lane_val(bh, k - kWave + 3), 3);
            }
            for (; k < K; ++k) tail(lane_val(bh, k - kWave), k);
            // (E_re + i O_re) + i (E_im + i O_im): re = E_re - O_im, im = E_im + O_re
            const double r = fma(sgn, other(O), E);
            // times e^{i c dt / ns}
            v = fma(ph.re, r, sgn * ph.im * other(r));
            if (j >= rows) v = 0.0;
        }
        if (j < rows) {
            double* F = reinterpret_cast<double*>(&S.F[j]);
            if (k1 && j == M + 1) F[imh ? 0 : 1] = imh ? v : -v;   // x_{M+1} = -i y_{M+1}
            else F[imh ? 1 : 0] = v;
        }
        terms = K * ns;
#ifdef OOPS_PHASE_TIMERS
        if (tid == 0) S.tser += clock64() - ts0;
#endif
    }
    __syncthreads();
    return terms;
}

// One exponential: F = exp(tau H_aug) e1 (or F with from_f).
template <int M, typename SM>
__device__ int cheb_coeffs(SM& S, int mb, int k1, double tau, double beta_m, bool from_f = false) {
    cheb_setup<M>(S, mb, k1, 1, [&](int) { return tau; });
    return cheb_series<M>(S, mb, k1, beta_m, 0, tau, from_f);
}

// The exponentials of one trial step's walk: F walks from e_1 at t0 through
// the nint output times otau[k0 .. k0 + nint) to t1, emit(i) running (every
// thread) while F holds the coefficients of output k0 + i; on return F holds
// those of t1.  The Bessel tables of up to kChebTables intervals are set up
// together.  Counts the exponentials in nexph; returns the series terms.
template <int M, typename SM, typename Emit>
__device__ int cheb_walk(SM& S, int mb, int k1, double beta_m, double t0, double t1, const double* otau, int k0,
                         int nint, int& nexph, Emit emit) {
    auto dt_of = [&](int i) -> double {   // interval i: to output k0 + i, the last one to t1
        const double from = i == 0 ? t0 : otau[k0 + i - 1];
        const double to = i < nint ? otau[k0 + i] : t1;
        return to - from;
    };
    int terms = 0;
    for (int i = 0; i <= nint; ++i) {
        if (i % kChebTables == 0)
            cheb_setup<M>(S, mb, k1, min(kChebTables, nint + 1 - i), [&](int q) { return dt_of(i + q); });
        if (i > 0) __syncthreads();   // every thread has read F
        ++nexph;
        terms += cheb_series<M>(S, mb, k1, beta_m, i % kChebTables, dt_of(i), i > 0);
        if (i < nint) {
#ifdef OOPS_PHASE_TIMERS
            const unsigned long long te0 = clock64();
#endif
            emit(i);
#ifdef OOPS_PHASE_TIMERS
            if (threadIdx.x == 0) S.temit += clock64() - te0;
#endif
        }
    }
    return terms;
}

}  // namespace oops
