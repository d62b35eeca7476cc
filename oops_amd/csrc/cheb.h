// Chebyshev-Bessel evaluation of the small exponential of the augmented
// Lanczos tridiagonal matrix, shared by the Lanczos kernels (cce_lanczos.hip,
// cce_medium.hip).  The workgroup's LDS struct provides d, e (the tridiagonal),
// F (the result) and bes (scratch for the Bessel values).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "devcommon.h"

namespace oops {

// Neighbour exchange across the wave (DPP wave shifts, 64-bit as two halves):
// lane j receives lane j-1's value (shr) or lane j+1's (shl); 0 at the ends.
__device__ __forceinline__ double wave_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double wave_shl1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Coefficients of one trial step tau: F = exp(tau H_aug) e1 for the
// augmented matrix of the Lanczos tridiagonal T (mb x mb) plus, with the
// correction (k1 = 2), rows M (coupling -i beta_m to row M-1) and M+1 (1 from
// row M).  Wave 0 runs the Chebyshev-Bessel series on X = -i H_aug:
//   (X x)_j = -(d_j x_j + e_{j-1} x_{j-1} + e_j x_{j+1})   j < mb
//   (X x)_M = -beta_m x_{M-1},   (X x)_{M+1} = -i x_M.
// Lane j owns row j; neighbours come by DPP wave shifts.  Long steps are split
// into ns equal substeps of tau r / ns <= 64.  J_k(x) for k < 128 lives in the
// lanes (J_k in lane k of `bl`, J_{64+k} in lane k of `bh`) and is read with
// readlane; the series stops at the first k > x with |J_k| < 1e-15 (the tail
// decays faster than geometrically there).  Returns the number of terms.
// from_f: start from the vector in F (exp(tau H_aug) F, e.g. to walk a step's
// output times one interval at a time) instead of e_1.
template <int M, typename SM>
__device__ int cheb_coeffs(SM& S, int mb, int k1, double tau, double beta_m, bool from_f = false) {
    static_assert(M + 2 <= kWave, "one lane per row of the augmented matrix");
    const int tid = threadIdx.x;
    int terms = 0;
    if (tid < kWave) {
#ifdef OOPS_PHASE_TIMERS
        const unsigned long long tb0 = clock64();
#endif
        const int lane = tid;
        const int j = tid & 31;          // row of the augmented matrix
        const bool imh = tid >= 32;      // this lane carries the imaginary part of row j
        const int rows = k1 ? M + 2 : mb;
        const double dj = j < mb ? S.d[j] : 0.0;
        const double ej = j + 1 < mb ? S.e[j] : 0.0;                 // coupling to row j+1 inside T
        const double ejm = (j >= 1 && j < mb) ? S.e[j - 1] : 0.0;    // coupling to row j-1 inside T
        // Gershgorin interval of T, then of X = -T (+ 0 for the augmentation)
        double lo = j < mb ? dj - fabs(ejm) - fabs(ej) : 1e300;
        double hi = j < mb ? dj + fabs(ejm) + fabs(ej) : -1e300;
        lo = wave_reduce_dpp(lo, [](double a, double b) { return fmin(a, b); });
        hi = wave_reduce_dpp(hi, [](double a, double b) { return fmax(a, b); });
        double xlo = -hi, xhi = -lo;
        if (k1) { xlo = fmin(xlo, 0.0); xhi = fmax(xhi, 0.0); }
        const double c = 0.5 * (xlo + xhi);
        const double r = fmax(0.5 * (xhi - xlo), 1e-300);
        const double ir = 1.0 / r;
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
        const double xfull = tau * r;
        // substeps of tau r / ns <= 64 (the Bessel values below stay within
        // 2e-15 of J_k for x <= 64 with at most 128 recurrence steps)
        const int ns = xfull > 64.0 ? (int)ceil(xfull / 64.0) : 1;
        const double x = xfull / ns;
        // Bessel J_k(x) by backward recurrence (Miller) from N, normalised with
        // J_0 + 2 sum J_2k = 1.  Every lane runs the recurrence (four steps per
        // iteration, the parity of k fixed); lane 0 stores J_k in LDS and every
        // lane then keeps J_j (bl) and J_{64+j} (bh).
        double bl = 0.0, bh = 0.0;
        int K = 2;
        if (x < 1e-12) {   // exp(i x s) = 1 + i x s to double precision
            bl = lane == 0 ? 1.0 : (lane == 1 ? 0.5 * x : 0.0);
        } else {
            const float xf = (float)fmax(x, 1.0);
            int N = x < 1e-3 ? 20 : (int)ceilf((float)x + 10.0f * __expf(__logf(xf) * (1.0f / 3.0f)) + 32.0f);
            N = min((N + 3) & ~3, 2 * kWave);
            double jp = 0.0, jc = 1e-280, even = 0.0;
            const double tx = 2.0 / x;
            for (int k = N; k >= 4; k -= 4) {
                const double j1 = (double)k * tx * jc - jp;          // J_{k-1}
                const double j2 = (double)(k - 1) * tx * j1 - jc;    // J_{k-2}
                const double j3 = (double)(k - 2) * tx * j2 - j1;    // J_{k-3}
                const double j4 = (double)(k - 3) * tx * j3 - j2;    // J_{k-4}
                if (lane == 0) {
                    S.bes[k - 1] = j1;
                    S.bes[k - 2] = j2;
                    S.bes[k - 3] = j3;
                    S.bes[k - 4] = j4;
                }
                even += j2 + j4;
                jp = j3;
                jc = j4;
                if (fabs(jc) > 1e150) {   // rescale everything so far (four steps grow < 1e130)
                    jp *= 1e-150;
                    jc *= 1e-150;
                    even *= 1e-150;
                    if (lane == 0)
                        for (int q = k - 4; q < N; ++q) S.bes[q] *= 1e-150;
                }
            }
            const double inorm = 1.0 / (2.0 * even - jc);   // jc = J_0
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            bl = lane < N ? S.bes[lane] * inorm : 0.0;
            bh = lane + kWave < N ? S.bes[lane + kWave] * inorm : 0.0;
            const unsigned long long sl = __ballot((double)lane > x && fabs(bl) < 1e-15);
            const unsigned long long sh = __ballot((double)(lane + kWave) > x && fabs(bh) < 1e-15);
            K = sl ? __ffsll((long long)sl) - 1 : (sh ? kWave + __ffsll((long long)sh) - 1 : N);
            if (K < 2) K = 2;
        }
#ifdef OOPS_PHASE_TIMERS
        if (j == 0) S.tbes += clock64() - tb0;
#endif
        // 2 J_k for k >= 1 (J_0 as is): the series coefficients
        if (lane != 0) bl *= 2.0;
        bh *= 2.0;
        auto bes = [&](int k) -> double { return k < kWave ? lane_val(bl, k) : lane_val(bh, k - kWave); };
        auto apply2 = [&](double t) -> double {   // 2 S t (one real component per lane)
            return fma(sdiag, t, fma(slo, wave_shr1(t), sup * wave_shl1(t)));
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
        const cd ph = ool_expi(c * (tau / ns));    // e^{i c tau / ns}
        for (int sub = 0; sub < ns; ++sub) {
            // sum_k c_k i^k T_k v = E + i O with E the even, O the odd terms
            // (signs of i^k folded in), T_{k+1} = 2 S T_k - T_{k-1}
            double tm = v;                        // T_0(S) v
            double tk = 0.5 * apply2(v);          // T_1(S) v
            double E = bes(0) * tm;
            double O = bes(1) * tk;
            auto step = [&](int k, int R) {       // i^k = (1, i, -1, -i)[R]
                const double tn = apply2(tk) - tm;
                tm = tk;
                tk = tn;
                const double bk = bes(k);
                if (R == 0) E = fma(bk, tk, E);
                else if (R == 1) O = fma(bk, tk, O);
                else if (R == 2) E = fma(-bk, tk, E);
                else O = fma(-bk, tk, O);
            };
            int k = 2;
            for (; k + 3 < K; k += 4) {
                step(k, 2);
                step(k + 1, 3);
                step(k + 2, 0);
                step(k + 3, 1);
            }
            for (; k < K; ++k) {   // the tail, i^k by selects (no indexed accumulator)
                const double tn = apply2(tk) - tm;
                tm = tk;
                tk = tn;
                const int R = k & 3;
                const double bk = (R & 2) ? -bes(k) : bes(k);
                const double eb = (R & 1) ? 0.0 : bk;
                const double ob = (R & 1) ? bk : 0.0;
                E = fma(eb, tk, E);
                O = fma(ob, tk, O);
            }
            // (E_re + i O_re) + i (E_im + i O_im): re = E_re - O_im, im = E_im + O_re
            const double r = fma(sgn, other(O), E);
            // times e^{i c tau / ns}
            v = fma(ph.re, r, sgn * ph.im * other(r));
            if (j >= rows) v = 0.0;
        }
        if (j < rows) {
            double* F = reinterpret_cast<double*>(&S.F[j]);
            if (k1 && j == M + 1) F[imh ? 0 : 1] = imh ? v : -v;   // x_{M+1} = -i y_{M+1}
            else F[imh ? 1 : 0] = v;
        }
        terms = K * ns;
    }
    __syncthreads();
    return terms;
}

}  // namespace oops
