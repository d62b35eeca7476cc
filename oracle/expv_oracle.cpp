// ============================================================================
// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain CPU restatement of the reference's Krylov-expv path, used only by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker (or timed CPU baseline).  The product path never links or calls it.
//
// Every routine below restates a reference routine from its published
// algorithm, in code written for this project:
//   orc_dlaran   <- src/expv/source/expokit/DLARAN.f:3-91 (LAPACK DLARAN,
//                   48-bit MCG, multiplier 33952834046453)
//   orc_zgpadm   <- src/expv/source/expokit/expokit.f:762-934 (ZGPADM, the
//                   irreducible diagonal Pade(ideg) + scaling and squaring of
//                   Sidje, ACM TOMS 24(1):130-156, 1998; Moler & Van Loan)
//   orc_hamvec   <- src/expv/source/hamvec_func3.cpp:307-389 with the
//                   in-place strided local apply of kron_func_v3 (:73-96)
//   orc_expv     <- src/expv/source/hamvec_zgexpv_w_mkl_profile.f:4-538
//                   (Lanczos + power-iteration norm estimate :210-265,
//                   coefficient switch to -i*c :269-272, first step :329-334,
//                   Arnoldi with MGS :357-386, 2-corrected augmentation :401,
//                   error estimate / rejection :431-472, update :476-494,
//                   multiple output times :339-517), entered the way
//                   MatExpVector::runInexplicit does (src/source/math/MatExp.cpp:216-365)
//
// Decisions where the reference is undefined:
//   * sgn = +1 (the reference reads an uninitialised `t`, :16,:320);
//   * the Lanczos step subtracts bjm*qjm with a purely real alpha (the
//     reference passes a REAL*8 to ZAXPY's COMPLEX*16 alpha, :233);
//   * DCMPLX(DLARAN, DLARAN) draws the real part first;
//   * the iseed state is owned by the caller (the reference SAVEs it across
//     calls, :182), the power iteration and the step loop are capped.
// ============================================================================
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <atomic>
#include <limits>

typedef std::complex<double> cplx;

namespace {

// ---------------------------------------------------------------- helpers
inline double znrm2(int64_t n, const cplx* x) {
    // scaled sum of squares (the BLAS DZNRM2 contract: no overflow)
    double scale = 0.0, ssq = 1.0;
    for (int64_t i = 0; i < n; ++i) {
        const double parts[2] = {x[i].real(), x[i].imag()};
        for (double p : parts) {
            if (p != 0.0) {
                double a = std::fabs(p);
                if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
                else { ssq += (a / scale) * (a / scale); }
            }
        }
    }
    return scale * std::sqrt(ssq);
}

inline cplx zdotc(int64_t n, const cplx* x, const cplx* y) {
    cplx s(0.0, 0.0);
    for (int64_t i = 0; i < n; ++i) s += std::conj(x[i]) * y[i];
    return s;
}

// Fortran 10.0d0**k for integer k (exact powers, reciprocal for k<0)
inline double pow10i(int k) {
    double r = 1.0;
    int a = k < 0 ? -k : k;
    for (int i = 0; i < a; ++i) r *= 10.0;
    return k < 0 ? 1.0 / r : r;
}

inline double fnint(double x) {  // Fortran NINT: round half away from zero
    return x >= 0.0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5);
}

// Expokit's 2-significant-digit rounding of a step size (:333-334, :454-455, :487-488)
inline double round_step(double t) {
    // A zero error estimate makes the proposed step +inf (t*tol/0); the
    // reference then feeds LOG10(inf) to NINT, which is undefined.  Keep it
    // +inf so the next step is clamped to the next output time.
    if (!(t > 0.0) || !std::isfinite(t)) return std::numeric_limits<double>::infinity();
    const double sqr1 = std::sqrt(0.1);
    double p1 = pow10i((int)fnint(std::log10(t) - sqr1) - 1);
    return std::trunc(t / p1 + 0.55) * p1;
}

// column-major complex matmul C = A*B, all n x n with leading dims
void zgemm_nn(int n, const cplx* A, int lda, const cplx* B, int ldb, cplx* C, int ldc) {
    for (int j = 0; j < n; ++j) {
        for (int i = 0; i < n; ++i) C[i + (int64_t)ldc * j] = 0.0;
        for (int l = 0; l < n; ++l) {
            const cplx b = B[l + (int64_t)ldb * j];
            for (int i = 0; i < n; ++i) C[i + (int64_t)ldc * j] += A[i + (int64_t)lda * l] * b;
        }
    }
}

// Solve A X = B (n x n, nrhs columns) by LU with partial pivoting; A, B overwritten.
int zgesv(int n, int nrhs, cplx* A, int lda, cplx* B, int ldb) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = std::abs(A[k + (int64_t)lda * k]);
        for (int i = k + 1; i < n; ++i) {
            double a = std::abs(A[i + (int64_t)lda * k]);
            if (a > best) { best = a; p = i; }
        }
        if (best == 0.0) return k + 1;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k + (int64_t)lda * j], A[p + (int64_t)lda * j]);
            for (int j = 0; j < nrhs; ++j) std::swap(B[k + (int64_t)ldb * j], B[p + (int64_t)ldb * j]);
        }
        const cplx piv = A[k + (int64_t)lda * k];
        for (int i = k + 1; i < n; ++i) {
            const cplx f = A[i + (int64_t)lda * k] / piv;
            A[i + (int64_t)lda * k] = f;
            for (int j = k + 1; j < n; ++j) A[i + (int64_t)lda * j] -= f * A[k + (int64_t)lda * j];
            for (int j = 0; j < nrhs; ++j) B[i + (int64_t)ldb * j] -= f * B[k + (int64_t)ldb * j];
        }
    }
    for (int j = 0; j < nrhs; ++j) {
        for (int i = n - 1; i >= 0; --i) {
            cplx s = B[i + (int64_t)ldb * j];
            for (int l = i + 1; l < n; ++l) s -= A[i + (int64_t)lda * l] * B[l + (int64_t)ldb * j];
            B[i + (int64_t)ldb * j] = s / A[i + (int64_t)lda * i];
        }
    }
    return 0;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- DLARAN
// Uniform (0,1) from x <- x*33952834046453 mod 2^48 held as four 12-bit limbs
// iseed[0] (most significant) .. iseed[3] (least significant).
double orc_dlaran(int32_t* iseed) {
    const uint64_t mult = 33952834046453ull;
    const uint64_t mask48 = (1ull << 48) - 1;
    uint64_t x = ((uint64_t)iseed[0] << 36) | ((uint64_t)iseed[1] << 24) |
                 ((uint64_t)iseed[2] << 12) | (uint64_t)iseed[3];
    // 48x48-bit product modulo 2^48 without overflow: split the multiplier
    uint64_t lo = (x * (mult & 0xFFFFFFull)) & mask48;
    uint64_t hi = ((x * (mult >> 24)) & ((1ull << 24) - 1)) << 24;
    x = (lo + hi) & mask48;
    iseed[0] = (int32_t)((x >> 36) & 4095);
    iseed[1] = (int32_t)((x >> 24) & 4095);
    iseed[2] = (int32_t)((x >> 12) & 4095);
    iseed[3] = (int32_t)(x & 4095);
    return std::ldexp((double)x, -48);
}

// ---------------------------------------------------------------- ZGPADM
// exp(t*H) for the leading mx x mx block of H (leading dim ldh), written
// column-major into out (leading dim mx).  Returns 0, or <0 on a null H /
// singular denominator (the reference stops the process there).
int orc_zgpadm(int ideg, int mx, double t, const cplx* H, int ldh, cplx* out, int* ns_out) {
    const int m = mx;
    const int64_t mm = (int64_t)m * m;
    double hnorm = 0.0;
    for (int i = 0; i < m; ++i) {
        double rs = 0.0;
        for (int j = 0; j < m; ++j) rs += std::abs(H[i + (int64_t)ldh * j]);
        hnorm = std::max(hnorm, rs);
    }
    hnorm = std::fabs(t * hnorm);
    if (hnorm == 0.0) return -1;
    int ns = std::max(0, (int)(std::log(hnorm) / std::log(2.0)) + 2);
    const double scale = t / std::ldexp(1.0, ns);
    const double scale2 = scale * scale;
    std::vector<double> c(ideg + 1);
    c[0] = 1.0;
    for (int k = 1; k <= ideg; ++k)
        c[k] = c[k - 1] * (double)(ideg + 1 - k) / (double)(k * (2 * ideg + 1 - k));
    std::vector<cplx> Hd(mm), H2(mm), P(mm), Q(mm), T(mm);
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) Hd[i + (int64_t)m * j] = H[i + (int64_t)ldh * j];
    zgemm_nn(m, Hd.data(), m, Hd.data(), m, H2.data(), m);
    for (auto& x : H2) x *= scale2;
    // Horner in H2: the polynomial holding c[ideg] is "even" (degree ideg),
    // the other holds c[ideg-1]; they alternate down to c[0].
    std::fill(P.begin(), P.end(), cplx(0.0));
    std::fill(Q.begin(), Q.end(), cplx(0.0));
    for (int i = 0; i < m; ++i) { P[i + (int64_t)m * i] = c[ideg - 1]; Q[i + (int64_t)m * i] = c[ideg]; }
    bool q_turn = true;  // which polynomial takes the next coefficient
    for (int k = ideg - 1; k > 0; --k) {
        std::vector<cplx>& X = q_turn ? Q : P;
        zgemm_nn(m, X.data(), m, H2.data(), m, T.data(), m);
        for (int i = 0; i < m; ++i) T[i + (int64_t)m * i] += c[k - 1];
        X.swap(T);
        q_turn = !q_turn;
    }
    // the polynomial that did NOT take c[0] is the odd part: multiply by scale*H
    std::vector<cplx>& odd = q_turn ? Q : P;
    std::vector<cplx>& even = q_turn ? P : Q;
    zgemm_nn(m, odd.data(), m, Hd.data(), m, T.data(), m);
    for (auto& x : T) x *= scale;
    odd.swap(T);
    // R = I + 2 (even - odd)^{-1} odd  ==  (even + odd)/(even - odd)
    std::vector<cplx> D(mm), N(odd);
    for (int64_t i = 0; i < mm; ++i) D[i] = even[i] - odd[i];
    if (zgesv(m, m, D.data(), m, N.data(), m) != 0) return -2;
    for (auto& x : N) x *= 2.0;
    for (int i = 0; i < m; ++i) N[i + (int64_t)m * i] += 1.0;
    // The polynomial holding c[0] is always the even part here, so N is
    // +r(X) for every ideg (ZGPADM's sign flip for odd ideg does not arise).
    for (int k = 0; k < ns; ++k) {
        zgemm_nn(m, N.data(), m, N.data(), m, T.data(), m);
        N.swap(T);
    }
    std::memcpy(out, N.data(), sizeof(cplx) * mm);
    if (ns_out) *ns_out = ns;
    return 0;
}

// ---------------------------------------------------------------- matvec
// w = sum_t coeff[t] * (kron_b A_{t,b}) v in the reference's flattened
// term format (MatExp.cpp:228-262); spin 0 is the slowest Kronecker factor.
void orc_hamvec(int32_t nspin, int32_t nterm, const cplx* coeff, const int64_t* nbody,
                const int64_t* pos_off, const int64_t* pos, const int64_t* dims,
                const int64_t* mat_off, const cplx* mats, int64_t n, const int64_t* spin_dim,
                const cplx* v, cplx* w, cplx* w_med) {
    std::vector<int64_t> mlst(nspin), nlst(nspin);
    for (int s = 0; s < nspin; ++s) {
        int64_t a = 1, b = 1;
        for (int j = 0; j < s; ++j) a *= spin_dim[j];
        for (int j = s + 1; j < nspin; ++j) b *= spin_dim[j];
        mlst[s] = a; nlst[s] = b;
    }
    std::fill(w, w + n, cplx(0.0));
    std::vector<cplx> xs, ys;
    for (int t = 0; t < nterm; ++t) {
        std::memcpy(w_med, v, sizeof(cplx) * n);
        for (int64_t b = 0; b < nbody[t]; ++b) {
            const int64_t idx = pos_off[t] + nbody[t] - 1 - b;  // bodies in reverse
            const int64_t s = pos[idx], d = dims[idx];
            const cplx* A = mats + mat_off[idx];
            const int64_t mm_ = mlst[s], nn_ = nlst[s];
            xs.resize(d); ys.resize(d);
            for (int64_t k = 0; k < mm_ * nn_; ++k) {
                const int64_t base = (k / nn_) * d * nn_ + (k % nn_);
                for (int64_t q = 0; q < d; ++q) xs[q] = w_med[base + nn_ * q];
                for (int64_t p = 0; p < d; ++p) {
                    cplx r(0.0);
                    for (int64_t q = 0; q < d; ++q) r += xs[q] * A[p + d * q];
                    ys[p] = r;
                }
                for (int64_t p = 0; p < d; ++p) w_med[base + nn_ * p] = ys[p];
            }
        }
        const cplx c = coeff[t];
        for (int64_t i = 0; i < n; ++i) w[i] += c * w_med[i];
    }
}

struct OrcExpvStats {
    int64_t nmult, nexph, nscale, nstep, nreject, ibrkflag, mbrkdwn, iflag;
    double step_min, step_max, x_error, s_error, tbrkdwn, t_now, hump, scaled_norm, anorm;
};

// ---------------------------------------------------------------- expv
// w_seq[:, k] = exp(-i*H*ta[k]) v for ascending ta, H = sum_t coeff[t] kron A.
// Returns the iflag (0 ok, <0 bad input, 1 step cap reached).
int orc_expv(int32_t nspin, int32_t nterm, const double* coeff, const int64_t* nbody,
             const int64_t* pos_off, const int64_t* pos, const int64_t* dims,
             const int64_t* mat_off, const cplx* mats, int64_t n, const int64_t* spin_dim,
             const cplx* v, int32_t klim, int32_t nt, const double* ta, int32_t m, double tol,
             cplx* w_seq, int32_t* iseed, OrcExpvStats* st) {
    const int ideg = 6;
    const double delta = 1.2, gamma = 0.9;
    const int64_t max_steps = 50000000;
    OrcExpvStats S;
    std::memset(&S, 0, sizeof(S));
    if (m >= n || m <= 0 || m > 50) { S.iflag = -3; if (st) *st = S; return -3; }
    std::vector<cplx> cz(nterm), w(n), wmed(n);
    auto matvec = [&](const cplx* x, cplx* y) {
        orc_hamvec(nspin, nterm, cz.data(), nbody, pos_off, pos, dims, mat_off, mats, n,
                   spin_dim, x, y, wmed.data());
    };
    // ---- norm estimate: Lanczos on +H, then power iteration on T_k
    for (int t = 0; t < nterm; ++t) cz[t] = cplx(coeff[t], 0.0);
    int k = klim;
    if (k > 0) k = (int)std::min<int64_t>(k, n);
    double anorm = 0.0;
    {
        std::vector<cplx> qj(n), qjm(n, cplx(0.0));
        std::vector<double> tk((size_t)k * k, 0.0);
        for (int64_t i = 0; i < n; ++i) {
            double re = orc_dlaran(iseed);
            double im = orc_dlaran(iseed);
            qj[i] = cplx(re, im);
        }
        double bj = znrm2(n, qj.data());
        for (auto& x : qj) x *= 1.0 / bj;
        double bjm = 0.0;
        for (int j = 1; j <= k; ++j) {
            matvec(qj.data(), w.data());
            const cplx aj = zdotc(n, qj.data(), w.data());
            tk[(size_t)(j - 1) * k + (j - 1)] = aj.real();
            for (int64_t i = 0; i < n; ++i) w[i] -= aj * qj[i];
            for (int64_t i = 0; i < n; ++i) w[i] -= bjm * qjm[i];
            bj = znrm2(n, w.data());
            if (j == k) break;
            tk[(size_t)(j - 1) * k + j] = bj;
            tk[(size_t)j * k + (j - 1)] = bj;
            if (bj < tol) break;
            qjm = qj;
            for (int64_t i = 0; i < n; ++i) qj[i] = w[i] * (1.0 / bj);
            bjm = bj;
        }
        std::vector<double> qr(k), wr(k);
        for (int j = 0; j < k; ++j) { qr[j] = orc_dlaran(iseed); wr[j] = 0.0; }
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
    }
    S.anorm = anorm;
    for (int t = 0; t < nterm; ++t) cz[t] = cplx(0.0, -coeff[t]);

    // ---- Krylov integration
    int k1 = 2;
    const int mh = m + 2;
    std::vector<cplx> V((size_t)n * (m + 1)), Hm((size_t)mh * mh), F((size_t)mh * mh);
    int64_t ibrkflag = 0, mbrkdwn = m, nmult = 0, nreject = 0, nexph = 0, nscale = 0, nstep = 0;
    double tbrkdwn = 0.0, step_min = ta && nt > 0 ? std::fabs(ta[nt - 1]) : 0.0, step_max = 0.0;
    double s_error = 0.0, x_error = 0.0, t_now = 0.0, t_new = 0.0;
    double eps = 2.220446049250313e-16;
    if (tol <= eps) tol = std::sqrt(eps);
    const double rndoff = eps * anorm;
    const double break_tol = 1.0e-7;
    const double sgn = 1.0;
    std::memcpy(w.data(), v, sizeof(cplx) * n);
    double beta = znrm2(n, w.data());
    const double vnorm = beta;
    double hump = beta;
    double xm = 1.0 / (double)m;
    {
        double p2 = tol * std::pow((m + 1) / 2.72, m + 1) * std::sqrt(2.0 * 3.14 * (m + 1));
        t_new = (1.0 / anorm) * std::pow(p2 / (4.0 * beta * anorm), xm);
        t_new = round_step(t_new);
    }
    double avnorm = 0.0, err_loc = 0.0;
    int iflag = 0;
    for (int tt = 0; tt < nt; ++tt) {
        const double t_out = ta[tt];
        while (t_now < t_out) {
            if (nstep >= max_steps) { iflag = 1; break; }
            ++nstep;
            // Each step starts a fresh Arnoldi process.  The reference sets
            // k1 = 2 / mbrkdwn = m only once per call (:284, :292) and never
            // restores them after a happy breakdown (:374-376), so every later
            // step of a multi-output call runs with err_loc = tol and its step
            // size collapses; restoring them per step is the intended
            // single-output Expokit behaviour.
            k1 = 2;
            mbrkdwn = m;
            double t_step = std::min(t_out - t_now, t_new);
            const double p1 = 1.0 / beta;
            for (int64_t i = 0; i < n; ++i) V[i] = p1 * w[i];
            std::fill(Hm.begin(), Hm.end(), cplx(0.0));
            bool broke = false;
            for (int j = 1; j <= m; ++j) {
                ++nmult;
                cplx* vj = &V[(size_t)n * (j - 1)];
                cplx* vn = &V[(size_t)n * j];
                matvec(vj, vn);
                for (int i = 1; i <= j; ++i) {
                    const cplx* vi = &V[(size_t)n * (i - 1)];
                    const cplx hij = zdotc(n, vi, vn);
                    for (int64_t l = 0; l < n; ++l) vn[l] -= hij * vi[l];
                    Hm[(size_t)(j - 1) * mh + (i - 1)] = hij;
                }
                const double hj1j = znrm2(n, vn);
                if (hj1j <= break_tol) {
                    k1 = 0; ibrkflag = 1; mbrkdwn = j; tbrkdwn = t_now;
                    t_step = t_out - t_now;
                    broke = true;
                    break;
                }
                Hm[(size_t)(j - 1) * mh + j] = hj1j;
                for (int64_t l = 0; l < n; ++l) vn[l] *= 1.0 / hj1j;
            }
            if (!broke) {
                ++nmult;
                std::vector<cplx> tmp(n);
                matvec(&V[(size_t)n * m], tmp.data());
                avnorm = znrm2(n, tmp.data());
            }
            Hm[(size_t)m * mh + m + 1] = 1.0;
            int ireject = 0;
            int mx;
            for (;;) {
                ++nexph;
                mx = (int)(mbrkdwn + k1);
                int ns = 0;
                if (orc_zgpadm(ideg, mx, sgn * t_step, Hm.data(), mh, F.data(), &ns) != 0) {
                    S.iflag = -4; if (st) *st = S; return -4;
                }
                nscale += ns;
                if (k1 == 0) {
                    err_loc = tol;
                } else {
                    const double q1 = std::abs(F[m]) * beta;
                    const double q2 = std::abs(F[m + 1]) * beta * avnorm;
                    if (q1 > 10.0 * q2) { err_loc = q2; xm = 1.0 / (double)m; }
                    else if (q1 > q2) { err_loc = (q1 * q2) / (q1 - q2); xm = 1.0 / (double)m; }
                    else { err_loc = q1; xm = 1.0 / (double)(m - 1); }
                }
                if (k1 != 0 && err_loc > delta * t_step * tol) {
                    t_step = gamma * t_step * std::pow(t_step * tol / err_loc, xm);
                    t_step = round_step(t_step);
                    ++ireject; ++nreject;
                    continue;
                }
                break;
            }
            // w = beta * V(:,1:mx) * F(1:mx, 1)
            mx = (int)(mbrkdwn + std::max(0, k1 - 1));
            for (int64_t i = 0; i < n; ++i) {
                cplx s(0.0);
                for (int j = 0; j < mx; ++j) s += V[(size_t)n * j + i] * F[j];
                w[i] = beta * s;
            }
            beta = znrm2(n, w.data());
            hump = std::max(hump, beta);
            t_new = gamma * t_step * std::pow(t_step * tol / err_loc, xm);
            t_new = round_step(t_new);
            err_loc = std::max(err_loc, rndoff);
            t_now += t_step;
            step_min = std::min(step_min, t_step);
            step_max = std::max(step_max, t_step);
            s_error += err_loc;
            x_error = std::max(x_error, err_loc);
        }
        std::memcpy(w_seq + (size_t)n * tt, w.data(), sizeof(cplx) * n);
        if (iflag) break;
    }
    S.nmult = nmult; S.nexph = nexph; S.nscale = nscale; S.nstep = nstep; S.nreject = nreject;
    S.ibrkflag = ibrkflag; S.mbrkdwn = mbrkdwn; S.iflag = iflag;
    S.step_min = step_min; S.step_max = step_max; S.x_error = x_error; S.s_error = s_error;
    S.tbrkdwn = tbrkdwn; S.t_now = sgn * t_now; S.hump = hump / vnorm; S.scaled_norm = beta / vnorm;
    if (st) *st = S;
    return iflag;
}

// ---------------------------------------------------------------- batches
// One packed operator in the reference's flattened term format.
struct OrcTermTable {
    int32_t nspin, nterm;
    const double* coeff;
    const int64_t* nbody;
    const int64_t* pos_off;
    const int64_t* pos;
    const int64_t* dims;
    const int64_t* mat_off;
    const cplx* mats;
    const int64_t* spin_dim;
    int64_t n;
};

// w(t_k) = exp(-i H (t_k - t_{k-1})) w(t_{k-1}) with a dense Pade(6)
// exponential, recomputed only when the output spacing changes.
int orc_dense_sweep(const OrcTermTable& T, const cplx* v, int32_t nt, const double* ta, cplx* out) {
    const int64_t n = T.n;
    std::vector<cplx> H((size_t)n * n), e(n, cplx(0.0)), col(n), wm(n), cz(T.nterm), U((size_t)n * n), X((size_t)n * n);
    for (int t = 0; t < T.nterm; ++t) cz[t] = cplx(T.coeff[t], 0.0);
    for (int64_t j = 0; j < n; ++j) {
        e[j] = 1.0;
        orc_hamvec(T.nspin, T.nterm, cz.data(), T.nbody, T.pos_off, T.pos, T.dims, T.mat_off, T.mats, n,
                   T.spin_dim, e.data(), col.data(), wm.data());
        e[j] = 0.0;
        for (int64_t i = 0; i < n; ++i) H[i + n * j] = cplx(0.0, -1.0) * col[i];
    }
    std::vector<cplx> w(v, v + n), w2(n);
    double tprev = 0.0, dprev = -1.0;
    for (int k = 0; k < nt; ++k) {
        const double d = ta[k] - tprev;
        tprev = ta[k];
        if (d > 0.0) {
            if (d != dprev) {
                int ns = 0;
                if (orc_zgpadm(6, (int)n, d, H.data(), (int)n, U.data(), &ns) != 0) return -4;
                dprev = d;
            }
            for (int64_t i = 0; i < n; ++i) {
                cplx s(0.0);
                for (int64_t j = 0; j < n; ++j) s += U[i + n * j] * w[j];
                w2[i] = s;
            }
            w.swap(w2);
        }
        std::memcpy(out + (size_t)n * k, w.data(), sizeof(cplx) * n);
    }
    return 0;
}

// Hahn-echo coherence of a batch of clusters by two Krylov sweeps each:
//   L[c, k] = Re( scale[c] * < expv(opB[c], y0[c], ta[k]), expv(opA[c], u0[c], ta[k]) > )
// opB must already be the bra-side generator (-lvB^dagger).  Clusters with
// n <= m use orc_dense_sweep (the reference's driver stops there).  One cluster per
// thread, nthreads threads (the reference runs one cluster per MPI rank,
// src/source/app/cce.cpp:54-70).  Each expv starts from iseed = (0,0,0,5).
int orc_hahn_batch(int32_t nclu, const OrcTermTable* opA, const OrcTermTable* opB,
                   const cplx* const* u0, const cplx* const* y0, const double* scale,
                   int32_t klim, int32_t nt, const double* ta, int32_t m, double tol,
                   double* L, OrcExpvStats* stats, int32_t nthreads) {
    std::atomic<int> next(0), err(0);
    auto worker = [&]() {
        for (;;) {
            const int c = next.fetch_add(1);
            if (c >= nclu) break;
            const OrcTermTable& A = opA[c];
            const OrcTermTable& B = opB[c];
            const int64_t n = A.n;
            std::vector<cplx> ua((size_t)n * nt), yb((size_t)n * nt);
            int32_t seed[4] = {0, 0, 0, 5};
            OrcExpvStats sa, sb;
            std::memset(&sa, 0, sizeof(sa));
            std::memset(&sb, 0, sizeof(sb));
            int fa, fb;
            if (n <= m) {
                // the reference stops for m >= n; use the exact dense exponential
                fa = orc_dense_sweep(A, u0[c], nt, ta, ua.data());
                fb = orc_dense_sweep(B, y0[c], nt, ta, yb.data());
            } else {
                fa = orc_expv(A.nspin, A.nterm, A.coeff, A.nbody, A.pos_off, A.pos, A.dims,
                              A.mat_off, A.mats, n, A.spin_dim, u0[c], klim, nt, ta, m, tol,
                              ua.data(), seed, &sa);
                seed[0] = 0; seed[1] = 0; seed[2] = 0; seed[3] = 5;
                fb = orc_expv(B.nspin, B.nterm, B.coeff, B.nbody, B.pos_off, B.pos, B.dims,
                              B.mat_off, B.mats, n, B.spin_dim, y0[c], klim, nt, ta, m, tol,
                              yb.data(), seed, &sb);
            }
            if (fa != 0 || fb != 0) err.store(fa != 0 ? fa : fb);
            for (int k = 0; k < nt; ++k) {
                const cplx d = zdotc(n, &yb[(size_t)n * k], &ua[(size_t)n * k]);
                L[(size_t)c * nt + k] = (scale[c] * d).real();
            }
            if (stats) { stats[2 * c] = sa; stats[2 * c + 1] = sb; }
        }
    };
    std::vector<std::thread> th;
    if (nthreads < 1) nthreads = 1;
    for (int i = 0; i < nthreads; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
    return err.load();
}

// Dense exp(-i*dt*A) powers, the reference's production per-cluster propagator
// (src/source/quantum/QuantumEvolution/QuantumEvolutionAlgorithm.cpp:60-83):
//   state_i = prod_j U_j^i v,  U_j = exp(-i * seg[j] * dt * A_j),  i = 0..nt-1
// with the U_j from orc_zgpadm and the powers accumulated by repeated products.
// ops: nseg dense d x d column-major matrices. out: d x nt column-major.
int orc_piecewise_dense(int32_t d, int32_t nseg, const cplx* ops, const double* seg,
                        double dt, const cplx* v, int32_t nt, cplx* out) {
    const int64_t dd = (int64_t)d * d;
    std::vector<cplx> U((size_t)nseg * dd), P((size_t)nseg * dd), T(dd), A(dd);
    for (int j = 0; j < nseg; ++j) {
        // MatExp(m, -i*seg*dt).pade_exp_mat() = zgpadm(t=1, H = prefactor*m)
        const cplx pf(0.0, -seg[j] * dt);
        for (int64_t i = 0; i < dd; ++i) A[i] = pf * ops[(size_t)j * dd + i];
        int ns = 0;
        if (orc_zgpadm(6, d, 1.0, A.data(), d, &U[(size_t)j * dd], &ns) != 0) return -1;
    }
    P = U;
    std::memcpy(out, v, sizeof(cplx) * d);
    std::vector<cplx> s(d), s2(d);
    for (int i = 1; i < nt; ++i) {
        std::memcpy(s.data(), v, sizeof(cplx) * d);
        for (int j = 0; j < nseg; ++j) {
            const cplx* Pj = &P[(size_t)j * dd];
            for (int r = 0; r < d; ++r) {
                cplx acc(0.0);
                for (int c = 0; c < d; ++c) acc += Pj[r + (int64_t)d * c] * s[c];
                s2[r] = acc;
            }
            s.swap(s2);
            zgemm_nn(d, &U[(size_t)j * dd], d, &P[(size_t)j * dd], d, T.data(), d);
            std::memcpy(&P[(size_t)j * dd], T.data(), sizeof(cplx) * dd);
        }
        std::memcpy(out + (size_t)d * i, s.data(), sizeof(cplx) * d);
    }
    return 0;
}

}  // extern "C"
