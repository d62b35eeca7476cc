// Factored propagator powers for the CPMG-N (and Hahn) coherence of
// unpolarised 13C clusters of up to four spins on uniform time grids.
//
// Without dephasing the Liouvillians of the reference's ensemble CCE
// (ensemble_cce.cpp:41-42: lvA = H0# - H1b, lvB = H1# - H0b, the SHARP/FLAT
// expansion of KronProd.cpp:9-28) act on the density matrix from both sides:
//   exp(-i lvA s) rho = e^{i H1 s} rho e^{-i H0 s},
//   exp(-i lvB s) rho = e^{i H0 s} rho e^{-i H1 s},
// so the dense Liouville propagator the reference's production path applies
// (PiecewiseFullMatrixVectorEvolution, QuantumEvolutionAlgorithm.cpp:60-83)
// is the Kronecker product of two D x D Hilbert propagators.  The CPMG-N
// sequence (Pulse_Interval / riffle, misc.cpp:59-84: N + 1 segments with
// fractions 1/(2N) at both ends and 1/N inside, lvA on even segments) maps
// rho0 = I / D to U rho0 W^+ with
//   U = V1(s_N) .. V0(s_1) V1(s_0)  (even N; V_a(s) = e^{i H_a s}),
// W the same chain with H0 and H1 exchanged, and the reference's
// L = Re(D <rho0, rho(t)>) (ensemble_cce.cpp:46-55) becomes
//   L(t) = Re tr(W^+ U) / D.
// On a uniform grid t_k = k dt every segment is a power of P_a = V_a(delta),
// delta = dt / (2N) (the Hahn echo: dt), so one output time costs a few D x D
// products: the running powers Q_a = P_a^k, their squares, the cells
//   C = Q1 Q0^2 Q1 and C' = Q0 Q1^2 Q0,
// and for even N  U = C^{N/2}, W = C'^{N/2};  for odd N
// U = Q0 Q1 C^{(N-1)/2}, W = Q1 Q0 C'^{(N-1)/2}.
//
// One 256-thread workgroup per cluster (work queue): couplings from the
// coordinates, H0 and H1 in LDS (thread (r, c) sums its entry over the
// one-spin and pair terms), P_a by Pade(6) with scaling and squaring, then the
// products one entry per thread.  Everything in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "couplings.h"
#include "devcommon.h"
#include "oops_kernels.h"

namespace oops {

namespace {

constexpr int kFT = 256;      // threads per workgroup: one entry of a 16 x 16 matrix each
constexpr int kFD = 16;       // largest Hilbert dimension (four spin-1/2)

struct FacSmem {
    cd M[2][kFD * kFD];       // i H_a
    cd P[2][kFD * kFD];       // P_a = exp(delta i H_a)
    cd Q[2][kFD * kFD];       // Q_a = P_a^k
    cd Q2[2][kFD * kFD];      // Q_a^2
    cd C[2][kFD * kFD];       // cells (ket side, bra side)
    cd X[2][kFD * kFD];       // chains U (ket side) and W (bra side)
    cd T[4][kFD * kFD];       // temporaries; the Pade workspace at set-up
    double r[4][3];
    double J[6][9];
    double f[2][4][3];        // NV field on each spin for both NV states
    double red[160];
    int ints[8];
};

// Entry (a, b) of the spin-1/2 matrix S_c (c = x, y, z), basis (up, down).
__device__ __forceinline__ cd spin_half(int c, int a, int b) {
    if (c == 0) return mkcd(a != b ? 0.5 : 0.0, 0.0);
    if (c == 1) return mkcd(0.0, a == b ? 0.0 : (a == 0 ? -0.5 : 0.5));
    return mkcd(a == b ? (a == 0 ? 0.5 : -0.5) : 0.0, 0.0);
}

// Entry (row, col) of H_a for k spins (spin 0 the most significant digit):
// sum_s sum_c (zee_c + f_a[s]_c) S_c^(s) + sum_{s<t} sum_pq J_st[3p+q] S_p^(s) S_q^(t)
__device__ cd ham_entry(const FactoredArgs& a, const FacSmem& S, int which, int k, int row, int col) {
    cd h = mkcd(0.0, 0.0);
    const int diff = row ^ col;
    for (int s = 0; s < k; ++s) {
        const int sh = k - 1 - s;
        if (diff & ~(1 << sh)) continue;   // other spins must agree
        const int rs = (row >> sh) & 1, cs = (col >> sh) & 1;
        for (int c = 0; c < 3; ++c) {
            const double coef = a.zee[c] + S.f[which][s][c];
            h = cfma(mkcd(coef, 0.0), spin_half(c, rs, cs), h);
        }
    }
    int p = 0;
    for (int s = 0; s < k; ++s)
        for (int t = s + 1; t < k; ++t, ++p) {
            const int ss = k - 1 - s, st = k - 1 - t;
            if (diff & ~((1 << ss) | (1 << st))) continue;
            const int rs = (row >> ss) & 1, cs = (col >> ss) & 1;
            const int rt = (row >> st) & 1, ct = (col >> st) & 1;
            for (int pa = 0; pa < 3; ++pa) {
                const cd x = spin_half(pa, rs, cs);
                if (x.re == 0.0 && x.im == 0.0) continue;
                for (int qb = 0; qb < 3; ++qb)
                    h = cfma(cmul(mkcd(S.J[p][3 * pa + qb], 0.0), x), spin_half(qb, rt, ct), h);
            }
        }
    return h;
}

// C = A B for d x d matrices (column-major, leading dimension d), one entry
// per thread; ends with a barrier.
__device__ __forceinline__ void mm(cd* C, const cd* A, const cd* B, int d) {
    const int tid = threadIdx.x;
    if (tid < d * d) {
        const int i = tid % d, j = tid / d;
        cd s0 = mkcd(0.0, 0.0), s1 = mkcd(0.0, 0.0);
        int l = 0;
        for (; l + 1 < d; l += 2) {
            s0 = cfma(A[i + d * l], B[l + d * j], s0);
            s1 = cfma(A[i + d * (l + 1)], B[l + 1 + d * j], s1);
        }
        if (l < d) s0 = cfma(A[i + d * l], B[l + d * j], s0);
        C[tid] = cadd(s0, s1);
    }
    __syncthreads();
}

// C^e (e >= 1) by binary powering over the three buffers b0, b1, b2 (none
// of them C); returns the buffer holding the result.
__device__ cd* mpow(cd* b0, cd* b1, cd* b2, const cd* C, int e, int d) {
    int top = 0;
    while ((e >> (top + 1)) != 0) ++top;
    cd* buf[3] = {b0, b1, b2};
    int ci = 0;
    const int tid = threadIdx.x;
    if (tid < d * d) buf[0][tid] = C[tid];
    __syncthreads();
    for (int b = top - 1; b >= 0; --b) {
        const int si = (ci + 1) % 3;
        mm(buf[si], buf[ci], buf[ci], d);
        if ((e >> b) & 1) {
            const int mi = (si + 1) % 3;   // the third buffer (neither ci nor si)
            mm(buf[mi], buf[si], C, d);
            ci = mi;
        } else {
            ci = si;
        }
    }
    return buf[ci];
}

__global__ __launch_bounds__(kFT) void cce_factored_kernel(FactoredArgs a) {
    __shared__ FacSmem S;
    const int tid = threadIdx.x;
    const int k = a.k;
    const int d = 1 << k;
    const int dd = d * d;
    const int N = a.npulse;
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const int nt = a.item_nt[item];
        const double* tau = a.tau + a.item_tau_off[item];
        double* L = a.L + a.item_L_off[item];
        // couplings: pair tensors and the NV field of both states on every spin
        if (tid < k)
            for (int c = 0; c < 3; ++c) S.r[tid][c] = a.coords[3 * a.clusters[(int64_t)item * k + tid] + c];
        __syncthreads();
        const int npair = k * (k - 1) / 2;
        if (tid < npair) {
            int i = 0, q = tid;
            while (q >= k - 1 - i) { q -= k - 1 - i; ++i; }
            dipole9(S.r[i], S.r[i + 1 + q], a.gamma_c, a.gamma_c, S.J[tid]);
        } else if (tid >= 64 && tid < 64 + k) {
            const int s = tid - 64;
            double dp[9];
            dipole9(S.r[s], a.center, a.gamma_c, a.gamma_e, dp);
            for (int rr = 0; rr < 3; ++rr) {
                double v0 = 0.0, v1 = 0.0;
                for (int c = 0; c < 3; ++c) {
                    v0 += dp[rr + 3 * c] * a.sv0[c];
                    v1 += dp[rr + 3 * c] * a.sv1[c];
                }
                S.f[0][s][rr] = v0;
                S.f[1][s][rr] = v1;
            }
        }
        __syncthreads();
        // M_a = i H_a (column-major)
        if (tid < dd) {
            const int row = tid % d, col = tid / d;
            for (int w = 0; w < 2; ++w) {
                const cd h = ham_entry(a, S, w, k, row, col);
                S.M[w][tid] = mkcd(-h.im, h.re);
            }
        }
        __syncthreads();
        // uniform grid t_j = j dt (checked on the host); segment unit delta
        const double dt = nt >= 2 ? tau[nt - 1] / (double)(nt - 1) : 0.0;
        const double delta = N == 1 ? dt : dt / (2.0 * N);
        int fail = 0;
        for (int w = 0; w < 2 && nt >= 2; ++w) {
            int ns = 0;
            const int which = lds_pade6<kFT>(S.M[w], d, d, delta, S.T[0], S.T[1], S.T[2], S.T[3], S.red, S.ints, ns);
            if (which == -1) {   // H = 0: P = I
                if (tid < dd) S.P[w][tid] = mkcd((tid % d) == (tid / d) ? 1.0 : 0.0, 0.0);
            } else if (which < 0) {
                fail = 2;
            } else if (tid < dd) {
                S.P[w][tid] = S.T[which][tid];
            }
            __syncthreads();
        }
        if (tid < dd)
            for (int w = 0; w < 2; ++w) S.Q[w][tid] = mkcd((tid % d) == (tid / d) ? 1.0 : 0.0, 0.0);
        __syncthreads();
        const int half = N / 2;                 // cell power
        const int e = (N & 1) ? (N - 1) / 2 : half;
        for (int j = 0; j < nt && !fail; ++j) {
            if (j > 0) {
                // Q_a <- Q_a P_a
                mm(S.T[0], S.Q[0], S.P[0], d);
                mm(S.T[1], S.Q[1], S.P[1], d);
                if (tid < dd) {
                    S.Q[0][tid] = S.T[0][tid];
                    S.Q[1][tid] = S.T[1][tid];
                }
                __syncthreads();
            }
            const cd* U;
            const cd* W;
            if (N == 1) {
                mm(S.X[0], S.Q[0], S.Q[1], d);
                mm(S.X[1], S.Q[1], S.Q[0], d);
                U = S.X[0];
                W = S.X[1];
            } else {
                mm(S.Q2[0], S.Q[0], S.Q[0], d);
                mm(S.Q2[1], S.Q[1], S.Q[1], d);
                mm(S.T[0], S.Q[1], S.Q2[0], d);
                mm(S.C[0], S.T[0], S.Q[1], d);     // Q1 Q0^2 Q1
                mm(S.T[0], S.Q[0], S.Q2[1], d);
                mm(S.C[1], S.T[0], S.Q[0], d);     // Q0 Q1^2 Q0
                const cd* Ue = e >= 1 ? mpow(S.X[0], S.T[0], S.T[1], S.C[0], e, d) : nullptr;
                const cd* We = e >= 1 ? mpow(S.X[1], S.T[2], S.T[3], S.C[1], e, d) : nullptr;
                if (N & 1) {
                    // U = Q0 Q1 C^e, W = Q1 Q0 C'^e (e >= 1 for odd N >= 3)
                    mm(S.Q2[0], S.Q[0], S.Q[1], d);
                    mm(S.Q2[1], S.Q[1], S.Q[0], d);
                    cd* u2 = Ue == S.X[0] ? S.T[0] : S.X[0];
                    cd* w2 = We == S.X[1] ? S.T[2] : S.X[1];
                    mm(u2, S.Q2[0], Ue, d);
                    mm(w2, S.Q2[1], We, d);
                    U = u2;
                    W = w2;
                } else {
                    U = Ue;
                    W = We;
                }
            }
            // L = Re tr(W^+ U) / D = Re sum conj(W_rc) U_rc / D
            double v[1] = {0.0};
            if (tid < dd) v[0] = fma(W[tid].re, U[tid].re, W[tid].im * U[tid].im);
            block_sum<1, kFT>(v, S.red, S.red + 128);
            if (tid == 0) L[j] = S.red[128] / (double)d;
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

}  // namespace

int launch_factored(const FactoredArgs& a, int grid, hipStream_t stream) {
    if (a.k < 1 || a.k > 4 || a.npulse < 1) return -11;
    hipLaunchKernelGGL(cce_factored_kernel, dim3(grid), dim3(kFT), 0, stream, a);
    return (int)hipGetLastError();
}

}  // namespace oops
