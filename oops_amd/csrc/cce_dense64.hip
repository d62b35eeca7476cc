// Dense batched propagator path on the matrix cores for n = 64 (Liouville
// space of 3 spin-1/2): the reference's production per-cluster method,
// PiecewiseFullMatrixVectorEvolution (QuantumEvolutionAlgorithm.cpp:60-83),
// for a Hahn echo or CPMG-N on a uniform time grid t_k = k dt:
//   U_g = exp(-i L_g dt / (2N))           g in {A = lvA, B = lvB}
//   P_g(k) = U_g^k                        (one product per output time)
//   s(t_k) = P_B^{e_N} .. P_A^{e_0} u0    (segment j: generator A for even j,
//                                          B for odd j, applied once at the
//                                          ends (fraction 1/(2N)), twice inside)
//   L(t_k) = Re(scale <y0, s(t_k)>)
// A Hahn echo (N = 1) needs no powers: L(t_k) = Re(scale <b_k, a_k>) with
// a_k = U_A a_{k-1}, b_k = exp(-i opB dt) b_{k-1} (opB = -lvB^dagger, so
// exp(-i opB dt) = U_B^H), one 64 x 64 matrix-vector product per side and
// output time; the powers are formed for CPMG-N, whose segments interleave.
// All complex products are FP64 MFMA (v_mfma_f64_16x16x4f64): complex
// C = A B as four real products, one 16x16 output tile per (row block,
// column block), 16 rows per wave.  The exponential is a degree-18 Taylor
// polynomial of X = -i L dt/(2N) / 2^s (||X||_inf <= 1/2, truncation below
// 1e-22) followed by s squarings.  One 256-thread workgroup per cluster; the
// powers live in LDS (each wave owns 16 rows of P_A and P_B), the generator,
// Taylor and squaring matrices in a per-workgroup device workspace.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "krylov_common.h"

namespace oops {

namespace {

constexpr int kD = 64;             // state dimension handled here
constexpr int kDD = kD * kD;
constexpr int kTaylor = 18;
typedef double d4 __attribute__((ext_vector_type(4)));

// Per-workgroup device workspace (doubles): re/im planes of
//   Xc (generator, column-major), Tr/Tc x 2 (Taylor and squaring ping-pong,
//   row- and column-major copies), UAr, UBr (exponentials, row-major).
constexpr int kPlanes = 2;
constexpr int64_t kWsMats = 1 + 4 + 2;
constexpr int64_t kWsDoubles = kWsMats * kPlanes * kDD;

struct Dense64Smem {
    double pst[2][4][2][kD * 16];  // [A/B][wave][re/im][k*16 + i]: rows 16w+i of P_g, column k
    double s[2][2][kD];            // [ping-pong][re/im]
    double hb[2][2][kD];           // Hahn echo: the bra-side state [ping-pong][re/im]
    double y0[2][kD];
    OpLds op;
    double red[8];
    double scal[4];
    int ints[8];
};

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Rows 16w..16w+15 of C = A B (64 x 64 complex).  aload(i, k, re, im) gives
// A(16w + i, k); B is row-major in global memory.  Lane l holds output
// (16w + 4r + l/16, 16J + l%16) in cre[J][r] / cim[J][r].
template <class AF>
__device__ __forceinline__ void mm_rowblock(AF aload, const double* __restrict__ Bre, const double* __restrict__ Bim,
                                            d4 cre[4], d4 cim[4]) {
    const int l = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        cre[J] = d4{0.0, 0.0, 0.0, 0.0};
        cim[J] = d4{0.0, 0.0, 0.0, 0.0};
    }
#pragma unroll 2
    for (int ks = 0; ks < kD / 4; ++ks) {
        const int k = 4 * ks + (l >> 4);
        double ar, ai;
        aload(l & 15, k, ar, ai);
        const double nai = -ai;
#pragma unroll
        for (int J = 0; J < 4; ++J) {
            const int j = 16 * J + (l & 15);
            const double br = Bre[k * kD + j], bi = Bim[k * kD + j];
            cre[J] = mfma(ar, br, cre[J]);
            cre[J] = mfma(nai, bi, cre[J]);
            cim[J] = mfma(ar, bi, cim[J]);
            cim[J] = mfma(ai, br, cim[J]);
        }
    }
}

// Store a row block result (optionally scaled and plus the identity) row- and/or
// column-major into global planes.
__device__ __forceinline__ void store_block(const d4 cre[4], const d4 cim[4], int w, double scale, bool plus_eye,
                                            double* rr, double* ri, double* cr, double* ci) {
    const int l = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int J = 0; J < 4; ++J)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 16 * w + 4 * r + (l >> 4), j = 16 * J + (l & 15);
            const double vr = cre[J][r] * scale + (plus_eye && i == j ? 1.0 : 0.0);
            const double vi = cim[J][r] * scale;
            if (rr) { rr[i * kD + j] = vr; ri[i * kD + j] = vi; }
            if (cr) { cr[i + kD * j] = vr; ci[i + kD * j] = vi; }
        }
}

// U = exp(-i H tau) for the operator in S.op (row-major planes Ur/Ui), all
// four waves cooperating; ws: Xc, T0r, T0c, T1r, T1c planes.
__device__ int dense_expm(Dense64Smem& S, double tau, double* ws, double* Ur, double* Ui) {
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    double* Xr = ws;
    double* Xi = ws + kDD;
    double* Tr[2] = {ws + 2 * kDD, ws + 6 * kDD};
    double* Ti[2] = {ws + 3 * kDD, ws + 7 * kDD};
    double* Tcr[2] = {ws + 4 * kDD, ws + 8 * kDD};
    double* Tci[2] = {ws + 5 * kDD, ws + 9 * kDD};
    // dense H by rows (thread i < 64 owns row i): H x for x = e_c, read off entry (i, c)
    // from the ELL tables directly: row i couples to base(i) + offs[e] with vals[e].
    double rowsum = 0.0;
    if (tid < kD) {
        const DevOpHeader& h = S.op.hdr;
        for (int c = 0; c < kD; ++c) { Xr[tid + kD * c] = 0.0; Xi[tid + kD * c] = 0.0; }
        Xr[tid + kD * tid] += h.c0re;
        Xi[tid + kD * tid] += h.c0im;
        const int ld = h.ld_bits, lmask = (1 << ld) - 1, top = h.nspin - 1;
        for (int Lc = 0; Lc < h.nloc; ++Lc) {
            const DevOpLocal lo = S.op.locs[Lc];
            const int shs = ld * (top - lo.s);
            const int ls = (tid >> shs) & lmask;
            int base = tid - (ls << shs);
            int row = ls;
            if (lo.nb == 2) {
                const int sht = ld * (top - lo.t);
                const int lt = (tid >> sht) & lmask;
                base -= lt << sht;
                row = (ls << ld) | lt;
            }
            const int e0 = lo.ent + row * lo.R;
            for (int e = 0; e < lo.R; ++e) {
                const cd v = S.op.vals[e0 + e];
                const int c = base + S.op.offs[e0 + e];
                Xr[tid + kD * c] += v.re;
                Xi[tid + kD * c] += v.im;
            }
        }
        for (int c = 0; c < kD; ++c) rowsum += sqrt(Xr[tid + kD * c] * Xr[tid + kD * c] + Xi[tid + kD * c] * Xi[tid + kD * c]);
    }
    if (tid < kWave) {
        double m = rowsum;
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1) m = fmax(m, __shfl_xor(m, mask, kWave));
        if (tid == 0) S.scal[0] = m;
    }
    __syncthreads();
    const double nrm = S.scal[0] * fabs(tau);
    if (!(nrm < 1e300)) return -3;
    int sq = 0;
    while (ldexp(nrm, -sq) > 0.5 && sq < 60) ++sq;
    const double f = tau * ldexp(1.0, -sq);
    // X = -i H f  (in place, column-major): (a + ib)(-i f) = f b - i f a
    for (int idx = tid; idx < kDD; idx += kNT) {
        const double a = Xr[idx], b = Xi[idx];
        Xr[idx] = f * b;
        Xi[idx] = -f * a;
    }
    // T = I + X / K
    for (int idx = tid; idx < kDD; idx += kNT) {
        const int i = idx % kD, j = idx / kD;  // column-major index of X
        const double vr = Xr[idx] / kTaylor + (i == j ? 1.0 : 0.0), vi = Xi[idx] / kTaylor;
        Tr[0][i * kD + j] = vr; Ti[0][i * kD + j] = vi;
        Tcr[0][idx] = vr; Tci[0][idx] = vi;
    }
    __syncthreads();
    int cur = 0;
    d4 cre[4], cim[4];
    for (int k = kTaylor - 1; k >= 1; --k) {
        // T <- I + X T / k
        mm_rowblock([&](int i, int kk, double& ar, double& ai) {
            ar = Xr[(16 * w + i) + kD * kk];
            ai = Xi[(16 * w + i) + kD * kk];
        }, Tr[cur], Ti[cur], cre, cim);
        store_block(cre, cim, w, 1.0 / k, true, Tr[cur ^ 1], Ti[cur ^ 1], Tcr[cur ^ 1], Tci[cur ^ 1]);
        __syncthreads();
        cur ^= 1;
    }
    for (int q = 0; q < sq; ++q) {
        const double* Acr = Tcr[cur];
        const double* Aci = Tci[cur];
        mm_rowblock([&](int i, int kk, double& ar, double& ai) {
            ar = Acr[(16 * w + i) + kD * kk];
            ai = Aci[(16 * w + i) + kD * kk];
        }, Tr[cur], Ti[cur], cre, cim);
        store_block(cre, cim, w, 1.0, false, Tr[cur ^ 1], Ti[cur ^ 1], Tcr[cur ^ 1], Tci[cur ^ 1]);
        __syncthreads();
        cur ^= 1;
    }
    for (int idx = tid; idx < kDD; idx += kNT) { Ur[idx] = Tr[cur][idx]; Ui[idx] = Ti[cur][idx]; }
    __syncthreads();
    return sq;
}

}  // namespace

__global__ __launch_bounds__(kNT) void cce_dense64_kernel(KrylovArgs a, double* __restrict__ work) {
    __shared__ Dense64Smem S;
    const int tid = threadIdx.x;
    const int w = tid / kWave, l = tid & (kWave - 1);
    double* ws = work + (int64_t)blockIdx.x * kWsDoubles;
    double* UAr = ws + 10 * kDD;
    double* UAi = ws + 11 * kDD;
    double* UBr = ws + 12 * kDD;
    double* UBi = ws + 13 * kDD;
    const OpSrc src{a.hdr, a.loc, a.val, a.off};
    for (;;) {
        if (tid == 0) S.ints[7] = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = S.ints[7];
        __syncthreads();
        if (item >= a.nitems) break;
        const ItemSeq q = item_seq(a, item);
        const int64_t voff = a.vec_off[item];
        const int N = q.npulse;
        const double dt = q.nt > 1 ? q.tau[1] - q.tau[0] : 0.0;
        // Hahn echo (N = 1): the item's times are the sweep times t/2, so each
        // half of the sequence lasts tau_k = k dt; CPMG-N: times are t_k and the
        // end segments last t_k / (2N).
        const double delta = N == 1 ? dt : dt / (2.0 * N);
        // Hahn echo: L(t_k) = Re(scale <(U_B^k)^H y0, U_A^k u0>) with
        // (U_B)^H = exp(-i opB dt) (opB = -lvB^dagger), so both sides advance by
        // one matrix-vector product per output time instead of a power
        const bool hahn_vec = N == 1;
        int fail = 0;
        // exponentials of lvA and lvB (Hahn items carry -lvB^dagger: lvB = -(opB)^H)
        for (int g = 0; g < 2; ++g) {
            int opi = g == 0 ? a.item_opA[item] : (N > 1 ? a.item_opC[item] : a.item_opB[item]);
            op_load_lds<kNT, false>(src, opi, S.op);
            __syncthreads();
            const int sq = dense_expm(S, delta, ws, g == 0 ? UAr : UBr, g == 0 ? UAi : UBi);
            if (sq < 0) fail = 2;
            if (g == 1 && N == 1 && !hahn_vec) {
                // exp(-i lvB d) with lvB = -(opB)^H equals exp(-i (-opB) d)^H ... computed
                // here as exp(+i opB^H d): conjugate-transpose of exp(-i opB d).
                for (int idx = tid; idx < kDD; idx += kNT) {
                    const int i = idx / kD, j = idx % kD;
                    if (i < j) {
                        const double r1 = UBr[i * kD + j], i1 = UBi[i * kD + j];
                        const double r2 = UBr[j * kD + i], i2 = UBi[j * kD + i];
                        UBr[i * kD + j] = r2; UBi[i * kD + j] = -i2;
                        UBr[j * kD + i] = r1; UBi[j * kD + i] = -i1;
                    } else if (i == j) {
                        UBi[i * kD + i] = -UBi[i * kD + i];
                    }
                }
            }
            __syncthreads();
        }
        const double scale = a.scale[item];
        if (hahn_vec) {
            // U_A and exp(-i opB dt) into the power slots (rows 16w+i, column k)
            for (int g = 0; g < 2; ++g) {
                const double* ur = g == 0 ? UAr : UBr;
                const double* ui = g == 0 ? UAi : UBi;
                for (int idx = l; idx < kD * 16; idx += kWave) {
                    const int k = idx / 16, i = idx % 16;
                    S.pst[g][w][0][idx] = ur[(16 * w + i) * kD + k];
                    S.pst[g][w][1][idx] = ui[(16 * w + i) * kD + k];
                }
            }
            if (tid < kD) {
                S.s[0][0][tid] = a.u0[voff + tid].x;
                S.s[0][1][tid] = a.u0[voff + tid].y;
                S.hb[0][0][tid] = a.y0[voff + tid].x;
                S.hb[0][1][tid] = a.y0[voff + tid].y;
            }
            __syncthreads();
            int cur = 0;
            for (int k = 0; k < q.nt && !fail; ++k) {
                if (k > 0) {
                    // ket s <- U_A s and bra hb <- exp(-i opB dt) hb, rows 16w..16w+15
                    const int i = l & 15, kb = (l >> 4) * 16;
                    double ar = 0.0, ai = 0.0, br = 0.0, bi = 0.0;
#pragma unroll 4
                    for (int kk = kb; kk < kb + 16; ++kk) {
                        const double mr = S.pst[0][w][0][kk * 16 + i], mi = S.pst[0][w][1][kk * 16 + i];
                        const double xr = S.s[cur][0][kk], xi = S.s[cur][1][kk];
                        ar = fma(mr, xr, fma(-mi, xi, ar));
                        ai = fma(mr, xi, fma(mi, xr, ai));
                        const double nr = S.pst[1][w][0][kk * 16 + i], ni = S.pst[1][w][1][kk * 16 + i];
                        const double yr = S.hb[cur][0][kk], yi = S.hb[cur][1][kk];
                        br = fma(nr, yr, fma(-ni, yi, br));
                        bi = fma(nr, yi, fma(ni, yr, bi));
                    }
#pragma unroll
                    for (int m = 16; m <= 32; m <<= 1) {
                        ar += __shfl_xor(ar, m, kWave);
                        ai += __shfl_xor(ai, m, kWave);
                        br += __shfl_xor(br, m, kWave);
                        bi += __shfl_xor(bi, m, kWave);
                    }
                    if (l < 16) {
                        S.s[cur ^ 1][0][16 * w + i] = ar;
                        S.s[cur ^ 1][1][16 * w + i] = ai;
                        S.hb[cur ^ 1][0][16 * w + i] = br;
                        S.hb[cur ^ 1][1][16 * w + i] = bi;
                    }
                    __syncthreads();
                    cur ^= 1;
                }
                if (tid < kWave) {
                    double v = fma(S.hb[cur][0][tid], S.s[cur][0][tid], S.hb[cur][1][tid] * S.s[cur][1][tid]);
#pragma unroll
                    for (int mask = 32; mask >= 1; mask >>= 1) v += __shfl_xor(v, mask, kWave);
                    if (tid == 0) q.L[k] = scale * v;
                }
            }
            if (tid == 0) {
                int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
                for (int z = 0; z < kStatsPerItem; ++z) st[z] = 0;
                st[8] = fail;
            }
            __syncthreads();
            continue;
        }
        // P_A = P_B = I (each wave its 16 rows), s = u0, y0
        for (int g = 0; g < 2; ++g)
            for (int idx = l; idx < kD * 16; idx += kWave) {
                const int k = idx / 16, i = idx % 16;
                S.pst[g][w][0][idx] = (k == 16 * w + i) ? 1.0 : 0.0;
                S.pst[g][w][1][idx] = 0.0;
            }
        if (tid < kD) {
            S.y0[0][tid] = a.y0[voff + tid].x;
            S.y0[1][tid] = a.y0[voff + tid].y;
        }
        __syncthreads();
        for (int k = 0; k < q.nt && !fail; ++k) {
            if (k > 0) {
                // P_g <- P_g U_g for both generators (rows of this wave only)
                for (int g = 0; g < 2; ++g) {
                    double* pr = S.pst[g][w][0];
                    double* pi = S.pst[g][w][1];
                    d4 cre[4], cim[4];
                    mm_rowblock([&](int i, int kk, double& ar, double& ai) {
                        ar = pr[kk * 16 + i];
                        ai = pi[kk * 16 + i];
                    }, g == 0 ? UAr : UBr, g == 0 ? UAi : UBi, cre, cim);
#pragma unroll
                    for (int J = 0; J < 4; ++J)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int i = 4 * r + (l >> 4), j = 16 * J + (l & 15);
                            pr[j * 16 + i] = cre[J][r];
                            pi[j * 16 + i] = cim[J][r];
                        }
                }
            }
            // s = chain of the segments applied to u0
            int cur = 0;
            if (tid < kD) {
                S.s[0][0][tid] = a.u0[voff + tid].x;
                S.s[0][1][tid] = a.u0[voff + tid].y;
            }
            __syncthreads();
            for (int j = 0; j <= N; ++j) {
                const int g = j & 1;
                const int reps = (j == 0 || j == N) ? 1 : 2;
                for (int rep = 0; rep < reps; ++rep) {
                    const double* pr = S.pst[g][w][0];
                    const double* pi = S.pst[g][w][1];
                    const int i = l & 15, kb = (l >> 4) * 16;
                    double sr = 0.0, si = 0.0;
#pragma unroll 4
                    for (int kk = kb; kk < kb + 16; ++kk) {
                        const double mr = pr[kk * 16 + i], mi = pi[kk * 16 + i];
                        const double xr = S.s[cur][0][kk], xi = S.s[cur][1][kk];
                        sr = fma(mr, xr, fma(-mi, xi, sr));
                        si = fma(mr, xi, fma(mi, xr, si));
                    }
                    sr += __shfl_xor(sr, 16, kWave);
                    si += __shfl_xor(si, 16, kWave);
                    sr += __shfl_xor(sr, 32, kWave);
                    si += __shfl_xor(si, 32, kWave);
                    if (l < 16) {
                        S.s[cur ^ 1][0][16 * w + i] = sr;
                        S.s[cur ^ 1][1][16 * w + i] = si;
                    }
                    __syncthreads();
                    cur ^= 1;
                }
            }
            if (tid < kWave) {
                double v = fma(S.y0[0][tid], S.s[cur][0][tid], S.y0[1][tid] * S.s[cur][1][tid]);
#pragma unroll
                for (int mask = 32; mask >= 1; mask >>= 1) v += __shfl_xor(v, mask, kWave);
                if (tid == 0) q.L[k] = scale * v;
            }
            __syncthreads();
        }
        if (tid == 0) {
            int32_t* st = a.stats + (int64_t)item * kStatsPerItem;
            for (int z = 0; z < kStatsPerItem; ++z) st[z] = 0;
            st[8] = fail;
        }
        __syncthreads();
    }
}

int64_t dense64_workspace_doubles(int grid) { return (int64_t)grid * kWsDoubles; }

int launch_dense64(const KrylovArgs& a, double* work, int grid, hipStream_t stream) {
    if (a.n != kD) return -13;
    hipLaunchKernelGGL(cce_dense64_kernel, dim3(grid), dim3(kNT), 0, stream, a, work);
    return (int)hipGetLastError();
}

}  // namespace oops
