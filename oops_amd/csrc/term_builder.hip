// Device-side term builder: the packed operator tables of a batch of 13C
// clusters, built from the bath coordinates on the GPU.
//
// Restates, per cluster and generator, what the host pipeline does term by
// term (cce_host.cpp build_problem + capi.cpp compile_operator), after the
// reference's operator construction:
//   dipolar tensor, Zeeman and NV-field coefficients   source/misc/misc.cpp:11-57
//   term order and the exact "!= 0" filter             SpinInteractionComponent.cpp:184-304,
//                                                      SpinInteraction.cpp:29-65
//   Hamiltonian assembly                               source/app/ensemble_cce.cpp:57-71
//   lvA = H0# - H1b + G, lvB = H1# - H0b + G            ensemble_cce.cpp:41-42, 85-91
//   SHARP / FLAT expansion                             source/kron/KronProd.cpp:9-28, 84-107
//   negated adjoint of the bra-side generators         QuantumOperator.cpp:33-49
// and merges the terms of every spin and spin pair into one local operator,
// stored as ELL rows with flip-mask slots.
//
// Every term of a spin-1/2 Liouvillian is a coefficient times entries of
// SHARP/FLAT-expanded spin matrices, which are 0, +-1/2, +-i/2 (one spin) or
// their products +-1/4, +-i/4 (pairs): each term's contribution is exact, so
// a merged entry summed over the same terms in the same order equals the
// host's bit for bit.  Coefficients are computed with the host's operation
// order and without contraction into fused multiply-adds.
//
// One 256-thread workgroup per (cluster, generator): the threads build the
// dense merged local matrices in LDS (each entry summed by one thread in term
// order), then thread 0 compacts them as the host compiler does (zero entries
// skipped, rows padded to 4/8/16 entries, flip-mask slots numbered in order of
// first appearance).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "couplings.h"
#include "term_builder.h"

#pragma clang fp contract(off)

namespace oops {

namespace {

struct cdx {
    double re, im;
};
__device__ __forceinline__ cdx mk(double a, double b) { cdx r; r.re = a; r.im = b; return r; }
// std::complex<double> multiplication for finite operands: (ac - bd) + (ad + bc) i
__device__ __forceinline__ cdx mul(cdx a, cdx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ cdx add(cdx a, cdx b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cdx conjx(cdx a) { return mk(a.re, -a.im); }
__device__ __forceinline__ bool nonzero(cdx a) { return a.re != 0.0 || a.im != 0.0; }

constexpr int kMaxK = 4;
constexpr int kMaxPairs = kMaxK * (kMaxK - 1) / 2;

struct BuildSmem {
    cdx single[kMaxK][16];          // merged 4 x 4 one-spin operators (column-major)
    cdx pair[kMaxPairs][256];       // merged 16 x 16 pair operators (column-major)
    double J[kMaxPairs][9];         // dipolar tensors
    double f0[kMaxK][3], f1[kMaxK][3];  // NV fields for the two states
    double r[kMaxK][3];
    int32_t mask64[128], mask32[128];
};

// Term coefficient and matrix of the one-spin term list of generator `which`
// for spin s: entry e of the 4 x 4 merged matrix accumulates, in the host's
// order, the h_sharp Zeeman (x, y, z) and field terms, then the h_flat ones
// (negated), then the dephasing term; for the negated-adjoint generators the
// same list with negated coefficients and adjoint matrices.
__device__ cdx single_entry(const TermBuilderArgs& a, const BuildSmem& S, int which, int s, int e) {
    // which 0: lvA = h0# - h1b, 1: -(lvB)^dagger, 2: lvB = h1# - h0b, 3: -(lvA)^dagger
    const bool base_a = which == 0 || which == 3;   // the underlying list is lvA
    const bool adj = which == 1 || which == 3;
    const double* fs = base_a ? S.f0[s] : S.f1[s];  // field of the sharp-side state
    const double* ff = base_a ? S.f1[s] : S.f0[s];  // field of the flat-side state
    // adjoint: entry (r, c) of A^dagger is conj(A(c, r))
    const int ee = adj ? ((e % 4) * 4 + e / 4) : e;
    cdx acc = mk(0.0, 0.0);
    for (int side = 0; side < 2; ++side) {
        const double sign = side == 0 ? 1.0 : -1.0;
        const double* f = side == 0 ? fs : ff;
        const double2* M = side == 0 ? a.sharp : a.flat;   // [3][16]
        for (int part = 0; part < 2; ++part)               // Zeeman, then the NV field
            for (int c = 0; c < 3; ++c) {
                const double coef = part == 0 ? a.zee[c] : f[c];
                if (coef == 0.0) continue;                 // the reference's != 0 filter
                double cc = coef * sign;                   // expand(.., FLAT, -1)
                if (adj) cc = -cc;                         // negated adjoint
                const double2 m = M[16 * c + ee];
                const cdx mm = adj ? conjx(mk(m.x, m.y)) : mk(m.x, m.y);
                acc = add(acc, mul(mk(cc, 0.0), mm));
            }
    }
    if (a.deph_rate != 0.0) {
        double cc = a.deph_rate;
        if (adj) cc = -cc;
        const double2 m = a.dm[ee];
        const cdx mm = adj ? conjx(mk(m.x, m.y)) : mk(m.x, m.y);
        acc = add(acc, mul(mk(cc, 0.0), mm));
    }
    return acc;
}

// Entry e of the merged 16 x 16 operator of pair p: the nine h_sharp dipolar
// terms J_ab SHARP(S_a) (x) SHARP(S_b), then the nine h_flat ones (negated);
// the dipolar part does not depend on the NV state, so lvA and lvB share it.
__device__ cdx pair_entry(const TermBuilderArgs& a, const BuildSmem& S, int which, int p, int e) {
    const bool adj = which == 1 || which == 3;
    int r = e % 16, c = e / 16;
    if (adj) { const int t = r; r = c; c = t; }
    const int ra = r / 4, rb = r % 4, ca = c / 4, cb = c % 4;
    cdx acc = mk(0.0, 0.0);
    for (int side = 0; side < 2; ++side) {
        const double2* M = side == 0 ? a.sharp : a.flat;
        const double sign = side == 0 ? 1.0 : -1.0;
        for (int pa = 0; pa < 3; ++pa)
            for (int qb = 0; qb < 3; ++qb) {
                const double J = S.J[p][3 * pa + qb];
                if (J == 0.0) continue;
                double cc = J * sign;
                if (adj) cc = -cc;
                const double2 x = M[16 * pa + ra + 4 * ca], y = M[16 * qb + rb + 4 * cb];
                cdx A = mk(x.x, x.y), B = mk(y.x, y.y);
                if (adj) { A = conjx(A); B = conjx(B); }
                acc = add(acc, mul(mul(mk(cc, 0.0), A), B));
            }
    }
    return acc;
}

}  // namespace

__global__ __launch_bounds__(256) void term_builder_kernel(TermBuilderArgs a) {
    __shared__ BuildSmem S;
    const int tid = threadIdx.x;
    const int op = blockIdx.x;               // item * nops + which
    const int item = op / a.nops, which = op % a.nops;
    const int k = a.k;
    const int P = k * (k - 1) / 2;
    if (tid < k)
        for (int c = 0; c < 3; ++c) S.r[tid][c] = a.coords[3 * a.clusters[(int64_t)item * k + tid] + c];
    __syncthreads();
    // couplings: pair tensors, and the NV field of both states on every spin
    // (f[r] = ((0 + dp[r] s0) + dp[r+3] s1) + dp[r+6] s2, misc.cpp:50-57)
    if (tid < P) {
        int i = 0, j = 1, q = tid;
        while (q >= k - 1 - i) { q -= k - 1 - i; ++i; }
        j = i + 1 + q;
        dipole9(S.r[i], S.r[j], a.gamma_c, a.gamma_c, S.J[tid]);
    } else if (tid >= 64 && tid < 64 + k) {
        const int s = tid - 64;
        double dp[9];
        dipole9(S.r[s], a.center, a.gamma_c, a.gamma_e, dp);
        for (int rr = 0; rr < 3; ++rr) {
            double v0 = 0.0, v1 = 0.0;
            for (int c = 0; c < 3; ++c) {
                v0 = v0 + dp[rr + 3 * c] * a.sv0[c];
                v1 = v1 + dp[rr + 3 * c] * a.sv1[c];
            }
            S.f0[s][rr] = v0;
            S.f1[s][rr] = v1;
        }
    }
    __syncthreads();
    for (int e = tid; e < k * 16; e += blockDim.x) S.single[e / 16][e % 16] = single_entry(a, S, which, e / 16, e % 16);
    for (int e = tid; e < P * 256; e += blockDim.x) S.pair[e / 256][e % 256] = pair_entry(a, S, which, e / 256, e % 256);
    __syncthreads();
    if (tid != 0) return;

    // compaction (compile_operator): locals in order, rows, nonzero columns ascending
    DevOpHeader* H = a.hdr + op;
    DevOpLocal* L = a.loc + (int64_t)op * a.loc_stride;
    double2* V = a.val + (int64_t)op * a.ent_stride;
    int32_t* O = a.off + (int64_t)op * a.ent_stride;
    uint8_t* E = a.eslot + (int64_t)op * a.ent_stride;
    const int ld = 2, ns = k;
    int nloc = 0, nent = 0, max_row = 0, n64 = 1, n32 = 0, err = 0, ent2_first = 0;
    bool herm = true;
    S.mask64[0] = 0;
    auto stride = [&](int s) { return 1 << (ld * (ns - 1 - s)); };
    auto slot_of = [&](int mask) -> int {
        int fields = 0;
        for (int q = 0; q < ns; ++q)
            if ((mask >> (ld * (ns - 1 - q))) & 3) ++fields;
        int32_t* lst = fields <= 1 ? S.mask64 : S.mask32;
        int& cnt = fields <= 1 ? n64 : n32;
        int id = -1;
        for (int q = 0; q < cnt && q < 128; ++q)
            if (lst[q] == mask) { id = q; break; }
        if (id < 0) {
            id = cnt++;
            if (id < 128) lst[id] = mask;
        }
        if (id > 127) return 0;
        return fields <= 1 ? id : (id | 128);
    };
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) ent2_first = nent;
        const int nlocal = pass == 0 ? k : P;
        for (int l = 0; l < nlocal; ++l) {
            const int rows = pass == 0 ? 4 : 16;
            const cdx* M = pass == 0 ? S.single[l] : S.pair[l];
            int s = l, u = l;
            if (pass == 1) {
                int q = l;
                s = 0;
                while (q >= k - 1 - s) { q -= k - 1 - s; ++s; }
                u = s + 1 + q;
            }
            int R = 0;
            double mx = 0.0;
            for (int r = 0; r < rows; ++r) {
                int cnt = 0;
                for (int c = 0; c < rows; ++c)
                    if (nonzero(M[r + rows * c])) ++cnt;
                R = R > cnt ? R : cnt;
            }
            if (R == 0) continue;
            R = R <= 4 ? 4 : (R <= 8 ? 8 : 16);
            if (nloc >= a.loc_stride || nent + rows * R > a.ent_stride) { err = 1; break; }
            DevOpLocal lo;
            lo.nb = (int16_t)(pass + 1);
            lo.s = (int16_t)s;
            lo.t = (int16_t)u;
            lo.R = (int16_t)R;
            lo.ent = nent;
            lo.rows = rows;
            for (int r = 0; r < rows; ++r) {
                int cnt = 0;
                for (int c = 0; c < rows; ++c) {
                    const cdx v = M[r + rows * c];
                    if (!nonzero(v)) continue;
                    int off, mask;
                    if (pass == 0) {
                        off = c * stride(s);
                        mask = (r ^ c) * stride(s);
                    } else {
                        off = (c / 4) * stride(s) + (c % 4) * stride(u);
                        mask = ((r / 4) ^ (c / 4)) * stride(s) + ((r % 4) ^ (c % 4)) * stride(u);
                    }
                    V[nent] = make_double2(v.re, v.im);
                    O[nent] = off;
                    E[nent] = (uint8_t)slot_of(mask);
                    ++nent;
                    ++cnt;
                }
                for (; cnt < R; ++cnt) {
                    V[nent] = make_double2(0.0, 0.0);
                    O[nent] = 0;
                    E[nent] = 0;
                    ++nent;
                }
            }
            for (int q = 0; q < rows * rows; ++q) mx = fmax(mx, sqrt(M[q].re * M[q].re + M[q].im * M[q].im));
            for (int r = 0; r < rows && herm; ++r)
                for (int c = 0; c <= r; ++c) {
                    const cdx x = M[r + rows * c], y = M[c + rows * r];
                    const double dr = x.re - y.re, di = x.im + y.im;
                    if (sqrt(dr * dr + di * di) > 1e-13 * mx) { herm = false; break; }
                }
            L[nloc++] = lo;
            max_row = max_row > R ? max_row : R;
        }
        if (err) break;
    }
    DevOpHeader h;
    h.n = 1 << (2 * k);
    h.nspin = k;
    h.ld_bits = 2;
    h.nloc = nloc;
    h.nent = nent;
    h.loc_off = op * a.loc_stride;
    h.ent_off = op * a.ent_stride;
    h.max_row = max_row;
    h.c0re = 0.0;
    h.c0im = 0.0;
    h.ent2_first = ent2_first;
    h.pad0 = 0;
    h.pad1 = 0.0;
    *H = h;
    const int s64 = n64 > 128 || n32 > 128 ? -1 : n64;
    for (int q = 0; q < kTBSlotMaskStride; ++q) {
        int32_t mk = 0;
        if (q < kTBSlot64 && q < n64 && q < 128) mk = S.mask64[q];
        if (q >= kTBSlot64 && q - kTBSlot64 < kTBSlot32 && q - kTBSlot64 < n32) mk = S.mask32[q - kTBSlot64];
        a.smask[(int64_t)op * kTBSlotMaskStride + q] = mk;
    }
    for (int q = 0; q < kTBMedMaskStride; ++q) {
        int32_t mk = 0;
        if (q < kTBMedSlot64 && s64 > 0 && q < n64) mk = S.mask64[q];
        if (q >= kTBMedSlot64 && q - kTBMedSlot64 < n32 && q - kTBMedSlot64 < 128) mk = S.mask32[q - kTBMedSlot64];
        a.mmask[(int64_t)op * kTBMedMaskStride + q] = mk;
    }
    int32_t* sm = a.summary + (int64_t)op * kTBSummary;
    sm[0] = nloc;
    sm[1] = nent;
    sm[2] = max_row;
    sm[3] = s64;
    sm[4] = n32;
    sm[5] = herm ? 1 : 0;
    sm[6] = err;
    sm[7] = 0;
}

int launch_term_builder(const TermBuilderArgs& a, int nops_total, hipStream_t stream) {
    if (a.k < 1 || a.k > kMaxK) return -11;
    hipLaunchKernelGGL(term_builder_kernel, dim3(nops_total), dim3(256), 0, stream, a);
    return (int)hipGetLastError();
}

}  // namespace oops
