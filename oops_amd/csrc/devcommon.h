// Device-side building blocks shared by the CCE expv kernels (gfx950).
//
//  * complex FP64 arithmetic on a plain {re, im} pair;
//  * block reductions built on wavefront butterflies (64-lane waves);
//  * the compiled-operator matvec: every merged 1-body / 2-body local
//    operator stored as a padded row table (ELL) whose entries carry the
//    gather offset of the source element, so one thread computes one output
//    element of  y = c0*x + sum_ops (I x A_op x I) x  from an LDS copy of x;
//  * Pade(6) + scaling and squaring of a small dense matrix held in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oops_device_types.h"

namespace oops {

struct cd {
    double re, im;
};

__device__ __forceinline__ cd mkcd(double a, double b) { cd r; r.re = a; r.im = b; return r; }

// Scalar FP64 library functions called out of line.  Inlined into a kernel
// with long-lived loops, their polynomial constants (two VGPRs each) are
// hoisted to the kernel entry and stay live everywhere; these are called a
// few times per Krylov step at most.
static __device__ __noinline__ double ool_pow(double x, double y) { return pow(x, y); }
static __device__ __noinline__ double ool_log(double x) { return log(x); }
static __device__ __noinline__ double ool_log10(double x) { return log10(x); }
static __device__ __noinline__ cd ool_expi(double x) {   // (cos x, sin x)
    double sn, cs;
    sincos(x, &sn, &cs);
    return mkcd(cs, sn);
}
__device__ __forceinline__ cd cadd(cd a, cd b) { return mkcd(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cd csub(cd a, cd b) { return mkcd(a.re - b.re, a.im - b.im); }
__device__ __forceinline__ cd cscale(cd a, double s) { return mkcd(a.re * s, a.im * s); }
__device__ __forceinline__ cd cmul(cd a, cd b) {
    return mkcd(fma(a.re, b.re, -a.im * b.im), fma(a.re, b.im, a.im * b.re));
}
// a*b + c
__device__ __forceinline__ cd cfma(cd a, cd b, cd c) {
    return mkcd(fma(a.re, b.re, fma(-a.im, b.im, c.re)), fma(a.re, b.im, fma(a.im, b.re, c.im)));
}
// conj(a)*b + c
__device__ __forceinline__ cd cjfma(cd a, cd b, cd c) {
    return mkcd(fma(a.re, b.re, fma(a.im, b.im, c.re)), fma(a.re, b.im, fma(-a.im, b.re, c.im)));
}
// c - a*b
__device__ __forceinline__ cd cfms(cd a, cd b, cd c) {
    return mkcd(fma(-a.re, b.re, fma(a.im, b.im, c.re)), fma(-a.re, b.im, fma(-a.im, b.re, c.im)));
}
__device__ __forceinline__ double cabs2(cd a) { return fma(a.re, a.re, a.im * a.im); }
// -i * a
__device__ __forceinline__ cd cmulmi(cd a) { return mkcd(a.im, -a.re); }
__device__ __forceinline__ cd cdiv(cd a, cd b) {
    // Smith's algorithm
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, d = b.re + b.im * r;
        return mkcd((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    }
    const double r = b.re / b.im, d = b.im + b.re * r;
    return mkcd((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}
__device__ __forceinline__ double cabs(cd a) { return hypot(a.re, a.im); }

constexpr int kWave = 64;

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
constexpr int pow2ceil(int v) { int p = 1; while (p < v) p *= 2; return p; }

// ---------------------------------------------------------------------------
// Block sum of P doubles per thread (P a power of two, P <= 64); value k of a
// thread is produced on demand by gen(k), so at most P/2 partial sums are live
// in registers at any time.
// Stage s of a halving butterfly exchanges half of the remaining values with
// the lane at distance 32>>s, so after log2(P) stages lane L holds the partial
// sum of value (L >> (6-log2 P)); a plain butterfly over the remaining low
// lane bits completes the wave sum.  Wave partials are combined through LDS.
// On return out[0..P) holds the block totals (visible to every thread).
template <int P, int NT, typename Gen>
__device__ __forceinline__ void block_sum_gen(Gen gen, double* ws, double* out) {
    static_assert((P & (P - 1)) == 0 && P <= 64, "P must be a power of two <= 64");
    constexpr int LP = ilog2c(P);
    constexpr int NW = NT / kWave;
    constexpr int H0 = P > 1 ? P / 2 : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    double v[H0];
    if constexpr (P == 1) {
        v[0] = gen(0);
    } else {
        const bool up = (lane & 32) != 0;
#pragma unroll
        for (int k = 0; k < H0; ++k) {
            const double a = gen(k), b = gen(k + H0);
            const double send = up ? a : b;
            const double keep = up ? b : a;
            v[k] = keep + __shfl_xor(send, 32, kWave);
        }
#pragma unroll
        for (int st = 1; st < LP; ++st) {
            const int mask = 32 >> st;
            const int half = P >> (st + 1);
            const bool upm = (lane & mask) != 0;
#pragma unroll
            for (int k = 0; k < half; ++k) {
                const double a = v[k], b = v[k + half];
                const double send = upm ? a : b;
                const double keep = upm ? b : a;
                v[k] = keep + __shfl_xor(send, mask, kWave);
            }
        }
    }
#pragma unroll
    for (int mask = (32 >> LP); mask >= 1; mask >>= 1) v[0] += __shfl_xor(v[0], mask, kWave);
    constexpr int GROUP = kWave / P;  // lanes sharing one value index
    const int idx = lane >> (6 - LP);
    if (NW == 1) {
        if ((lane & (GROUP - 1)) == 0) out[idx] = v[0];
        __syncthreads();
    } else {
        if ((lane & (GROUP - 1)) == 0) ws[wave * P + idx] = v[0];
        __syncthreads();
        if ((int)threadIdx.x < P) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += ws[w * P + threadIdx.x];
            out[threadIdx.x] = s;
        }
        __syncthreads();
    }
}

// Value k (compile-time after unrolling) of a wave-distributed vector whose
// lane L holds entry L: a VGPR -> SGPR lane read, no memory traffic.
__device__ __forceinline__ double lane_val(double x, int k) {
    const unsigned long long bits = __double_as_longlong(x);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)(bits & 0xffffffffull), k);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(bits >> 32), k);
    return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}

// Same reduction with a single barrier: every wave leaves its P partial sums
// in one of two alternating LDS slots, each lane adds the NW partials of the
// value it owns, and BlockTotals::get broadcasts a total from its owner lane
// with a lane read.  A slot is rewritten two reductions later, after every
// wave has passed the barrier of the reduction in between.
template <int NT>
struct BlockRed {
    double* ws;  // 3 * (NT/64) * 64 doubles: two partial slots + per-wave totals
    int slot;
    __device__ explicit BlockRed(double* w) : ws(w), slot(0) {}
};

template <int NT, int P>
struct BlockTotals {
    const double* vals;  // this wave's copy of the P block totals (LDS)
    __device__ __forceinline__ double get(int k) const { return vals[k]; }
};

template <int P, int NT, typename Gen>
__device__ __forceinline__ BlockTotals<NT, P> block_sum1(BlockRed<NT>& br, Gen gen) {
    static_assert((P & (P - 1)) == 0 && P <= 64, "P must be a power of two <= 64");
    constexpr int LP = ilog2c(P);
    constexpr int H0 = P > 1 ? P / 2 : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    double v[H0];
    if constexpr (P == 1) {
        v[0] = gen(0);
    } else {
        const bool up = (lane & 32) != 0;
#pragma unroll
        for (int k = 0; k < H0; ++k) {
            const double a = gen(k), b = gen(k + H0);
            const double send = up ? a : b;
            const double keep = up ? b : a;
            v[k] = keep + __shfl_xor(send, 32, kWave);
            // generate operands pair by pair instead of all P at once
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int st = 1; st < LP; ++st) {
            const int mask = 32 >> st;
            const int half = P >> (st + 1);
            const bool upm = (lane & mask) != 0;
#pragma unroll
            for (int k = 0; k < half; ++k) {
                const double a = v[k], b = v[k + half];
                const double send = upm ? a : b;
                const double keep = upm ? b : a;
                v[k] = keep + __shfl_xor(send, mask, kWave);
            }
        }
    }
#pragma unroll
    for (int mask = (32 >> LP); mask >= 1; mask >>= 1) v[0] += __shfl_xor(v[0], mask, kWave);
    constexpr int GROUP = kWave / P;
    const int idx = lane >> (6 - LP);
    constexpr int NW = NT / kWave;
    BlockTotals<NT, P> t;
    double* mine = br.ws + 2 * NW * 64 + wave * 64;  // per-wave copy of the totals
    double tot = v[0];
    if constexpr (NW > 1) {
        double* part = br.ws + br.slot * NW * 64;
        if ((lane & (GROUP - 1)) == 0) part[wave * P + idx] = v[0];
        br.slot ^= 1;
        __syncthreads();
        tot = part[idx];
#pragma unroll
        for (int w = 1; w < NW; ++w) tot += part[w * P + idx];
    }
    // lanes of one wave exchange through their own LDS copy: no block barrier
    if ((lane & (GROUP - 1)) == 0) mine[idx] = tot;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    t.vals = mine;
    return t;
}

// Wave sum of a double without LDS traffic: DPP row_mirror, row_half_mirror
// and two quad permutations sum each 16-lane row (every lane of the row ends
// with the row total, all lanes adding in an order that gives bit-identical
// results), then gfx950's permlane16/permlane32 swaps add the rows.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <bool HALF32>
__device__ __forceinline__ double swap_add_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = (int)b, hi = (int)(b >> 32);
    const auto l = HALF32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                          : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto h = HALF32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                          : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const double e = __longlong_as_double(((long long)h[0] << 32) | (unsigned)l[0]);
    const double o = __longlong_as_double(((long long)h[1] << 32) | (unsigned)l[1]);
    return e + o;
}

// Generic wave reduction with the same DPP/permlane pattern (op must be
// commutative and associative on the values: min, max).
template <typename Op>
__device__ __forceinline__ double wave_reduce_dpp(double v, Op op) {
    v = op(v, dpp_f64<0x140>(v));
    v = op(v, dpp_f64<0x141>(v));
    v = op(v, dpp_f64<0x4E>(v));
    v = op(v, dpp_f64<0xB1>(v));
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const long long b = __double_as_longlong(v);
        const int lo = (int)b, hi = (int)(b >> 32);
        const auto l = h ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                         : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto u = h ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                         : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        v = op(__longlong_as_double(((long long)u[0] << 32) | (unsigned)l[0]),
               __longlong_as_double(((long long)u[1] << 32) | (unsigned)l[1]));
    }
    return v;
}

__device__ __forceinline__ double wave_sum_dpp(double v) {
    v += dpp_f64<0x140>(v);   // row_mirror: lane i + lane 15-i
    v += dpp_f64<0x141>(v);   // row_half_mirror: lane i + lane 7-i (within 8)
    v += dpp_f64<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_f64<0xB1>(v);    // quad_perm [1,0,3,2]
    v = swap_add_f64<false>(v);
    return swap_add_f64<true>(v);
}

// Block sums of P values with one barrier; every lane gets all P totals in
// registers (bit-identical across the block).
template <int P>
struct LaneTotals {
    double v[P];
    __device__ __forceinline__ double get(int k) const { return v[k]; }
};

// STRIDE: doubles per wave in a partial slot (64 matches block_sum1's
// layout, so the two may alternate on one BlockRed; a kernel that only uses
// this reduction may pass P-sized slots: 2 * NW * STRIDE doubles of LDS).
template <int P, int NT, int STRIDE = 64, typename Gen>
__device__ __forceinline__ LaneTotals<P> block_sum_lanes(BlockRed<NT>& br, Gen gen) {
    static_assert(P <= STRIDE, "slot too small");
    constexpr int NW = NT / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    LaneTotals<P> t;
#pragma unroll
    for (int k = 0; k < P; ++k) t.v[k] = wave_sum_dpp(gen(k));
    double* part = br.ws + br.slot * NW * STRIDE;
#pragma unroll
    for (int k = 0; k < P; ++k)
        if (lane == k) part[wave * P + k] = t.v[k];
    br.slot ^= 1;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double s = part[k];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += part[w * P + k];
        t.v[k] = s;
    }
    return t;
}

template <int P, int NT>
__device__ __forceinline__ void block_sum(double (&v)[P], double* ws, double* out) {
    block_sum_gen<P, NT>([&](int k) { return v[k]; }, ws, out);
}

// Block max of one double (values >= 0).
template <int NT>
__device__ __forceinline__ double block_max1(double v, double* ws) {
    constexpr int NW = NT / kWave;
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) v = fmax(v, __shfl_xor(v, mask, kWave));
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    double r = ws[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) r = fmax(r, ws[w]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Compiled-operator matvec for output element i (i < n) from xs (LDS).
// Returns  c0*x_i + sum_ops sum_e val[row(i), e] * x[base(i) + off[row(i), e]].
// Row widths are padded on the host to multiples of 4 (4, 8 or 16 entries).
// One ELL row of R entries (R a multiple of 4), processed four entries at a
// time: the offsets arrive as one 16-byte load and the four gathers are
// independent, while the live set stays small.
__device__ __forceinline__ cd ell_row(const cd* __restrict__ xs, int base, const cd* __restrict__ vals,
                                      const int32_t* __restrict__ offs, int e0, int R, cd acc) {
#pragma unroll 1
    for (int e = 0; e < R; e += 4) {
        const int4 q = *reinterpret_cast<const int4*>(offs + e0 + e);
        const cd x0 = xs[base + q.x], x1 = xs[base + q.y], x2 = xs[base + q.z], x3 = xs[base + q.w];
        acc = cfma(vals[e0 + e], x0, acc);
        acc = cfma(vals[e0 + e + 1], x1, acc);
        acc = cfma(vals[e0 + e + 2], x2, acc);
        acc = cfma(vals[e0 + e + 3], x3, acc);
    }
    return acc;
}

__device__ __forceinline__ cd op_apply_elem(int i, const cd* __restrict__ xs, const DevOpHeader& h,
                                            const DevOpLocal* __restrict__ locs,
                                            const cd* __restrict__ vals, const int32_t* __restrict__ offs) {
    cd acc = cmul(mkcd(h.c0re, h.c0im), xs[i]);
    const int ld = h.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = h.nspin - 1;
    for (int L = 0; L < h.nloc; ++L) {
        const DevOpLocal lo = locs[L];
        const int shs = ld * (top - lo.s);
        const int ls = (i >> shs) & lmask;
        int base = i - (ls << shs);
        int row = ls;
        if (lo.nb == 2) {
            const int sht = ld * (top - lo.t);
            const int lt = (i >> sht) & lmask;
            base -= lt << sht;
            row = (ls << ld) | lt;
        }
        acc = ell_row(xs, base, vals, offs, lo.ent + row * lo.R, lo.R, acc);
    }
    return acc;
}

// Mixed-precision variant: 1-body operators (Zeeman, hyperfine field,
// dephasing; the large part of the operator) in FP64 from xs, 2-body
// couplings in complex FP32 from the FP32 copy xs32 with FP32 values vals32
// (entry e of a 2-body operator at vals32[e - ent2_first]).
__device__ __forceinline__ float2 ell_row32(const float2* __restrict__ xs32, int base,
                                            const float2* __restrict__ v32, const int32_t* __restrict__ offs,
                                            int e0, int e0v, int R, float2 acc) {
#pragma unroll 1
    for (int e = 0; e < R; e += 4) {
        const int4 q = *reinterpret_cast<const int4*>(offs + e0 + e);
        const float2 xv[4] = {xs32[base + q.x], xs32[base + q.y], xs32[base + q.z], xs32[base + q.w]};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float2 vv = v32[e0v + e + u];
            acc.x = fmaf(vv.x, xv[u].x, fmaf(-vv.y, xv[u].y, acc.x));
            acc.y = fmaf(vv.x, xv[u].y, fmaf(vv.y, xv[u].x, acc.y));
        }
    }
    return acc;
}

__device__ __forceinline__ cd op_apply_elem_mixed(int i, const cd* __restrict__ xs, const float2* __restrict__ xs32,
                                                  const DevOpHeader& h, const DevOpLocal* __restrict__ locs,
                                                  const cd* __restrict__ vals, const int32_t* __restrict__ offs) {
    cd acc = cmul(mkcd(h.c0re, h.c0im), xs[i]);
    float2 acc32 = make_float2(0.f, 0.f);
    const float2* v32 = reinterpret_cast<const float2*>(vals + h.ent2_first);
    const int ld = h.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = h.nspin - 1;
    for (int L = 0; L < h.nloc; ++L) {
        const DevOpLocal lo = locs[L];
        const int shs = ld * (top - lo.s);
        const int ls = (i >> shs) & lmask;
        int base = i - (ls << shs);
        int row = ls;
        if (lo.nb == 2) {
            const int sht = ld * (top - lo.t);
            const int lt = (i >> sht) & lmask;
            base -= lt << sht;
            row = (ls << ld) | lt;
            const int e0 = lo.ent + row * lo.R;
            acc32 = ell_row32(xs32, base, v32, offs, e0, e0 - h.ent2_first, lo.R, acc32);
        } else {
            acc = ell_row(xs, base, vals, offs, lo.ent + row * lo.R, lo.R, acc);
        }
    }
    return mkcd(acc.re + (double)acc32.x, acc.im + (double)acc32.y);
}

// Cooperative copy of one compiled operator into LDS.
template <int NT>
__device__ __forceinline__ void op_load(const DevOpHeader* __restrict__ hdr_g, const DevOpLocal* __restrict__ loc_g,
                                        const double2* __restrict__ val_g, const int32_t* __restrict__ off_g,
                                        int op, DevOpHeader& h, DevOpLocal* locs, cd* vals, int32_t* offs) {
    h = hdr_g[op];
    const int tid = threadIdx.x;
    for (int k = tid; k < h.nloc; k += NT) locs[k] = loc_g[h.loc_off + k];
    for (int k = tid; k < h.nent; k += NT) {
        const double2 v = val_g[h.ent_off + k];
        vals[k] = mkcd(v.x, v.y);
        offs[k] = off_g[h.ent_off + k];
    }
}

// ---------------------------------------------------------------------------
// Small dense complex matrices in LDS, column-major with leading dim ld.
template <int NT>
__device__ __forceinline__ void lds_matmul(cd* C, int ldc, const cd* A, int lda, const cd* B, int ldb, int m,
                                           cd alpha) {
    for (int idx = threadIdx.x; idx < m * m; idx += NT) {
        const int i = idx % m, j = idx / m;
        cd s = mkcd(0.0, 0.0);
        for (int l = 0; l < m; ++l) s = cfma(A[i + lda * l], B[l + ldb * j], s);
        C[i + ldc * j] = cmul(alpha, s);
    }
    __syncthreads();
}

// exp(t*H) for the leading m x m block of H (leading dim ldh) by the
// irreducible diagonal Pade(6) approximant with scaling and squaring.
// Workspace: four m x m matrices w0..w3 (leading dim m).  Returns the
// buffer index (0..3) holding the result, or -1 for a null H / -2 for a
// singular denominator.  ns receives the number of squarings.
template <int NT>
__device__ int lds_pade6(const cd* H, int ldh, int m, double t, cd* w0, cd* w1, cd* w2, cd* w3, double* ws,
                         int* sh_int, int& ns) {
    const int tid = threadIdx.x;
    // infinity norm (max row sum of |H|)
    double rs = 0.0;
    for (int i = tid; i < m; i += NT) {
        double s = 0.0;
        for (int j = 0; j < m; ++j) s += cabs(H[i + ldh * j]);
        rs = fmax(rs, s);
    }
    const double hnorm = fabs(t * block_max1<NT>(rs, ws));
    if (hnorm == 0.0) return -1;
    if (!(hnorm < 1e300)) return -3;  // non-finite input
    ns = (int)(ool_log(hnorm) / 0.6931471805599453) + 2;
    if (ns < 0) ns = 0;
    const double scale = t / ldexp(1.0, ns);
    // Pade(6) coefficients by the recurrence c_k = c_{k-1} (7-k) / (k (13-k))
    double c[7];
    c[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= 6; ++k) c[k] = c[k - 1] * (double)(7 - k) / (double)(k * (13 - k));
    cd* buf[4] = {w0, w1, w2, w3};
    // X2 = scale^2 H H  -> buf 0
    lds_matmul<NT>(w0, m, H, ldh, H, ldh, m, mkcd(scale * scale, 0.0));
    // even part starts at c6 I (buf 1), odd part at c5 I (buf 2)
    for (int idx = tid; idx < m * m; idx += NT) {
        const bool dg = (idx % m) == (idx / m);
        w1[idx] = mkcd(dg ? c[6] : 0.0, 0.0);
        w2[idx] = mkcd(dg ? c[5] : 0.0, 0.0);
    }
    __syncthreads();
    int ie = 1, io = 2, it = 3;
    bool even_turn = true;
    for (int k = 5; k > 0; --k) {
        const int src = even_turn ? ie : io;
        lds_matmul<NT>(buf[it], m, buf[src], m, w0, m, m, mkcd(1.0, 0.0));
        for (int i = tid; i < m; i += NT) buf[it][i + m * i].re += c[k - 1];
        __syncthreads();
        if (even_turn) { const int tmp = ie; ie = it; it = tmp; }
        else { const int tmp = io; io = it; it = tmp; }
        even_turn = !even_turn;
    }
    // odd <- scale * odd * H  (into the free buffer, then X2 slot is free)
    lds_matmul<NT>(buf[it], m, buf[io], m, H, ldh, m, mkcd(scale, 0.0));
    {
        const int tmp = io; io = it; it = tmp;
    }
    // D = even - odd  (in the even buffer), N = odd  (RHS)
    cd* D = buf[ie];
    cd* N = buf[io];
    for (int idx = tid; idx < m * m; idx += NT) D[idx] = csub(D[idx], N[idx]);
    __syncthreads();
    // LU with partial pivoting, applied to N
    for (int k = 0; k < m; ++k) {
        if (tid == 0) {
            int p = k;
            double best = cabs(D[k + m * k]);
            for (int i = k + 1; i < m; ++i) {
                const double a = cabs(D[i + m * k]);
                if (a > best) { best = a; p = i; }
            }
            sh_int[0] = best == 0.0 ? -1 : p;
        }
        __syncthreads();
        const int p = sh_int[0];
        if (p < 0) return -2;
        if (p != k) {
            for (int j = tid; j < 2 * m; j += NT) {
                cd* M_ = j < m ? D : N;
                const int jj = j < m ? j : j - m;
                const cd a = M_[k + m * jj];
                M_[k + m * jj] = M_[p + m * jj];
                M_[p + m * jj] = a;
            }
            __syncthreads();
        }
        const cd piv = D[k + m * k];
        for (int i = k + 1 + tid; i < m; i += NT) D[i + m * k] = cdiv(D[i + m * k], piv);
        __syncthreads();
        const int rows = m - k - 1;
        const int cols = (m - k - 1) + m;
        for (int idx = tid; idx < rows * cols; idx += NT) {
            const int i = k + 1 + idx % rows;
            const int jj = idx / rows;
            const cd f = D[i + m * k];
            if (jj < m - k - 1) {
                const int j = k + 1 + jj;
                D[i + m * j] = cfms(f, D[k + m * j], D[i + m * j]);
            } else {
                const int j = jj - (m - k - 1);
                N[i + m * j] = cfms(f, N[k + m * j], N[i + m * j]);
            }
        }
        __syncthreads();
    }
    for (int j = tid; j < m; j += NT) {
        for (int i = m - 1; i >= 0; --i) {
            cd s = N[i + m * j];
            for (int l = i + 1; l < m; ++l) s = cfms(D[i + m * l], N[l + m * j], s);
            N[i + m * j] = cdiv(s, D[i + m * i]);
        }
    }
    __syncthreads();
    // R = I + 2 N
    for (int idx = tid; idx < m * m; idx += NT) {
        const bool dg = (idx % m) == (idx / m);
        N[idx] = mkcd(2.0 * N[idx].re + (dg ? 1.0 : 0.0), 2.0 * N[idx].im);
    }
    __syncthreads();
    int ir = io;
    int free_a = ie;  // D is no longer needed
    for (int k = 0; k < ns; ++k) {
        lds_matmul<NT>(buf[free_a], m, buf[ir], m, buf[ir], m, m, mkcd(1.0, 0.0));
        const int tmp = ir; ir = free_a; free_a = tmp;
    }
    return ir;
}

// ---------------------------------------------------------------------------
// Faster Pade(6) + scaling and squaring for the Krylov kernel (same algorithm
// as lds_pade6, ZGPADM's): matrix products use 2x2 register tiles per thread,
// the denominator solve is Gauss-Jordan with implicit partial pivoting (the
// pivot of column k is the largest |D(i,k)| over unused rows, no row swaps, no
// back substitution), and the last squaring only forms the first column.
// H: m x m block of an LDS matrix with leading dim ld; W0..W3: ld x ld LDS
// workspaces; F receives exp(t H) e1 (m entries).  NT threads cooperate.
// Returns ns (>= 0) or a negative code (-1 null H, -2 singular, -3 non-finite).
// Thread (ii, jj) of the h x h grid (h = ceil(m / 2)) forms rows ii and
// ii + h of columns 2 jj and 2 jj + 1: the lanes of a wave read consecutive
// rows of A (16-byte steps; rows two apart put four lanes on the same LDS
// banks) and a few columns of B (broadcast reads).
template <int NT>
__device__ __forceinline__ void tile_matmul(cd* C, const cd* A, const cd* B, int m, int ld, double alpha) {
    const int tiles = (m + 1) >> 1;
    for (int t = threadIdx.x; t < tiles * tiles; t += NT) {
        const int i0 = t % tiles, j0 = 2 * (t / tiles);
        const int i1 = i0 + tiles < m ? i0 + tiles : i0;   // row of c10 / c11 (a repeat of i0 when odd m leaves none)
        cd c00 = mkcd(0.0, 0.0), c10 = c00, c01 = c00, c11 = c00;
        for (int l = 0; l < m; ++l) {
            const cd a0 = A[i0 + ld * l], a1 = A[i1 + ld * l];
            const cd b0 = B[l + ld * j0], b1 = B[l + ld * (j0 + 1)];
            c00 = cfma(a0, b0, c00);
            c10 = cfma(a1, b0, c10);
            c01 = cfma(a0, b1, c01);
            c11 = cfma(a1, b1, c11);
        }
        C[i0 + ld * j0] = cscale(c00, alpha);
        if (i0 + tiles < m) C[i1 + ld * j0] = cscale(c10, alpha);
        if (j0 + 1 < m) {
            C[i0 + ld * (j0 + 1)] = cscale(c01, alpha);
            if (i0 + tiles < m) C[i1 + ld * (j0 + 1)] = cscale(c11, alpha);
        }
    }
    __syncthreads();
}

// One Gauss-Jordan elimination step on [D | N] (D in W1, N in W2, m <= 64):
// every row i != p loses f_i = D(i,k) / D(p,k) times row p in the columns
// k+1.. of D and all columns of N.  Lane i of each wave takes row i (its f_i
// once), the NT / 64 waves the columns in turn: no index division, and the
// lanes of a wave touch consecutive rows.  Ends with a block barrier.
template <int NT>
__device__ __forceinline__ void gj_eliminate(cd* W1, cd* W2, int ld, int m, int k, int p, cd inv) {
    static_assert(NT % kWave == 0, "whole waves");
    constexpr int NG = NT / kWave;
    const int i = threadIdx.x & (kWave - 1);
    const int g = threadIdx.x / kWave;
    const int nd = m - k - 1;   // columns k+1 .. m-1 of D
    if (i < m && i != p) {
        const cd f = cmul(W1[i + ld * k], inv);
        for (int jj = g; jj < nd + m; jj += NG) {
            if (jj < nd) {
                const int j = k + 1 + jj;
                W1[i + ld * j] = cfms(f, W1[p + ld * j], W1[i + ld * j]);
            } else {
                const int j = jj - nd;
                W2[i + ld * j] = cfms(f, W2[p + ld * j], W2[i + ld * j]);
            }
        }
    }
    __syncthreads();
}

// Pade(6) rational part: R = r(H t / 2^ns) into W0 (W1..W3 are workspace);
// returns ns (the squarings still to do) or a negative code.
template <int NT>
__device__ int pade6_rational(const cd* H, int ld, int m, double t, cd* W0, cd* W1, cd* W2, cd* W3, int* piv_lds,
                              double* scal) {
    const int tid = threadIdx.x;
    // ||H||_inf with complex moduli (the reference's scaling rule)
    if (tid < kWave) {
        double rs = 0.0;
        if (tid < m)
            for (int j = 0; j < m; ++j) rs += cabs(H[tid + ld * j]);
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1) rs = fmax(rs, __shfl_xor(rs, mask, kWave));
        if (tid == 0) scal[0] = rs;
    }
    __syncthreads();
    const double hnorm = fabs(t * scal[0]);
    if (hnorm == 0.0) return -1;
    if (!(hnorm < 1e300)) return -3;
    int ns = (int)(ool_log(hnorm) / 0.6931471805599453) + 2;
    if (ns < 0) ns = 0;
    const double scale = t / ldexp(1.0, ns);
    double c[7];
    c[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= 6; ++k) c[k] = c[k - 1] * (double)(7 - k) / (double)(k * (13 - k));
    // W0 = X2 = scale^2 H H
    tile_matmul<NT>(W0, H, H, m, ld, scale * scale);
    // even part E (c6, c4, c2, c0) in W1, odd part O (c5, c3, c1) in W2
    for (int idx = tid; idx < m * m; idx += NT) {
        const int i = idx % m, j = idx / m;
        W1[i + ld * j] = cscale(W0[i + ld * j], c[6]);
        W2[i + ld * j] = cscale(W0[i + ld * j], c[5]);
        if (i == j) { W1[i + ld * j].re += c[4]; W2[i + ld * j].re += c[3]; }
    }
    __syncthreads();
    // E = (E X2 + c2) X2 + c0 ; O = O X2 + c1   (W3 scratch)
    tile_matmul<NT>(W3, W1, W0, m, ld, 1.0);
    for (int i = tid; i < m; i += NT) W3[i + ld * i].re += c[2];
    __syncthreads();
    tile_matmul<NT>(W1, W3, W0, m, ld, 1.0);
    for (int i = tid; i < m; i += NT) W1[i + ld * i].re += c[0];
    tile_matmul<NT>(W3, W2, W0, m, ld, 1.0);
    for (int i = tid; i < m; i += NT) W3[i + ld * i].re += c[1];
    __syncthreads();
    // zO = scale * O H -> W2 ; D = E - zO -> W1 ; N = zO (W2)
    tile_matmul<NT>(W2, W3, H, m, ld, scale);
    for (int idx = tid; idx < m * m; idx += NT) {
        const int i = idx % m, j = idx / m;
        W1[i + ld * j] = csub(W1[i + ld * j], W2[i + ld * j]);
    }
    if (tid < kWave) piv_lds[tid] = 0;  // used-row flags (m <= 64)
    __syncthreads();
    // Gauss-Jordan on [D | N] with implicit partial pivoting
    for (int k = 0; k < m; ++k) {
        if (tid < kWave) {
            double a = -1.0;
            if (tid < m && piv_lds[tid] == 0) a = cabs(W1[tid + ld * k]);
            int p = tid;
#pragma unroll
            for (int mask = 32; mask >= 1; mask >>= 1) {
                const double ao = __shfl_xor(a, mask, kWave);
                const int po = __shfl_xor(p, mask, kWave);
                if (ao > a || (ao == a && po < p)) { a = ao; p = po; }
            }
            if (tid == 0) piv_lds[64 + k] = (a > 0.0) ? p : -1;
        }
        __syncthreads();
        const int p = piv_lds[64 + k];
        if (p < 0) return -2;
        if (tid == 0) piv_lds[p] = 1;
        const cd inv = cdiv(mkcd(1.0, 0.0), W1[p + ld * k]);
        gj_eliminate<NT>(W1, W2, ld, m, k, p, inv);
    }
    // R = I + 2 Y, Y(k,:) = N(perm k,:) / D(perm k, k)   -> W0
    for (int idx = tid; idx < m * m; idx += NT) {
        const int k = idx % m, j = idx / m;
        const int pr = piv_lds[64 + k];
        const cd y = cdiv(W2[pr + ld * j], W1[pr + ld * k]);
        W0[k + ld * j] = mkcd(2.0 * y.re + (k == j ? 1.0 : 0.0), 2.0 * y.im);
    }
    __syncthreads();
    return ns;
}

template <int NT>
__device__ int pade6_first_column(const cd* H, int ld, int m, double t, cd* W0, cd* W1, cd* W2, cd* W3, cd* F,
                                  int* piv_lds, double* scal) {
    const int tid = threadIdx.x;
    const int ns = pade6_rational<NT>(H, ld, m, t, W0, W1, W2, W3, piv_lds, scal);
    if (ns < 0) return ns;
    // squarings; the last one only forms the first column
    cd* R = W0;
    cd* T = W3;
    for (int q = 0; q + 1 < ns; ++q) {
        tile_matmul<NT>(T, R, R, m, ld, 1.0);
        cd* tmp = R; R = T; T = tmp;
    }
    if (ns == 0) {
        for (int i = tid; i < m; i += NT) F[i] = R[i];
    } else {
        for (int i = tid; i < m; i += NT) {
            cd acc = mkcd(0.0, 0.0);
            for (int l = 0; l < m; ++l) acc = cfma(R[i + ld * l], R[l], acc);
            F[i] = acc;
        }
    }
    __syncthreads();
    return ns;
}

// Full exp(t H): Pade(6) and all ns squarings; *out points at the result
// (W0 or W3, leading dimension ld).
template <int NT>
__device__ int pade6_full(const cd* H, int ld, int m, double t, cd* W0, cd* W1, cd* W2, cd* W3, int* piv_lds,
                          double* scal, cd** out) {
    const int ns = pade6_rational<NT>(H, ld, m, t, W0, W1, W2, W3, piv_lds, scal);
    if (ns < 0) return ns;
    cd* R = W0;
    cd* T = W3;
    for (int q = 0; q < ns; ++q) {
        tile_matmul<NT>(T, R, R, m, ld, 1.0);
        cd* tmp = R; R = T; T = tmp;
    }
    *out = R;
    return ns;
}

// In-place variant of tile_matmul (C may alias A or B): every thread keeps
// its 2x2 tile in registers until all threads have finished reading.
// Needs ((m+1)/2)^2 <= NT, i.e. m <= 32 for 256 threads.
template <int NT>
__device__ __forceinline__ void tile_matmul_ip(cd* C, const cd* A, const cd* B, int m, int ld, double alpha) {
    const int tiles = (m + 1) >> 1;
    const int t = threadIdx.x;
    const bool active = t < tiles * tiles;
    const int i0 = t % tiles, j0 = 2 * (t / tiles);   // rows i0 and i0 + tiles, as in tile_matmul
    const int i1 = i0 + tiles < m ? i0 + tiles : i0;
    cd c00 = mkcd(0.0, 0.0), c10 = c00, c01 = c00, c11 = c00;
    if (active) {
        for (int l = 0; l < m; ++l) {
            const cd a0 = A[i0 + ld * l], a1 = A[i1 + ld * l];
            const cd b0 = B[l + ld * j0], b1 = B[l + ld * (j0 + 1)];
            c00 = cfma(a0, b0, c00);
            c10 = cfma(a1, b0, c10);
            c01 = cfma(a0, b1, c01);
            c11 = cfma(a1, b1, c11);
        }
    }
    __syncthreads();
    if (active) {
        C[i0 + ld * j0] = cscale(c00, alpha);
        if (i0 + tiles < m) C[i1 + ld * j0] = cscale(c10, alpha);
        if (j0 + 1 < m) {
            C[i0 + ld * (j0 + 1)] = cscale(c01, alpha);
            if (i0 + tiles < m) C[i1 + ld * (j0 + 1)] = cscale(c11, alpha);
        }
    }
    __syncthreads();
}

// pade6_first_column with three workspaces (W0 = X2 then R, W1 = even part
// then the denominator, W2 = odd part then the numerator), using in-place
// register-tiled products.  m <= 32.
template <int NT>
__device__ int pade6_first_column3(const cd* H, int ld, int m, double t, cd* W0, cd* W1, cd* W2, cd* F,
                                   int* piv_lds, double* scal) {
    const int tid = threadIdx.x;
    if (tid < kWave) {
        double rs = 0.0;
        if (tid < m)
            for (int j = 0; j < m; ++j) rs += cabs(H[tid + ld * j]);
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1) rs = fmax(rs, __shfl_xor(rs, mask, kWave));
        if (tid == 0) scal[0] = rs;
    }
    __syncthreads();
    const double hnorm = fabs(t * scal[0]);
    if (hnorm == 0.0) return -1;
    if (!(hnorm < 1e300)) return -3;
    int ns = (int)(ool_log(hnorm) / 0.6931471805599453) + 2;
    if (ns < 0) ns = 0;
    const double scale = t / ldexp(1.0, ns);
    double c[7];
    c[0] = 1.0;
#pragma unroll
    for (int k = 1; k <= 6; ++k) c[k] = c[k - 1] * (double)(7 - k) / (double)(k * (13 - k));
    tile_matmul<NT>(W0, H, H, m, ld, scale * scale);  // X2
    for (int idx = tid; idx < m * m; idx += NT) {
        const int i = idx % m, j = idx / m;
        W1[i + ld * j] = cscale(W0[i + ld * j], c[6]);
        W2[i + ld * j] = cscale(W0[i + ld * j], c[5]);
        if (i == j) { W1[i + ld * j].re += c[4]; W2[i + ld * j].re += c[3]; }
    }
    __syncthreads();
    tile_matmul_ip<NT>(W1, W0, W1, m, ld, 1.0);  // E = X2 (c6 X2 + c4)
    for (int i = tid; i < m; i += NT) W1[i + ld * i].re += c[2];
    __syncthreads();
    tile_matmul_ip<NT>(W1, W0, W1, m, ld, 1.0);  // E = X2 E + c0
    tile_matmul_ip<NT>(W2, W0, W2, m, ld, 1.0);  // O = X2 (c5 X2 + c3) + c1
    for (int i = tid; i < m; i += NT) {
        W1[i + ld * i].re += c[0];
        W2[i + ld * i].re += c[1];
    }
    __syncthreads();
    tile_matmul_ip<NT>(W2, W2, H, m, ld, scale);  // zO = scale O H
    for (int idx = tid; idx < m * m; idx += NT) {
        const int i = idx % m, j = idx / m;
        W1[i + ld * j] = csub(W1[i + ld * j], W2[i + ld * j]);  // D = E - zO
    }
    if (tid < kWave) piv_lds[tid] = 0;
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        if (tid < kWave) {
            double a = -1.0;
            if (tid < m && piv_lds[tid] == 0) a = cabs(W1[tid + ld * k]);
            int p = tid;
#pragma unroll
            for (int mask = 32; mask >= 1; mask >>= 1) {
                const double ao = __shfl_xor(a, mask, kWave);
                const int po = __shfl_xor(p, mask, kWave);
                if (ao > a || (ao == a && po < p)) { a = ao; p = po; }
            }
            if (tid == 0) piv_lds[64 + k] = (a > 0.0) ? p : -1;
        }
        __syncthreads();
        const int p = piv_lds[64 + k];
        if (p < 0) return -2;
        if (tid == 0) piv_lds[p] = 1;
        const cd inv = cdiv(mkcd(1.0, 0.0), W1[p + ld * k]);
        const int ncol = (m - k - 1) + m;
        for (int idx = tid; idx < m * ncol; idx += NT) {
            const int i = idx % m;
            const int jj = idx / m;
            if (i == p) continue;
            const cd f = cmul(W1[i + ld * k], inv);
            if (jj < m - k - 1) {
                const int j = k + 1 + jj;
                W1[i + ld * j] = cfms(f, W1[p + ld * j], W1[i + ld * j]);
            } else {
                const int j = jj - (m - k - 1);
                W2[i + ld * j] = cfms(f, W2[p + ld * j], W2[i + ld * j]);
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < m * m; idx += NT) {
        const int k = idx % m, j = idx / m;
        const int pr = piv_lds[64 + k];
        const cd y = cdiv(W2[pr + ld * j], W1[pr + ld * k]);
        W0[k + ld * j] = mkcd(2.0 * y.re + (k == j ? 1.0 : 0.0), 2.0 * y.im);
    }
    __syncthreads();
    for (int q = 0; q + 1 < ns; ++q) tile_matmul_ip<NT>(W0, W0, W0, m, ld, 1.0);
    if (ns == 0) {
        for (int i = tid; i < m; i += NT) F[i] = W0[i];
    } else {
        for (int i = tid; i < m; i += NT) {
            cd acc = mkcd(0.0, 0.0);
            for (int l = 0; l < m; ++l) acc = cfma(W0[i + ld * l], W0[l], acc);
            F[i] = acc;
        }
    }
    __syncthreads();
    return ns;
}

}  // namespace oops
