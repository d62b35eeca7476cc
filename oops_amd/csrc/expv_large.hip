// Krylov expv for one large state vector (any n up to 2^30), basis in HBM.
//
// Device side of MatExpVector::runInexplicitGPU / MAIN_CACHE
// (src/source/math/MatExp.cpp:367-521, src/expv/source/zgexpv_cache.f:4-690)
// and of hamvec_cuda3 (src/expv/source/hamvec_cuda3.cu:392-645), redesigned
// so that nothing but a few scalars per Krylov step crosses PCIe: the basis
// V (n x (m+1)), the iterate w and the Hessenberg matrix stay in device
// memory, the small exponential runs on the device, and the host keeps only
// Expokit's scalar step control.
//
// Kernels:
//   large_matvec     y = alpha * H * (x * xscale), optionally storing the
//                    scaled x as the next basis vector; one thread per
//                    element, the compiled operator (merged 1-/2-body ELL
//                    tables) in LDS, gathers from L2/HBM.
//   large_matvec_waves  the same for small states, the operators over the
//                    waves of a workgroup.
//   large_dots       partial sums of conj(V_l) . p for l <= J and |p|^2
//                    (grid.y = vector index), 256-thread block reductions.
//   large_finish     sums the partials, stores h into the Hessenberg column
//                    and the reciprocal norm for the next product.
//   large_gs_*       fused Gram-Schmidt passes over the whole basis.
//   large_pade       Pade(6) first column of exp(t H_aug) on one workgroup.
//   large_combine    w = beta * sum_j F_j V_j.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "devcommon.h"
#include "oops_kernels.h"
#include "oops_large.h"

namespace oops {

constexpr int kLT = 256;

// The local-operator list sits in LDS; the ELL rows (up to a few thousand
// entries for a 10-spin Liouvillian) are read through the L1/L2 caches.
struct LargeOpLds {
    DevOpHeader hdr;
    DevOpLocal locs[kLargeMaxLocals];
};

__device__ __forceinline__ void large_op_load(const LargeOpArgs& o, LargeOpLds& L) {
    if (threadIdx.x == 0) L.hdr = *o.hdr;
    const int nloc = o.hdr->nloc;
    for (int k = threadIdx.x; k < nloc; k += kLT) L.locs[k] = o.loc[k];
    __syncthreads();
}

// y = alpha * H (xscale * x); if xstore, xstore = xscale * x.
// xscale is read from device memory (the reciprocal norm of the previous
// Arnoldi step), alpha: 0 -> +1 (norm estimate), 1 -> -i.
// Basis element types: FP64 (double2) or FP32 (float2) storage.
__device__ __forceinline__ void vstore(double2* p, double re, double im) { *p = make_double2(re, im); }
__device__ __forceinline__ void vstore(float2* p, double re, double im) { *p = make_float2((float)re, (float)im); }
__device__ __forceinline__ double2 vload(const double2* p) { return *p; }
__device__ __forceinline__ double2 vload(const float2* p) {
    const float2 f = *p;
    return make_double2((double)f.x, (double)f.y);
}

template <typename VT>
__global__ __launch_bounds__(kLT) void large_matvec(LargeOpArgs o, const double2* __restrict__ x,
                                                    const double* __restrict__ xscale_p, double2* __restrict__ y,
                                                    VT* __restrict__ xstore, int64_t n, int alpha_mode) {
    __shared__ LargeOpLds L;
    large_op_load(o, L);
    const double xs = xscale_p ? *xscale_p : 1.0;
    const cd* xc = reinterpret_cast<const cd*>(x);
    const cd* vals = reinterpret_cast<const cd*>(o.val);
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        cd acc = op_apply_elem((int)i, xc, L.hdr, L.locs, vals, o.off);
        acc = cscale(acc, xs);
        if (alpha_mode == 1) acc = cmulmi(acc);
        y[i] = make_double2(acc.re, acc.im);
        if (xstore) {
            const double2 xv = x[i];
            vstore(xstore + i, xv.x * xs, xv.y * xs);
        }
    }
}

// The same product for states of up to 2^17 elements (too few to fill the
// device with one thread per element: a 14-spin Hilbert state has 16 384
// elements and ~100 local operators), with the local operators split over
// the NW waves of a workgroup: lane l of every wave takes element 64 b + l,
// wave w the operators w, w + NW, ..; the partial sums meet in LDS
// (summed in wave order).  All lanes of a wave read the tables of one local
// operator (at most d^2 rows) and gather from 64 neighbouring elements, so
// each load instruction touches a few cache lines (split over the lanes of
// one element instead, the 16 lanes would read 16 operators' tables and
// gather from 16 unrelated places: twice the time for a 14-spin state).
template <typename VT, int NW>
__global__ __launch_bounds__(NW* kWave) void large_matvec_waves(LargeOpArgs o, const double2* __restrict__ x,
                                                                const double* __restrict__ xscale_p,
                                                                double2* __restrict__ y, VT* __restrict__ xstore,
                                                                int64_t n, int alpha_mode) {
    __shared__ LargeOpLds L;
    __shared__ cd part[NW][kWave];
    if (threadIdx.x == 0) L.hdr = *o.hdr;
    const int nloc = o.hdr->nloc;
    for (int k = threadIdx.x; k < nloc; k += NW * kWave) L.locs[k] = o.loc[k];
    __syncthreads();
    const double xs = xscale_p ? *xscale_p : 1.0;
    const cd* xc = reinterpret_cast<const cd*>(x);
    const cd* vals = reinterpret_cast<const cd*>(o.val);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int ld = L.hdr.ld_bits;
    const int lmask = (1 << ld) - 1;
    const int top = L.hdr.nspin - 1;
    // the bound is uniform over the workgroup (the barriers need every wave)
    for (int64_t i0 = (int64_t)blockIdx.x * kWave; i0 < n; i0 += (int64_t)gridDim.x * kWave) {
        const int64_t i = i0 + lane;
        const bool live = i < n;
        const int ii = (int)i;
        cd acc = (w == 0 && live) ? cmul(mkcd(L.hdr.c0re, L.hdr.c0im), xc[ii]) : mkcd(0.0, 0.0);
        if (live)
            for (int Lo = w; Lo < L.hdr.nloc; Lo += NW) {
                const DevOpLocal lo = L.locs[Lo];
                const int shs = ld * (top - lo.s);
                const int ls = (ii >> shs) & lmask;
                int base = ii - (ls << shs);
                int row = ls;
                if (lo.nb == 2) {
                    const int sht = ld * (top - lo.t);
                    const int lt = (ii >> sht) & lmask;
                    base -= lt << sht;
                    row = (ls << ld) | lt;
                }
                acc = ell_row(xc, base, vals, o.off, lo.ent + row * lo.R, lo.R, acc);
            }
        part[w][lane] = acc;
        __syncthreads();
        if (w == 0 && live) {
            cd s = part[0][lane];
#pragma unroll
            for (int q = 1; q < NW; ++q) s = cadd(s, part[q][lane]);
            s = cscale(s, xs);
            if (alpha_mode == 1) s = cmulmi(s);
            y[i] = make_double2(s.re, s.im);
            if (xstore) {
                const double2 xv = x[i];
                vstore(xstore + i, xv.x * xs, xv.y * xs);
            }
        }
        __syncthreads();   // part is rewritten for the next elements
    }
}

// Tiled product: workgroup b owns the tile whose block-spin digits are the
// digits of b; tile element e (tile-spin digits) sits at global index
// base(b) | scatter(e).  The tile is read once (coalesced in runs of the
// lowest spin's digits, which every pass keeps in G), the pass's local
// operators are applied from LDS, and y is written (pass 0, which also
// stores the scaled basis vector) or accumulated.
constexpr int kTT = 1024;   // threads of the tiled product (four elements each)

struct LargeTileLds {
    cd xs[1 << kLargeTileBits];
    cd vals[kLargeTileMaxEnt];
    int32_t offs[kLargeTileMaxEnt];
    DevOpHeader hdr;
    DevOpLocal locs[kLargeTileMaxLocals];
    int32_t map[kLargeTileMapStride];
};

template <typename VT>
__global__ __launch_bounds__(kTT) void large_tile_matvec(LargeTilePass tp, const double2* __restrict__ x,
                                                         const double* __restrict__ xscale_p,
                                                         double2* __restrict__ y, VT* __restrict__ xstore,
                                                         int alpha_mode, int accumulate) {
    __shared__ LargeTileLds S;
    const int tid = threadIdx.x;
    if (tid == 0) S.hdr = *tp.op.hdr;
    if (tid < kLargeTileMapStride) S.map[tid] = tp.map[tid];
    const int nloc = tp.op.hdr->nloc;
    const int nent = tp.op.hdr->nent;
    for (int k = tid; k < nloc; k += kTT) S.locs[k] = tp.op.loc[k];
    for (int k = tid; k < nent; k += kTT) {
        const double2 v = tp.op.val[k];
        S.vals[k] = mkcd(v.x, v.y);
        S.offs[k] = tp.op.off[k];
    }
    __syncthreads();
    const int ld = tp.ld, ng = tp.ng, nb = tp.nb;
    const int64_t lmask = (1 << ld) - 1;
    const int tsize = 1 << (ld * ng);
    int64_t base = 0;
    {
        int64_t b = blockIdx.x;
        for (int q = nb - 1; q >= 0; --q) {
            base |= (b & lmask) << S.map[16 + q];
            b >>= ld;
        }
    }
    auto gidx = [&](int e) -> int64_t {
        int64_t g = base;
        for (int q = ng - 1; q >= 0; --q) {
            g |= ((int64_t)e & lmask) << S.map[q];
            e >>= ld;
        }
        return g;
    };
    const double xsc = xscale_p ? *xscale_p : 1.0;
    for (int e = tid; e < tsize; e += kTT) {
        const int64_t g = gidx(e);
        const double2 xv = x[g];
        S.xs[e] = mkcd(xv.x, xv.y);
        if (xstore) vstore(xstore + g, xv.x * xsc, xv.y * xsc);
    }
    __syncthreads();
    const int top = ng - 1;
    const int lm = (1 << ld) - 1;
    for (int e = tid; e < tsize; e += kTT) {
        // the pass's local operators, tables entry-major (entry e of row r at
        // ent + e * rows + r)
        cd acc = cmul(mkcd(S.hdr.c0re, S.hdr.c0im), S.xs[e]);
        for (int L = 0; L < S.hdr.nloc; ++L) {
            const DevOpLocal lo = S.locs[L];
            const int shs = ld * (top - lo.s);
            const int ls = (e >> shs) & lm;
            int base = e - (ls << shs);
            int row = ls;
            if (lo.nb == 2) {
                const int sht = ld * (top - lo.t);
                const int lt = (e >> sht) & lm;
                base -= lt << sht;
                row = (ls << ld) | lt;
            }
            const int k0 = lo.ent + row;
            for (int q = 0; q < lo.R; ++q) {
                const int k = k0 + q * lo.rows;
                acc = cfma(S.vals[k], S.xs[base + S.offs[k]], acc);
            }
        }
        acc = cscale(acc, xsc);
        if (alpha_mode == 1) acc = cmulmi(acc);
        const int64_t g = gidx(e);
        if (accumulate) {
            const double2 yo = y[g];
            acc = mkcd(acc.re + yo.x, acc.im + yo.y);
        }
        y[g] = make_double2(acc.re, acc.im);
    }
}

// Two-axis product (see LargeAxisPass).  spread() deposits the bits of v at
// the even bit positions (v < 2^16).
__device__ __forceinline__ int64_t spread_bits(int64_t v) {
    v = (v | (v << 8)) & 0x00FF00FFLL;
    v = (v | (v << 4)) & 0x0F0F0F0FLL;
    v = (v | (v << 2)) & 0x33333333LL;
    v = (v | (v << 1)) & 0x55555555LL;
    return v;
}

constexpr int kAT = 1024;   // threads of the two-axis product

template <typename VT>
__global__ __launch_bounds__(kAT) void large_axis_matvec(LargeAxisPass ap, const double2* __restrict__ x,
                                                         const double* __restrict__ xscale_p,
                                                         double2* __restrict__ y, VT* __restrict__ xstore,
                                                         int alpha_mode, int accumulate) {
    __shared__ cd xs[kLargeAxisMaxTile];
    const int tid = threadIdx.x;
    const int k = ap.k, bb = ap.bbits;
    const int D = 1 << k, B = 1 << bb;
    const int tsize = D << bb;
    const int64_t hi = (int64_t)blockIdx.x;      // the other axis's digits above the tile's
    // tile element u (consecutive u: increasing global index) -> axis value a,
    // other-axis low value b, global index
    auto decode = [&](int u, int& a, int& b) {
        const int low = u & ((1 << (2 * bb)) - 1);
        int ev = 0, od = 0;
        for (int q = 0; q < bb; ++q) {
            ev |= ((low >> (2 * q)) & 1) << q;
            od |= ((low >> (2 * q + 1)) & 1) << q;
        }
        const int up = u >> (2 * bb);
        // side 0: even global bits are the ket (axis) digits, odd the bra (b);
        // side 1: odd bits are the bra (axis) digits, even the ket (b)
        if (ap.side == 0) { a = ev | (up << bb); b = od; }
        else { a = od | (up << bb); b = ev; }
    };
    auto gidx = [&](int a, int b) -> int64_t {
        const int64_t other = (hi << bb) | b;
        const int64_t ket = ap.side == 0 ? a : other;
        const int64_t bra = ap.side == 0 ? other : a;
        return spread_bits(ket) | (spread_bits(bra) << 1);
    };
    const double xsc = xscale_p ? *xscale_p : 1.0;
    for (int u = tid; u < tsize; u += kAT) {
        int a, b;
        decode(u, a, b);
        const int64_t g = gidx(a, b);
        const double2 xv = x[g];
        xs[a * B + b] = mkcd(xv.x, xv.y);
        if (xstore) vstore(xstore + g, xv.x * xsc, xv.y * xsc);
    }
    __syncthreads();
    // rows of the axis operator: thread r handles rows r, r + kAT, ..
    constexpr int kMaxRows = kLargeAxisMaxTile / kAT;   // rows x B <= tile / kAT per thread
    cd acc[kMaxRows];
    const int rows_per = (D + kAT - 1) / kAT;
#pragma unroll
    for (int q = 0; q < kMaxRows; ++q) acc[q] = mkcd(0.0, 0.0);
    // accumulator q = (row slot rq, column b): rq * B + b < kMaxRows
    // the row's entries in chunks of 8: the table loads of a chunk are issued
    // together (one cache-miss latency per chunk, not per entry)
    constexpr int kChunk = 8;
    for (int rq = 0; rq < rows_per; ++rq) {
        const int r = tid + rq * kAT;
        if (r >= D) break;
        for (int j0 = 0; j0 < ap.W; j0 += kChunk) {
            double2 v[kChunk];
            int c[kChunk];
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const bool in = j0 + u < ap.W;
                const int64_t at = (int64_t)(in ? j0 + u : 0) * D + r;
                v[u] = ap.val[at];
                c[u] = ap.col[at];
                if (!in) v[u] = make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int u = 0; u < kChunk; ++u) {
                const cd vv = mkcd(v[u].x, v[u].y);
#pragma unroll
                for (int q = 0; q < kMaxRows; ++q) {
                    const int bq = q - rq * B;
                    if (bq >= 0 && bq < B) acc[q] = cfma(vv, xs[c[u] * B + bq], acc[q]);
                }
            }
        }
    }
    __syncthreads();   // every thread has read xs
    for (int rq = 0; rq < rows_per; ++rq) {
        const int r = tid + rq * kAT;
        if (r >= D) break;
#pragma unroll
        for (int q = 0; q < kMaxRows; ++q) {
            const int bq = q - rq * B;
            if (bq >= 0 && bq < B) {
                cd v = cscale(acc[q], xsc);
                if (alpha_mode == 1) v = cmulmi(v);
                xs[r * B + bq] = v;
            }
        }
    }
    __syncthreads();
    for (int u = tid; u < tsize; u += kAT) {
        int a, b;
        decode(u, a, b);
        const int64_t g = gidx(a, b);
        cd v = xs[a * B + b];
        if (accumulate) {
            const double2 yo = y[g];
            v = mkcd(v.re + yo.x, v.im + yo.y);
        }
        y[g] = make_double2(v.re, v.im);
    }
}

// Same product for a general sparse matrix in CSR form (row_ptr[n+1], col,
// val): y = alpha * A (xscale * x), alpha_mode 0: +1, 1: -i.
__global__ __launch_bounds__(kLT) void large_csr_matvec(const int64_t* __restrict__ row_ptr,
                                                        const int32_t* __restrict__ col,
                                                        const double2* __restrict__ val, const double2* __restrict__ x,
                                                        const double* __restrict__ xscale_p, double2* __restrict__ y,
                                                        double2* __restrict__ xstore, int64_t n, int alpha_mode) {
    const double xs = xscale_p ? *xscale_p : 1.0;
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        cd acc = mkcd(0.0, 0.0);
        for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
            const double2 a = val[e], xv = x[col[e]];
            acc = cfma(mkcd(a.x, a.y), mkcd(xv.x, xv.y), acc);
        }
        acc = cscale(acc, xs);
        if (alpha_mode == 1) acc = cmulmi(acc);
        y[i] = make_double2(acc.re, acc.im);
        if (xstore) {
            const double2 xv = x[i];
            xstore[i] = make_double2(xv.x * xs, xv.y * xs);
        }
    }
}

// part[(l * gridDim.x + block) * 2 + {re, im}] = partial conj(V_l) . p,
// l = 0..J; l = J+1 (if with_norm) holds |p|^2 in the re slot.
__global__ __launch_bounds__(kLT) void large_dots(const double2* __restrict__ V, int64_t ldv,
                                                  const double2* __restrict__ p, int64_t n, int J, int with_norm,
                                                  double* __restrict__ part) {
    __shared__ double red[2 * (kLT / kWave)];
    const int l = blockIdx.y;
    double sr = 0.0, si = 0.0;
    if (l <= J) {
        const double2* v = V + (int64_t)l * ldv;
        for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
            const double2 a = v[i], b = p[i];
            sr = fma(a.x, b.x, fma(a.y, b.y, sr));
            si = fma(a.x, b.y, fma(-a.y, b.x, si));
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
            const double2 b = p[i];
            sr = fma(b.x, b.x, fma(b.y, b.y, sr));
        }
    }
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) {
        sr += __shfl_xor(sr, mask, kWave);
        si += __shfl_xor(si, mask, kWave);
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) { red[2 * wave] = sr; red[2 * wave + 1] = si; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kLT / kWave; ++w) { a += red[2 * w]; b += red[2 * w + 1]; }
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2] = a;
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2 + 1] = b;
    }
}

// Sum the partials of the Gram-Schmidt passes and of large_dots for
// Arnoldi column jcol (one block):
//   mode 0: after the first dots pass (nrows = jcol + 2): h(l, jcol) = coef_l
//           = sums for l <= jcol; row jcol + 1 is |p|^2, kept in scal[3];
//   mode 1: after the first update-and-norm pass (row 0 = |p'|^2): when
//           |p'| >= |p| / 2 (the first pass removed at most 3/4 of |p|^2:
//           the Daniel-Gragg-Kaufman-Stewart test) the second pass would
//           change p' by rounding only, so flags[2] = 1 (skip) and the norm
//           entry is set as in mode 2; otherwise flags[2] = 0;
//   mode 4: after the second dots pass (skipped with flags[2]):
//           h(l, jcol) += sums, coef_l = sums for l <= jcol;
//   mode 2: after the second update-and-norm pass (skipped with flags[2]):
//           h(jcol+1, jcol) = sqrt(|p|^2) unless it is <= break_tol (then
//           flags = {1, jcol+1} once); scal[0] = 1/h (0 after a
//           breakdown), scal[1] = h;
//   mode 3: norm (row 0 of part) into scal[2] (the extra product's norm).
constexpr int kFinishThreads = 1024;

// One wave per row: the lanes sum strided partials (independent loads), then
// a wave reduction; the block's waves take the rows in turn.
__device__ void finish_rows(const double* __restrict__ part, int nblk, int nrows, int jcol, int mode,
                            double2* __restrict__ hess, int ldh, double2* __restrict__ coef,
                            double* __restrict__ scal, int* __restrict__ flags, double break_tol) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    for (int l = wave; l < nrows; l += nw) {
        double a = 0.0, b = 0.0;
        // eight loads in flight per lane (the sums keep their order)
        const double2* pr = reinterpret_cast<const double2*>(part) + (int64_t)l * nblk;
        int k = lane;
        for (; k + 7 * kWave < nblk; k += 8 * kWave) {
            double2 v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = pr[k + u * kWave];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                a += v[u].x;
                b += v[u].y;
            }
        }
        for (; k < nblk; k += kWave) {
            a += pr[k].x;
            b += pr[k].y;
        }
#pragma unroll
        for (int mask = 32; mask >= 1; mask >>= 1) {
            a += __shfl_xor(a, mask, kWave);
            b += __shfl_xor(b, mask, kWave);
        }
        if (lane != 0) continue;
        if (mode == 1 || mode == 2) {
            if (mode == 1) {
                const int skip = a >= 0.25 * scal[3];
                flags[2] = skip;
                if (!skip) continue;
            }
            const double h = sqrt(a);
            const bool broke = flags[0] != 0 || h <= break_tol;
            if (!broke) hess[(jcol + 1) + (int64_t)ldh * jcol] = make_double2(h, 0.0);
            if (flags[0] == 0 && h <= break_tol) {
                flags[0] = 1;
                flags[1] = jcol + 1;  // basis size at the breakdown
            }
            scal[0] = broke ? 0.0 : 1.0 / h;
            scal[1] = h;
        } else if (mode == 3) {
            scal[2] = sqrt(a);
        } else if (l == jcol + 1) {
            if (mode == 0) scal[3] = a;   // |p|^2 before the first pass
        } else {
            coef[l] = make_double2(a, b);
            double2& hv = hess[l + (int64_t)ldh * jcol];
            if (mode == 0) hv = make_double2(a, b);
            else hv = make_double2(hv.x + a, hv.y + b);
        }
    }
}

__global__ __launch_bounds__(kFinishThreads) void large_finish(const double* __restrict__ part, int nblk, int nrows,
                                                               int jcol, int mode, double2* __restrict__ hess, int ldh,
                                                               double2* __restrict__ coef, double* __restrict__ scal,
                                                               int* __restrict__ flags, double break_tol) {
    if ((mode == 2 || mode == 4) && flags[2]) return;   // no second pass
    finish_rows(part, nblk, nrows, jcol, mode, hess, ldh, coef, scal, flags, break_tol);
}

// Classical Gram-Schmidt passes (one read of each basis vector per pass):
//   large_gs_dots:    part_l = V_l^H p (l <= J), part_{J+1} = |p|^2
//   large_gs_update:  p -= sum_l coef_l V_l, part_0 = |p|^2
// A pass launched with a skip flag returns at once while *skip is set (the
// second pass of a column that did not need it).  Partial sums are reduced
// per wave with shuffles, then across the block's waves in LDS.  MAXJ bounds
// J + 1.  Every thread issues all MAXJ basis loads of an element without a
// branch (indices past J repeat column J, a cache hit; their partial sums
// are not reduced, their update coefficients are zero), so the loads are in
// flight together instead of one at a time behind a per-column branch.

// Rows 0..nrows-1 from (sr, si) and row nrows = (x, 0); red holds MAXJ + 1
// rows per wave.
template <int MAXJ>
__device__ __forceinline__ void gs_block_reduce(const double (&sr)[MAXJ], const double (&si)[MAXJ], int nrows,
                                                double* red, double* __restrict__ part, double x) {
    constexpr int RS = MAXJ + 1;   // row stride of red
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int l = 0; l < MAXJ; ++l) {
        if (l < nrows) {
            double a = sr[l], b = si[l];
#pragma unroll
            for (int mask = 32; mask >= 1; mask >>= 1) {
                a += __shfl_xor(a, mask, kWave);
                b += __shfl_xor(b, mask, kWave);
            }
            if (lane == 0) {
                red[(wave * RS + l) * 2] = a;
                red[(wave * RS + l) * 2 + 1] = b;
            }
        }
    }
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) x += __shfl_xor(x, mask, kWave);
    if (lane == 0) {
        red[(wave * RS + nrows) * 2] = x;
        red[(wave * RS + nrows) * 2 + 1] = 0.0;
    }
    __syncthreads();
    for (int l = threadIdx.x; l <= nrows; l += kLT) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kLT / kWave; ++w) { a += red[(w * RS + l) * 2]; b += red[(w * RS + l) * 2 + 1]; }
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2] = a;
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2 + 1] = b;
    }
}

template <int MAXJ, typename VT>
__global__ __launch_bounds__(kLT) void large_gs_dots(const VT* __restrict__ V, int64_t ldv,
                                                     const double2* __restrict__ p, int64_t n, int J,
                                                     double* __restrict__ part, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double red[2 * (MAXJ + 1) * (kLT / kWave)];
    double sr[MAXJ], si[MAXJ];
    double nr = 0.0;   // |p|^2
#pragma unroll
    for (int l = 0; l < MAXJ; ++l) { sr[l] = 0.0; si[l] = 0.0; }
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        const double2 b = p[i];
        nr = fma(b.x, b.x, fma(b.y, b.y, nr));
#pragma unroll
        for (int l = 0; l < MAXJ; ++l) {
            const double2 a = vload(V + (int64_t)min(l, J) * ldv + i);
            sr[l] = fma(a.x, b.x, fma(a.y, b.y, sr[l]));
            si[l] = fma(a.x, b.y, fma(-a.y, b.x, si[l]));
        }
    }
    gs_block_reduce<MAXJ>(sr, si, J + 1, red, part, nr);
}

template <int MAXJ, typename VT>
__global__ __launch_bounds__(kLT) void large_gs_update(const VT* __restrict__ V, int64_t ldv,
                                                       double2* __restrict__ p, int64_t n, int J,
                                                       const double2* __restrict__ coef,
                                                       double* __restrict__ part, const int* __restrict__ skip) {
    if (skip && *skip) return;
    __shared__ double red[2 * (kLT / kWave)];
    __shared__ double2 c[MAXJ];
    for (int l = threadIdx.x; l < MAXJ; l += kLT) c[l] = l <= J ? coef[l] : make_double2(0.0, 0.0);
    __syncthreads();
    double nr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        cd acc = mkcd(p[i].x, p[i].y);
#pragma unroll
        for (int l = 0; l < MAXJ; ++l) {
            const double2 v = vload(V + (int64_t)min(l, J) * ldv + i);
            acc = cfms(mkcd(c[l].x, c[l].y), mkcd(v.x, v.y), acc);
        }
        p[i] = make_double2(acc.re, acc.im);
        nr = fma(acc.re, acc.re, fma(acc.im, acc.im, nr));
    }
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) nr += __shfl_xor(nr, mask, kWave);
    if (lane == 0) red[wave] = nr;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int w = 0; w < kLT / kWave; ++w) a += red[w];
        part[(int64_t)blockIdx.x * 2] = a;
        part[(int64_t)blockIdx.x * 2 + 1] = 0.0;
    }
}

// The dots pass with the basis range split over S lane groups: the S lanes
// q, q + 64/S, .. of a wave own the same element, group g the vectors
// l in [g L, (g + 1) L), L = ceil((J + 1) / S), so each lane holds L partial
// sums and loads (long bases keep four or five waves per SIMD with all loads
// of a lane in flight).  Per wave 64/S elements; |p|^2 is summed by group 0.
template <int S, int HJ>
__device__ __forceinline__ void gs_split_reduce(const double (&sr)[HJ], const double (&si)[HJ], int L, int nrows,
                                                double* red, double* __restrict__ part, double x) {
    constexpr int GL = kWave / S;      // lanes per group
    constexpr int RS = S * HJ + 1;     // row stride of red (the last row: x, row nrows of part)
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int g = lane / GL;
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) x += __shfl_xor(x, mask, kWave);
    if (lane == 0) {
        red[(wave * RS + nrows) * 2] = x;
        red[(wave * RS + nrows) * 2 + 1] = 0.0;
    }
#pragma unroll
    for (int q = 0; q < HJ; ++q) {
        double a = sr[q], b = si[q];
#pragma unroll
        for (int mask = GL / 2; mask >= 1; mask >>= 1) {
            a += __shfl_xor(a, mask, kWave);
            b += __shfl_xor(b, mask, kWave);
        }
        const int l = g * L + q;
        if ((lane & (GL - 1)) == 0 && q < L && l < nrows) {
            red[(wave * RS + l) * 2] = a;
            red[(wave * RS + l) * 2 + 1] = b;
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l <= nrows; l += kLT) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kLT / kWave; ++w) { a += red[(w * RS + l) * 2]; b += red[(w * RS + l) * 2 + 1]; }
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2] = a;
        part[((int64_t)l * gridDim.x + blockIdx.x) * 2 + 1] = b;
    }
}

template <int S, int HJ, typename VT>
__global__ __launch_bounds__(kLT) void large_gs_dots_split(const VT* __restrict__ V, int64_t ldv,
                                                           const double2* __restrict__ p, int64_t n, int J,
                                                           double* __restrict__ part, const int* __restrict__ skip) {
    if (skip && *skip) return;
    constexpr int GL = kWave / S;
    __shared__ double red[2 * (S * HJ + 1) * (kLT / kWave)];
    const int lane = threadIdx.x & (kWave - 1);
    const int g = lane / GL;
    const int L = (J + S) / S;                       // ceil((J + 1) / S) <= HJ
    const int lb = g * L, le = min(lb + L, J + 1);   // this group's vectors (le <= lb: none)
    const int lmax = max(le - 1, 0);
    double sr[HJ], si[HJ];
#pragma unroll
    for (int q = 0; q < HJ; ++q) { sr[q] = 0.0; si[q] = 0.0; }
    const int64_t w0 = ((int64_t)blockIdx.x * kLT + threadIdx.x - lane) / S;   // this wave's first element
    double nr = 0.0;   // |p|^2, counted by group 0 (each element once)
    for (int64_t i = w0 + (lane & (GL - 1)); i < n; i += (int64_t)gridDim.x * (kLT / S)) {
        const double2 b = p[i];
        if (g == 0) nr = fma(b.x, b.x, fma(b.y, b.y, nr));
#pragma unroll
        for (int q = 0; q < HJ; ++q) {
            const double2 a = vload(V + (int64_t)min(lb + q, lmax) * ldv + i);
            sr[q] = fma(a.x, b.x, fma(a.y, b.y, sr[q]));
            si[q] = fma(a.x, b.y, fma(-a.y, b.x, si[q]));
        }
    }
    gs_split_reduce<S, HJ>(sr, si, L, J + 1, red, part, nr);
}

// F = exp(t H(0:mx,0:mx)) e1 for the augmented Hessenberg matrix (Pade(6),
// one workgroup).  status[0] = number of squarings or a negative error.
struct LargePadeSmem {
    cd W[4][(kLargeMaxM + 2) * (kLargeMaxM + 3)];  // one spare column per matrix (tiled reads)
    cd H[(kLargeMaxM + 2) * (kLargeMaxM + 3)];     // the input matrix, copied in once
    cd F[kLargeMaxM + 2];
    int piv[128];
    double scal[16];
};

// H (mx x mx, leading dimension ldh <= kLargeMaxM + 2) from device memory
// into LDS: the norm, the two products with H and the solve then read LDS
// instead of waiting on device-memory latency in their inner loops.
__device__ __forceinline__ const cd* pade_input(const double2* __restrict__ hess, int ldh, int mx,
                                                LargePadeSmem& S) {
    for (int idx = threadIdx.x; idx < mx * mx; idx += kLT) {
        const int i = idx % mx, j = idx / mx;
        const double2 v = hess[i + (int64_t)ldh * j];
        S.H[i + ldh * j] = mkcd(v.x, v.y);
    }
    __syncthreads();
    return S.H;
}

__global__ __launch_bounds__(kLT) void large_pade(const double2* __restrict__ hess, int ldh, int mx, double t,
                                                  double2* __restrict__ F, int* __restrict__ status) {
    extern __shared__ unsigned char dyn[];
    LargePadeSmem& S = *reinterpret_cast<LargePadeSmem*>(dyn);
    const cd* H = pade_input(hess, ldh, mx, S);
    const int ns = pade6_first_column<kLT>(H, ldh, mx, t, S.W[0], S.W[1], S.W[2], S.W[3], S.F, S.piv, S.scal);
    for (int i = threadIdx.x; i < mx; i += kLT) F[i] = make_double2(S.F[i].re, S.F[i].im);
    if (threadIdx.x == 0) status[0] = ns;
}

// Full exp(t H) (m x m, leading dimension ldh, column-major) by Pade(6) with
// scaling and squaring (ZGPADM, expokit.f:762-934), one workgroup; out has
// leading dimension m.  status[0] = squarings or a negative error.
__global__ __launch_bounds__(kLT) void large_pade_full(const double2* __restrict__ hess, int ldh, int mx, double t,
                                                       double2* __restrict__ out, int* __restrict__ status) {
    extern __shared__ unsigned char dyn[];
    LargePadeSmem& S = *reinterpret_cast<LargePadeSmem*>(dyn);
    const cd* H = pade_input(hess, ldh, mx, S);
    cd* R = nullptr;
    const int ns = pade6_full<kLT>(H, ldh, mx, t, S.W[0], S.W[1], S.W[2], S.W[3], S.piv, S.scal, &R);
    if (ns >= 0)
        for (int idx = threadIdx.x; idx < mx * mx; idx += kLT) {
            const int i = idx % mx, j = idx / mx;
            out[idx] = make_double2(R[i + ldh * j].re, R[i + ldh * j].im);
        }
    if (threadIdx.x == 0) status[0] = ns;
}

// w = beta * sum_{j<mx} F_j V_j ; partial |w|^2 per block into part
template <typename VT>
__global__ __launch_bounds__(kLT) void large_combine(const VT* __restrict__ V, int64_t ldv, int mx,
                                                     const double2* __restrict__ F, double beta,
                                                     double2* __restrict__ w, int64_t n, double* __restrict__ part) {
    __shared__ double2 c[kLargeMaxM + 2];
    __shared__ double red[kLT / kWave];
    for (int l = threadIdx.x; l < mx; l += kLT) c[l] = make_double2(F[l].x * beta, F[l].y * beta);
    __syncthreads();
    double s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        cd acc = mkcd(0.0, 0.0);
        for (int l = 0; l < mx; ++l) {
            const double2 v = vload(V + (int64_t)l * ldv + i);
            acc = cfma(mkcd(c[l].x, c[l].y), mkcd(v.x, v.y), acc);
        }
        w[i] = make_double2(acc.re, acc.im);
        s2 = fma(acc.re, acc.re, fma(acc.im, acc.im, s2));
    }
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) s2 += __shfl_xor(s2, mask, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = s2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0;
        for (int q = 0; q < kLT / kWave; ++q) a += red[q];
        part[blockIdx.x] = a;
    }
}

// y_i = a x_i + b y_i (complex a, b), small BLAS-1 helper (scaled copies).
// x may alias y; y is not read when b == 0.
__global__ __launch_bounds__(kLT) void large_axpby(double2 a, const double2* x, double2 b, double2* y, int64_t n) {
    const bool read_y = b.x != 0.0 || b.y != 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kLT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLT) {
        const cd xv = mkcd(x[i].x, x[i].y);
        cd r = cmul(mkcd(a.x, a.y), xv);
        if (read_y) r = cadd(r, cmul(mkcd(b.x, b.y), mkcd(y[i].x, y[i].y)));
        y[i] = make_double2(r.re, r.im);
    }
}

// Launch wrappers -------------------------------------------------------------
int large_grid(int64_t n, int num_cu) {
    const int64_t want = (n + kLT - 1) / kLT;
    const int64_t cap = (int64_t)num_cu * 8;
    return (int)(want < cap ? want : cap);
}

void launch_large_matvec(const LargeOpArgs& o, const double2* x, const double* xscale, double2* y, void* xstore,
                         int64_t n, int alpha_mode, int grid, hipStream_t s, bool fp32) {
    if (n <= (int64_t(1) << 17)) {
        // 64 elements per 1024-thread workgroup, the operators over its 16 waves
        constexpr int NW = 16;
        const int g = (int)std::min<int64_t>((n + kWave - 1) / kWave, 4096);
        if (fp32)
            hipLaunchKernelGGL((large_matvec_waves<float2, NW>), dim3(g), dim3(NW * kWave), 0, s, o, x, xscale, y,
                               (float2*)xstore, n, alpha_mode);
        else
            hipLaunchKernelGGL((large_matvec_waves<double2, NW>), dim3(g), dim3(NW * kWave), 0, s, o, x, xscale, y,
                               (double2*)xstore, n, alpha_mode);
        return;
    }
    if (fp32)
        hipLaunchKernelGGL(large_matvec<float2>, dim3(grid), dim3(kLT), 0, s, o, x, xscale, y, (float2*)xstore, n,
                           alpha_mode);
    else
        hipLaunchKernelGGL(large_matvec<double2>, dim3(grid), dim3(kLT), 0, s, o, x, xscale, y, (double2*)xstore, n,
                           alpha_mode);
}

void launch_large_axis_matvec(const LargeAxisPass& ap, const double2* x, const double* xscale, double2* y,
                              void* xstore, int alpha_mode, int accumulate, hipStream_t s, bool fp32) {
    if (fp32)
        hipLaunchKernelGGL(large_axis_matvec<float2>, dim3(ap.ntiles), dim3(kAT), 0, s, ap, x, xscale, y,
                           (float2*)xstore, alpha_mode, accumulate);
    else
        hipLaunchKernelGGL(large_axis_matvec<double2>, dim3(ap.ntiles), dim3(kAT), 0, s, ap, x, xscale, y,
                           (double2*)xstore, alpha_mode, accumulate);
}

void launch_large_tile_matvec(const LargeTilePass& tp, const double2* x, const double* xscale, double2* y,
                              void* xstore, int alpha_mode, int accumulate, hipStream_t s, bool fp32) {
    if (fp32)
        hipLaunchKernelGGL(large_tile_matvec<float2>, dim3(tp.ntiles), dim3(kTT), 0, s, tp, x, xscale, y,
                           (float2*)xstore, alpha_mode, accumulate);
    else
        hipLaunchKernelGGL(large_tile_matvec<double2>, dim3(tp.ntiles), dim3(kTT), 0, s, tp, x, xscale, y,
                           (double2*)xstore, alpha_mode, accumulate);
}

void launch_large_dots(const double2* V, int64_t ldv, const double2* p, int64_t n, int J, int with_norm, double* part,
                       int nblk, hipStream_t s) {
    hipLaunchKernelGGL(large_dots, dim3(nblk, J + 1 + (with_norm ? 1 : 0)), dim3(kLT), 0, s, V, ldv, p, n, J,
                       with_norm, part);
}

void launch_large_finish(const double* part, int nblk, int nrows, int jcol, int mode, double2* hess, int ldh,
                         double2* coef, double* scal, int* flags, double break_tol, hipStream_t s) {
    hipLaunchKernelGGL(large_finish, dim3(1), dim3(kFinishThreads), 0, s, part, nblk, nrows, jcol, mode, hess, ldh,
                       coef, scal, flags, break_tol);
}

// Gram-Schmidt passes; the register footprint follows the number of basis
// vectors (J + 1 <= 8, 16, 24, 32 or 40).
template <int MAXJ>
static void gs_dots_j(const void* V, int64_t ldv, const double2* p, int64_t n, int J, double* part, int nblk,
                      hipStream_t s, bool fp32, const int* skip) {
    if (fp32)
        hipLaunchKernelGGL((large_gs_dots<MAXJ, float2>), dim3(nblk), dim3(kLT), 0, s, (const float2*)V, ldv, p, n, J,
                           part, skip);
    else
        hipLaunchKernelGGL((large_gs_dots<MAXJ, double2>), dim3(nblk), dim3(kLT), 0, s, (const double2*)V, ldv, p, n,
                           J, part, skip);
}

template <int MAXJ>
static void gs_update_j(const void* V, int64_t ldv, double2* p, int64_t n, int J, const double2* coef, double* part,
                        int nblk, hipStream_t s, bool fp32, const int* skip) {
    if (fp32)
        hipLaunchKernelGGL((large_gs_update<MAXJ, float2>), dim3(nblk), dim3(kLT), 0, s, (const float2*)V, ldv, p, n,
                           J, coef, part, skip);
    else
        hipLaunchKernelGGL((large_gs_update<MAXJ, double2>), dim3(nblk), dim3(kLT), 0, s, (const double2*)V, ldv, p,
                           n, J, coef, part, skip);
}

template <int S, int HJ>
static void gs_dots_s(const void* V, int64_t ldv, const double2* p, int64_t n, int J, double* part, int nblk,
                      hipStream_t s, bool fp32, const int* skip) {
    if (fp32)
        hipLaunchKernelGGL((large_gs_dots_split<S, HJ, float2>), dim3(nblk), dim3(kLT), 0, s, (const float2*)V, ldv,
                           p, n, J, part, skip);
    else
        hipLaunchKernelGGL((large_gs_dots_split<S, HJ, double2>), dim3(nblk), dim3(kLT), 0, s, (const double2*)V, ldv,
                           p, n, J, part, skip);
}

void launch_large_gs_dots(const void* V, int64_t ldv, const double2* p, int64_t n, int J, double* part, int nblk,
                          hipStream_t s, bool fp32, const int* skip) {
    static_assert(kLargeMaxM <= 40, "the split dots pass holds at most 4 x 10 vectors");
    if (J < 8) gs_dots_j<8>(V, ldv, p, n, J, part, nblk, s, fp32, skip);
    else if (J < 16) gs_dots_j<16>(V, ldv, p, n, J, part, nblk, s, fp32, skip);
    else if (J < 24) gs_dots_s<2, 12>(V, ldv, p, n, J, part, nblk, s, fp32, skip);
    else gs_dots_s<4, 10>(V, ldv, p, n, J, part, nblk, s, fp32, skip);
}

void launch_large_gs_update(const void* V, int64_t ldv, double2* p, int64_t n, int J, const double2* coef,
                            double* part, int nblk, hipStream_t s, bool fp32, const int* skip) {
    if (J < 8) gs_update_j<8>(V, ldv, p, n, J, coef, part, nblk, s, fp32, skip);
    else if (J < 16) gs_update_j<16>(V, ldv, p, n, J, coef, part, nblk, s, fp32, skip);
    else if (J < 24) gs_update_j<24>(V, ldv, p, n, J, coef, part, nblk, s, fp32, skip);
    else if (J < 32) gs_update_j<32>(V, ldv, p, n, J, coef, part, nblk, s, fp32, skip);
    else gs_update_j<kLargeMaxM>(V, ldv, p, n, J, coef, part, nblk, s, fp32, skip);
}

int launch_large_pade(const double2* hess, int ldh, int mx, double t, double2* F, int* status, hipStream_t s) {
    const size_t bytes = sizeof(LargePadeSmem);
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void*)large_pade, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
            hipSuccess)
            return -1;
        attr_set = true;
    }
    hipLaunchKernelGGL(large_pade, dim3(1), dim3(kLT), bytes, s, hess, ldh, mx, t, F, status);
    return 0;
}

void launch_large_csr_matvec(const int64_t* row_ptr, const int32_t* col, const double2* val, const double2* x,
                             const double* xscale, double2* y, double2* xstore, int64_t n, int alpha_mode, int grid,
                             hipStream_t s) {
    hipLaunchKernelGGL(large_csr_matvec, dim3(grid), dim3(kLT), 0, s, row_ptr, col, val, x, xscale, y, xstore, n,
                       alpha_mode);
}

int launch_large_pade_full(const double2* hess, int ldh, int mx, double t, double2* out, int* status, hipStream_t s) {
    const size_t bytes = sizeof(LargePadeSmem);
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void*)large_pade_full, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes) != hipSuccess)
            return -1;
        attr_set = true;
    }
    hipLaunchKernelGGL(large_pade_full, dim3(1), dim3(kLT), bytes, s, hess, ldh, mx, t, out, status);
    return 0;
}

__global__ __launch_bounds__(kLT) void large_diag_add(double2* A, int m, double2 v) {
    const int i = blockIdx.x * kLT + threadIdx.x;
    if (i < m) {
        double2* p = A + (int64_t)i * (m + 1);
        *p = make_double2(p->x + v.x, p->y + v.y);
    }
}

void launch_large_diag_add(double2* A, int m, double2 v, hipStream_t s) {
    hipLaunchKernelGGL(large_diag_add, dim3((m + kLT - 1) / kLT), dim3(kLT), 0, s, A, m, v);
}

void launch_large_combine(const void* V, int64_t ldv, int mx, const double2* F, double beta, double2* w, int64_t n,
                          double* part, int grid, hipStream_t s, bool fp32) {
    if (fp32)
        hipLaunchKernelGGL(large_combine<float2>, dim3(grid), dim3(kLT), 0, s, (const float2*)V, ldv, mx, F, beta, w,
                           n, part);
    else
        hipLaunchKernelGGL(large_combine<double2>, dim3(grid), dim3(kLT), 0, s, (const double2*)V, ldv, mx, F, beta,
                           w, n, part);
}

void launch_large_axpby(double2 a, const double2* x, double2 b, double2* y, int64_t n, int grid, hipStream_t s) {
    hipLaunchKernelGGL(large_axpby, dim3(grid), dim3(kLT), 0, s, a, x, b, y, n);
}

}  // namespace oops
