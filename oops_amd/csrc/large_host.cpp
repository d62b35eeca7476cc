// Host driver of the large-state Krylov path (kernels in expv_large.hip).
//
// Restates the control flow of the reference's multi-output Krylov driver
// (src/expv/source/hamvec_zgexpv_w_mkl_profile.f:4-538; its GPU twin
// src/expv/source/zgexpv_cache.f:4-690) with every vector on the device:
// norm estimate (Lanczos k + power iteration, :210-265), first step
// (:329-334), m Arnoldi steps + one extra product, augmented Hessenberg,
// Pade(6) on the device, local error estimate, rejection (gamma 0.9,
// delta 1.2), update, 2-significant-digit step rounding, output at every
// requested time.  Only scalars (breakdown flag, norms, two entries of the
// small exponential) come back per Krylov step.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/oops_expv.h"
#include "host_common.h"
#include "oops_large.h"

using namespace oops;

namespace {

typedef std::complex<double> cplx;

double round_step(double t) {
    if (!(t > 0.0) || !std::isfinite(t)) return std::numeric_limits<double>::infinity();
    const double sqr1 = std::sqrt(0.1);
    const double x = std::log10(t) - sqr1;
    const int k = (int)(x >= 0.0 ? std::floor(x + 0.5) : -std::floor(-x + 0.5)) - 1;
    double p = 1.0;
    for (int i = 0; i < std::abs(k); ++i) p *= 10.0;
    if (k < 0) p = 1.0 / p;
    return std::trunc(t / p + 0.55) * p;
}

struct LargeRun {
    oops_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    CompiledOp op;
    DevBuf hdr, loc, val, off;
    LargeOpArgs args{};
    // general sparse operator (CSR) instead of the Kronecker tables
    bool csr = false;
    DevBuf rp, ci, cv;
    // basis stored in complex FP32 (the FP32 variant; Kronecker operators only)
    bool fp32 = false;
    int64_t n = 0;
    int64_t ldv = 0;   // basis column stride (elements)
    int m = 0, ldh = 0, grid = 0, nblk = 0, nblk_dots = 0;
    DevBuf V, p0, p1, w, v0, part, hess, coef, scal, flags, F, status, wpart;
    DevBuf wk, wkpart;   // a state at an output time inside a step (interior-output mode)
    // tiled product (large Kronecker operators): one entry per pass
    struct Tile {
        DevBuf hdr, loc, val, off, map;
        LargeTilePass pass{};
    };
    std::vector<Tile> tiles;
    // two-axis product (Liouville-space operators split into ket / bra sides)
    struct Axis {
        DevBuf val, col;
        LargeAxisPass pass{};
    };
    std::vector<Axis> axis;
    // local dimensions that are not one power of two (spin-1: 3, its
    // Liouville space: 9, mixed spins): every spin is embedded in the next
    // power of two; pos[i] is the device index of caller index i
    bool padded = false;
    int64_t n_logical = 0;
    std::vector<int64_t> pos;
    int rc = 0;
};

// Pass plan of the tiled product.  The lowest spins filling kLargeTileBits
// bits form pass 0; the other passes hold every spin outside that set (the
// "outer" spins), the lowest spin (coalesced runs) and a chunk of the
// remaining inner spins, so that every pair (outer, inner) and every local
// operator among the outer spins lies inside some pass.  Each local operator
// goes to the first pass containing its spins.  Returns false (flat ELL
// product) when the state is small or the outer spins do not fit a tile.
struct TilePlan {
    std::vector<int> G;   // spins of the tile, slowest first
    CompiledOp op;        // the pass's local operators in tile digits
};

static bool plan_tiles(const CompiledOp& op, std::vector<TilePlan>& plan) {
    const int ns = op.nspin, ld = op.ld_bits;
    if (ld < 1 || ns * ld < 19 || ns * ld > 30) return false;
    const int tmax = kLargeTileBits / ld;
    const int nouter = ns - tmax;
    if (nouter < 1 || nouter + 2 > tmax || nouter > 32) return false;
    std::vector<std::vector<int>> G;
    {
        std::vector<int> g0;
        for (int s = nouter; s < ns; ++s) g0.push_back(s);
        G.push_back(g0);
        const int chunk = tmax - nouter - 1;
        for (int s0 = nouter; s0 < ns - 1; s0 += chunk) {
            std::vector<int> g;
            for (int s = 0; s < nouter; ++s) g.push_back(s);
            for (int s = s0; s < std::min(s0 + chunk, ns - 1); ++s) g.push_back(s);
            g.push_back(ns - 1);
            G.push_back(g);
        }
    }
    const int64_t lmask = (1LL << ld) - 1;
    auto gpos = [&](int s) { return ld * (ns - 1 - s); };
    plan.assign(G.size(), TilePlan{});
    std::vector<int> done(op.locs.size(), 0);
    for (size_t p = 0; p < G.size(); ++p) {
        const std::vector<int>& g = G[p];
        std::vector<int> tq(ns, -1);
        for (size_t q = 0; q < g.size(); ++q) tq[g[q]] = (int)q;
        const int ng = (int)g.size();
        auto tpos = [&](int q) { return ld * (ng - 1 - q); };
        CompiledOp& T = plan[p].op;
        plan[p].G = g;
        T.n = 1LL << (ld * ng);
        T.nspin = ng;
        T.ld_bits = ld;
        T.c0 = p == 0 ? op.c0 : std::complex<double>(0.0, 0.0);
        for (int pass_nb = 1; pass_nb <= 2; ++pass_nb) {   // 1-body entries first
            if (pass_nb == 2) T.ent2_first = (int32_t)T.val.size();
            for (size_t L = 0; L < op.locs.size(); ++L) {
                const DevOpLocal& lo = op.locs[L];
                if (done[L] || lo.nb != pass_nb) continue;
                if (tq[lo.s] < 0 || (lo.nb == 2 && tq[lo.t] < 0)) continue;
                done[L] = 1;
                DevOpLocal t = lo;
                t.s = (int16_t)tq[lo.s];
                t.t = (int16_t)(lo.nb == 2 ? tq[lo.t] : tq[lo.s]);
                t.ent = (int32_t)T.val.size();
                // entry-major (entry e of every row, then e + 1): lanes whose
                // rows differ read consecutive table slots (no LDS bank conflicts)
                for (int64_t e = 0; e < lo.R; ++e)
                    for (int64_t row = 0; row < lo.rows; ++row) {
                        const int64_t k = lo.ent + row * lo.R + e;
                        const int64_t off = op.off[k];
                        int64_t o = ((off >> gpos(lo.s)) & lmask) << tpos(t.s);
                        if (lo.nb == 2) o |= ((off >> gpos(lo.t)) & lmask) << tpos(t.t);
                        T.val.push_back(op.val[k]);
                        T.off.push_back((int32_t)o);
                    }
                T.locs.push_back(t);
                T.max_row = std::max(T.max_row, (int)lo.R);
            }
        }
        if ((int)T.locs.size() > kLargeTileMaxLocals || (int)T.val.size() > kLargeTileMaxEnt) return false;
    }
    for (int d : done)
        if (!d) return false;
    // every global index appears in exactly one (tile, element) of every pass
    const int64_t n = op.n;
    for (const TilePlan& tp : plan) {
        int64_t bits = 0;
        for (int s : tp.G) bits |= lmask << gpos(s);
        if ((int64_t)tp.op.n * (n / tp.op.n) != n) return false;
        if (__builtin_popcountll((unsigned long long)bits) != ld * (int)tp.G.size()) return false;
    }
    return true;
}

template <typename T>
int upload(oops_ctx* ctx, DevBuf& b, const std::vector<T>& v);

// Device form of a compiled operator's tables: header, local operators, ELL
// values and offsets (at least one of each).
struct OpImage {
    DevOpHeader hdr;
    std::vector<DevOpLocal> locs;
    std::vector<double> vals;
    std::vector<int32_t> offs;
};

static OpImage op_image(const CompiledOp& T) {
    OpImage im;
    std::memset(&im.hdr, 0, sizeof(im.hdr));
    im.hdr.n = (int32_t)T.n;
    im.hdr.nspin = T.nspin;
    im.hdr.ld_bits = T.ld_bits;
    im.hdr.nloc = (int32_t)T.locs.size();
    im.hdr.nent = (int32_t)T.val.size();
    im.hdr.max_row = T.max_row;
    im.hdr.c0re = T.c0.real();
    im.hdr.c0im = T.c0.imag();
    im.hdr.ent2_first = T.ent2_first;
    for (auto& z : T.val) { im.vals.push_back(z.real()); im.vals.push_back(z.imag()); }
    if (im.vals.empty()) { im.vals.push_back(0.0); im.vals.push_back(0.0); }
    im.offs = T.off;
    if (im.offs.empty()) im.offs.push_back(0);
    im.locs = T.locs;
    if (im.locs.empty()) im.locs.push_back(DevOpLocal{});
    return im;
}

// Bit positions of a pass's tile spins ([0,16)) and block spins ([16,48)).
static std::vector<int32_t> tile_map(const CompiledOp& op, const std::vector<int>& g) {
    const int ns = op.nspin, ld = op.ld_bits;
    std::vector<int32_t> map(kLargeTileMapStride, 0);
    std::vector<int> in(ns, 0);
    for (size_t q = 0; q < g.size(); ++q) { map[q] = ld * (ns - 1 - g[q]); in[g[q]] = 1; }
    int nb = 0;
    for (int s2 = 0; s2 < ns; ++s2)
        if (!in[s2]) map[16 + nb++] = ld * (ns - 1 - s2);
    return map;
}

// The no-allocation array: a fixed device array that oops_hamvec_noalloc
// stages an operator's tables in instead of allocating (each device holds
// its own copy until the process ends).  Calls on one device share it.
constexpr size_t kNoAllocBytes = size_t(16) << 20;   // 16 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static int setup_tiles(oops_ctx* ctx, LargeRun& R) {
    std::vector<TilePlan> plan;
    if (!R.axis.empty() || !plan_tiles(R.op, plan)) return OOPS_OK;
    const int ns = R.op.nspin, ld = R.op.ld_bits;
    R.tiles.resize(plan.size());
    for (size_t p = 0; p < plan.size(); ++p) {
        LargeRun::Tile& t = R.tiles[p];
        const OpImage im = op_image(plan[p].op);
        const std::vector<int>& g = plan[p].G;
        const std::vector<int32_t> map = tile_map(R.op, g);
        const int nb = ns - (int)g.size();
        int rc = upload(ctx, t.hdr, std::vector<DevOpHeader>{im.hdr});
        rc = rc ? rc : upload(ctx, t.loc, im.locs);
        rc = rc ? rc : upload(ctx, t.val, im.vals);
        rc = rc ? rc : upload(ctx, t.off, im.offs);
        rc = rc ? rc : upload(ctx, t.map, map);
        if (rc) return rc;
        t.pass.op.hdr = (const DevOpHeader*)t.hdr.p;
        t.pass.op.loc = (const DevOpLocal*)t.loc.p;
        t.pass.op.val = (const double2*)t.val.p;
        t.pass.op.off = (const int32_t*)t.off.p;
        t.pass.map = (const int32_t*)t.map.p;
        t.pass.ng = (int32_t)g.size();
        t.pass.nb = nb;
        t.pass.ld = ld;
        t.pass.ntiles = (int32_t)(R.op.n >> (ld * (int)g.size()));
    }
    return OOPS_OK;
}

// y = alpha Op (xscale x), alpha_mode 0: +1, 1: -i.  A Kronecker term table
// carries the Hamiltonian H (the driver works on -iH); a CSR matrix carries
// the generator A itself (the driver works on A, as krylov_zgexpv does).
void matvec(LargeRun& R, const double2* x, const double* xs, double2* y, void* xstore, int alpha_mode) {
    if (R.csr)
        launch_large_csr_matvec((const int64_t*)R.rp.p, (const int32_t*)R.ci.p, (const double2*)R.cv.p, x, xs, y,
                                (double2*)xstore, R.n, alpha_mode, R.grid, R.s);
    else if (!R.axis.empty())
        for (size_t p = 0; p < R.axis.size(); ++p)
            launch_large_axis_matvec(R.axis[p].pass, x, xs, y, p == 0 ? xstore : nullptr, alpha_mode, p > 0 ? 1 : 0,
                                     R.s, R.fp32);
    else if (!R.tiles.empty())
        for (size_t p = 0; p < R.tiles.size(); ++p)
            launch_large_tile_matvec(R.tiles[p].pass, x, xs, y, p == 0 ? xstore : nullptr, alpha_mode, p > 0 ? 1 : 0,
                                     R.s, R.fp32);
    else
        launch_large_matvec(R.args, x, xs, y, xstore, R.n, alpha_mode, R.grid, R.s, R.fp32);
}

// Basis column j.
void* vcol(LargeRun& R, int j) {
    return R.fp32 ? (void*)((float2*)R.V.p + (size_t)j * R.ldv) : (void*)((double2*)R.V.p + (size_t)j * R.ldv);
}
int arnoldi_mode(const LargeRun& R) { return R.csr ? 0 : 1; }

template <typename T>
int upload(oops_ctx* ctx, DevBuf& b, const std::vector<T>& v) {
    b.release();
    if (v.empty()) return OOPS_OK;
    int rc = hip_check(ctx, hipMalloc(&b.p, v.size() * sizeof(T)), "hipMalloc");
    if (rc) return rc;
    b.bytes = v.size() * sizeof(T);
    return hip_check(ctx, hipMemcpy(b.p, v.data(), b.bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int alloc(oops_ctx* ctx, DevBuf& b, size_t bytes) {
    b.release();
    int rc = hip_check(ctx, hipMalloc(&b.p, bytes), "hipMalloc");
    if (!rc) b.bytes = bytes;
    return rc;
}

int setup_buffers(oops_ctx* ctx, int m, LargeRun& R);

int setup_compiled(oops_ctx* ctx, int m, LargeRun& R);

// Ket / bra split of a Liouville-space term table (every spin of local
// dimension 4, local index ket + 2 bra): a 4 x 4 body matrix kron(I2, B) acts
// on the ket digit with B, kron(A, I2) on the bra digit with A (the FLAT and
// SHARP expansions, KronProd.cpp:9-28); the identity is both.  Each term must
// act on one side only.  HK / HB come back as sparse rows (D = 2^k rows,
// column -> value, exact zeros dropped); false when some term mixes the sides
// or has another form (dephasing), or the table is not a spin-1/2 Liouvillian.
static bool axis_split(const oops_term_table* T, const cplx* ccoeff, std::vector<std::map<int32_t, cplx>> rows[2]) {
    const int k = T->nspin;
    if (k < 1 || k > 15) return false;
    for (int s = 0; s < k; ++s)
        if (T->spin_dim[s] != 4) return false;
    const int D = 1 << k;
    const cplx* mats = reinterpret_cast<const cplx*>(T->mats);
    rows[0].assign(D, {});
    rows[1].assign(D, {});
    cplx c0(0.0, 0.0);
    for (int t = 0; t < T->nterm; ++t) {
        cplx c = ccoeff ? ccoeff[t] : cplx(T->coeff[t], 0.0);
        const int64_t nb = T->nbody[t];
        const int64_t p0 = T->pos_offset[t];
        if (nb > 2 || nb < 0) return false;
        int side = -1;
        int sp[2];
        cplx m2[2][4];   // 2 x 2 column-major per remaining body
        int nbody = 0;
        for (int64_t b = 0; b < nb; ++b) {
            if (T->dims[p0 + b] != 4) return false;
            const cplx* M = mats + T->mat_offset[p0 + b];
            auto at = [&](int r, int cc) { return M[r + 4 * cc]; };
            bool ket = true, bra = true;
            for (int r1 = 0; r1 < 2; ++r1)
                for (int c1 = 0; c1 < 2; ++c1)
                    for (int r2 = 0; r2 < 2; ++r2)
                        for (int c2 = 0; c2 < 2; ++c2) {
                            const cplx v = at(2 * r1 + r2, 2 * c1 + c2);
                            if (v != (r1 == c1 ? at(r2, c2) : cplx(0.0))) ket = false;
                            if (v != (r2 == c2 ? at(2 * r1, 2 * c1) : cplx(0.0))) bra = false;
                        }
            if (ket && bra) {   // f times the identity (ket- and bra-form at once): fold f into c
                if (at(0, 0) != at(1, 1) || at(1, 0) != 0.0 || at(0, 1) != 0.0) return false;
                c *= at(0, 0);
                continue;
            }
            const int sd = ket ? 0 : (bra ? 1 : -2);
            if (sd < 0) return false;
            if (side >= 0 && side != sd) return false;
            side = sd;
            sp[nbody] = (int)T->pos[p0 + b];
            for (int r = 0; r < 2; ++r)
                for (int cc = 0; cc < 2; ++cc)
                    m2[nbody][r + 2 * cc] = sd == 0 ? at(r, cc) : at(2 * r, 2 * cc);
            ++nbody;
        }
        if (nbody == 0) {
            c0 += c;
            continue;
        }
        if (sp[0] < 0 || sp[0] >= k || (nbody == 2 && (sp[1] < 0 || sp[1] >= k || sp[1] == sp[0]))) return false;
        auto& R = rows[side];
        for (int r = 0; r < D; ++r) {
            const int sh0 = k - 1 - sp[0];
            const int r0 = (r >> sh0) & 1;
            if (nbody == 1) {
                for (int c0b = 0; c0b < 2; ++c0b) {
                    const cplx v = c * m2[0][r0 + 2 * c0b];
                    if (v == 0.0) continue;
                    R[r][(r & ~(1 << sh0)) | (c0b << sh0)] += v;
                }
            } else {
                const int sh1 = k - 1 - sp[1];
                const int r1 = (r >> sh1) & 1;
                for (int a = 0; a < 2; ++a)
                    for (int b2 = 0; b2 < 2; ++b2) {
                        const cplx v = c * m2[0][r0 + 2 * a] * m2[1][r1 + 2 * b2];
                        if (v == 0.0) continue;
                        R[r][(r & ~((1 << sh0) | (1 << sh1))) | (a << sh0) | (b2 << sh1)] += v;
                    }
            }
        }
    }
    if (c0 != 0.0)
        for (int r = 0; r < D; ++r) rows[0][r][r] += c0;
    return true;
}

// The two-axis product for spin-1/2 Liouvillians of 8-12 spins whose terms
// split into ket and bra sides; OOPS_LARGE_AXIS=0 keeps the tiled / flat
// products.
static int setup_axis(oops_ctx* ctx, const oops_term_table* T, const cplx* ccoeff, LargeRun& R) {
    R.axis.clear();
    if (const char* e = std::getenv("OOPS_LARGE_AXIS"))
        if (e[0] == '0') return OOPS_OK;
    const int k = T->nspin;
    if (k < 8 || k > 12) return OOPS_OK;
    std::vector<std::map<int32_t, cplx>> rows[2];
    if (!axis_split(T, ccoeff, rows)) return OOPS_OK;
    const int D = 1 << k;
    const int bbits = std::min(2, 12 - k);   // tile D x 2^bbits <= 4096 elements
    R.axis.resize(2);
    for (int side = 0; side < 2; ++side) {
        int W = 1;
        for (const auto& row : rows[side]) W = std::max(W, (int)row.size());
        std::vector<double> val(2 * (size_t)W * D, 0.0);
        std::vector<int32_t> col((size_t)W * D, 0);
        for (int r = 0; r < D; ++r) {
            int j = 0;
            for (const auto& e : rows[side][r]) {
                val[2 * ((size_t)j * D + r)] = e.second.real();
                val[2 * ((size_t)j * D + r) + 1] = e.second.imag();
                col[(size_t)j * D + r] = e.first;
                ++j;
            }
            for (; j < W; ++j) col[(size_t)j * D + r] = r;
        }
        LargeRun::Axis& A = R.axis[side];
        int rc = upload(ctx, A.val, val);
        rc = rc ? rc : upload(ctx, A.col, col);
        if (rc) return rc;
        A.pass.val = (const double2*)A.val.p;
        A.pass.col = (const int32_t*)A.col.p;
        A.pass.W = W;
        A.pass.k = k;
        A.pass.bbits = bbits;
        A.pass.side = side;
        A.pass.ntiles = 1 << (k - bbits);
    }
    return OOPS_OK;
}

// A term table whose spins do not share one power-of-two local dimension
// (spin-1 spins: 3, or 9 in Liouville space; mixed spin kinds) embedded in the
// local dimension P = the next power of two of the largest: every local matrix
// becomes the top-left block of a P x P zero matrix, and the caller's index
// (mixed radix dims[s], spin 0 slowest) maps to sum_s i_s P^(k-1-s).  The
// padding states are never reached: the operator has zero rows and columns
// there and the start vector is zero there.
struct PaddedTable {
    oops_term_table T{};
    std::vector<double> coeff, mats;
    std::vector<int64_t> nbody, pos_offset, pos, dims, mat_offset, spin_dim;
    std::vector<int64_t> map;   // caller index -> padded index
    int64_t n_logical = 0;
};

static bool needs_padding(const oops_term_table* T) {
    if (!T || T->nspin <= 0) return false;
    const int64_t d0 = T->spin_dim[0];
    for (int s = 0; s < T->nspin; ++s)
        if (T->spin_dim[s] != d0) return true;
    return d0 < 2 || (d0 & (d0 - 1)) != 0;
}

static int pad_table(const oops_term_table* T, PaddedTable& P, std::string& err) {
    const int k = T->nspin;
    int64_t dmax = 1;
    for (int s = 0; s < k; ++s) {
        if (T->spin_dim[s] < 1) { err = "bad local dimension"; return OOPS_ERR_ARG; }
        dmax = std::max(dmax, (int64_t)T->spin_dim[s]);
    }
    int64_t Pd = 2;
    while (Pd < dmax) Pd *= 2;
    int64_t nlog = 1, npad = 1;
    for (int s = 0; s < k; ++s) {
        nlog *= T->spin_dim[s];
        npad *= Pd;
        if (npad > (int64_t(1) << 30)) { err = "state too large after padding"; return OOPS_ERR_UNSUPPORTED; }
    }
    P.n_logical = nlog;
    const int64_t nb_total = T->pos_offset[T->nterm];
    P.coeff.assign(T->coeff, T->coeff + T->nterm);
    P.nbody.assign(T->nbody, T->nbody + T->nterm);
    P.pos_offset.assign(T->pos_offset, T->pos_offset + T->nterm + 1);
    P.pos.assign(T->pos, T->pos + nb_total);
    P.dims.assign((size_t)nb_total, Pd);
    P.mat_offset.assign((size_t)nb_total + 1, 0);
    P.mats.assign((size_t)2 * nb_total * Pd * Pd, 0.0);
    for (int64_t b = 0; b < nb_total; ++b) {
        const int64_t d = T->dims[b];
        const int s = (int)T->pos[b];
        if (s < 0 || s >= k || d != T->spin_dim[s]) { err = "body dimension differs from its spin's"; return OOPS_ERR_ARG; }
        P.mat_offset[b] = b * Pd * Pd;
        P.mat_offset[b + 1] = (b + 1) * Pd * Pd;
        const double* src = T->mats + 2 * T->mat_offset[b];
        double* dst = P.mats.data() + 2 * P.mat_offset[b];
        for (int64_t c = 0; c < d; ++c)
            for (int64_t r = 0; r < d; ++r) {
                dst[2 * (r + Pd * c)] = src[2 * (r + d * c)];
                dst[2 * (r + Pd * c) + 1] = src[2 * (r + d * c) + 1];
            }
    }
    P.spin_dim.assign((size_t)k, Pd);
    P.T.nspin = k;
    P.T.nterm = T->nterm;
    P.T.coeff = P.coeff.data();
    P.T.nbody = P.nbody.data();
    P.T.pos_offset = P.pos_offset.data();
    P.T.pos = P.pos.data();
    P.T.dims = P.dims.data();
    P.T.mat_offset = P.mat_offset.data();
    P.T.mats = P.mats.data();
    P.T.spin_dim = P.spin_dim.data();
    P.map.resize((size_t)nlog);
    for (int64_t i = 0; i < nlog; ++i) {
        int64_t q = i, g = 0, stride = 1;
        for (int s = k - 1; s >= 0; --s) {
            const int64_t d = T->spin_dim[s];
            g += (q % d) * stride;
            q /= d;
            stride *= Pd;
        }
        P.map[(size_t)i] = g;
    }
    return OOPS_OK;
}

int setup(oops_ctx* ctx, const oops_term_table* T, int m, LargeRun& R, const std::complex<double>* ccoeff = nullptr) {
    std::string err;
    if (needs_padding(T)) {
        PaddedTable P;
        int rc = pad_table(T, P, err);
        if (rc) return set_err(ctx, rc, err);
        R.padded = true;
        R.n_logical = P.n_logical;
        R.pos = std::move(P.map);
        rc = compile_operator(&P.T, R.op, err, kLargeMaxLocals, ccoeff, kLargeMaxRow);
        if (rc) return set_err(ctx, rc, err);
        if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
        return setup_compiled(ctx, m, R);
    }
    int rc = compile_operator(T, R.op, err, kLargeMaxLocals, ccoeff, kLargeMaxRow);
    if (rc) return set_err(ctx, rc, err);
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    rc = setup_axis(ctx, T, ccoeff, R);
    if (rc) return rc;
    return setup_compiled(ctx, m, R);
}

// Upload R.op (already compiled) and allocate the Krylov buffers.
int setup_compiled(oops_ctx* ctx, int m, LargeRun& R) {
    int rc = OOPS_OK;
    R.ctx = ctx;
    R.n = R.op.n;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    const OpImage im = op_image(R.op);
    rc = upload(ctx, R.hdr, std::vector<DevOpHeader>{im.hdr});
    rc = rc ? rc : upload(ctx, R.loc, im.locs);
    rc = rc ? rc : upload(ctx, R.val, im.vals);
    rc = rc ? rc : upload(ctx, R.off, im.offs);
    if (rc) return rc;
    R.args.hdr = (const DevOpHeader*)R.hdr.p;
    R.args.loc = (const DevOpLocal*)R.loc.p;
    R.args.val = (const double2*)R.val.p;
    R.args.off = (const int32_t*)R.off.p;
    rc = setup_tiles(ctx, R);
    if (rc) return rc;
    return setup_buffers(ctx, m, R);
}

// CSR operator from row-sorted (row, col, value) triples, 0-based.
int setup_csr(oops_ctx* ctx, int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& ci,
              const std::vector<double>& cv, int m, LargeRun& R) {
    R.ctx = ctx;
    R.n = n;
    R.csr = true;
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    int rc = upload(ctx, R.rp, rp);
    rc = rc ? rc : upload(ctx, R.ci, ci.empty() ? std::vector<int32_t>{0} : ci);
    rc = rc ? rc : upload(ctx, R.cv, cv.empty() ? std::vector<double>{0.0, 0.0} : cv);
    if (rc) return rc;
    return setup_buffers(ctx, m, R);
}

int setup_buffers(oops_ctx* ctx, int m, LargeRun& R) {
    int rc = OOPS_OK;
    R.m = m;
    R.ldh = m + 2;
    R.grid = large_grid(R.n, ctx->num_cu);
    // blocks of the Gram-Schmidt and dot-product passes: the full grid (up to
    // eight per CU), so the long dots passes keep enough loads in flight
    R.nblk = R.grid;
    // the dots passes' partial sums are reduced row by row in one block: on
    // full grids half the blocks there halve that reduction (and the passes
    // themselves run as fast or faster: 2048 -> 1024 blocks for 2^20 elements)
    R.nblk_dots = R.grid >= 1024 ? R.grid / 2 : R.grid;
    const size_t vb = sizeof(double2) * (size_t)R.n;
    // basis columns 4352 B (17 x 256 B) apart beyond a power of two: columns
    // of 2^k elements would start at addresses equal modulo every power of two
    // up to their size, so the loads of one element across the basis would
    // meet in the same memory channels
    const size_t esz = R.fp32 ? sizeof(float2) : sizeof(double2);
    R.ldv = R.n + (R.n >= (1 << 16) ? (int64_t)(4352 / esz) : 0);
    rc = rc ? rc : alloc(ctx, R.V, esz * (size_t)R.ldv * (m + 1));
    rc = rc ? rc : alloc(ctx, R.p0, vb);
    rc = rc ? rc : alloc(ctx, R.p1, vb);
    rc = rc ? rc : alloc(ctx, R.w, vb);
    rc = rc ? rc : alloc(ctx, R.v0, vb);
    rc = rc ? rc : alloc(ctx, R.part, sizeof(double) * 2 * (size_t)R.nblk * (m + 4));
    // one spare column: the 2x2-tiled products of the Pade kernel read one
    // element past an odd-sized matrix (into discarded accumulators)
    rc = rc ? rc : alloc(ctx, R.hess, sizeof(double2) * R.ldh * (R.ldh + 1));
    rc = rc ? rc : alloc(ctx, R.coef, sizeof(double2) * (m + 2));
    rc = rc ? rc : alloc(ctx, R.scal, sizeof(double) * 8);
    rc = rc ? rc : alloc(ctx, R.flags, sizeof(int) * 4);
    rc = rc ? rc : alloc(ctx, R.F, sizeof(double2) * (m + 2));
    rc = rc ? rc : alloc(ctx, R.status, sizeof(int) * 4);
    rc = rc ? rc : alloc(ctx, R.wpart, sizeof(double) * R.grid);
    return rc;
}

// complex dot <a, b> (conj a) on the device, returned to the host
int dev_dot(LargeRun& R, const double2* a, const double2* b, cplx& out) {
    launch_large_dots(a, 0, b, R.n, 0, 0, (double*)R.part.p, R.nblk, R.s);
    std::vector<double> h(2 * R.nblk);
    int rc = hip_check(R.ctx, hipMemcpyAsync(h.data(), R.part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, R.s),
                       "D2H dot");
    rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync dot");
    double re = 0.0, im = 0.0;
    for (int k = 0; k < R.nblk; ++k) { re += h[2 * k]; im += h[2 * k + 1]; }
    out = cplx(re, im);
    return rc;
}

int dev_norm(LargeRun& R, const double2* a, double& out) {
    launch_large_dots(a, 0, a, R.n, -1, 1, (double*)R.part.p, R.nblk, R.s);
    std::vector<double> h(2 * R.nblk);
    int rc = hip_check(R.ctx, hipMemcpyAsync(h.data(), R.part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, R.s),
                       "D2H norm");
    rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync norm");
    double s2 = 0.0;
    for (int k = 0; k < R.nblk; ++k) s2 += h[2 * k];
    out = std::sqrt(s2);
    return rc;
}

// Lanczos + power iteration norm estimate on +H (reference :210-265)
int norm_estimate(LargeRun& R, int klim, double tol, double& anorm) {
    const int64_t n = R.n;
    const int64_t nl = R.padded ? R.n_logical : n;   // the draws follow the caller's indices
    int k = klim;
    if (k > nl) k = (int)nl;
    std::vector<double> rnd(2 * nl + k);
    int32_t seed[4] = {0, 0, 0, 5};
    dlaran_fill(seed, (int64_t)rnd.size(), rnd.data());
    if (R.padded) {
        std::vector<double> full(2 * n + k, 0.0);
        for (int64_t i = 0; i < nl; ++i) {
            full[2 * R.pos[(size_t)i]] = rnd[2 * i];
            full[2 * R.pos[(size_t)i] + 1] = rnd[2 * i + 1];
        }
        for (int j = 0; j < k; ++j) full[2 * n + j] = rnd[2 * nl + j];
        rnd.swap(full);
    }
    double2* q = (double2*)R.p0.p;
    double2* qm = (double2*)R.p1.p;
    double2* wv = (double2*)R.w.p;
    int rc = hip_check(R.ctx, hipMemcpyAsync(q, rnd.data(), sizeof(double2) * n, hipMemcpyHostToDevice, R.s), "H2D q");
    rc = rc ? rc : hip_check(R.ctx, hipMemsetAsync(qm, 0, sizeof(double2) * n, R.s), "memset");
    if (rc) return rc;
    double bj = 0.0;
    rc = dev_norm(R, q, bj);
    if (rc) return rc;
    launch_large_axpby(make_double2(1.0 / bj, 0.0), q, make_double2(0.0, 0.0), q, n, R.grid, R.s);
    std::vector<double> tk((size_t)k * k, 0.0);
    double bjm = 0.0;
    for (int j = 1; j <= k; ++j) {
        matvec(R, q, nullptr, wv, nullptr, 0);
        cplx aj;
        rc = dev_dot(R, q, wv, aj);
        if (rc) return rc;
        tk[(size_t)(j - 1) * k + (j - 1)] = aj.real();
        launch_large_axpby(make_double2(-aj.real(), -aj.imag()), q, make_double2(1.0, 0.0), wv, n, R.grid, R.s);
        launch_large_axpby(make_double2(-bjm, 0.0), qm, make_double2(1.0, 0.0), wv, n, R.grid, R.s);
        rc = dev_norm(R, wv, bj);
        if (rc) return rc;
        if (j == k) break;
        tk[(size_t)(j - 1) * k + j] = bj;
        tk[(size_t)j * k + (j - 1)] = bj;
        if (bj < tol) break;
        std::swap(q, qm);  // qm <- q
        launch_large_axpby(make_double2(1.0 / bj, 0.0), wv, make_double2(0.0, 0.0), q, n, R.grid, R.s);
        bjm = bj;
    }
    std::vector<double> qr(k), wr(k);
    for (int j = 0; j < k; ++j) qr[j] = rnd[2 * n + j];
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
    return OOPS_OK;
}

struct Counters {
    int64_t nmult = 0, nexph = 0, nscale = 0, nstep = 0, nreject = 0, ibrk = 0;
    double anorm = 0.0, kernel_ms = 0.0;
};

// The Expokit integration.  out_mode 0: w_seq (n x nt) to the host;
// 1: overlaps <v0, w(t_k)> to the host (2 doubles per time); 2: w(t_k) into
// the device array dev (n x nt); 3: overlaps <w(t_k), dev(:, k)> to the host.
// interior: the steps are not cut at the output times (they run to
// min(t_new, last output time) under Expokit's error control) and every output
// time inside a step is evaluated from that step's basis (one small
// exponential and one combination each).  Off: the reference's schedule, a
// step ends at every output time (zgexpv_cache.f:418-440).
int run_expv(LargeRun& R, const double* v_host, int nt, const double* ta, int klim, double tol, int out_mode,
             double* out, Counters& C, double2* dev = nullptr, bool interior = false) {
    const int m = R.m;
    const int64_t n = R.n;
    const double delta = 1.2, gamma = 0.9, break_tol = 1.0e-7;
    const int64_t max_steps = max_steps_limit();
    int rc = norm_estimate(R, klim, tol, C.anorm);
    if (rc) return rc;
    const double anorm = C.anorm;
    rc = hip_check(R.ctx, hipMemcpyAsync(R.w.p, v_host, sizeof(double2) * n, hipMemcpyHostToDevice, R.s), "H2D v");
    rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(R.v0.p, v_host, sizeof(double2) * n, hipMemcpyHostToDevice, R.s),
                             "H2D v0");
    if (rc) return rc;
    double beta = 0.0;
    rc = dev_norm(R, (const double2*)R.w.p, beta);
    if (rc) return rc;
    const double eps = 2.220446049250313e-16;
    if (tol <= eps) tol = std::sqrt(eps);
    double xm = 1.0 / (double)m;
    double t_new;
    {
        const double p2 = tol * std::pow((m + 1) / 2.72, m + 1) * std::sqrt(2.0 * 3.14 * (m + 1));
        t_new = round_step((1.0 / anorm) * std::pow(p2 / (4.0 * beta * anorm), xm));
    }
    double t_now = 0.0;
    double2* hess = (double2*)R.hess.p;
    double* scal = (double*)R.scal.p;
    // timing events, destroyed on every return
    struct EventGuard {
        hipEvent_t e = nullptr;
        ~EventGuard() {
            if (e) (void)hipEventDestroy(e);
        }
    } g0, g1;
    rc = hip_check(R.ctx, hipEventCreate(&g0.e), "hipEventCreate");
    rc = rc ? rc : hip_check(R.ctx, hipEventCreate(&g1.e), "hipEventCreate");
    if (rc) return rc;
    const hipEvent_t e0 = g0.e, e1 = g1.e;
    float ms_acc = 0.0f;
    if (interior) {
        rc = alloc(R.ctx, R.wk, sizeof(double2) * (size_t)n);
        rc = rc ? rc : alloc(R.ctx, R.wkpart, sizeof(double) * (size_t)R.grid);
        if (rc) return rc;
    }
    // the output action at time index tt on the state wv
    auto emit = [&](int tt, const double2* wv) -> int {
        int e = OOPS_OK;
        if (out_mode == 2) {
            e = hip_check(R.ctx, hipMemcpyAsync(dev + (size_t)n * tt, wv, sizeof(double2) * n, hipMemcpyDeviceToDevice,
                                                R.s), "D2D w");
        } else if (out_mode == 3) {
            cplx ov;
            e = dev_dot(R, wv, dev + (size_t)n * tt, ov);
            out[2 * tt] = ov.real();
            out[2 * tt + 1] = ov.imag();
        } else if (out) {
            if (out_mode == 0) {
                e = hip_check(R.ctx, hipMemcpyAsync(out + 2 * (size_t)n * tt, wv, sizeof(double2) * n,
                                                    hipMemcpyDeviceToHost, R.s), "D2H w");
                e = e ? e : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync w");
            } else {
                cplx ov;
                e = dev_dot(R, (const double2*)R.v0.p, wv, ov);
                out[2 * tt] = ov.real();
                out[2 * tt + 1] = ov.imag();
            }
        }
        return e;
    };
    // one Arnoldi cycle: m products with their Gram-Schmidt passes and the extra product of the
    // error estimate, all on R.s
    auto arnoldi_launches = [&]() {
        double2* bufs[2] = {(double2*)R.p0.p, (double2*)R.p1.p};
        const double2* x = (const double2*)R.w.p;
        for (int j = 0; j < m; ++j) {
            double2* p = bufs[j & 1];
            matvec(R, x, scal, p, vcol(R, j), arnoldi_mode(R));
            // classical Gram-Schmidt, one read of V per pass: dots and |p|,
            // update and |p'|; then, only when the first update cancelled
            // more than 3/4 of |p|^2 (decided on the device by the finish
            // kernel; the skipped kernels return at once), dots and update
            // again
            const int* skip = (const int*)R.flags.p + 2;
            double* part = (double*)R.part.p;
            double2* coef = (double2*)R.coef.p;
            launch_large_gs_dots(R.V.p, R.ldv, p, n, j, part, R.nblk_dots, R.s, R.fp32);
            launch_large_finish(part, R.nblk_dots, j + 2, j, 0, hess, R.ldh, coef, scal, (int*)R.flags.p, break_tol,
                                R.s);
            launch_large_gs_update(R.V.p, R.ldv, p, n, j, coef, part, R.nblk, R.s, R.fp32);
            launch_large_finish(part, R.nblk, 1, j, 1, hess, R.ldh, coef, scal, (int*)R.flags.p, break_tol, R.s);
            launch_large_gs_dots(R.V.p, R.ldv, p, n, j, part, R.nblk_dots, R.s, R.fp32, skip);
            launch_large_finish(part, R.nblk_dots, j + 2, j, 4, hess, R.ldh, coef, scal, (int*)R.flags.p, break_tol,
                                R.s);
            launch_large_gs_update(R.V.p, R.ldv, p, n, j, coef, part, R.nblk, R.s, R.fp32, skip);
            launch_large_finish(part, R.nblk, 1, j, 2, hess, R.ldh, coef, scal, (int*)R.flags.p, break_tol, R.s);
            x = p;
        }
        // extra product for the error estimate: V[m] = p / h(m, m-1), p' = -iH V[m]
        double2* pe = bufs[m & 1];
        matvec(R, x, scal, pe, vcol(R, m), arnoldi_mode(R));
        launch_large_dots(pe, 0, pe, n, -1, 1, (double*)R.part.p, R.nblk, R.s);
        launch_large_finish((double*)R.part.p, R.nblk, 1, 0, 3, hess, R.ldh, nullptr, scal, (int*)R.flags.p,
                            break_tol, R.s);
    };
    int tt = 0;   // next output time
    // one Krylov step from t_now toward t_out
    auto step = [&](double t_out) -> int {
        {
            if (C.nstep >= max_steps) return set_err(R.ctx, OOPS_ERR_NUMERIC, "step limit reached");
            ++C.nstep;
            double t_step = std::min(t_out - t_now, t_new);
            int k1 = 2, mb = m;
            (void)hipEventRecord(e0, R.s);
            (void)hipMemsetAsync(hess, 0, sizeof(double2) * R.ldh * (R.ldh + 1), R.s);
            (void)hipMemsetAsync(R.flags.p, 0, sizeof(int) * 4, R.s);
            const double inv_beta = 1.0 / beta;
            (void)hipMemcpyAsync(scal, &inv_beta, sizeof(double), hipMemcpyHostToDevice, R.s);
            arnoldi_launches();
            int fl[4];
            double sc[4];
            rc = hip_check(R.ctx, hipMemcpyAsync(fl, R.flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, R.s), "D2H");
            rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(sc, scal, sizeof(double) * 4, hipMemcpyDeviceToHost, R.s), "D2H");
            rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync arnoldi");
            if (rc) return rc;
            double avnorm = sc[2];
            if (fl[0]) {
                // happy breakdown at column mb: the products after it ran on a
                // zero vector and are not counted
                k1 = 0;
                mb = fl[1];
                t_step = t_out - t_now;
                C.ibrk = 1;
                C.nmult += mb;
            } else {
                C.nmult += m + 1;
            }
            const double2 one = make_double2(1.0, 0.0);
            (void)hipMemcpyAsync(hess + (m + 1) + (size_t)R.ldh * m, &one, sizeof(double2), hipMemcpyHostToDevice, R.s);
            double err_loc = 0.0;
            int mx;
            std::vector<double2> Fh(m + 2);
            for (;;) {
                ++C.nexph;
                mx = mb + k1;
                if (launch_large_pade(hess, R.ldh, mx, t_step, (double2*)R.F.p, (int*)R.status.p, R.s) != 0)
                    return set_err(R.ctx, OOPS_ERR_HIP, "Pade kernel launch");
                int st = 0;
                rc = hip_check(R.ctx, hipMemcpyAsync(Fh.data(), R.F.p, sizeof(double2) * mx, hipMemcpyDeviceToHost, R.s), "D2H F");
                rc = rc ? rc : hip_check(R.ctx, hipMemcpyAsync(&st, R.status.p, sizeof(int), hipMemcpyDeviceToHost, R.s), "D2H st");
                rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync pade");
                if (rc) return rc;
                if (st < 0) return set_err(R.ctx, OOPS_ERR_NUMERIC, "small exponential failed");
                C.nscale += st;
                if (k1 == 0) {
                    err_loc = tol;
                } else {
                    const double p1 = std::abs(cplx(Fh[m].x, Fh[m].y)) * beta;
                    const double p2 = std::abs(cplx(Fh[m + 1].x, Fh[m + 1].y)) * beta * avnorm;
                    if (p1 > 10.0 * p2) { err_loc = p2; xm = 1.0 / (double)m; }
                    else if (p1 > p2) { err_loc = (p1 * p2) / (p1 - p2); xm = 1.0 / (double)m; }
                    else { err_loc = p1; xm = 1.0 / (double)(m - 1); }
                }
                if (k1 != 0 && err_loc > delta * t_step * tol) {
                    t_step = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
                    ++C.nreject;
                    if (!std::isfinite(t_step)) return set_err(R.ctx, OOPS_ERR_NUMERIC, "step size collapsed");
                    continue;
                }
                break;
            }
            mx = mb + std::max(0, k1 - 1);
            launch_large_combine(R.V.p, R.ldv, mx, (double2*)R.F.p, beta, (double2*)R.w.p, n, (double*)R.wpart.p, R.grid,
                                 R.s, R.fp32);
            (void)hipEventRecord(e1, R.s);
            std::vector<double> wp(R.grid);
            rc = hip_check(R.ctx, hipMemcpyAsync(wp.data(), R.wpart.p, sizeof(double) * R.grid, hipMemcpyDeviceToHost, R.s),
                           "D2H beta");
            rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync combine");
            if (rc) return rc;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_acc += ms;
            if (interior && tt < nt && ta[tt] < (t_step >= t_out - t_now ? t_out : t_now + t_step)) {
                // output times inside the step, from the same basis (their
                // exponentials, combinations and overlaps count as device time)
                const double t1 = t_step >= t_out - t_now ? t_out : t_now + t_step;
                (void)hipEventRecord(e0, R.s);
                for (; tt < nt && ta[tt] < t1; ++tt) {
                    ++C.nexph;
                    if (launch_large_pade(hess, R.ldh, mb + k1, ta[tt] - t_now, (double2*)R.F.p, (int*)R.status.p,
                                          R.s) != 0)
                        return set_err(R.ctx, OOPS_ERR_HIP, "Pade kernel launch");
                    launch_large_combine(R.V.p, R.ldv, mx, (double2*)R.F.p, beta, (double2*)R.wk.p, n,
                                         (double*)R.wkpart.p, R.grid, R.s, R.fp32);
                    int st = 0;
                    rc = hip_check(R.ctx, hipMemcpyAsync(&st, R.status.p, sizeof(int), hipMemcpyDeviceToHost, R.s),
                                   "D2H st");
                    rc = rc ? rc : emit(tt, (const double2*)R.wk.p);
                    rc = rc ? rc : hip_check(R.ctx, hipStreamSynchronize(R.s), "sync interior");
                    if (rc) return rc;
                    if (st < 0) return set_err(R.ctx, OOPS_ERR_NUMERIC, "small exponential failed");
                    C.nscale += st;
                }
                (void)hipEventRecord(e1, R.s);
                rc = hip_check(R.ctx, hipEventSynchronize(e1), "sync interior");
                if (rc) return rc;
                float mi = 0.f;
                (void)hipEventElapsedTime(&mi, e0, e1);
                ms_acc += mi;
            }
            double s2 = 0.0;
            for (double v : wp) s2 += v;
            beta = std::sqrt(s2);
            t_new = round_step(gamma * t_step * std::pow(t_step * tol / err_loc, xm));
            if (interior && t_step >= t_out - t_now) t_now = t_out;
            else t_now += t_step;
        }
        return OOPS_OK;
    };
    if (interior) {
        const double t_last = ta[nt - 1];
        for (; tt < nt && ta[tt] <= t_now; ++tt)
            if ((rc = emit(tt, (const double2*)R.w.p)) != 0) return rc;
        while (t_now < t_last) {
            if ((rc = step(t_last)) != 0) return rc;
            for (; tt < nt && ta[tt] <= t_now; ++tt)
                if ((rc = emit(tt, (const double2*)R.w.p)) != 0) return rc;
        }
    } else {
        for (tt = 0; tt < nt; ++tt) {
            while (t_now < ta[tt])
                if ((rc = step(ta[tt])) != 0) return rc;
            if ((rc = emit(tt, (const double2*)R.w.p)) != 0) return rc;
        }
    }
    C.kernel_ms = ms_acc;
    return OOPS_OK;
}

}  // namespace

namespace oops {

// One Hahn-echo cluster on the single-state path (states too large for the
// batched kernels): the ket sweep's states are kept on the device, the bra
// sweep forms L(tau_k) = Re(scale <bra(tau_k), ket(tau_k)>) at each output.
int large_hahn_item(oops_ctx* ctx, const CompiledOp& A, const CompiledOp& B, const double* u0, const double* y0,
                    double scale, int nt, const double* tau, int m, int klim, double tol, double* L, int32_t* stats) {
    LargeRun RA, RB;
    RA.op = A;
    RB.op = B;
    int rc = setup_compiled(ctx, m, RA);
    rc = rc ? rc : setup_compiled(ctx, m, RB);
    if (rc) return rc;
    if (m >= RA.n) return set_err(ctx, OOPS_ERR_ARG, "the Krylov dimension must be smaller than the state dimension");
    DevBuf ket;
    rc = alloc(ctx, ket, sizeof(double2) * (size_t)RA.n * nt);
    if (rc) return rc;
    std::vector<double> ov(2 * (size_t)nt);
    Counters CA, CB;
    rc = hip_check(ctx, hipStreamCreateWithFlags(&RA.s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    if (rc) return rc;
    RB.s = RA.s;
    rc = run_expv(RA, u0, nt, tau, klim, tol, 2, nullptr, CA, (double2*)ket.p);
    rc = rc ? rc : run_expv(RB, y0, nt, tau, klim, tol, 3, ov.data(), CB, (double2*)ket.p);
    (void)hipStreamDestroy(RA.s);
    if (rc) return rc;
    for (int k = 0; k < nt; ++k) L[k] = scale * ov[2 * k];
    if (stats) {
        stats[0] = (int32_t)CA.nstep; stats[1] = (int32_t)CA.nmult; stats[2] = (int32_t)CA.nreject;
        stats[3] = (int32_t)CA.nexph;
        stats[4] = (int32_t)CB.nstep; stats[5] = (int32_t)CB.nmult; stats[6] = (int32_t)CB.nreject;
        stats[7] = (int32_t)CB.nexph;
        stats[8] = (int32_t)((CA.ibrk << 8) | (CB.ibrk << 9));
        stats[9] = (int32_t)(CA.nscale + CB.nscale);
        std::memcpy(stats + 10, &CA.anorm, sizeof(double));
        std::memcpy(stats + 12, &CB.anorm, sizeof(double));
    }
    return OOPS_OK;
}

// hipGetSymbolAddress of the no-allocation array, once per device (with that
// device current); the host-side name `scratch` is not a device address.
int noalloc_array_address(int device, void** ptr) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto f = addr.find(device);
    if (f != addr.end()) {
        *ptr = f->second;
        return OOPS_OK;
    }
    if (hipSetDevice(device) != hipSuccess) return OOPS_ERR_HIP;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return OOPS_ERR_HIP;
    addr[device] = p;
    *ptr = p;
    return OOPS_OK;
}

}  // namespace oops

// Host evaluation of the compiled product (ELL rows of the merged local
// operators), element i of y = H x over a state of op.n elements whose index
// is taken through `gidx` (identity for the flat product, the tile scatter for
// a pass).
template <typename Index>
static std::complex<double> host_elem(const CompiledOp& op, int64_t i, const std::complex<double>* x, Index gidx,
                                      bool entry_major = false) {
    const int ld = op.ld_bits;
    const int64_t lmask = (1LL << ld) - 1;
    const int top = op.nspin - 1;
    std::complex<double> acc = op.c0 * x[gidx(i)];
    for (const DevOpLocal& lo : op.locs) {
        const int shs = ld * (top - lo.s);
        int64_t row = (i >> shs) & lmask;
        int64_t base = i - (row << shs);
        if (lo.nb == 2) {
            const int sht = ld * (top - lo.t);
            const int64_t lt = (i >> sht) & lmask;
            base -= lt << sht;
            row = (row << ld) | lt;
        }
        for (int e = 0; e < lo.R; ++e) {
            const int64_t k = entry_major ? lo.ent + (int64_t)e * lo.rows + row : lo.ent + row * lo.R + e;
            acc += op.val[k] * x[gidx(base + op.off[k])];
        }
    }
    return acc;
}

extern "C" {

int oops_tile_plan(const oops_term_table* T, int32_t* npasses, int32_t mode, const double* v, double* w) {
    if (!T || !npasses) return OOPS_ERR_ARG;
    if (mode == 3) {
        // the two-axis product on the host through the kernel's tile mapping
        std::vector<std::map<int32_t, cplx>> rows[2];
        const int k = T->nspin;
        if (k < 1 || k > 12 || !axis_split(T, nullptr, rows)) { *npasses = 0; return OOPS_ERR_UNSUPPORTED; }
        *npasses = 2;
        if (!v || !w) return OOPS_OK;
        const cplx* x = reinterpret_cast<const cplx*>(v);
        cplx* y = reinterpret_cast<cplx*>(w);
        const int64_t n = 1LL << (2 * k);
        const int bb = std::max(0, std::min(2, 12 - k));
        const int D = 1 << k, B = 1 << bb;
        auto spread = [](int64_t q) {
            int64_t r = 0;
            for (int j = 0; j < 16; ++j) r |= ((q >> j) & 1) << (2 * j);
            return r;
        };
        std::vector<uint8_t> seen(n);
        std::vector<cplx> xs((size_t)D * B);
        for (int side = 0; side < 2; ++side) {
            std::fill(seen.begin(), seen.end(), 0);
            for (int64_t hi = 0; hi < (int64_t)(1 << (k - bb)); ++hi) {
                auto decode = [&](int u, int& a, int& b) {
                    const int low = u & ((1 << (2 * bb)) - 1);
                    int ev = 0, od = 0;
                    for (int q = 0; q < bb; ++q) {
                        ev |= ((low >> (2 * q)) & 1) << q;
                        od |= ((low >> (2 * q + 1)) & 1) << q;
                    }
                    const int up = u >> (2 * bb);
                    if (side == 0) { a = ev | (up << bb); b = od; }
                    else { a = od | (up << bb); b = ev; }
                };
                auto gidx = [&](int a, int b) {
                    const int64_t other = (hi << bb) | b;
                    const int64_t ket = side == 0 ? a : other, bra = side == 0 ? other : a;
                    return spread(ket) | (spread(bra) << 1);
                };
                for (int u = 0; u < D * B; ++u) {
                    int a, b;
                    decode(u, a, b);
                    const int64_t g = gidx(a, b);
                    if (g < 0 || g >= n || seen[g]) return OOPS_ERR_NUMERIC;   // not a bijection
                    seen[g] = 1;
                    xs[(size_t)a * B + b] = x[g];
                }
                for (int a = 0; a < D; ++a)
                    for (int b = 0; b < B; ++b) {
                        cplx acc(0.0);
                        for (const auto& e : rows[side][a]) acc += e.second * xs[(size_t)e.first * B + b];
                        const int64_t g = gidx(a, b);
                        y[g] = side == 0 ? acc : y[g] + acc;
                    }
            }
        }
        return OOPS_OK;
    }
    CompiledOp op;
    std::string err;
    // local dimensions that are not one power of two are padded to one: the
    // product runs in the padded layout and is read back in the caller's
    PaddedTable P;
    const bool pad = needs_padding(T);
    int rc = pad ? pad_table(T, P, err) : OOPS_OK;
    rc = rc ? rc : compile_operator(pad ? &P.T : T, op, err, kLargeMaxLocals, nullptr, kLargeMaxRow);
    if (rc) return rc;
    std::vector<TilePlan> plan;
    const bool tiled = plan_tiles(op, plan);
    *npasses = tiled ? (int32_t)plan.size() : 0;
    if (mode == 0 || !v || !w) return OOPS_OK;
    const int64_t n = op.n;
    std::vector<cplx> xp, yp;
    const cplx* x = reinterpret_cast<const cplx*>(v);
    cplx* y = reinterpret_cast<cplx*>(w);
    if (pad) {
        xp.assign((size_t)n, cplx(0.0));
        yp.assign((size_t)n, cplx(0.0));
        for (int64_t i = 0; i < P.n_logical; ++i) xp[(size_t)P.map[(size_t)i]] = x[i];
        x = xp.data();
        y = yp.data();
    }
    if (mode == 2 || !tiled) {
        for (int64_t i = 0; i < n; ++i) y[i] = host_elem(op, i, x, [](int64_t g) { return g; });
    } else {
        // the kernel's mapping: tile b, element e -> base(b) | scatter(e)
        const int ns = op.nspin, ld = op.ld_bits;
        const int64_t lmask = (1LL << ld) - 1;
        std::vector<uint8_t> seen(n);
        for (size_t p = 0; p < plan.size(); ++p) {
            const std::vector<int>& g = plan[p].G;
            std::vector<int> in(ns, 0), blk;
            for (int s2 : g) in[s2] = 1;
            for (int s2 = 0; s2 < ns; ++s2)
                if (!in[s2]) blk.push_back(s2);
            const int ng = (int)g.size(), nb = (int)blk.size();
            const int64_t tsize = 1LL << (ld * ng), ntiles = n >> (ld * ng);
            std::fill(seen.begin(), seen.end(), 0);
            for (int64_t b = 0; b < ntiles; ++b) {
                int64_t base = 0, bb = b;
                for (int q = nb - 1; q >= 0; --q) { base |= (bb & lmask) << (ld * (ns - 1 - blk[q])); bb >>= ld; }
                auto scatter = [&](int64_t e) {
                    int64_t gi = base;
                    for (int q = ng - 1; q >= 0; --q) { gi |= (e & lmask) << (ld * (ns - 1 - g[q])); e >>= ld; }
                    return gi;
                };
                for (int64_t e = 0; e < tsize; ++e) {
                    const int64_t gi = scatter(e);
                    if (gi < 0 || gi >= n || seen[gi]) return OOPS_ERR_NUMERIC;   // not a bijection
                    seen[gi] = 1;
                    const cplx a = host_elem(plan[p].op, e, x, scatter, true);
                    y[gi] = p == 0 ? a : y[gi] + a;
                }
            }
        }
    }
    if (pad)
        for (int64_t i = 0; i < P.n_logical; ++i) reinterpret_cast<cplx*>(w)[i] = yp[(size_t)P.map[(size_t)i]];
    return OOPS_OK;
}

int oops_hamvec_noalloc(oops_ctx* ctx, const oops_term_table* H, const double* x, double* y) {
    if (!ctx || !H || !x || !y) return set_err(ctx, OOPS_ERR_ARG, "bad arguments");
    if (!ctx->noalloc) return set_err(ctx, OOPS_ERR_HIP, "no-allocation array not available on this device");
    CompiledOp op;
    std::string err;
    int rc = compile_operator(H, op, err, kLargeMaxLocals, nullptr, kLargeMaxRow);
    if (rc) return set_err(ctx, rc, err);
    // host image of the array: the tables oops_hamvec uploads, 256-byte aligned
    std::vector<unsigned char> img;
    auto put = [&](const void* src, size_t bytes) {
        const size_t at = (img.size() + 255) & ~size_t(255);
        img.resize(at + bytes);
        std::memcpy(img.data() + at, src, bytes);
        return at;
    };
    struct Table { size_t hdr, loc, val, off, map; };
    auto put_op = [&](const CompiledOp& o) {
        const OpImage im = op_image(o);
        Table t{};
        t.hdr = put(&im.hdr, sizeof(DevOpHeader));
        t.loc = put(im.locs.data(), sizeof(DevOpLocal) * im.locs.size());
        t.val = put(im.vals.data(), sizeof(double) * im.vals.size());
        t.off = put(im.offs.data(), sizeof(int32_t) * im.offs.size());
        return t;
    };
    std::vector<TilePlan> plan;
    const bool tiled = plan_tiles(op, plan);
    std::vector<Table> tabs;
    std::vector<std::vector<int32_t>> maps;
    if (tiled) {
        for (const TilePlan& tp : plan) {
            tabs.push_back(put_op(tp.op));
            maps.push_back(tile_map(op, tp.G));
            tabs.back().map = put(maps.back().data(), sizeof(int32_t) * maps.back().size());
        }
    } else {
        tabs.push_back(put_op(op));
    }
    if (img.size() > kNoAllocBytes)
        return set_err(ctx, OOPS_ERR_CAPACITY, "operator tables of " + std::to_string(img.size()) +
                                                   " bytes exceed the no-allocation array (" +
                                                   std::to_string(kNoAllocBytes) + " bytes)");
    if (hipSetDevice(ctx->device) != hipSuccess) return set_err(ctx, OOPS_ERR_HIP, "hipSetDevice");
    unsigned char* base = static_cast<unsigned char*>(ctx->noalloc);
    rc = hip_check(ctx, hipMemcpyAsync(base, img.data(), img.size(), hipMemcpyHostToDevice, ctx->stream),
                   "H2D no-allocation tables");
    if (rc) return rc;
    auto args_of = [&](const Table& t) {
        LargeOpArgs a;
        a.hdr = (const DevOpHeader*)(base + t.hdr);
        a.loc = (const DevOpLocal*)(base + t.loc);
        a.val = (const double2*)(base + t.val);
        a.off = (const int32_t*)(base + t.off);
        return a;
    };
    if (tiled) {
        const int ld = op.ld_bits;
        for (size_t p = 0; p < plan.size(); ++p) {
            LargeTilePass tp{};
            tp.op = args_of(tabs[p]);
            tp.map = (const int32_t*)(base + tabs[p].map);
            tp.ng = (int32_t)plan[p].G.size();
            tp.nb = op.nspin - tp.ng;
            tp.ld = ld;
            tp.ntiles = (int32_t)(op.n >> (ld * tp.ng));
            launch_large_tile_matvec(tp, (const double2*)x, nullptr, (double2*)y, nullptr, 0, p > 0 ? 1 : 0,
                                     ctx->stream);
        }
    } else {
        launch_large_matvec(args_of(tabs[0]), (const double2*)x, nullptr, (double2*)y, nullptr, op.n, 0,
                            large_grid(op.n, ctx->num_cu), ctx->stream);
    }
    rc = hip_check(ctx, hipGetLastError(), "no-allocation product launch");
    // the array is reused by the next call: wait for this one
    return rc ? rc : hip_check(ctx, hipStreamSynchronize(ctx->stream), "no-allocation product");
}

int oops_hamvec(oops_ctx* ctx, const oops_term_table* H, const double* v, double* w) {
    if (!ctx || !H || !v || !w) return OOPS_ERR_ARG;
    LargeRun R;
    int rc = setup(ctx, H, 1, R);
    if (rc) return rc;
    R.s = nullptr;
    std::vector<double> vp, wp;
    if (R.padded) {   // caller's layout <-> the padded device layout
        vp.assign(2 * (size_t)R.n, 0.0);
        wp.assign(2 * (size_t)R.n, 0.0);
        for (int64_t i = 0; i < R.n_logical; ++i) {
            vp[2 * R.pos[(size_t)i]] = v[2 * i];
            vp[2 * R.pos[(size_t)i] + 1] = v[2 * i + 1];
        }
    }
    rc = hip_check(ctx, hipMemcpy(R.p0.p, R.padded ? vp.data() : v, sizeof(double2) * R.n, hipMemcpyHostToDevice),
                   "H2D v");
    if (rc) return rc;
    matvec(R, (const double2*)R.p0.p, nullptr, (double2*)R.p1.p, nullptr, 0);
    rc = hip_check(ctx, hipGetLastError(), "matvec launch");
    rc = rc ? rc : hip_check(ctx, hipMemcpy(R.padded ? wp.data() : w, R.p1.p, sizeof(double2) * R.n,
                                            hipMemcpyDeviceToHost), "D2H w");
    if (!rc && R.padded)
        for (int64_t i = 0; i < R.n_logical; ++i) {
            w[2 * i] = wp[2 * R.pos[(size_t)i]];
            w[2 * i + 1] = wp[2 * R.pos[(size_t)i] + 1];
        }
    return rc;
}

int oops_expv(oops_ctx* ctx, const oops_term_table* H, const double* v, int32_t nt, const double* ta, int32_t klim,
              int32_t m, double tol, int32_t out_mode, double* out, double* stats) {
    return oops_expv_ex(ctx, H, v, nt, ta, klim, m, tol, 0, out_mode, out, stats);
}

int oops_expv_ex(oops_ctx* ctx, const oops_term_table* H, const double* v, int32_t nt, const double* ta,
                 int32_t klim, int32_t m, double tol, int32_t precision, int32_t out_mode, double* out,
                 double* stats) {
    if (!ctx || !H || !v || nt <= 0 || !ta) return set_err(ctx, OOPS_ERR_ARG, "bad arguments");
    if (precision & ~3)
        return set_err(ctx, OOPS_ERR_ARG, "precision: bit 0 FP32 basis, bit 1 output times inside steps; no other bits");
    if (out_mode != 0 && out_mode != 1) return set_err(ctx, OOPS_ERR_ARG, "out_mode must be 0 or 1");
    if (m < 2 || m > kLargeMaxM) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "m must be in [2, 40] on the large path");
    if (klim < 1 || klim > 64) return set_err(ctx, OOPS_ERR_ARG, "klim must be in [1, 64]");
    for (int k = 1; k < nt; ++k)
        if (!(ta[k] >= ta[k - 1])) return set_err(ctx, OOPS_ERR_ARG, "output times must be ascending");
    LargeRun R;
    R.fp32 = (precision & 1) != 0;
    int rc = setup(ctx, H, m, R);
    if (rc) return rc;
    if (m >= (R.padded ? R.n_logical : R.n))
        return set_err(ctx, OOPS_ERR_ARG, "the Krylov dimension must be smaller than the state dimension");
    std::vector<double> vp, outp;
    if (R.padded) {   // caller's layout <-> the padded device layout
        vp.assign(2 * (size_t)R.n, 0.0);
        for (int64_t i = 0; i < R.n_logical; ++i) {
            vp[2 * R.pos[(size_t)i]] = v[2 * i];
            vp[2 * R.pos[(size_t)i] + 1] = v[2 * i + 1];
        }
        if (out_mode == 0) outp.assign(2 * (size_t)R.n * nt, 0.0);
    }
    rc = hip_check(ctx, hipStreamCreateWithFlags(&R.s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    if (rc) return rc;
    Counters C;
    rc = run_expv(R, R.padded ? vp.data() : v, nt, ta, klim, tol, out_mode,
                  (R.padded && out_mode == 0) ? outp.data() : out, C, nullptr, (precision & 2) != 0);
    (void)hipStreamDestroy(R.s);
    if (!rc && R.padded && out_mode == 0 && out)
        for (int q = 0; q < nt; ++q)
            for (int64_t i = 0; i < R.n_logical; ++i) {
                out[2 * ((size_t)q * R.n_logical + i)] = outp[2 * ((size_t)q * R.n + R.pos[(size_t)i])];
                out[2 * ((size_t)q * R.n_logical + i) + 1] = outp[2 * ((size_t)q * R.n + R.pos[(size_t)i]) + 1];
            }
    if (stats) {
        stats[0] = (double)C.nmult;
        stats[1] = (double)C.nexph;
        stats[2] = (double)C.nscale;
        stats[3] = (double)C.nstep;
        stats[4] = (double)C.nreject;
        stats[5] = (double)C.ibrk;
        stats[6] = C.anorm;
        stats[7] = C.kernel_ms;
    }
    return rc;
}

// Fortran-ABI twin of the reference's MAIN_CACHE (all arguments by reference,
// size_t integers as MatExpVector::runInexplicitGPU passes them,
// src/source/math/MatExp.cpp:487-511).  maxThreadsPerBlock / maxGridSize are
// accepted and ignored (launch shapes are chosen here).  Runs on device 0.
void main_cache_(size_t* nspin, size_t* nTerm, double* coeff_lst, size_t* nbody_lst, size_t* pos_i_idx,
                 size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
                 size_t* nspin_dim, double* v, size_t* pos_idx, size_t* mat_idx, size_t* k, size_t* maxThreadsPerBlock,
                 size_t* maxGridSize, size_t* tn, double* ta, size_t* m, double* tol, size_t* itrace, double* w_seq,
                 size_t* w_seq_len) {
    (void)pos_idx; (void)mat_idx; (void)maxThreadsPerBlock; (void)maxGridSize; (void)itrace; (void)ham_dim;
    (void)w_seq_len;
    oops_term_table T;
    T.nspin = (int32_t)*nspin;
    T.nterm = (int32_t)*nTerm;
    T.coeff = coeff_lst;
    T.nbody = reinterpret_cast<const int64_t*>(nbody_lst);
    T.pos_offset = reinterpret_cast<const int64_t*>(pos_i_idx);
    T.pos = reinterpret_cast<const int64_t*>(pos_i_lst);
    T.dims = reinterpret_cast<const int64_t*>(dim_i_lst);
    T.mat_offset = reinterpret_cast<const int64_t*>(mat_i_idx);
    T.mats = mat_i_lst;
    T.spin_dim = reinterpret_cast<const int64_t*>(nspin_dim);
    oops_ctx* ctx = nullptr;
    if (oops_ctx_create(0, &ctx) != OOPS_OK) return;
    oops_expv(ctx, &T, v, (int32_t)*tn, ta, (int32_t)*k, (int32_t)*m, *tol, 0, w_seq, nullptr);
    oops_ctx_destroy(ctx);
}

// ---------------------------------------------------------------------------
// Link-compatible symbols of the reference's expv library (expv/include/*.h).
// Integers are by reference where the reference's Fortran ABI has them so;
// complex arrays are interleaved (re, im) doubles.  They run on device 0
// through one process-wide context.

static oops_ctx* abi_ctx() {
    static oops_ctx* ctx = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (oops_ctx_create(0, &ctx) != OOPS_OK) ctx = nullptr;
    });
    return ctx;
}

static oops_term_table term_table(size_t nspin, size_t nterm, const double* coeff, const size_t* nbody,
                                  const size_t* pos_i_idx, const size_t* pos_i_lst, const size_t* dim_i_lst,
                                  const size_t* mat_i_idx, const double* mats, const size_t* nspin_dim) {
    oops_term_table T;
    T.nspin = (int32_t)nspin;
    T.nterm = (int32_t)nterm;
    T.coeff = coeff;
    T.nbody = reinterpret_cast<const int64_t*>(nbody);
    T.pos_offset = reinterpret_cast<const int64_t*>(pos_i_idx);
    T.pos = reinterpret_cast<const int64_t*>(pos_i_lst);
    T.dims = reinterpret_cast<const int64_t*>(dim_i_lst);
    T.mat_offset = reinterpret_cast<const int64_t*>(mat_i_idx);
    T.mats = mats;
    T.spin_dim = reinterpret_cast<const int64_t*>(nspin_dim);
    return T;
}

// MAIN_MKL (main_mkl.h:10): main_cache_ without the launch-shape arguments.
void main_mkl_(size_t* nspin, size_t* nTerm, double* coeff_lst, size_t* nbody_lst, size_t* pos_i_idx,
               size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
               size_t* nspin_dim, double* v, size_t* pos_idx, size_t* mat_idx, size_t* k, size_t* tn, double* ta,
               size_t* m, double* tol, size_t* itrace, double* w_seq, size_t* w_seq_len) {
    (void)pos_idx; (void)mat_idx; (void)itrace; (void)ham_dim; (void)w_seq_len;
    oops_ctx* ctx = abi_ctx();
    if (!ctx) return;
    oops_term_table T = term_table(*nspin, *nTerm, coeff_lst, nbody_lst, pos_i_idx, pos_i_lst, dim_i_lst, mat_i_idx,
                                   mat_i_lst, nspin_dim);
    oops_expv(ctx, &T, v, (int32_t)*tn, ta, (int32_t)*k, (int32_t)*m, *tol, 0, w_seq, nullptr);
}

static int csr_expv(int64_t n, const std::vector<int64_t>& rp, const std::vector<int32_t>& ci,
                    const std::vector<double>& cv, const double* v, int tn, const double* ta, double* w_seq, int klim,
                    int m, double tol) {
    oops_ctx* ctx = abi_ctx();
    if (!ctx) return OOPS_ERR_HIP;
    if (m < 2 || m > kLargeMaxM || m >= n) return set_err(ctx, OOPS_ERR_ARG, "m must be in [2, 40] and below n");
    if (klim < 1 || klim > 64) klim = 10;
    for (int k = 1; k < tn; ++k)
        if (!(ta[k] >= ta[k - 1])) return set_err(ctx, OOPS_ERR_ARG, "output times must be ascending");
    LargeRun R;
    int rc = setup_csr(ctx, n, rp, ci, cv, m, R);
    if (rc) return rc;
    rc = hip_check(ctx, hipStreamCreateWithFlags(&R.s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    if (rc) return rc;
    Counters C;
    rc = run_expv(R, v, tn, ta, klim, tol, 0, w_seq, C);
    (void)hipStreamDestroy(R.s);
    return rc;
}

// krylov_zgexpv (krylov_expv.h:6): w(t_k) = exp(t_k A) v for a dense n x n
// column-major A (the caller folds the -i prefactor into A, MatExp.cpp:108).
// Returns 0 or a negative status (the reference returns Expokit's iflag).
int krylov_zgexpv(const int n, const double* a, const double* v, const int tn, const double* ta, double* w_seq,
                  const int klim, const int m, const double tol, const int itrace) {
    (void)itrace;
    if (n <= 0 || !a || !v || tn <= 0 || !ta || !w_seq) return OOPS_ERR_ARG;
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> ci;
    std::vector<double> cv;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            const double re = a[2 * ((size_t)i + (size_t)n * j)], im = a[2 * ((size_t)i + (size_t)n * j) + 1];
            if (re == 0.0 && im == 0.0) continue;
            ci.push_back(j);
            cv.push_back(re);
            cv.push_back(im);
        }
        rp[i + 1] = (int64_t)ci.size();
    }
    return csr_expv(n, rp, ci, cv, v, tn, ta, w_seq, klim, m, tol);
}

// krylov_zcooexpv (krylov_expv.h:9): the same for a sparse A in 1-based COO
// (ia rows, ja columns, MatExp.cpp:188-196); duplicates are summed.
int krylov_zcooexpv(const int n, const int nz, const int* ia, const int* ja, const double* a, const double* v,
                    const int tn, const double* ta, double* w_seq, const int klim, const int m, const double tol,
                    const int itrace) {
    (void)itrace;
    if (n <= 0 || nz < 0 || (nz > 0 && (!ia || !ja || !a)) || !v || tn <= 0 || !ta || !w_seq) return OOPS_ERR_ARG;
    std::vector<std::vector<std::pair<int32_t, std::complex<double>>>> rows(n);
    for (int e = 0; e < nz; ++e) {
        const int i = ia[e] - 1, j = ja[e] - 1;
        if (i < 0 || i >= n || j < 0 || j >= n) return OOPS_ERR_ARG;
        rows[i].push_back({j, std::complex<double>(a[2 * e], a[2 * e + 1])});
    }
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> ci;
    std::vector<double> cv;
    for (int i = 0; i < n; ++i) {
        auto& r = rows[i];
        std::sort(r.begin(), r.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
        for (size_t q = 0; q < r.size(); ++q) {
            if (!ci.empty() && (int64_t)ci.size() > rp[i] && ci.back() == r[q].first) {
                cv[cv.size() - 2] += r[q].second.real();
                cv[cv.size() - 1] += r[q].second.imag();
                continue;
            }
            ci.push_back(r[q].first);
            cv.push_back(r[q].second.real());
            cv.push_back(r[q].second.imag());
        }
        rp[i + 1] = (int64_t)ci.size();
    }
    return csr_expv(n, rp, ci, cv, v, tn, ta, w_seq, klim, m, tol);
}

static int hamvec_complex(oops_ctx* ctx, const oops_term_table& T, const std::complex<double>* cc,
                          const double2* dv, double2* dw, hipStream_t st) {
    LargeRun R;
    int rc = setup(ctx, &T, 1, R, cc);
    if (rc) return rc;
    R.s = st;
    if (R.padded) {
        // the caller's device vectors are in its own layout: through the host
        // into the padded layout and back
        std::vector<double> hv(2 * (size_t)R.n_logical), vp(2 * (size_t)R.n, 0.0), wp(2 * (size_t)R.n);
        rc = hip_check(ctx, hipMemcpy(hv.data(), dv, sizeof(double2) * R.n_logical, hipMemcpyDeviceToHost), "D2H v");
        if (rc) return rc;
        for (int64_t i = 0; i < R.n_logical; ++i) {
            vp[2 * R.pos[(size_t)i]] = hv[2 * i];
            vp[2 * R.pos[(size_t)i] + 1] = hv[2 * i + 1];
        }
        rc = hip_check(ctx, hipMemcpy(R.p0.p, vp.data(), sizeof(double2) * R.n, hipMemcpyHostToDevice), "H2D v");
        if (rc) return rc;
        matvec(R, (const double2*)R.p0.p, nullptr, (double2*)R.p1.p, nullptr, 0);
        rc = hip_check(ctx, hipGetLastError(), "matvec launch");
        rc = rc ? rc : hip_check(ctx, hipMemcpy(wp.data(), R.p1.p, sizeof(double2) * R.n, hipMemcpyDeviceToHost), "D2H w");
        if (rc) return rc;
        for (int64_t i = 0; i < R.n_logical; ++i) {
            hv[2 * i] = wp[2 * R.pos[(size_t)i]];
            hv[2 * i + 1] = wp[2 * R.pos[(size_t)i] + 1];
        }
        return hip_check(ctx, hipMemcpy(dw, hv.data(), sizeof(double2) * R.n_logical, hipMemcpyHostToDevice), "H2D w");
    }
    matvec(R, dv, nullptr, dw, nullptr, 0);
    rc = hip_check(ctx, hipGetLastError(), "matvec launch");
    return rc ? rc : hip_check(ctx, hipStreamSynchronize(st), "sync matvec");
}

// hamvec_func3_ (hamvec_func3.cpp:399): w = sum_t coeff_zplx[t] (x)_b A_tb v,
// host vectors; w_med is not needed.
void hamvec_func3_(int* nspin, int* nTerm, double* coeff_lst_zplx, size_t* nbody_lst, size_t* pos_i_idx,
                   size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx, double* mat_i_lst, size_t* ham_dim,
                   size_t* nspin_dim, size_t* nspin_m_lst, size_t* nspin_n_lst, double* v, double* w, double* w_med) {
    (void)nspin_m_lst; (void)nspin_n_lst; (void)w_med;
    oops_ctx* ctx = abi_ctx();
    if (!ctx) return;
    std::vector<double> ones(*nTerm, 1.0);
    oops_term_table T = term_table(*nspin, *nTerm, ones.data(), nbody_lst, pos_i_idx, pos_i_lst, dim_i_lst, mat_i_idx,
                                   mat_i_lst, nspin_dim);
    const size_t n = *ham_dim;
    DevBuf dv, dw;
    if (alloc(ctx, dv, sizeof(double2) * n) || alloc(ctx, dw, sizeof(double2) * n)) return;
    if (hipMemcpy(dv.p, v, sizeof(double2) * n, hipMemcpyHostToDevice) != hipSuccess) return;
    if (hamvec_complex(ctx, T, reinterpret_cast<const std::complex<double>*>(coeff_lst_zplx), (const double2*)dv.p,
                       (double2*)dw.p, ctx->stream))
        return;
    (void)hipMemcpy(w, dw.p, sizeof(double2) * n, hipMemcpyDeviceToHost);
}

// hamvec_cuda3_init_ / _term_ (hamvec_cuda3.cu:661-665): the handle is an
// oops_ctx on device 0 (the reference stores a cuBLAS handle there).
void hamvec_cuda3_init_(size_t* handle) {
    oops_ctx* c = nullptr;
    if (handle) *handle = oops_ctx_create(0, &c) == OOPS_OK ? (size_t)c : 0;
}

void hamvec_cuda3_term_(size_t* handle) {
    if (handle && *handle) {
        oops_ctx_destroy((oops_ctx*)*handle);
        *handle = 0;
    }
}

// hamvec_cuda3_ (hamvec_cuda3.cu:661): device pointers passed as size_t; the
// local matrices are device memory (dev_mat_i_lst), coefficients on the host.
void hamvec_cuda3_(size_t* handle, int* nspin, int* nTerm, double* coeff_lst_zplx, size_t* nbody_lst,
                   size_t* pos_i_idx, size_t* pos_i_lst, size_t* dim_i_lst, size_t* mat_i_idx,
                   size_t* dev_mat_i_lst, size_t* ham_dim, size_t* nspin_dim, size_t* nspin_m_lst,
                   size_t* nspin_n_lst, size_t* dev_v, size_t* dev_w, size_t* dev_w_med, size_t* dev_coeff,
                   size_t* maxThreadsPerBlock, size_t* maxGridSize) {
    (void)ham_dim; (void)nspin_m_lst; (void)nspin_n_lst; (void)dev_w_med; (void)dev_coeff;
    (void)maxThreadsPerBlock; (void)maxGridSize;
    oops_ctx* ctx = handle ? (oops_ctx*)*handle : nullptr;
    if (!ctx) return;
    size_t nbodies = 0;
    for (int t = 0; t < *nTerm; ++t) nbodies += nbody_lst[t];
    const size_t nmat = mat_i_idx[nbodies];
    std::vector<double> mats(2 * std::max<size_t>(nmat, 1));
    if (nmat && hipMemcpy(mats.data(), (const void*)*dev_mat_i_lst, sizeof(double2) * nmat, hipMemcpyDeviceToHost) !=
                    hipSuccess)
        return;
    std::vector<double> ones(*nTerm, 1.0);
    oops_term_table T = term_table(*nspin, *nTerm, ones.data(), nbody_lst, pos_i_idx, pos_i_lst, dim_i_lst, mat_i_idx,
                                   mats.data(), nspin_dim);
    hamvec_complex(ctx, T, reinterpret_cast<const std::complex<double>*>(coeff_lst_zplx), (const double2*)*dev_v,
                   (double2*)*dev_w, ctx->stream);
}

}  // extern "C"

// rocBLAS products and the rocSOLVER LU solve of the dense Pade path,
// resolved when it is first used (no link-time dependency: a process that
// already holds a copy of these libraries, PyTorch's for one, shares it).
namespace {
struct DenseLib {
    bool ok = false;
    rocblas_handle h = nullptr;
    rocblas_status (*create)(rocblas_handle*) = nullptr;
    rocblas_status (*set_stream)(rocblas_handle, hipStream_t) = nullptr;
    rocblas_status (*zgemm)(rocblas_handle, rocblas_operation, rocblas_operation, rocblas_int, rocblas_int,
                            rocblas_int, const rocblas_double_complex*, const rocblas_double_complex*, rocblas_int,
                            const rocblas_double_complex*, rocblas_int, const rocblas_double_complex*,
                            rocblas_double_complex*, rocblas_int) = nullptr;
    rocblas_status (*zaxpy)(rocblas_handle, rocblas_int, const rocblas_double_complex*,
                            const rocblas_double_complex*, rocblas_int, rocblas_double_complex*,
                            rocblas_int) = nullptr;
    rocblas_status (*zdscal)(rocblas_handle, rocblas_int, const double*, rocblas_double_complex*,
                             rocblas_int) = nullptr;
    rocblas_status (*zgesv)(rocblas_handle, rocblas_int, rocblas_int, rocblas_double_complex*, rocblas_int,
                            rocblas_int*, rocblas_double_complex*, rocblas_int, rocblas_int*) = nullptr;
};

template <typename F>
bool resolve(void* lib, const char* name, F& f) {
    f = lib ? reinterpret_cast<F>(dlsym(lib, name)) : nullptr;
    return f != nullptr;
}

// One handle on the given context's stream (the process-wide context of the
// reference ABI symbols).
DenseLib* dense_lib(oops_ctx* ctx) {
    static DenseLib L;
    static std::once_flag once;
    std::call_once(once, [ctx] {
        void* b = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
        if (!b) b = dlopen("librocblas.so", RTLD_NOW | RTLD_GLOBAL);
        void* s = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL);
        if (!s) s = dlopen("librocsolver.so", RTLD_NOW | RTLD_GLOBAL);
        L.ok = resolve(b, "rocblas_create_handle", L.create) && resolve(b, "rocblas_set_stream", L.set_stream) &&
               resolve(b, "rocblas_zgemm", L.zgemm) && resolve(b, "rocblas_zaxpy", L.zaxpy) &&
               resolve(b, "rocblas_zdscal", L.zdscal) && resolve(s, "rocsolver_zgesv", L.zgesv);
        L.ok = L.ok && L.create(&L.h) == rocblas_status_success &&
               L.set_stream(L.h, ctx->stream) == rocblas_status_success;
    });
    return L.ok ? &L : nullptr;
}
}  // namespace

// exp(t H) of a dense m x m complex matrix (column-major, leading dimension
// ldh) by ZGPADM's Pade(6) with scaling and squaring (expokit.f:762-934),
// step for step: the infinity norm and squaring count, the Horner recurrence
// in H^2 for the numerator and denominator, the final product with H, the
// solve (Q - P) X = P, E = +-(I + 2X) and ns squarings; products on rocBLAS,
// the LU solve on rocSOLVER, on the context's stream.  out: m*m complex.
static int pade_dense(oops_ctx* ctx, int m, double t, const double* H, int ldh, double* out, int* ns_out) {
    DenseLib* L = dense_lib(ctx);
    if (!L) return set_err(ctx, OOPS_ERR_UNSUPPORTED, "rocBLAS / rocSOLVER not available for the dense Pade path");
    typedef rocblas_double_complex zc;
    const size_t mm = (size_t)m * m;
    double hnorm = 0.0;
    for (int i = 0; i < m; ++i) {
        double row = 0.0;
        for (int j = 0; j < m; ++j) row += std::hypot(H[2 * (i + (size_t)ldh * j)], H[2 * (i + (size_t)ldh * j) + 1]);
        hnorm = std::max(hnorm, row);
    }
    hnorm = std::fabs(t * hnorm);
    if (hnorm == 0.0) return set_err(ctx, OOPS_ERR_ARG, "null H");
    const int ideg = 6;
    const int ns = std::max(0, (int)(std::log(hnorm) / std::log(2.0)) + 2);
    const double scale = t / std::ldexp(1.0, ns);
    double coef[ideg + 1];
    coef[0] = 1.0;
    for (int k = 1; k <= ideg; ++k) coef[k] = coef[k - 1] * (double)(ideg + 1 - k) / (double)(k * (2 * ideg + 1 - k));
    std::vector<double> hp(2 * mm);
    for (int j = 0; j < m; ++j)
        std::memcpy(&hp[2 * (size_t)m * j], &H[2 * (size_t)ldh * j], sizeof(double) * 2 * m);
    DevBuf dH, dH2, b0, b1, b2, piv, info;
    int rc = upload(ctx, dH, hp);
    for (DevBuf* d : {&dH2, &b0, &b1, &b2}) rc = rc ? rc : alloc(ctx, *d, sizeof(zc) * mm);
    rc = rc ? rc : alloc(ctx, piv, sizeof(rocblas_int) * m);
    rc = rc ? rc : alloc(ctx, info, sizeof(rocblas_int));
    if (rc) return rc;
    auto Z = [](const DevBuf& d) { return (zc*)d.p; };
    const zc one{1.0, 0.0}, zero{0.0, 0.0}, mone{-1.0, 0.0}, s2{scale * scale, 0.0}, sc{scale, 0.0};
    bool ok = true;
    auto blas = [&](rocblas_status st) { ok = ok && st == rocblas_status_success; };
    const rocblas_operation N = rocblas_operation_none;
    blas(L->zgemm(L->h, N, N, m, m, m, &s2, Z(dH), m, Z(dH), m, &zero, Z(dH2), m));   // (scale H)^2
    zc* p = Z(b0);
    zc* q = Z(b1);
    zc* f = Z(b2);
    ok = ok && hipMemsetAsync(p, 0, sizeof(zc) * mm, ctx->stream) == hipSuccess &&
         hipMemsetAsync(q, 0, sizeof(zc) * mm, ctx->stream) == hipSuccess;
    launch_large_diag_add((double2*)p, m, make_double2(coef[ideg - 1], 0.0), ctx->stream);
    launch_large_diag_add((double2*)q, m, make_double2(coef[ideg], 0.0), ctx->stream);
    bool odd = true;
    for (int k = ideg - 1; k > 0; --k) {   // Horner in H^2, numerator and denominator alternately
        zc* used = odd ? q : p;
        blas(L->zgemm(L->h, N, N, m, m, m, &one, used, m, Z(dH2), m, &zero, f, m));
        launch_large_diag_add((double2*)f, m, make_double2(coef[k - 1], 0.0), ctx->stream);
        if (odd) q = f;
        else p = f;
        f = used;
        odd = !odd;
    }
    if (odd) {
        blas(L->zgemm(L->h, N, N, m, m, m, &sc, q, m, Z(dH), m, &zero, f, m));
        std::swap(q, f);
    } else {
        blas(L->zgemm(L->h, N, N, m, m, m, &sc, p, m, Z(dH), m, &zero, f, m));
        std::swap(p, f);
    }
    blas(L->zaxpy(L->h, (rocblas_int)mm, &mone, p, 1, q, 1));                          // Q - P
    blas(L->zgesv(L->h, m, m, q, m, (rocblas_int*)piv.p, p, m, (rocblas_int*)info.p));   // X = (Q - P)^-1 P
    const double two = 2.0, neg = -1.0;
    blas(L->zdscal(L->h, (rocblas_int)mm, &two, p, 1));
    launch_large_diag_add((double2*)p, m, make_double2(1.0, 0.0), ctx->stream);      // I + 2X
    zc* res = p;
    if (ns == 0 && odd) {
        blas(L->zdscal(L->h, (rocblas_int)mm, &neg, p, 1));
    } else {
        bool parity = true;
        for (int k = 0; k < ns; ++k) {   // squarings, alternating between p and q
            zc* get = parity ? p : q;
            zc* put = parity ? q : p;
            blas(L->zgemm(L->h, N, N, m, m, m, &one, get, m, get, m, &zero, put, m));
            res = put;
            parity = !parity;
        }
    }
    if (!ok) return set_err(ctx, OOPS_ERR_HIP, "rocBLAS / rocSOLVER call failed");
    rocblas_int st = 0;
    rc = hip_check(ctx, hipMemcpyAsync(out, res, sizeof(zc) * mm, hipMemcpyDeviceToHost, ctx->stream), "copy");
    rc = rc ? rc : hip_check(ctx, hipMemcpyAsync(&st, info.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream), "copy");
    rc = rc ? rc : hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc) return rc;
    if (st != 0) return set_err(ctx, OOPS_ERR_NUMERIC, "singular Pade denominator");
    *ns_out = ns;
    return OOPS_OK;
}

extern "C" {

// zgpadm_ (expokit.h:6): exp(t H) of an m x m matrix by Pade(ideg = 6) with
// scaling and squaring on the device; the result is written to wsp(1..m*m)
// (*iexph = 1), *ns squarings; *iflag < 0 on error.  m <= 42 runs in one
// workgroup (the Krylov driver's small exponential); larger m (the dense
// propagators of MatExp::pade_exp_mat, MatExp.cpp:29-77) on rocBLAS /
// rocSOLVER.
void zgpadm_(int* ideg, int* m, double* t, double* H, int* ldh, double* wsp, int* lwsp, int* ipiv, int* iexph,
             int* ns, int* iflag) {
    (void)ipiv;
    *iflag = 0;
    const int mm = *m;
    if (*ideg != 6 || mm < 1 || *ldh < mm || *lwsp < mm * mm) { *iflag = -1; return; }
    oops_ctx* ctx = abi_ctx();
    if (!ctx) { *iflag = -2; return; }
    if (mm > kLargeMaxM + 2) {
        const int rc = pade_dense(ctx, mm, *t, H, *ldh, wsp, ns);
        if (rc) { *iflag = rc == OOPS_ERR_NUMERIC ? -8 : -2; return; }
        *iexph = 1;
        return;
    }
    std::vector<double> h((size_t)2 * mm * (mm + 1), 0.0);  // one spare column (tiled reads)
    for (int j = 0; j < mm; ++j)
        for (int i = 0; i < mm; ++i) {
            h[2 * (i + (size_t)mm * j)] = H[2 * (i + (size_t)(*ldh) * j)];
            h[2 * (i + (size_t)mm * j) + 1] = H[2 * (i + (size_t)(*ldh) * j) + 1];
        }
    DevBuf dh, dout, dst;
    if (upload(ctx, dh, h) || alloc(ctx, dout, sizeof(double2) * mm * mm) || alloc(ctx, dst, sizeof(int) * 4)) {
        *iflag = -2;
        return;
    }
    if (launch_large_pade_full((const double2*)dh.p, mm, mm, *t, (double2*)dout.p, (int*)dst.p, ctx->stream) != 0) {
        *iflag = -2;
        return;
    }
    int st = 0;
    if (hipMemcpyAsync(wsp, dout.p, sizeof(double2) * mm * mm, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(&st, dst.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        *iflag = -2;
        return;
    }
    if (st < 0) { *iflag = st == -2 ? -8 : -9; return; }
    *ns = st;
    *iexph = 1;
}

}  // extern "C"
