// Launch interface of the large-state Krylov path (expv_large.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oops_device_types.h"

namespace oops {

constexpr int kLargeMaxM = 40;      // Krylov dimension limit of the large path (LDS of the Pade kernel)
constexpr int kLargeMaxLocals = 1024;  // merged local operators of one operator (LDS of the product kernel)
constexpr int kLargeMaxRow = 256;      // widest merged local-operator row (padded spin-1 pairs: 81 entries)

struct LargeOpArgs {
    const DevOpHeader* hdr;  // device copy of the header
    const DevOpLocal* loc;
    const double2* val;
    const int32_t* off;
};

// Tiled product for large states: the vector is cut into tiles of
// 2^kLargeTileBits elements whose index bits are the digits of a set G of
// spins (the rest of the index fixed per workgroup); a tile is staged in LDS
// and every local operator acting inside G is applied there.  A few passes
// with different G (each one launch) cover all local operators; the first
// writes y, the others add to it.
constexpr int kLargeTileBits = 12;
constexpr int kLargeTileMapStride = 48;  // int32 per pass: [0,16) tile spins' bit positions, [16,48) block spins'
constexpr int kLargeTileMaxLocals = 256; // local operators of one pass (LDS)
constexpr int kLargeTileMaxEnt = 2560;   // ELL entries of one pass (LDS)

struct LargeTilePass {
    LargeOpArgs op;       // tile operator (local operators of this pass, offsets in tile digits)
    const int32_t* map;   // kLargeTileMapStride ints (device)
    int32_t ng;           // tile spins
    int32_t nb;           // block spins
    int32_t ld;           // bits per local digit
    int32_t ntiles;       // 2^(ld * nb)
};

void launch_large_tile_matvec(const LargeTilePass& tp, const double2* x, const double* xscale, double2* y,
                              void* xstore, int alpha_mode, int accumulate, hipStream_t s, bool fp32 = false);

// Two-axis product of a Liouville-space operator whose every term acts on the
// ket digits only or on the bra digits only (the SHARP / FLAT expansions of
// Hamiltonians): the state index holds spin s's ket digit at bit 2(k-1-s) and
// its bra digit at bit 2(k-1-s)+1, and L x = HK along the ket axis + HB along
// the bra axis, HK and HB sparse D x D (D = 2^k) in column-major ELL form
// (entry j of row r at j * D + r; padding entries have value 0 and col = r).
// Pass `side` (0: ket axis, 1: bra axis) stages tiles of D x 2^bbits elements
// (the whole axis times the other axis's lowest bbits digits) in LDS; pass 0
// writes y and stores the scaled basis vector, pass 1 adds to y.
constexpr int kLargeAxisMaxTile = 4096;   // elements of one staged tile (64 KB FP64)
struct LargeAxisPass {
    const double2* val;   // W x D
    const int32_t* col;   // W x D
    int32_t W;
    int32_t k;            // spins (D = 2^k)
    int32_t bbits;        // other-axis digits per tile
    int32_t side;         // 0: ket axis, 1: bra axis
    int32_t ntiles;       // 2^(k - bbits)
};
void launch_large_axis_matvec(const LargeAxisPass& ap, const double2* x, const double* xscale, double2* y,
                              void* xstore, int alpha_mode, int accumulate, hipStream_t s, bool fp32 = false);

int large_grid(int64_t n, int num_cu);
// xstore (basis column, may be NULL) is float2 when fp32, else double2.
void launch_large_matvec(const LargeOpArgs& o, const double2* x, const double* xscale, double2* y, void* xstore,
                         int64_t n, int alpha_mode, int grid, hipStream_t s, bool fp32 = false);
void launch_large_dots(const double2* V, int64_t ldv, const double2* p, int64_t n, int J, int with_norm, double* part,
                       int nblk, hipStream_t s);
void launch_large_finish(const double* part, int nblk, int nrows, int jcol, int mode, double2* hess, int ldh,
                         double2* coef, double* scal, int* flags, double break_tol, hipStream_t s);
// Classical Gram-Schmidt passes: part_l = V_l^H p (l <= J < kLargeMaxM) and
// part_{J+1} = |p|^2; p -= sum_l coef_l V_l and part_0 = |p|^2.  Launched
// with a skip flag, a pass does nothing while *skip is set.  part rows as for
// large_dots.
void launch_large_gs_dots(const void* V, int64_t ldv, const double2* p, int64_t n, int J, double* part, int nblk,
                          hipStream_t s, bool fp32 = false, const int* skip = nullptr);
void launch_large_gs_update(const void* V, int64_t ldv, double2* p, int64_t n, int J, const double2* coef,
                            double* part, int nblk, hipStream_t s, bool fp32 = false, const int* skip = nullptr);
int launch_large_pade(const double2* hess, int ldh, int mx, double t, double2* F, int* status, hipStream_t s);
void launch_large_combine(const void* V, int64_t ldv, int mx, const double2* F, double beta, double2* w, int64_t n,
                          double* part, int grid, hipStream_t s, bool fp32 = false);
void launch_large_csr_matvec(const int64_t* row_ptr, const int32_t* col, const double2* val, const double2* x,
                             const double* xscale, double2* y, double2* xstore, int64_t n, int alpha_mode, int grid,
                             hipStream_t s);
// exp(t H) for H m x m (m <= kLargeMaxM + 2), full matrix, leading dimension m in out
int launch_large_pade_full(const double2* hess, int ldh, int mx, double t, double2* out, int* status, hipStream_t s);
// A(i, i) += v for an m x m column-major matrix (leading dimension m).
void launch_large_diag_add(double2* A, int m, double2 v, hipStream_t s);
void launch_large_axpby(double2 a, const double2* x, double2 b, double2* y, int64_t n, int grid, hipStream_t s);

}  // namespace oops
