// Launch interface between the host runtime (capi.cpp) and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oops_device_types.h"

namespace oops {

constexpr int kMaxLocals = 64;      // merged local operators per compiled operator
constexpr int kDenseMaxEnt = 1024;  // ELL entries per operator on the dense path
constexpr int kKrylovMaxEnt = 1024; // ELL entries per operator on the batched Krylov path
constexpr int kDenseNMAX = 16;      // largest state handled by the dense path
constexpr int kStatsPerItem = 24;   // int32 slots of per-cluster counters
// Flip-mask (slot) form of the product on the mixed Lanczos path: per element
// kSlot64 FP64 coefficients (masks changing at most one spin's digits; for
// spin-1/2 Liouvillians of k <= 4 spins: the diagonal and 2k ket/bra flips)
// and kSlot32 FP32 coefficients (two-spin flips: 2 k(k-1)/2 <= 12).
constexpr int kMediumMaxEnt = 2560;  // ELL entries per operator on the n = 1024 / 4096 kernel
// flip-mask form of the n = 1024 / 4096 kernel: per operator kMedMaskStride
// masks, kMedSlot64 FP64-class (one spin flipped) then kMedSlot32 FP32-class
constexpr int kMedSlot64 = 16;
constexpr int kMedSlot32 = 48;
constexpr int kMedMaskStride = kMedSlot64 + kMedSlot32;
constexpr int kSlot64 = 9;
constexpr int kSlot32 = 12;
constexpr int kSlotMaskStride = 32;                            // int32 masks per operator
constexpr int kSlotOpD2 = kSlot64 * 256 + kSlot32 * 256 / 2;   // double2 per operator table (n <= 256)

// Bit of item_npulse marking an item with two forward chains (single-sample
// form): the ket side runs opA / opC segments and the bra side opB / opD
// segments forward from u0 / y0, L = Re(scale <bra, ket>).  The pulse count is
// the low 16 bits.
constexpr int32_t kTwoChainFlag = 1 << 16;

// One bucket of Hahn-echo clusters sharing the state dimension n.
struct KrylovArgs {
    const DevOpHeader* hdr;
    const DevOpLocal* loc;
    const double2* val;
    const int32_t* off;
    const int32_t* item_opA;   // operator index of sweep A (ket side, lvA)
    const int32_t* item_opB;   // operator index of sweep B (bra side, already -lvB^dagger)
    const int32_t* item_opC;   // CPMG-N only: lvB (ket side), or NULL
    const int32_t* item_opD;   // CPMG-N only: -lvA^dagger (bra side), or NULL
    const int32_t* item_npulse;  // pulses N per item (NULL: 1, the Hahn echo) | kTwoChainFlag
    const int32_t* item_nt;      // output times per item (NULL: nt for every item)
    const int64_t* item_tau_off; // offset of the item's times in tau (NULL: 0)
    const int64_t* item_L_off;   // offset of the item's L in L (NULL: item * nt)
    const int64_t* vec_off;    // offset of the item's u0 / y0 in the vector pool
    const double2* u0;
    const double2* y0;
    const double* scale;       // observable prefactor (the Hilbert dimension D)
    const double* tau;         // ascending times: sweep times t/2 (Hahn), full times t (CPMG-N)
    const double* rnd;         // DLARAN draws: 2n start-vector values + klim power-iteration values
    double* L;                 // nitems x nt
    int32_t* stats;            // nitems x kStatsPerItem
    int32_t* queue;            // work counter (zeroed by the launcher)
    const uint8_t* eslot;      // slot of every ELL entry (Lanczos slot kernel), or NULL
    const int32_t* smask;      // kSlotMaskStride masks per operator: kSlot64 FP64-class, then kSlot32
    double2* slotws;           // per-workgroup coefficient tables: 4 x kSlotOpD2 double2 per workgroup
    double2* sweepws;          // Hahn state store, or NULL: slot kernel 256 x ws_nt per workgroup (sweep A);
                               // medium kernel n x ws_nt per item and sweep
    int32_t nitems;
    int32_t n;
    int32_t nt;
    int32_t klim;
    int32_t max_steps;
    int32_t ws_nt;             // output times per workgroup that sweepws holds
    double tol;
    // flip-mask form of the n = 1024 / 4096 kernel (NULL medtab: ELL tables in LDS)
    const int32_t* medmask;    // kMedMaskStride masks per operator
    double2* medtab;           // per workgroup 4 generators x (ms64 FP64 + ms32 FP32 slots) x n coefficients
    int32_t ms64;              // slots per generator in the bucket (at most kMedSlot64 / kMedSlot32)
    int32_t ms32;
};

int launch_hahn_bucket(const KrylovArgs& a, int m, int precision, int grid, hipStream_t stream);
// Mixed-precision Lanczos kernel for buckets whose generators are all Hermitian
// (n > m); the caller zeroes the work counter first.
int launch_hahn_lanczos(const KrylovArgs& a, int m, int grid, hipStream_t stream);
// Slot form of the Lanczos kernel: every operator of the bucket has at most
// kSlot64 / kSlot32 slots; a.slotws holds grid * 4 * slot_table_double2(fp64)
// double2.  fp64 false: pair coefficients and the basis in FP32 (two
// workgroups per CU); true: everything in FP64, the same algorithm, schedule
// and tolerance (one workgroup per CU).  The caller zeroes the work counter.
int launch_hahn_lanczos_slot(const KrylovArgs& a, int m, int grid, hipStream_t stream, bool fp64 = false);
int64_t slot_table_double2(bool fp64);
// Mixed Lanczos kernel for n = 1024 and 4096 (1024 threads per cluster, the
// product input and the active generator's tables in LDS, the FP32 basis in
// `work`: medium_workspace_float2(grid, m, n) float2).  Hermitian generators,
// at most kMediumMaxEnt ELL entries and kMaxLocals local operators each.
int launch_hahn_medium(const KrylovArgs& a, float2* work, int m, int grid, hipStream_t stream);
int64_t medium_workspace_float2(int grid, int m, int n);
void launch_zero_counter(int32_t* counter, hipStream_t stream);
// Dense FP64-MFMA propagator kernel for n = 64 buckets whose items all have
// uniform time grids starting at 0 (Hahn echo or CPMG-N); `work` holds
// dense64_workspace_doubles(grid) doubles.  The caller zeroes the counter.
int launch_dense64(const KrylovArgs& a, double* work, int grid, hipStream_t stream);
int64_t dense64_workspace_doubles(int grid);
int krylov_entry_capacity(int m);

// Factored propagator powers (cce_factored.hip): CPMG-N / Hahn coherence of
// unpolarised clusters of k <= 4 13C spins without dephasing, on uniform time
// grids from 0; one 256-thread workgroup per cluster (work queue).
struct FactoredArgs {
    const double* coords;       // every bath spin, x y z (Angstrom)
    const int32_t* clusters;    // k spin indices per item
    int32_t nitems;
    int32_t k;
    int32_t npulse;             // N (1: Hahn echo, segments of tau each)
    double zee[3];              // -gamma_C B
    double center[3];           // NV electron position
    double gamma_c, gamma_e;
    double sv0[3], sv1[3];      // <S> of the two NV states
    const int32_t* item_nt;
    const int64_t* item_tau_off;
    const int64_t* item_L_off;
    const double* tau;
    double* L;
    int32_t* stats;             // nitems x kStatsPerItem (slot 8: failure flags)
    int32_t* queue;             // work counter, zeroed by the caller
};
int launch_factored(const FactoredArgs& a, int grid, hipStream_t stream);

}  // namespace oops
