// Interface of the device-side term builder (term_builder.hip): packed
// operator tables of 13C clusters of up to four spin-1/2 spins in Liouville
// space, built on the GPU in the layout of the host compiler.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oops_device_types.h"
#include "oops_kernels.h"

namespace oops {

constexpr int kTBSlot64 = kSlot64;
constexpr int kTBSlot32 = kSlot32;
constexpr int kTBSlotMaskStride = kSlotMaskStride;
constexpr int kTBMedSlot64 = kMedSlot64;
constexpr int kTBMedMaskStride = kMedMaskStride;
constexpr int kTBSummary = 8;   // int32 per operator: nloc, nent, max_row, nslot64, nslot32, hermitian, error, 0

struct TermBuilderArgs {
    const double* coords;     // every bath spin, x y z (Angstrom)
    const int32_t* clusters;  // k spin indices per item
    int32_t nitems;
    int32_t k;                // spins per cluster (1..4)
    int32_t nops;             // generators per item: 2 (lvA, -lvB^dagger) or 4 (+ lvB, -lvA^dagger)
    double zee[3];            // -gamma_C B_a (the Zeeman coefficients)
    double center[3];         // NV electron position
    double gamma_c, gamma_e;
    double sv0[3], sv1[3];    // <S> of the NV states state0 / state1
    double deph_rate;         // 2 pi x dephasing rate (0: no dephasing terms)
    double2 dm[16];           // dephasing superoperator i (sigma^T (x) sigma - I), column-major
    double2 sharp[48];        // SHARP(S_x), SHARP(S_y), SHARP(S_z): 3 x 16 column-major
    double2 flat[48];         // FLAT(S_x), FLAT(S_y), FLAT(S_z)
    // outputs, operator o = item * nops + which at fixed strides
    DevOpHeader* hdr;
    DevOpLocal* loc;          // loc_stride per operator
    double2* val;             // ent_stride per operator
    int32_t* off;
    uint8_t* eslot;
    int32_t* smask;           // kSlotMaskStride per operator
    int32_t* mmask;           // kMedMaskStride per operator
    int32_t* summary;         // kTBSummary per operator
    int32_t loc_stride;
    int32_t ent_stride;
};

// One workgroup per operator (nops_total = nitems * nops).
int launch_term_builder(const TermBuilderArgs& a, int nops_total, hipStream_t stream);

}  // namespace oops
