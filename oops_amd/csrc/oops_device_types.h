// Compiled-operator layout shared by the host compiler (capi.cpp) and the
// device kernels.  One operator  H = c0*I + sum_ops (I x A_op x I)  acting
// on a state of n = d^nspin entries (d = 2^ld_bits, spin 0 slowest).
#pragma once
#include <stdint.h>

namespace oops {

struct DevOpHeader {      // 64 bytes
    int32_t n;            // state dimension
    int32_t nspin;        // number of local factors
    int32_t ld_bits;      // log2 of the local dimension
    int32_t nloc;         // merged local operators
    int32_t nent;         // ELL entries of all local operators
    int32_t loc_off;      // first DevOpLocal of this operator
    int32_t ent_off;      // first entry of this operator
    int32_t max_row;      // largest R over the local operators
    double c0re, c0im;    // 0-body (identity) coefficient
    int32_t ent2_first;   // first entry of the 2-body operators (1-body entries come first)
    int32_t pad0;
    double pad1;
};

struct DevOpLocal {       // 16 bytes
    int16_t nb;           // 1 or 2 bodies
    int16_t s;            // first spin (slower)
    int16_t t;            // second spin (faster), == s for 1-body
    int16_t R;            // entries per row (padded)
    int32_t ent;          // first entry, relative to the operator's block
    int32_t rows;         // d or d*d
};

}  // namespace oops
