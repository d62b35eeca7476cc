// Host-side helpers shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/oops_cce.h"
#include "../../include/oops_expv.h"
#include "oops_device_types.h"

// One device, its stream and the last error text.
struct oops_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    int lds_bytes = 0;
    std::string name;
    std::string err;
    void* noalloc = nullptr;   // device address of the no-allocation array on this device
    int wg_per_cu[16] = {};    // persistent batch grids: workgroups per CU by kernel kind (0: the default)
};

namespace oops {

// A term table compiled into merged local operators with ELL row tables.
struct CompiledOp {
    int64_t n = 0;
    int nspin = 0;
    int ld_bits = 0;
    int max_row = 0;
    int32_t ent2_first = 0;
    bool hermitian = false;  // every merged local operator is Hermitian
    std::complex<double> c0{0.0, 0.0};
    std::vector<DevOpLocal> locs;
    std::vector<std::complex<double>> val;
    std::vector<int32_t> off;
    // Flip-mask slots: every entry moves element i to i ^ mask (the local
    // digits are bit fields).  Slot 0 is mask 0; FP64-class slots have masks
    // that change at most one spin's digits, FP32-class slots the others.
    // eslot[e] = slot of entry e (bit 7 set: FP32 class); nslot64 < 0 when
    // either class exceeds 127 slots.
    std::vector<uint8_t> eslot;
    int nslot64 = 0, nslot32 = 0;
    std::vector<int32_t> mask64, mask32;
};

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// One item of a batch (Hahn items have npulse = 1, nt = 0, t = NULL).
// chains 1: one pulse sequence U_N..U_0 split between the ket side (u0, lvA /
// lvB segments) and the bra side (y0, adjoint segments), the ensemble form.
// chains 2: two sequences run forward from u0 and y0, the ket side with opA
// (even segments) / opC (odd), the bra side with opB (even) / opD (odd), and
// L = Re(scale <bra, ket>): the single-sample form (opB = opC, opD = opA).
struct EchoIn {
    const oops_term_table *opA, *opB, *opC, *opD;
    const double *u0, *y0;
    double scale;
    int32_t npulse, nt;
    const double* t;
    int32_t chains = 1;
};

// Compile, bucket and upload items (capi.cpp); per_item: times carried by the items.
int batch_create(oops_ctx* ctx, int32_t nitems, const EchoIn* items, bool per_item, int32_t m, int32_t klim,
                 double tol, int32_t precision, oops_hahn_batch** out);

// Items of every cluster of an enumeration for one pulse sequence and time
// grid, built on `threads` host threads (cce_host.cpp).  `hold` owns the term
// tables and vectors the items point to.
struct CceItems {
    std::shared_ptr<void> hold;
    std::vector<EchoIn> items;
    std::vector<double> t;
};
int cce_build_items(const double* coords, const oops_clusters* C, const oops_cce_params* P, int32_t npulse,
                    int32_t nt, const double* t, int threads, CceItems& out);

// Clusters of one size for the device term builder (term_builder.hip).
struct BuiltGroup {
    int k = 0;
    std::vector<int32_t> spins;    // k spin indices per cluster
    std::vector<int64_t> items;    // item index of each cluster in the batch
};
struct TermBuilderArgs;
// Batch whose operator tables the device builds from coordinates: groups by
// cluster size, constants (couplings aside) in `consts`, rho0 = vec(I / 2^k)
// per group (rho[k]); npulse >= 2: CPMG-N (four generators per item).
int batch_create_built(oops_ctx* ctx, const double* coords, int32_t nspin, const std::vector<BuiltGroup>& groups,
                       int64_t nitems, const TermBuilderArgs& consts, const std::vector<std::vector<double>>& rho,
                       int32_t npulse, bool per_item, int32_t nt, const double* t, int32_t m, int32_t klim, double tol,
                       int32_t precision, oops_hahn_batch** out);

// Inputs of batch_create_built for every cluster of C (cce_host.cpp); false
// when the device term builder does not take them.
bool cce_builder_inputs(const oops_clusters* C, const oops_cce_params* P, std::vector<BuiltGroup>& groups,
                        int64_t& nitems, TermBuilderArgs& consts, std::vector<std::vector<double>>& rho);

// ccoeff (optional, nterm complex): complex multipliers replacing T->coeff.
// One Hahn-echo item on the single-state path (large_host.cpp), FP64;
// L[0..nt) and 24 stats slots as the batched kernels write them.
int large_hahn_item(oops_ctx* ctx, const CompiledOp& A, const CompiledOp& B, const double* u0, const double* y0,
                    double scale, int nt, const double* tau, int m, int klim, double tol, double* L, int32_t* stats);

// max_row: widest merged local-operator row (ELL rows are padded to 4, 8 or
// 16 entries, wider ones to a multiple of 4; the batched kernels take 16).
int compile_operator(const oops_term_table* T, CompiledOp& out, std::string& err, int max_locals = 64,
                     const std::complex<double>* ccoeff = nullptr, int max_row = 16);
int set_err(oops_ctx* ctx, int code, const std::string& msg);
// Device address of the no-allocation array (large_host.cpp) on the current
// device, looked up once per device.
// The factored propagator powers take the group of k-spin clusters of an
// unpolarised batch with N pulses (capi.cpp).
bool factored_applies(int k, int npulse, double deph_rate);
int noalloc_array_address(int device, void** ptr);
// Step limit of every Krylov integration (environment OOPS_MAX_STEPS, else 5e7).
int32_t max_steps_limit();
int hip_check(oops_ctx* ctx, hipError_t e, const char* what);
void dlaran_fill(int32_t* iseed, int64_t count, double* out);

}  // namespace oops
