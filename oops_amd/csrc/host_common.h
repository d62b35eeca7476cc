// Host-side helpers shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <string>
#include <vector>

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
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// max_locals: capacity of the kernel's local-operator table (LDS).
int compile_operator(const oops_term_table* T, CompiledOp& out, std::string& err, int max_locals = 64);
int set_err(oops_ctx* ctx, int code, const std::string& msg);
int hip_check(oops_ctx* ctx, hipError_t e, const char* what);
void dlaran_fill(int32_t* iseed, int64_t count, double* out);

}  // namespace oops
