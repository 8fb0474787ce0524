// kc_host.h — host-side helpers shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "kc_device.h"

namespace kc {

constexpr uint32_t BLOCK = 256;
constexpr uint32_t DATA_PAD = 64;  // bytes readable past the last read byte (window over-read)

// Device allocation that only grows.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        release();
        const size_t want = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        else p = nullptr;
        return e;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

inline Modulus make_modulus(uint64_t size_bits) {
    Modulus M{};
    M.d = size_bits;
    M.pow2 = (size_bits & (size_bits - 1)) == 0;
    M.mask = size_bits - 1;
    M.magic = M.pow2 ? 0 : (uint64_t)(((unsigned __int128)1 << 64) / size_bits);
    return M;
}

inline uint32_t grid_for(uint64_t work, uint32_t block = BLOCK, uint32_t cap = 16384) {
    const uint64_t g = (work + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, uint64_t n, T v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

inline std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

}  // namespace kc

struct kc_ctx;

namespace kc {
// One library's solid set after kc_count (kc_engine.hip): its size, k, and the context's
// device and stream.
struct SolidSet {
    int device;
    hipStream_t stream;
    uint32_t k;
    uint64_t n;
};
int solid_set(kc_ctx* ctx, uint32_t lib, SolidSet* s);
// Its k-mers in ascending order as vw value words each (V, lowest word first) and their
// counts, written on the context's stream into device buffers of SolidSet::n entries;
// idx and tmp are scratch.
int solid_values(kc_ctx* ctx, uint32_t lib, uint32_t vw, uint64_t* d_vals, uint32_t* d_counts, DevBuf& idx,
                 DevBuf& tmp);
// Record an error on the context (kc_last_error) and return code.
int ctx_fail(kc_ctx* ctx, int code, const std::string& msg);
// GRAPH_SCRATCH growable device buffers of the context that the unitig builder reuses
// from call to call (released with the context).
constexpr int GRAPH_SCRATCH = 32;
DevBuf* graph_scratch(kc_ctx* ctx);

// is_solid_kmer256 (bloom.rs:111-126) for n device-resident keys (kc_filter.hip).
hipError_t query_solid(const uint64_t* d_keys4, uint64_t n, const Modulus& M, const uint64_t* d_solid, uint8_t* d_out,
                       hipStream_t st);
}  // namespace kc
