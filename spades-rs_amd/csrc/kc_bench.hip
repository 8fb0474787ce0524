// kc_bench.hip — device microbenchmarks that give the measured ceilings the engine's
// kernels are compared with: random atomics on an array the size of a count table, random
// 64-byte reads, and a streaming copy (include/kc_bench.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/kc_bench.h"

namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// n_ops random operations on `words` 64-bit words; op i touches word mix64(i) % words.
__global__ void k_random(uint64_t* __restrict__ a, uint64_t words, uint64_t n_ops, int kind,
                         unsigned long long* __restrict__ sink) {
    unsigned long long acc = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_ops;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t w = __umul64hi(mix64(i), words);
        if (kind == KC_BENCH_ATOMIC_ADD_U32) {
            (void)__hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(a + w), 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        } else if (kind == KC_BENCH_ATOMIC_OR_U64) {
            (void)__hip_atomic_fetch_or(a + w, 1ull << (i & 63), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (kind == KC_BENCH_ATOMIC_ADD_U32_WG) {
            (void)__hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(a + w), 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (kind == KC_BENCH_ATOMIC_ADD_U32_RET) {
            acc += __hip_atomic_fetch_add(reinterpret_cast<uint32_t*>(a + w), 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        } else {
            acc += a[w];
        }
    }
    if (acc == 0x5A5A5A5A5A5A5A5Aull) *sink = acc;  // keeps the loads and returned values
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_zero(uint64_t* __restrict__ a, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

}  // namespace

extern "C" int kc_bench_device(int32_t device, int kind, uint64_t array_bytes, uint64_t n_ops, int reps,
                               double* ms_out) {
    if (!ms_out || array_bytes < 64 || reps < 1) return -1;
    if (kind < KC_BENCH_ATOMIC_ADD_U32 || kind > KC_BENCH_ATOMIC_ADD_U32_RET) return -1;
    if (hipSetDevice(device) != hipSuccess) return -3;
    const uint64_t words = array_bytes / 8;
    const bool copy = kind == KC_BENCH_COPY;
    const uint64_t bytes = words * 8 * (copy ? 2 : 1);
    void* p = nullptr;
    unsigned long long* sink = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return -2;
    if (hipMalloc(&sink, 8) != hipSuccess) {
        (void)hipFree(p);
        return -2;
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess)
        rc = -3;
    double best = 1e300;
    if (!rc) {
        k_zero<<<16384, 256, 0, st>>>(reinterpret_cast<uint64_t*>(p), bytes / 8);
        for (int r = 0; r <= reps && !rc; ++r) {  // run 0 warms up
            if (hipEventRecord(e0, st) != hipSuccess) rc = -3;
            if (copy) {
                const uint64_t n16 = words * 8 / 16;
                k_copy<<<16384, 256, 0, st>>>(reinterpret_cast<const uint4*>(p),
                                             reinterpret_cast<uint4*>(reinterpret_cast<uint8_t*>(p) + words * 8), n16);
            } else {
                k_random<<<16384, 256, 0, st>>>(reinterpret_cast<uint64_t*>(p), words, n_ops, kind, sink);
            }
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess)
                rc = -3;
            float ms = 0;
            if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -3;
            if (!rc && r) best = std::min(best, (double)ms);
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(sink);
    (void)hipFree(p);
    if (!rc) *ms_out = best;
    return rc;
}
