// kc_hammer.h — Hammer single-base correction (hammer.rs:34-120) on gfx950.
//
// One read per lane.  The lane's read is copied packed into its own LDS slice; windows
// are re-extracted from LDS with the same bit extraction as the counting kernels, and
// every solidity test is is_solid_kmer256 against the frozen solid tier.  The control
// flow is the reference's: windows in order over the progressively corrected read; for
// a non-solid window every offset and every other base in ACGT order is scored by the
// number of solid windows covering the position, and the first strictly larger gain wins.
#pragma once

#include "kc_device.h"

namespace kc {

constexpr uint32_t HAMMER_MAX_LEN = 1024;
constexpr uint32_t HAMMER_LANE_WORDS = HAMMER_MAX_LEN / 16 + 6;  // packed read + over-read pad

// Solid tier of one library as Hammer sees it.
struct LibBloom {
    const uint64_t* solid;
    Modulus M;
};

template <int KW>
__device__ __forceinline__ bool lds_window_solid(const uint32_t* rd, uint32_t s, uint32_t k,
                                                 const uint64_t* __restrict__ solid, const Modulus& M) {
    const Kv v = extract_be<KW>(rd + (s >> 4), (2 * s) & 31, k);
    return bloom_is_solid<KW>(canonical<KW>(v, k), k, solid, M);
}

__device__ __forceinline__ uint32_t lds_get_base(const uint8_t* rb, uint32_t pos) {
    return (rb[pos >> 2] >> (6 - 2 * (pos & 3))) & 3;
}

__device__ __forceinline__ void lds_set_base(uint8_t* rb, uint32_t pos, uint32_t c) {
    const uint32_t sh = 6 - 2 * (pos & 3);
    rb[pos >> 2] = (uint8_t)((rb[pos >> 2] & ~(3u << sh)) | (c << sh));
}

// Corrects the reads active[0..n_active) of data into corr (corr must already hold a copy
// of data).  read_lib (nullable) gives each read's library; n_corrected[lib] counts the
// reads modified (stride corrected_stride); too_long is set for reads longer than HAMMER_MAX_LEN (left unchanged).
template <int KW>
__global__ void __launch_bounds__(64) k_hammer(const uint8_t* __restrict__ data, uint8_t* __restrict__ corr,
                                               const uint32_t* __restrict__ active, uint64_t n_active,
                                               const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ lengths,
                                               const uint32_t* __restrict__ read_lib, const LibBloom* __restrict__ libs,
                                               uint32_t k, unsigned long long* __restrict__ n_corrected,
                                               size_t corrected_stride, uint32_t* __restrict__ too_long) {
    __shared__ uint32_t lds[64 * HAMMER_LANE_WORDS];
    uint32_t* rd = lds + threadIdx.x * HAMMER_LANE_WORDS;
    uint8_t* rb = reinterpret_cast<uint8_t*>(rd);
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n_active;
         a += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = active[a];
        const uint32_t L = lengths[r];
        if (L > HAMMER_MAX_LEN) {
            atomicOr(too_long, 1u);
            continue;
        }
        const uint32_t lib = read_lib ? read_lib[r] : 0;
        const uint64_t* __restrict__ solid = libs[lib].solid;
        const Modulus M = libs[lib].M;
        const uint64_t off = offsets[r];
        const uint32_t nb = (L + 3) / 4;
        for (uint32_t j = 0; j < HAMMER_LANE_WORDS; ++j) rd[j] = 0;
        for (uint32_t j = 0; j < nb; ++j) rb[j] = data[off + j];
        bool modified = false;
        for (uint32_t i = 0; i + k <= L; ++i) {
            if (lds_window_solid<KW>(rd, i, k, solid, M)) continue;
            bool have = false;
            uint32_t best_off = 0, best_c = 0, max_gain = 0;
            for (uint32_t o = 0; o < k; ++o) {
                const uint32_t pos = i + o;
                const uint32_t orig = lds_get_base(rb, pos);
                const uint32_t start = pos >= k - 1 ? pos - (k - 1) : 0;
                const uint32_t end = min(pos, L - k);
                uint32_t base_solid = 0;
                for (uint32_t s = start; s <= end; ++s)
                    if (s != i && lds_window_solid<KW>(rd, s, k, solid, M)) ++base_solid;
                for (uint32_t c = 0; c < 4; ++c) {
                    if (c == orig) continue;
                    lds_set_base(rb, pos, c);
                    if (lds_window_solid<KW>(rd, i, k, solid, M)) {
                        uint32_t sc = 1;
                        for (uint32_t s = start; s <= end; ++s)
                            if (s != i && lds_window_solid<KW>(rd, s, k, solid, M)) ++sc;
                        const uint32_t gain = sc > base_solid ? sc - base_solid : 0;
                        if (gain > max_gain) {
                            max_gain = gain;
                            have = true;
                            best_off = o;
                            best_c = c;
                        }
                    }
                    lds_set_base(rb, pos, orig);
                }
            }
            if (have) {
                lds_set_base(rb, i + best_off, best_c);
                modified = true;
            }
        }
        if (modified) {
            for (uint32_t j = 0; j < nb; ++j) corr[off + j] = rb[j];
            atomicAdd(&n_corrected[(size_t)lib * corrected_stride], 1ull);
        }
    }
}

}  // namespace kc
