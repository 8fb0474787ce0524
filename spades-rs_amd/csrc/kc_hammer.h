// kc_hammer.h — Hammer single-base correction (hammer.rs:34-120) on gfx950.
//
// One read per wave (a 64-thread workgroup).  The read's packed bases are staged in LDS a
// chunk of HAMMER_CHUNK bases at a time, next to one byte per window holding whether that
// window of the current (progressively corrected) read is Bloom-solid.  The reference's
// control flow is kept exactly, with the work of each step spread over the lanes:
//   - windows are visited in order; a solid window is skipped (the flags give the next
//     non-solid one by ballot);
//   - for a non-solid window i, the 3k candidates (offset o, base c != original) are
//     tested 64 at a time, candidate q = 3o + rank of c among the other bases in ACGT
//     order, so q follows the reference's loop order;
//   - for each candidate whose window i is solid, the windows covering the position are
//     re-tested with the base replaced (one lane each) and the base count is the popcount
//     of their flags: gain = (1 + solid) - base, saturating at 0;
//   - the first strictly greater gain in q order wins (hammer.rs:98-101), and the applied
//     base updates the flags of the windows that cover it, so later windows see the
//     corrected prefix.
// Every solidity test is is_solid_kmer256 against the frozen solid tier.  A read longer
// than the chunk is corrected with the chunk sliding along it: bases before the first
// window any later step can read are final, so they are written back and dropped.
#pragma once

#include "kc_device.h"

namespace kc {

constexpr uint32_t HAMMER_CHUNK = 2048;        // bases of a read held in LDS at a time (multiple of 16)
constexpr uint32_t HAMMER_CHUNK_WORDS = HAMMER_CHUNK / 16;
constexpr uint32_t HAMMER_PAD_WORDS = 10;      // over-read of extract_be<4> past the last base
static_assert(HAMMER_CHUNK >= 3 * 127 + 32, "a chunk must hold the 3k - 2 bases one step reads, plus alignment");

// Solid tier of one library as Hammer sees it.
struct LibBloom {
    const uint64_t* solid;
    Modulus M;
};

__device__ __forceinline__ uint32_t lds_get_base(const uint8_t* rb, uint32_t pos) {
    return (rb[pos >> 2] >> (6 - 2 * (pos & 3))) & 3;
}

// The value with the base whose low bit is at `bit` (even) XORed with x (2 bits).
__device__ __forceinline__ Kv kv_flip_base(Kv v, uint32_t bit, uint32_t x) {
    const uint64_t m = (uint64_t)x << (bit & 63);
    const uint32_t w = bit >> 6;
    if (w == 0) v.lo ^= m;
    else if (w == 1) v.hi ^= m;
    else if (w == 2) v.w2 ^= m;
    else v.w3 ^= m;
    return v;
}

template <int KW>
__device__ __forceinline__ Kv lds_window(const uint32_t* rd, uint32_t ls, uint32_t k) {
    return extract_be<KW>(rd + (ls >> 4), (2 * ls) & 31, k);
}

template <int KW>
__device__ __forceinline__ bool hm_solid(const Kv& v, uint32_t k, const uint64_t* __restrict__ solid, const Modulus& M) {
    return bloom_is_solid_lazy<KW>(canonical<KW>(v, k), k, solid, M);
}

// Corrects the reads active[0..n_active) of data into corr (corr must already hold a copy
// of data; only modified reads are written).  Every active read has length >= k.
// read_lib (nullable) gives each read's library; n_corrected[lib * corrected_stride]
// counts the reads modified (hammer.rs:123-143); n_queries (nullable) receives the number
// of is_solid tests made.  When n_active_dev is set, the number of active reads is read
// from it instead of n_active.  Launch with 64 threads per workgroup.
template <int KW>
__global__ void __launch_bounds__(64) k_hammer(const uint8_t* __restrict__ data, uint8_t* __restrict__ corr,
                                               const uint32_t* __restrict__ active, uint64_t n_active,
                                               const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ lengths,
                                               const uint32_t* __restrict__ read_lib, const LibBloom* __restrict__ libs,
                                               uint32_t k, unsigned long long* __restrict__ n_corrected,
                                               size_t corrected_stride, unsigned long long* __restrict__ n_queries,
                                               const unsigned long long* __restrict__ n_active_dev) {
    __shared__ uint32_t s_rd[HAMMER_CHUNK_WORDS + HAMMER_PAD_WORDS];
    __shared__ uint8_t s_ok[HAMMER_CHUNK];  // s_ok[s - b0]: window s of the current read is solid
    uint8_t* rb = reinterpret_cast<uint8_t*>(s_rd);
    const uint32_t lane = threadIdx.x;
    const uint32_t nc = 3 * k;  // candidates of one window
    uint64_t nq = 0;            // is_solid tests of this wave (wave-uniform)
    if (n_active_dev) n_active = *n_active_dev;
    for (uint64_t a = blockIdx.x; a < n_active; a += gridDim.x) {
        const uint32_t r = active[a];
        const uint32_t L = lengths[r];
        const uint32_t lib = read_lib ? read_lib[r] : 0;
        const uint64_t* __restrict__ solid = libs[lib].solid;
        const Modulus M = libs[lib].M;
        const uint64_t off = offsets[r];
        const uint32_t nwin = L - k + 1;
        uint32_t b0 = 0, b1 = 0, wend = 0;  // chunk = bases [b0, b1); flags of windows [b0, wend)
        bool modified = false;

        // Stage the chunk starting at base nb0 (a multiple of 16): bases kept from the
        // current chunk move down, the rest come from data (unmodified past b1), then the
        // flags of every window inside the chunk are computed.
        auto stage = [&](uint32_t nb0) {
            const uint32_t nb1 = min(L, nb0 + HAMMER_CHUNK);
            uint32_t kept = 0;  // bases kept from the current chunk (a multiple of 16 when < the chunk)
            __syncthreads();
            if (b1 > nb0 && nb0 > b0) {
                kept = b1 - nb0;
                const uint32_t delta = (nb0 - b0) >> 4, kw = (kept + 15) >> 4;
                uint32_t v0 = 0, v1 = 0;
                if (lane < kw) v0 = s_rd[delta + lane];
                if (lane + 64 < kw) v1 = s_rd[delta + lane + 64];
                __syncthreads();
                if (lane < kw) s_rd[lane] = v0;
                if (lane + 64 < kw) s_rd[lane + 64] = v1;
            }
            const uint32_t kwords = kept >> 4;  // whole kept words (kept < chunk only after a shift)
            for (uint32_t w = kwords + lane; w < HAMMER_CHUNK_WORDS + HAMMER_PAD_WORDS; w += 64) s_rd[w] = 0;
            __syncthreads();
            const uint32_t first_byte = (nb0 + kept) >> 2, end_byte = (nb1 + 3) >> 2;
            for (uint32_t j = first_byte + lane; j < end_byte; j += 64) rb[j - (nb0 >> 2)] = data[off + j];
            __syncthreads();
            b0 = nb0;
            b1 = nb1;
            wend = min(nwin, nb1 >= k ? nb1 - k + 1 : 0u);
            for (uint32_t s = b0 + lane; s < wend; s += 64)
                s_ok[s - b0] = hm_solid<KW>(lds_window<KW>(s_rd, s - b0, k), k, solid, M);
            nq += wend > b0 ? wend - b0 : 0u;
            __syncthreads();
        };
        // Bases [b0, upto) back to corr (upto a multiple of 4, or L).
        auto write_back = [&](uint32_t upto) {
            const uint32_t nb = (upto - b0 + 3) >> 2;
            for (uint32_t j = lane; j < nb; j += 64) corr[off + (b0 >> 2) + j] = rb[j];
        };

        stage(0);
        uint32_t i = 0;
        while (i < nwin) {
            if (b1 < L && i + 2 * k - 2 >= b1) {  // the step at i reads bases up to i + 2k - 2
                const uint32_t nb0 = (i >= k - 1 ? i - (k - 1) : 0u) & ~15u;
                if (modified) write_back(nb0);
                stage(nb0);
                continue;
            }
            // next non-solid window at or after i among the flagged ones
            uint32_t found = ~0u;
            for (uint32_t s0 = i; s0 < wend; s0 += 64) {
                const uint32_t s = s0 + lane;
                const uint64_t m = __ballot(s < wend && !s_ok[s - b0]);
                if (m) {
                    found = s0 + (uint32_t)__ffsll((unsigned long long)m) - 1;
                    break;
                }
            }
            if (found == ~0u) {
                i = wend;  // past the flagged windows: slide (or done at nwin)
                continue;
            }
            i = found;
            if (b1 < L && i + 2 * k - 2 >= b1) continue;

            const Kv vi = lds_window<KW>(s_rd, i - b0, k);
            uint32_t best_gain = 0, best_pos = 0, best_c = 0;
            for (uint32_t q0 = 0; q0 < nc; q0 += 64) {
                const uint32_t q = q0 + lane;
                bool pass = false;
                if (q < nc) {
                    const uint32_t o = q / 3, j = q - 3 * o;
                    const uint32_t orig = lds_get_base(rb, i + o - b0);
                    const uint32_t c = j + (j >= orig);
                    pass = hm_solid<KW>(kv_flip_base(vi, 2 * (k - 1 - o), orig ^ c), k, solid, M);
                }
                uint64_t m = __ballot(pass);
                nq += min(64u, nc - q0);
                while (m) {
                    const uint32_t qq = q0 + (uint32_t)__ffsll((unsigned long long)m) - 1;
                    m &= m - 1;
                    const uint32_t o = qq / 3, j = qq - 3 * o, pos = i + o;
                    const uint32_t orig = lds_get_base(rb, pos - b0);
                    const uint32_t c = j + (j >= orig);
                    const uint32_t start = pos >= k - 1 ? pos - (k - 1) : 0u, end = min(pos, nwin - 1);
                    uint32_t n_cand = 0, n_base = 0;
                    for (uint32_t s0 = start; s0 <= end; s0 += 64) {
                        const uint32_t s = s0 + lane;
                        bool sc = false, sb = false;
                        if (s <= end && s != i) {
                            sb = s_ok[s - b0] != 0;
                            const Kv v = kv_flip_base(lds_window<KW>(s_rd, s - b0, k), 2 * (k - 1 - (pos - s)), orig ^ c);
                            sc = hm_solid<KW>(v, k, solid, M);
                        }
                        n_cand += (uint32_t)__popcll(__ballot(sc));
                        n_base += (uint32_t)__popcll(__ballot(sb));
                    }
                    nq += end - start;  // [start, end] holds i
                    const uint32_t sc_all = 1 + n_cand;
                    const uint32_t gain = sc_all > n_base ? sc_all - n_base : 0u;
                    if (gain > best_gain) {
                        best_gain = gain;
                        best_pos = pos;
                        best_c = c;
                    }
                }
            }
            if (best_gain) {
                const uint32_t lp = best_pos - b0;
                if (lane == 0) {
                    const uint32_t sh = 6 - 2 * (lp & 3);
                    rb[lp >> 2] = (uint8_t)((rb[lp >> 2] & ~(3u << sh)) | (best_c << sh));
                }
                __syncthreads();
                const uint32_t s_lo = best_pos >= k - 1 ? best_pos - (k - 1) : 0u, s_hi = min(best_pos, nwin - 1);
                for (uint32_t s = s_lo + lane; s <= s_hi; s += 64)
                    s_ok[s - b0] = hm_solid<KW>(lds_window<KW>(s_rd, s - b0, k), k, solid, M);
                nq += s_hi - s_lo + 1;
                __syncthreads();
                modified = true;
            }
            ++i;
        }
        if (modified) {
            write_back(b1);
            if (lane == 0) atomicAdd(&n_corrected[(size_t)lib * corrected_stride], 1ull);
        }
    }
    if (n_queries && lane == 0 && nq) atomicAdd(n_queries, (unsigned long long)nq);
}

// Workgroups for k_hammer over n_active reads (one wave each, looping over reads).
inline uint32_t hammer_grid(uint64_t n_active) {
    return (uint32_t)(n_active < 65536 ? (n_active ? n_active : 1) : 65536);
}

}  // namespace kc
