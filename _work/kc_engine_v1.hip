// kc_engine.hip — gfx950 k-mer spectrum engine behind include/kc.h.
//
// Stage 2 (assemble.rs:129-161) and Stage 3 (assemble.rs:166-292) of spades-rs on one
// MI355X.  The reference inserts every window into the two-tier filter in read order;
// this engine computes the same filter state from the distinct canonical k-mers:
//
//   * every occurrence is inserted into a device hash table that keeps, per canonical
//     k-mer, its occurrence count and its first event index t0 (global window number);
//   * fs[p] = min over (k-mer x, hash j) with p_j(x) = p of 4*t0(x) + j is the first
//     setter of Bloom bit p in the reference's sequential order (h1 -> h2 -> h3 within
//     an insert, bloom.rs:95-97);
//   * seen bit p is set iff fs[p] is finite (seen = OR over all occurrences);
//   * an insert of x at event t promotes iff each of its 3 seen bits was already set,
//     i.e. iff fs[p_j] < 4t + j for all j (bloom.rs:99).  For every occurrence after the
//     first this holds; for the first one it holds iff x is not the first setter of any
//     of its bits.  So the solid tier is the OR of the bits of every x with count >= 2
//     or with fs[p_j(x)] != 4*t0(x) + j for all j;
//   * the count table of assemble.rs:193-205 holds exactly the Bloom-solid distinct
//     k-mers with their multiplicity (the filter is frozen during Stage 3), so it is the
//     hash table filtered by is_solid (bloom.rs:111-126), then sorted.
//
// With error_correct, reads are first corrected against the frozen filter
// (hammer.rs:34-120, one read per lane) and the Bloom-solid windows of the corrected
// reads are counted in a second table pass.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "../../include/kc.h"
#include "kc_device.h"

using namespace kc;

namespace {

constexpr uint32_t TILE_WINDOWS = 1024;  // windows per work tile
constexpr uint32_t BLOCK = 256;
constexpr uint32_t HLL_P = 12;
constexpr uint32_t HLL_M = 1u << HLL_P;
constexpr uint32_t HIST_BINS = 1024;     // count histogram bins; the last bin holds >= 1023
constexpr uint32_t HAMMER_MAX_LEN = 1024;
constexpr uint32_t HAMMER_LANE_WORDS = HAMMER_MAX_LEN / 16 + 6;  // packed read + over-read pad
constexpr uint32_t DATA_PAD = 64;

// ---------------------------------------------------------------------------
// read bookkeeping
// ---------------------------------------------------------------------------

__global__ void k_nwin(const uint32_t* __restrict__ lengths, uint64_t n, uint32_t k, uint64_t* __restrict__ nwin,
                       uint8_t* __restrict__ flags, uint32_t* __restrict__ iota) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t L = lengths[r];
        const uint64_t w = L >= k ? (uint64_t)(L - k + 1) : 0;
        nwin[r] = w;
        flags[r] = w > 0;
        iota[r] = (uint32_t)r;
    }
}

// For every active read: its first event index and byte offset, and the first active read
// of every tile whose first window it holds.
__global__ void k_active(const uint32_t* __restrict__ active, uint64_t n_active, const uint64_t* __restrict__ prefix,
                         const uint64_t* __restrict__ nwin, const uint64_t* __restrict__ offsets,
                         uint64_t* __restrict__ act_ev, uint64_t* __restrict__ act_off, uint32_t* __restrict__ tile_first) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n_active;
         a += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = active[a];
        const uint64_t s = prefix[r], e = s + nwin[r];
        act_ev[a] = s;
        act_off[a] = offsets[r];
        for (uint64_t b = (s + TILE_WINDOWS - 1) / TILE_WINDOWS; b * TILE_WINDOWS < e; ++b) tile_first[b] = (uint32_t)a;
    }
}

// Iterates the windows of tile `tile`: calls f(v, event) with the forward k-mer value of
// each window and its global event index.  All threads of the block must call it.
template <bool WIDE, class F>
__device__ __forceinline__ void tile_windows(uint64_t tile, const uint8_t* __restrict__ data,
                                             const uint64_t* __restrict__ act_ev, const uint64_t* __restrict__ act_off,
                                             const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                             uint64_t n_occ, uint32_t k, uint64_t* s_ev, F&& f) {
    const uint64_t a0 = tile_first[tile];
    const uint64_t w0 = tile * TILE_WINDOWS;
    const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
    const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
    for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
    __syncthreads();
    for (uint64_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
        uint32_t lo = 0, hi = na - 1;  // largest i with s_ev[i] <= w
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_ev[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t s = (uint32_t)(w - s_ev[lo]);
        const Kv v = load_window<WIDE>(data, act_off[a0 + lo], s, k);
        f(v, w);
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// distinct-count estimate (HyperLogLog, 2^12 registers) to size the hash table
// ---------------------------------------------------------------------------

template <bool WIDE>
__global__ void __launch_bounds__(BLOCK) k_hll(const uint8_t* __restrict__ data, const uint64_t* __restrict__ act_ev,
                                               const uint64_t* __restrict__ act_off,
                                               const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                               uint64_t n_occ, uint64_t n_tiles, uint32_t k,
                                               uint32_t* __restrict__ regs) {
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    __shared__ uint32_t s_reg[HLL_M];
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) s_reg[i] = 0;
    __syncthreads();
    constexpr int NW = WIDE ? 2 : 1;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        tile_windows<WIDE>(tile, data, act_ev, act_off, tile_first, n_active, n_occ, k, s_ev,
                           [&](const Kv& v, uint64_t) {
                               uint64_t key[NW];
                               pack_key<NW>(canonical<WIDE>(v, k), key);
                               const uint64_t h = table_hash<NW>(key);
                               const uint32_t idx = (uint32_t)(h >> (64 - HLL_P));
                               const uint64_t w = h << HLL_P;
                               const uint32_t rho = w ? (uint32_t)__clzll((long long)w) + 1 : 64 - HLL_P + 1;
                               atomicMax(&s_reg[idx], rho);
                           });
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x)
        if (s_reg[i]) atomicMax(&regs[i], s_reg[i]);
}

// Sum of a per-thread value over the block, added to *total with one atomic per block.
__device__ __forceinline__ void block_add(uint64_t v, unsigned long long* total) {
    __shared__ unsigned long long s_part[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (uint32_t w = 0; w < blockDim.x / 64; ++w) t += s_part[w];
        if (t) atomicAdd(total, t);
    }
}

// ---------------------------------------------------------------------------
// table passes
// ---------------------------------------------------------------------------

template <int NW, typename Ev>
__global__ void k_table_clear(Slot<NW, Ev>* __restrict__ tab, uint64_t cap) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        Slot<NW, Ev> s;
#pragma unroll
        for (int w = 0; w < NW; ++w) s.k[w] = EMPTY_KEY;
        s.t0 = (Ev)~(Ev)0;
        s.cnt = 0;
        tab[i] = s;
    }
}

// Stage 2 insert (TRACK_T0, every window) or Stage 3 recount after Hammer (SOLID_ONLY:
// only windows whose canonical k-mer is Bloom-solid, assemble.rs:196).  Each thread
// takes WPT windows of the tile and issues their first bucket loads together, so a wave
// keeps WPT independent random table accesses in flight.
template <bool WIDE, typename Ev, bool TRACK_T0, bool SOLID_ONLY>
__global__ void __launch_bounds__(BLOCK) k_insert(const uint8_t* __restrict__ data, const uint64_t* __restrict__ act_ev,
                                                  const uint64_t* __restrict__ act_off,
                                                  const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                                  uint64_t n_occ, uint32_t k, Slot<WIDE ? 2 : 1, Ev>* __restrict__ tab,
                                                  uint64_t nbuckets, uint32_t max_probe,
                                                  const uint64_t* __restrict__ solid, Modulus M,
                                                  uint32_t* __restrict__ overflow) {
    constexpr int NW = WIDE ? 2 : 1;
    constexpr uint32_t WPT = TILE_WINDOWS / BLOCK;
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    const uint64_t tile = blockIdx.x;
    const uint64_t a0 = tile_first[tile];
    const uint64_t w0 = tile * TILE_WINDOWS;
    const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
    const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
    for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
    __syncthreads();
    uint64_t key[WPT][NW], bkt[WPT], pre[WPT][SPB];
    bool valid[WPT];
#pragma unroll
    for (uint32_t j = 0; j < WPT; ++j) {
        const uint64_t w = w0 + j * BLOCK + threadIdx.x;
        valid[j] = w < w1;
        if (!valid[j]) continue;
        uint32_t lo = 0, hi = na - 1;  // largest i with s_ev[i] <= w
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (s_ev[mid] <= w) lo = mid;
            else hi = mid - 1;
        }
        const Kv c = canonical<WIDE>(load_window<WIDE>(data, act_off[a0 + lo], (uint32_t)(w - s_ev[lo]), k), k);
        if (SOLID_ONLY && !bloom_is_solid(c, solid, M)) {
            valid[j] = false;
            continue;
        }
        pack_key<NW>(c, key[j]);
        bkt[j] = home_bucket<NW>(key[j], nbuckets);
    }
#pragma unroll
    for (uint32_t j = 0; j < WPT; ++j)
        if (valid[j]) bucket_load<NW, Ev>(tab, bkt[j], pre[j]);
#pragma unroll
    for (uint32_t j = 0; j < WPT; ++j) {
        if (!valid[j]) continue;
        const Ev t = (Ev)(w0 + j * BLOCK + threadIdx.x);
        if (!table_insert_from<NW, Ev, TRACK_T0>(tab, nbuckets, bkt[j], pre[j], key[j], t, max_probe))
            atomicOr(overflow, 1u);
    }
}

// fs[p] = min(4*t0 + j) over the distinct k-mers' bits; also counts occupied slots.
template <int NW, typename Ev>
__global__ void k_first_setter(const Slot<NW, Ev>* __restrict__ tab, uint64_t cap, Ev* __restrict__ fs, Modulus M,
                               unsigned long long* __restrict__ n_distinct) {
    uint64_t local = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        const Slot<NW, Ev> s = tab[i];
        if (s.k[0] == EMPTY_KEY) continue;
        ++local;
        const Kv v = unpack_key<NW>(s.k);
        uint64_t p[3];
        bloom_positions(v.lo, v.hi, 0, 0, M, p);
        const Ev base = (Ev)(s.t0 * 4);
#pragma unroll
        for (int j = 0; j < 3; ++j) atomicMin(&fs[p[j]], (Ev)(base + j));
    }
    block_add(local, n_distinct);
}

// seen[w] bit b = (fs[64w + b] is finite); clears the solid tier.  One wave per word.
template <typename Ev>
__global__ void k_seen_from_fs(const Ev* __restrict__ fs, uint64_t num_words, uint64_t* __restrict__ seen,
                               uint64_t* __restrict__ solid) {
    const uint64_t wave = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t lane = __lane_id();
    for (uint64_t w = wave; w < num_words; w += nwaves) {
        const Ev v = fs[w * 64 + lane];
        const uint64_t m = __ballot(v != (Ev)~(Ev)0);
        if (lane == 0) {
            seen[w] = m;
            solid[w] = 0;
        }
    }
}

// Solid tier: OR of the bits of every promoted distinct k-mer (bloom.rs:99-103).
template <int NW, typename Ev>
__global__ void k_promote(const Slot<NW, Ev>* __restrict__ tab, uint64_t cap, const Ev* __restrict__ fs,
                          uint64_t* __restrict__ solid, Modulus M) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
        const Slot<NW, Ev> s = tab[i];
        if (s.k[0] == EMPTY_KEY) continue;
        const Kv v = unpack_key<NW>(s.k);
        uint64_t p[3];
        bloom_positions(v.lo, v.hi, 0, 0, M, p);
        bool promote = s.cnt >= 2;
        if (!promote) {
            const Ev base = (Ev)(s.t0 * 4);
            promote = fs[p[0]] != base && fs[p[1]] != (Ev)(base + 1) && fs[p[2]] != (Ev)(base + 2);
        }
        if (promote) {
#pragma unroll
            for (int j = 0; j < 3; ++j)
                atomicOr(reinterpret_cast<unsigned long long*>(&solid[p[j] >> 6]), 1ull << (p[j] & 63));
        }
    }
}

// Gather the count table: occupied slots whose k-mer is Bloom-solid (or every occupied
// slot when the table was built from Bloom-solid windows only).  Each block takes chunks
// of BLOCK * COMPACT_ITEMS slots and reserves its output range with one atomic.
constexpr uint32_t COMPACT_ITEMS = 16;

template <int NW, typename Ev, bool CHECK_SOLID>
__global__ void __launch_bounds__(BLOCK) k_compact(const Slot<NW, Ev>* __restrict__ tab, uint64_t cap,
                                                   const uint64_t* __restrict__ solid, Modulus M,
                                                   uint64_t* __restrict__ out_a, uint64_t* __restrict__ out_b,
                                                   uint32_t* __restrict__ out_cnt,
                                                   unsigned long long* __restrict__ counter) {
    __shared__ uint32_t s_scan[BLOCK];
    __shared__ unsigned long long s_base;
    const uint64_t chunk = (uint64_t)BLOCK * COMPACT_ITEMS;
    const uint32_t tid = threadIdx.x;
    for (uint64_t c0 = blockIdx.x * chunk; c0 < cap; c0 += (uint64_t)gridDim.x * chunk) {
        uint32_t bits = 0, cnt = 0;
#pragma unroll 4
        for (uint32_t j = 0; j < COMPACT_ITEMS; ++j) {
            const uint64_t i = c0 + (uint64_t)j * BLOCK + tid;
            if (i >= cap) break;
            const Slot<NW, Ev> sl = tab[i];
            bool take = sl.k[0] != EMPTY_KEY;
            if (CHECK_SOLID && take) take = bloom_is_solid(unpack_key<NW>(sl.k), solid, M);
            bits |= (uint32_t)take << j;
            cnt += take;
        }
        s_scan[tid] = cnt;
        __syncthreads();
        for (uint32_t off = 1; off < BLOCK; off <<= 1) {
            const uint32_t v = tid >= off ? s_scan[tid - off] : 0;
            __syncthreads();
            s_scan[tid] += v;
            __syncthreads();
        }
        const uint32_t total = s_scan[BLOCK - 1];
        if (tid == 0) s_base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
        __syncthreads();
        uint64_t o = s_base + s_scan[tid] - cnt;
        while (bits) {
            const uint32_t j = __ffs(bits) - 1;
            bits &= bits - 1;
            const Slot<NW, Ev> sl = tab[c0 + (uint64_t)j * BLOCK + tid];
            out_a[o] = sl.k[0];
            if (NW == 2) out_b[o] = sl.k[NW - 1];
            out_cnt[o] = sl.cnt;
            ++o;
        }
        __syncthreads();
    }
}

__global__ void k_iota(uint32_t* __restrict__ x, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = (uint32_t)i;
}

template <typename T>
__global__ void k_gather(const T* __restrict__ src, const uint32_t* __restrict__ idx, T* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[idx[i]];
}

template <typename T>
__global__ void k_fill(T* __restrict__ x, uint64_t n, T v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// Histogram of min(count, HIST_BINS - 1) for determine_solid_kmer_cutoff (assemble.rs:692-735).
__global__ void __launch_bounds__(BLOCK) k_count_hist(const uint32_t* __restrict__ cnt, uint64_t n,
                                                      uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[HIST_BINS];
    for (uint32_t i = threadIdx.x; i < HIST_BINS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[min(cnt[i], HIST_BINS - 1)], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < HIST_BINS; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

__global__ void k_solid_flags(const uint32_t* __restrict__ cnt, uint64_t n, uint32_t cutoff, uint8_t* __restrict__ f,
                              unsigned long long* __restrict__ n_solid) {
    uint64_t local = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const bool s = cnt[i] >= cutoff;
        f[i] = s;
        local += s;
    }
    block_add(local, n_solid);
}

// ---------------------------------------------------------------------------
// Hammer (hammer.rs:34-120): one read per lane, the read held packed in LDS.
// ---------------------------------------------------------------------------

template <bool WIDE>
__device__ __forceinline__ bool lds_window_solid(const uint32_t* rd, uint32_t s, uint32_t k,
                                                 const uint64_t* __restrict__ solid, const Modulus& M) {
    const Kv v = extract_be<WIDE>(rd + (s >> 4), (2 * s) & 31, k);
    return bloom_is_solid(canonical<WIDE>(v, k), solid, M);
}

__device__ __forceinline__ uint32_t lds_get_base(const uint8_t* rb, uint32_t pos) {
    return (rb[pos >> 2] >> (6 - 2 * (pos & 3))) & 3;
}

__device__ __forceinline__ void lds_set_base(uint8_t* rb, uint32_t pos, uint32_t c) {
    const uint32_t sh = 6 - 2 * (pos & 3);
    rb[pos >> 2] = (uint8_t)((rb[pos >> 2] & ~(3u << sh)) | (c << sh));
}

template <bool WIDE>
__global__ void __launch_bounds__(64) k_hammer(const uint8_t* __restrict__ data, uint8_t* __restrict__ corr,
                                               const uint32_t* __restrict__ active, uint64_t n_active,
                                               const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ lengths,
                                               uint32_t k, const uint64_t* __restrict__ solid, Modulus M,
                                               unsigned long long* __restrict__ n_corrected, uint32_t* __restrict__ too_long) {
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
        const uint64_t off = offsets[r];
        const uint32_t nb = (L + 3) / 4;
        for (uint32_t j = 0; j < HAMMER_LANE_WORDS; ++j) rd[j] = 0;
        for (uint32_t j = 0; j < nb; ++j) rb[j] = data[off + j];
        bool modified = false;
        for (uint32_t i = 0; i + k <= L; ++i) {
            if (lds_window_solid<WIDE>(rd, i, k, solid, M)) continue;
            bool have = false;
            uint32_t best_off = 0, best_c = 0, max_gain = 0;
            for (uint32_t o = 0; o < k; ++o) {
                const uint32_t pos = i + o;
                const uint32_t orig = lds_get_base(rb, pos);
                const uint32_t start = pos >= k - 1 ? pos - (k - 1) : 0;
                const uint32_t end = min(pos, L - k);
                uint32_t base_solid = 0;
                for (uint32_t s = start; s <= end; ++s)
                    if (s != i && lds_window_solid<WIDE>(rd, s, k, solid, M)) ++base_solid;
                for (uint32_t c = 0; c < 4; ++c) {
                    if (c == orig) continue;
                    lds_set_base(rb, pos, c);
                    if (lds_window_solid<WIDE>(rd, i, k, solid, M)) {
                        uint32_t sc = 1;
                        for (uint32_t s = start; s <= end; ++s)
                            if (s != i && lds_window_solid<WIDE>(rd, s, k, solid, M)) ++sc;
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
            atomicAdd(n_corrected, 1ull);
        }
    }
}

// ---------------------------------------------------------------------------
// standalone filter: sequential insert semantics over a batch of events
// ---------------------------------------------------------------------------

__device__ __forceinline__ void key_positions(const uint64_t* __restrict__ keys4, uint64_t t, const Modulus& M,
                                              uint64_t p[3]) {
    const uint64_t* w = keys4 + 4 * t;
    bloom_positions(w[0], w[1], w[2], w[3], M, p);
}

__global__ void k_ev_first_setter(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, uint32_t* __restrict__ fs) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
#pragma unroll
        for (int j = 0; j < 3; ++j) atomicMin(&fs[p[j]], (uint32_t)(4 * t + j));
    }
}

// prev_j(t) = bit set before this batch, or first set by an earlier (event, hash) of it.
__global__ void k_ev_promote(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint32_t* __restrict__ fs,
                             const uint64_t* __restrict__ seen, uint8_t* __restrict__ prom) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
        bool all = true;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const bool before = (seen[p[j] >> 6] >> (p[j] & 63)) & 1;
            all = all && (before || fs[p[j]] < (uint32_t)(4 * t + j));
        }
        prom[t] = all;
    }
}

__global__ void k_ev_apply(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint8_t* __restrict__ prom,
                           uint64_t* __restrict__ seen, uint64_t* __restrict__ solid, uint32_t* __restrict__ fs) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            atomicOr(reinterpret_cast<unsigned long long*>(&seen[p[j] >> 6]), 1ull << (p[j] & 63));
            if (prom[t]) atomicOr(reinterpret_cast<unsigned long long*>(&solid[p[j] >> 6]), 1ull << (p[j] & 63));
            fs[p[j]] = ~0u;  // restore for the next batch
        }
    }
}

__global__ void k_query(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint64_t* __restrict__ solid,
                        uint8_t* __restrict__ out) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
        bool all = true;
#pragma unroll
        for (int j = 0; j < 3; ++j) all = all && ((solid[p[j] >> 6] >> (p[j] & 63)) & 1);
        out[t] = all;
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

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
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 256));
        if (e == hipSuccess) bytes = std::max<size_t>(n, 256);
        else p = nullptr;
        return e;
    }
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

Modulus make_modulus(uint64_t size_bits) {
    Modulus M{};
    M.d = size_bits;
    M.pow2 = (size_bits & (size_bits - 1)) == 0;
    M.mask = size_bits - 1;
    M.magic = M.pow2 ? 0 : (uint64_t)(((unsigned __int128)1 << 64) / size_bits);
    return M;
}

uint32_t grid_for(uint64_t work, uint32_t block = BLOCK, uint32_t cap = 16384) {
    uint64_t g = (work + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

enum KernelId {
    KID_NWIN, KID_SCAN, KID_SELECT, KID_ACTIVE, KID_HLL, KID_CLEAR, KID_INSERT, KID_FS_CLEAR, KID_FIRST_SETTER,
    KID_SEEN, KID_PROMOTE, KID_HAMMER, KID_RECOUNT, KID_COMPACT, KID_SORT, KID_HIST, KID_FLAGS, KID_COUNT
};
const char* const KERNEL_NAMES[KID_COUNT] = {
    "k_nwin", "scan_windows", "select_active", "k_active", "k_hll", "k_table_clear", "k_insert", "k_fs_clear",
    "k_first_setter", "k_seen_from_fs", "k_promote", "k_hammer", "k_recount", "k_compact", "radix_sort", "k_count_hist",
    "k_solid_flags"};

}  // namespace

#define KC_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail(_e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,    \
                                          std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// spectrum context
// ---------------------------------------------------------------------------

struct kc_ctx {
    kc_opts o{};
    hipStream_t st = nullptr;
    std::string err;
    // reads
    DevBuf d_data, d_off, d_len;
    size_t data_len = 0, n_reads = 0;
    // read bookkeeping
    DevBuf d_nwin, d_prefix, d_flags, d_iota, d_active, d_act_ev, d_act_off, d_tile_first;
    uint64_t n_occ = 0, n_active = 0, n_tiles = 0;
    // filter
    DevBuf d_seen, d_solid, d_fs;
    uint64_t num_words = 0;
    Modulus M{};
    // table
    DevBuf d_tab;
    uint64_t cap = 0;
    // count table outputs (sorted) + scratch
    DevBuf d_a, d_b, d_cnt, d_a2, d_b2, d_cnt2, d_idx, d_idx2, d_sflags;
    DevBuf d_corr, d_hll, d_hist, d_counters, d_temp, d_keys, d_out8;
    uint64_t n_counts = 0;
    bool wide = false, ev64 = false, built = false, counted = false;
    kc_info info{};
    // small pinned staging
    uint64_t* h_stage = nullptr;
    // timing
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    double k_ms[KID_COUNT] = {0};
    uint32_t k_launches[KID_COUNT] = {0};

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    bool timing() const { return (o.flags & KC_FLAG_TIMING) != 0; }
    hipEvent_t new_event() {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_pool.push_back(e);
        return e;
    }
    size_t mark_idx = 0;
    void mark_begin(int kid, hipEvent_t* b) {
        *b = nullptr;
        if (!timing()) return;
        if (mark_idx * 2 + 1 >= ev_pool.size()) {
            new_event();
            new_event();
        }
        *b = ev_pool[mark_idx * 2];
        (void)hipEventRecord(*b, st);
        (void)kid;
    }
    void mark_end(int kid, hipEvent_t b) {
        if (!timing() || !b) return;
        hipEvent_t e = ev_pool[mark_idx * 2 + 1];
        (void)hipEventRecord(e, st);
        marks.push_back({kid, {b, e}});
        ++mark_idx;
    }
    void reset_marks() {
        marks.clear();
        mark_idx = 0;
    }
    void collect_marks() {
        for (int i = 0; i < KID_COUNT; ++i) {
            k_ms[i] = 0;
            k_launches[i] = 0;
        }
        for (auto& m : marks) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, m.second.first, m.second.second) == hipSuccess) {
                k_ms[m.first] += ms;
                k_launches[m.first] += 1;
            }
        }
    }
};

#define TIMED(kid, launch)           \
    do {                             \
        hipEvent_t _b;               \
        mark_begin(kid, &_b);        \
        launch;                      \
        KC_TRY(hipGetLastError());   \
        mark_end(kid, _b);           \
    } while (0)

namespace {

struct Run {
    kc_ctx* c;
    explicit Run(kc_ctx* ctx) : c(ctx) {}
    int fail(int code, const std::string& m) { return c->fail(code, m); }
    void mark_begin(int kid, hipEvent_t* b) { c->mark_begin(kid, b); }
    void mark_end(int kid, hipEvent_t b) { c->mark_end(kid, b); }

    int prep() {
        hipStream_t st = c->st;
        const uint64_t n = c->n_reads;
        const uint32_t k = c->o.k;
        KC_TRY(c->d_nwin.ensure(n * 8 + 8));
        KC_TRY(c->d_prefix.ensure(n * 8 + 8));
        KC_TRY(c->d_flags.ensure(n + 1));
        KC_TRY(c->d_iota.ensure(n * 4 + 4));
        KC_TRY(c->d_active.ensure(n * 4 + 4));
        KC_TRY(c->d_counters.ensure(64 * 8));
        c->n_occ = c->n_active = c->n_tiles = 0;
        if (n == 0) return KC_OK;
        TIMED(KID_NWIN, (k_nwin<<<grid_for(n), BLOCK, 0, st>>>(c->d_len.as<uint32_t>(), n, k, c->d_nwin.as<uint64_t>(),
                                                                c->d_flags.as<uint8_t>(), c->d_iota.as<uint32_t>())));
        size_t tb = 0, tb2 = 0;
        KC_TRY(rocprim::exclusive_scan(nullptr, tb, c->d_nwin.as<uint64_t>(), c->d_prefix.as<uint64_t>(), (uint64_t)0,
                                       (size_t)n, rocprim::plus<uint64_t>(), st));
        KC_TRY(rocprim::select(nullptr, tb2, c->d_iota.as<uint32_t>(), c->d_flags.as<uint8_t>(),
                               c->d_active.as<uint32_t>(), c->d_counters.as<uint64_t>() + 8, (size_t)n, st));
        KC_TRY(c->d_temp.ensure(std::max(tb, tb2)));
        tb = c->d_temp.bytes;
        TIMED(KID_SCAN, KC_TRY(rocprim::exclusive_scan(c->d_temp.p, tb, c->d_nwin.as<uint64_t>(),
                                                        c->d_prefix.as<uint64_t>(), (uint64_t)0, (size_t)n,
                                                        rocprim::plus<uint64_t>(), st)));
        tb2 = c->d_temp.bytes;
        TIMED(KID_SELECT, KC_TRY(rocprim::select(c->d_temp.p, tb2, c->d_iota.as<uint32_t>(), c->d_flags.as<uint8_t>(),
                                                  c->d_active.as<uint32_t>(), c->d_counters.as<uint64_t>() + 8,
                                                  (size_t)n, st)));
        // n_occ = prefix[n-1] + nwin[n-1]; n_active = selected count
        KC_TRY(hipMemcpyAsync(c->h_stage, c->d_prefix.as<uint64_t>() + (n - 1), 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipMemcpyAsync(c->h_stage + 1, c->d_nwin.as<uint64_t>() + (n - 1), 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipMemcpyAsync(c->h_stage + 2, c->d_counters.as<uint64_t>() + 8, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        c->n_occ = c->h_stage[0] + c->h_stage[1];
        c->n_active = c->h_stage[2];
        c->n_tiles = (c->n_occ + TILE_WINDOWS - 1) / TILE_WINDOWS;
        if (c->n_occ == 0) return KC_OK;
        if (c->n_tiles > 0xFFFFFFFFull) return fail(KC_EUNSUPPORTED, "too many windows for one launch");
        KC_TRY(c->d_act_ev.ensure(c->n_active * 8 + 8));
        KC_TRY(c->d_act_off.ensure(c->n_active * 8 + 8));
        KC_TRY(c->d_tile_first.ensure(c->n_tiles * 4 + 4));
        TIMED(KID_ACTIVE, (k_active<<<grid_for(c->n_active), BLOCK, 0, st>>>(
                              c->d_active.as<uint32_t>(), c->n_active, c->d_prefix.as<uint64_t>(),
                              c->d_nwin.as<uint64_t>(), c->d_off.as<uint64_t>(), c->d_act_ev.as<uint64_t>(),
                              c->d_act_off.as<uint64_t>(), c->d_tile_first.as<uint32_t>())));
        return KC_OK;
    }

    template <bool WIDE>
    int estimate_distinct(const uint8_t* data, uint64_t* est) {
        hipStream_t st = c->st;
        KC_TRY(c->d_hll.ensure(HLL_M * 4));
        k_fill<uint32_t><<<grid_for(HLL_M), BLOCK, 0, st>>>(c->d_hll.as<uint32_t>(), HLL_M, 0u);
        KC_TRY(hipGetLastError());
        const uint32_t g = (uint32_t)std::min<uint64_t>(c->n_tiles, 2048);
        TIMED(KID_HLL, (k_hll<WIDE><<<g, BLOCK, 0, st>>>(data, c->d_act_ev.as<uint64_t>(), c->d_act_off.as<uint64_t>(),
                                                         c->d_tile_first.as<uint32_t>(), c->n_active, c->n_occ,
                                                         c->n_tiles, c->o.k, c->d_hll.as<uint32_t>())));
        std::vector<uint32_t> regs(HLL_M);
        KC_TRY(hipMemcpyAsync(regs.data(), c->d_hll.p, HLL_M * 4, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        double sum = 0;
        uint32_t zeros = 0;
        for (uint32_t r : regs) {
            sum += std::ldexp(1.0, -(int)r);
            zeros += r == 0;
        }
        const double m = HLL_M;
        double E = (0.7213 / (1.0 + 1.079 / m)) * m * m / sum;
        if (E <= 2.5 * m && zeros) E = m * std::log(m / zeros);
        *est = (uint64_t)std::ceil(E);
        return KC_OK;
    }

    template <int NW, typename Ev>
    int alloc_table(uint64_t distinct_est) {
        const double load = (c->o.table_load > 0.05f && c->o.table_load < 0.95f) ? c->o.table_load : 0.5;
        uint64_t cap = (uint64_t)std::ceil((double)distinct_est * 1.10 / load) + 1024;
        cap = std::min<uint64_t>(cap, c->n_occ * 2 + 1024);
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        cap = (cap + SPB - 1) / SPB * SPB;
        c->cap = cap;
        KC_TRY(c->d_tab.ensure(cap * sizeof(Slot<NW, Ev>)));
        return KC_OK;
    }

    template <bool WIDE, typename Ev>
    int bloom_build() {
        constexpr int NW = WIDE ? 2 : 1;
        hipStream_t st = c->st;
        uint64_t est = 0;
        int rc = estimate_distinct<WIDE>(c->d_data.as<uint8_t>(), &est);
        if (rc) return rc;
        for (int attempt = 0;; ++attempt) {
            rc = alloc_table<NW, Ev>(est << attempt);
            if (rc) return rc;
            TIMED(KID_CLEAR, (k_table_clear<NW, Ev><<<grid_for(c->cap), BLOCK, 0, st>>>(
                                 c->d_tab.as<Slot<NW, Ev>>(), c->cap)));
            k_fill<unsigned long long><<<1, 64, 0, st>>>(c->d_counters.as<unsigned long long>(), 16, 0ull);
            KC_TRY(hipGetLastError());
            TIMED(KID_INSERT, (k_insert<WIDE, Ev, true, false><<<(uint32_t)c->n_tiles, BLOCK, 0, st>>>(
                                  c->d_data.as<uint8_t>(), c->d_act_ev.as<uint64_t>(), c->d_act_off.as<uint64_t>(),
                                  c->d_tile_first.as<uint32_t>(), c->n_active, c->n_occ, c->o.k,
                                  c->d_tab.as<Slot<NW, Ev>>(), c->cap / slots_per_bucket<NW, Ev>(),
                                  (uint32_t)std::min<uint64_t>(c->cap / slots_per_bucket<NW, Ev>(), 1u << 20),
                                  nullptr, c->M, c->d_counters.as<uint32_t>())));
            KC_TRY(hipMemcpyAsync(c->h_stage, c->d_counters.p, 8, hipMemcpyDeviceToHost, st));
            KC_TRY(hipStreamSynchronize(st));
            if ((c->h_stage[0] & 0xFFFFFFFFull) == 0) break;
            if (attempt >= 4) return fail(KC_ENOMEM, "device hash table overflow");
        }
        // first setters -> seen tier -> solid tier
        const uint64_t size_bits = c->num_words * 64;
        KC_TRY(c->d_fs.ensure(size_bits * sizeof(Ev)));
        TIMED(KID_FS_CLEAR, (k_fill<Ev><<<grid_for(size_bits), BLOCK, 0, st>>>(c->d_fs.as<Ev>(), size_bits, (Ev)~(Ev)0)));
        TIMED(KID_FIRST_SETTER, (k_first_setter<NW, Ev><<<grid_for(c->cap), BLOCK, 0, st>>>(
                                    c->d_tab.as<Slot<NW, Ev>>(), c->cap, c->d_fs.as<Ev>(), c->M,
                                    c->d_counters.as<unsigned long long>() + 1)));
        TIMED(KID_SEEN, (k_seen_from_fs<Ev><<<grid_for(c->num_words * 64), BLOCK, 0, st>>>(
                            c->d_fs.as<Ev>(), c->num_words, c->d_seen.as<uint64_t>(), c->d_solid.as<uint64_t>())));
        TIMED(KID_PROMOTE, (k_promote<NW, Ev><<<grid_for(c->cap), BLOCK, 0, st>>>(
                               c->d_tab.as<Slot<NW, Ev>>(), c->cap, c->d_fs.as<Ev>(), c->d_solid.as<uint64_t>(), c->M)));
        KC_TRY(hipMemcpyAsync(c->h_stage, c->d_counters.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        c->info.n_distinct = c->h_stage[0];
        c->info.table_capacity = c->cap;
        return KC_OK;
    }

    // Sort (a[, b], cnt) ascending by key; a/b are the table key words.
    template <int NW>
    int sort_counts(uint64_t n) {
        hipStream_t st = c->st;
        if (n == 0) return KC_OK;
        const uint32_t k = c->o.k;
        size_t tb = 0;
        if (NW == 1) {
            KC_TRY(rocprim::radix_sort_pairs(nullptr, tb, c->d_a.as<uint64_t>(), c->d_a2.as<uint64_t>(),
                                             c->d_cnt.as<uint32_t>(), c->d_cnt2.as<uint32_t>(), (size_t)n, 0u, 2 * k, st));
            KC_TRY(c->d_temp.ensure(tb));
            tb = c->d_temp.bytes;
            TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_a.as<uint64_t>(),
                                                              c->d_a2.as<uint64_t>(), c->d_cnt.as<uint32_t>(),
                                                              c->d_cnt2.as<uint32_t>(), (size_t)n, 0u, 2 * k, st)));
            std::swap(c->d_a, c->d_a2);
            std::swap(c->d_cnt, c->d_cnt2);
            return KC_OK;
        }
        // LSD over the two words.  a, b, cnt are the unsorted columns and stay intact
        // until the final gathers.
        //   1. (b, iota) sorted by b        -> perm p1 in d_idx2 (sorted b discarded)
        //   2. a1 = a[p1]                    -> d_a2
        //   3. (a1, p1) stably sorted by a1  -> sorted a in d_b2, perm p2 in d_idx
        //   4. b[p2] -> d_a2, cnt[p2] -> d_cnt2
        KC_TRY(c->d_idx.ensure(n * 4));
        KC_TRY(c->d_idx2.ensure(n * 4));
        k_iota<<<grid_for(n), BLOCK, 0, st>>>(c->d_idx.as<uint32_t>(), n);
        KC_TRY(hipGetLastError());
        const uint32_t abits = 2 * k - 63;
        size_t need1 = 0, need2 = 0;
        KC_TRY(rocprim::radix_sort_pairs(nullptr, need1, c->d_b.as<uint64_t>(), c->d_b2.as<uint64_t>(),
                                         c->d_idx.as<uint32_t>(), c->d_idx2.as<uint32_t>(), (size_t)n, 0u, 63u, st));
        KC_TRY(rocprim::radix_sort_pairs(nullptr, need2, c->d_a2.as<uint64_t>(), c->d_b2.as<uint64_t>(),
                                         c->d_idx2.as<uint32_t>(), c->d_idx.as<uint32_t>(), (size_t)n, 0u, abits, st));
        KC_TRY(c->d_temp.ensure(std::max(need1, need2)));
        tb = c->d_temp.bytes;
        TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_b.as<uint64_t>(),
                                                          c->d_b2.as<uint64_t>(), c->d_idx.as<uint32_t>(),
                                                          c->d_idx2.as<uint32_t>(), (size_t)n, 0u, 63u, st)));
        TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_a.as<uint64_t>(),
                                                                           c->d_idx2.as<uint32_t>(),
                                                                           c->d_a2.as<uint64_t>(), n)));
        tb = c->d_temp.bytes;
        TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_a2.as<uint64_t>(),
                                                          c->d_b2.as<uint64_t>(), c->d_idx2.as<uint32_t>(),
                                                          c->d_idx.as<uint32_t>(), (size_t)n, 0u, abits, st)));
        TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_b.as<uint64_t>(),
                                                                           c->d_idx.as<uint32_t>(),
                                                                           c->d_a2.as<uint64_t>(), n)));
        TIMED(KID_SORT, (k_gather<uint32_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_cnt.as<uint32_t>(),
                                                                           c->d_idx.as<uint32_t>(),
                                                                           c->d_cnt2.as<uint32_t>(), n)));
        std::swap(c->d_a, c->d_b2);
        std::swap(c->d_b, c->d_a2);
        std::swap(c->d_cnt, c->d_cnt2);
        return KC_OK;
    }

    // Stage 3 after the Bloom build: optional Hammer + recount, compaction, sort, cutoff.
    template <bool WIDE, typename Ev>
    int count() {
        constexpr int NW = WIDE ? 2 : 1;
        hipStream_t st = c->st;
        unsigned long long* ctr = c->d_counters.as<unsigned long long>();
        k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr, 16, 0ull);
        KC_TRY(hipGetLastError());
        c->info.corrected_reads = 0;
        const bool ec = c->o.error_correct != 0;
        if (ec) {
            // corr = data; Hammer writes the reads it modifies
            KC_TRY(c->d_corr.ensure(c->d_data.bytes));
            KC_TRY(hipMemcpyAsync(c->d_corr.p, c->d_data.p, c->d_data.bytes, hipMemcpyDeviceToDevice, st));
            if (c->n_active) {
                TIMED(KID_HAMMER, (k_hammer<WIDE><<<grid_for(c->n_active, 64, 65536), 64, 0, st>>>(
                                      c->d_data.as<uint8_t>(), c->d_corr.as<uint8_t>(), c->d_active.as<uint32_t>(),
                                      c->n_active, c->d_off.as<uint64_t>(), c->d_len.as<uint32_t>(), c->o.k,
                                      c->d_solid.as<uint64_t>(), c->M, ctr + 2, c->d_counters.as<uint32_t>() + 6)));
            }
            // recount Bloom-solid windows of the corrected reads into a cleared table
            for (int attempt = 0;; ++attempt) {
                if (attempt) {
                    const uint64_t cap2 = c->cap * 2;  // a multiple of every bucket size
                    KC_TRY(c->d_tab.ensure(cap2 * sizeof(Slot<NW, uint32_t>)));
                    c->cap = cap2;
                }
                TIMED(KID_CLEAR, (k_table_clear<NW, uint32_t><<<grid_for(c->cap), BLOCK, 0, st>>>(
                                     c->d_tab.as<Slot<NW, uint32_t>>(), c->cap)));
                k_fill<uint32_t><<<1, 64, 0, st>>>(c->d_counters.as<uint32_t>(), 1, 0u);
                KC_TRY(hipGetLastError());
                if (c->n_tiles) {
                    TIMED(KID_RECOUNT, (k_insert<WIDE, uint32_t, false, true><<<(uint32_t)c->n_tiles, BLOCK, 0, st>>>(
                                           c->d_corr.as<uint8_t>(), c->d_act_ev.as<uint64_t>(),
                                           c->d_act_off.as<uint64_t>(), c->d_tile_first.as<uint32_t>(), c->n_active,
                                           c->n_occ, c->o.k, c->d_tab.as<Slot<NW, uint32_t>>(),
                                           c->cap / slots_per_bucket<NW, uint32_t>(),
                                           (uint32_t)std::min<uint64_t>(c->cap / slots_per_bucket<NW, uint32_t>(), 1u << 20),
                                           c->d_solid.as<uint64_t>(),
                                           c->M, c->d_counters.as<uint32_t>())));
                }
                KC_TRY(hipMemcpyAsync(c->h_stage, c->d_counters.p, 32, hipMemcpyDeviceToHost, st));
                KC_TRY(hipStreamSynchronize(st));
                if ((c->h_stage[0] & 0xFFFFFFFFull) == 0) break;
                if (attempt >= 4) return fail(KC_ENOMEM, "device hash table overflow (recount)");
            }
            if (c->h_stage[3] & 0xFFFFFFFFull)
                return fail(KC_EUNSUPPORTED, "error_correct supports reads of at most 1024 bases");
            c->info.corrected_reads = c->h_stage[2];
        }
        const uint64_t cap = c->cap;
        KC_TRY(c->d_a.ensure(cap * 8));
        KC_TRY(c->d_b.ensure(cap * 8));
        KC_TRY(c->d_cnt.ensure(cap * 4));
        KC_TRY(c->d_a2.ensure(cap * 8));
        KC_TRY(c->d_b2.ensure(cap * 8));
        KC_TRY(c->d_cnt2.ensure(cap * 4));
        if (ec) {
            TIMED(KID_COMPACT, (k_compact<NW, uint32_t, false><<<grid_for(cap), BLOCK, 0, st>>>(
                                   c->d_tab.as<Slot<NW, uint32_t>>(), cap, c->d_solid.as<uint64_t>(), c->M,
                                   c->d_a.as<uint64_t>(), c->d_b.as<uint64_t>(), c->d_cnt.as<uint32_t>(), ctr + 4)));
        } else {
            TIMED(KID_COMPACT, (k_compact<NW, Ev, true><<<grid_for(cap), BLOCK, 0, st>>>(
                                   c->d_tab.as<Slot<NW, Ev>>(), cap, c->d_solid.as<uint64_t>(), c->M,
                                   c->d_a.as<uint64_t>(), c->d_b.as<uint64_t>(), c->d_cnt.as<uint32_t>(), ctr + 4)));
        }
        KC_TRY(hipMemcpyAsync(c->h_stage, ctr + 4, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        const uint64_t n = c->h_stage[0];
        c->n_counts = n;
        int rc = sort_counts<NW>(n);
        if (rc) return rc;
        // cutoff (assemble.rs:692-735) from the count histogram
        KC_TRY(c->d_hist.ensure(HIST_BINS * 4));
        k_fill<uint32_t><<<grid_for(HIST_BINS), BLOCK, 0, st>>>(c->d_hist.as<uint32_t>(), HIST_BINS, 0u);
        KC_TRY(hipGetLastError());
        if (n) TIMED(KID_HIST, (k_count_hist<<<grid_for(n, BLOCK, 1024), BLOCK, 0, st>>>(c->d_cnt.as<uint32_t>(), n,
                                                                                        c->d_hist.as<uint32_t>())));
        std::vector<uint32_t> hist(HIST_BINS);
        KC_TRY(hipMemcpyAsync(hist.data(), c->d_hist.p, HIST_BINS * 4, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        uint32_t top = 0;
        const uint32_t cutoff = cutoff_from_hist(hist, n, c->o.min_coverage, &top);
        c->info.cutoff = cutoff;
        c->info.top_cov = top;
        KC_TRY(c->d_sflags.ensure(n + 1));
        k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr + 5, 1, 0ull);
        KC_TRY(hipGetLastError());
        if (n) TIMED(KID_FLAGS, (k_solid_flags<<<grid_for(n), BLOCK, 0, st>>>(c->d_cnt.as<uint32_t>(), n, cutoff,
                                                                            c->d_sflags.as<uint8_t>(), ctr + 5)));
        KC_TRY(hipMemcpyAsync(c->h_stage, ctr + 5, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        c->info.n_solid = c->h_stage[0];
        c->info.n_counts = n;
        return KC_OK;
    }

    // determine_solid_kmer_cutoff (assemble.rs:692-735) evaluated on a histogram of
    // min(count, HIST_BINS-1).  The 95th-percentile count `top` only enters through
    // top >= 30 and clamp(top/4, 10, 35), which the capped histogram decides exactly.
    static uint32_t cutoff_from_hist(const std::vector<uint32_t>& hist, uint64_t n, double min_cov, uint32_t* top_out) {
        *top_out = 0;
        if (min_cov > 5.0) return (uint32_t)std::max(min_cov * 0.35, 2.0);
        if (n <= 100) return 2;
        const uint64_t rank = n * 95 / 100;  // index into the ascending counts
        uint64_t cum = 0;
        uint32_t top = HIST_BINS - 1;
        for (uint32_t b = 0; b < HIST_BINS; ++b) {
            cum += hist[b];
            if (cum > rank) {
                top = b;
                break;
            }
        }
        *top_out = top;
        if (top < 30) return 2;
        const uint32_t mvs = std::min<uint32_t>(std::max<uint32_t>(top / 4, 10), 35);
        uint32_t valley = 2;
        uint64_t min_val = ~0ull;
        bool found = false;
        for (uint32_t cc = 2; cc <= mvs; ++cc) {
            const uint64_t h = hist[cc];
            if (h <= min_val) {
                min_val = h;
                valley = cc;
            } else if (h > min_val * 2 && cc > valley + 2) {
                found = true;
                break;
            }
        }
        return found ? valley : 2;
    }
};

}  // namespace

// ---------------------------------------------------------------------------
// C ABI: spectrum context
// ---------------------------------------------------------------------------

#define CTX_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t _e = (expr);                                                                     \
        if (_e != hipSuccess) return ctx->fail(_e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,     \
                                               std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

extern "C" {

int kc_create(kc_ctx** out, const kc_opts* opts) {
    if (!out || !opts) return KC_EINVAL;
    *out = nullptr;
    if (opts->k < 1 || opts->k > KC_MAX_K) return KC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_EHIP;
    if (opts->device < 0 || opts->device >= ndev) return KC_EINVAL;
    if (hipSetDevice(opts->device) != hipSuccess) return KC_EHIP;
    kc_ctx* ctx = new kc_ctx();
    ctx->o = *opts;
    if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_stage, 64 * sizeof(uint64_t), hipHostMallocDefault) != hipSuccess) {
        delete ctx;
        return KC_EHIP;
    }
    ctx->num_words = std::max<uint64_t>((opts->bloom_bits + 63) / 64, 1);
    ctx->M = make_modulus(ctx->num_words * 64);
    ctx->wide = opts->k > 32;
    *out = ctx;
    return KC_OK;
}

void kc_destroy(kc_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->o.device);
    (void)hipStreamSynchronize(ctx->st);
    DevBuf* bufs[] = {&ctx->d_data, &ctx->d_off, &ctx->d_len, &ctx->d_nwin, &ctx->d_prefix, &ctx->d_flags,
                      &ctx->d_iota, &ctx->d_active, &ctx->d_act_ev, &ctx->d_act_off, &ctx->d_tile_first,
                      &ctx->d_seen, &ctx->d_solid, &ctx->d_fs, &ctx->d_tab, &ctx->d_a, &ctx->d_b, &ctx->d_cnt,
                      &ctx->d_a2, &ctx->d_b2, &ctx->d_cnt2, &ctx->d_idx, &ctx->d_idx2, &ctx->d_sflags,
                      &ctx->d_corr, &ctx->d_hll, &ctx->d_hist, &ctx->d_counters, &ctx->d_temp, &ctx->d_keys,
                      &ctx->d_out8};
    for (DevBuf* b : bufs) b->release();
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

const char* kc_last_error(const kc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kc_reads_upload(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                    const uint32_t* lengths, size_t n_reads) {
    if (!ctx) return KC_EINVAL;
    if (n_reads && (!offsets || !lengths)) return ctx->fail(KC_EINVAL, "null offsets/lengths");
    if (data_len && !data) return ctx->fail(KC_EINVAL, "null data");
    if (n_reads > 0xFFFFFFFFull) return ctx->fail(KC_EUNSUPPORTED, "more than 2^32-1 reads in one context");
    for (size_t i = 0; i < n_reads; ++i) {
        const uint64_t nb = ((uint64_t)lengths[i] + 3) / 4;
        if (offsets[i] > data_len || nb > data_len - offsets[i])
            return ctx->fail(KC_EINVAL, "read " + std::to_string(i) + " extends past data_len");
    }
    CTX_TRY(hipSetDevice(ctx->o.device));
    CTX_TRY(ctx->d_data.ensure(data_len + DATA_PAD));
    CTX_TRY(ctx->d_off.ensure(n_reads * 8 + 8));
    CTX_TRY(ctx->d_len.ensure(n_reads * 4 + 4));
    CTX_TRY(hipMemsetAsync(ctx->d_data.p, 0, ctx->d_data.bytes, ctx->st));
    if (data_len) CTX_TRY(hipMemcpyAsync(ctx->d_data.p, data, data_len, hipMemcpyHostToDevice, ctx->st));
    if (n_reads) {
        CTX_TRY(hipMemcpyAsync(ctx->d_off.p, offsets, n_reads * 8, hipMemcpyHostToDevice, ctx->st));
        CTX_TRY(hipMemcpyAsync(ctx->d_len.p, lengths, n_reads * 4, hipMemcpyHostToDevice, ctx->st));
    }
    CTX_TRY(hipStreamSynchronize(ctx->st));
    ctx->data_len = data_len;
    ctx->n_reads = n_reads;
    ctx->built = ctx->counted = false;
    return KC_OK;
}

int kc_bloom_build(kc_ctx* ctx) {
    if (!ctx) return KC_EINVAL;
    CTX_TRY(hipSetDevice(ctx->o.device));
    ctx->reset_marks();
    ctx->built = ctx->counted = false;
    ctx->info = kc_info{};
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    CTX_TRY(ctx->d_seen.ensure(ctx->num_words * 8));
    CTX_TRY(ctx->d_solid.ensure(ctx->num_words * 8));
    Run run(ctx);
    int rc = run.prep();
    if (rc) return rc;
    if (ctx->n_occ == 0) {
        k_fill<uint64_t><<<grid_for(ctx->num_words), BLOCK, 0, ctx->st>>>(ctx->d_seen.as<uint64_t>(), ctx->num_words, 0ull);
        k_fill<uint64_t><<<grid_for(ctx->num_words), BLOCK, 0, ctx->st>>>(ctx->d_solid.as<uint64_t>(), ctx->num_words, 0ull);
        CTX_TRY(hipGetLastError());
        ctx->cap = 0;
    } else {
        ctx->ev64 = ctx->n_occ >= (1ull << 30) - 1;
        if (!ctx->wide) rc = ctx->ev64 ? run.bloom_build<false, uint64_t>() : run.bloom_build<false, uint32_t>();
        else rc = ctx->ev64 ? run.bloom_build<true, uint64_t>() : run.bloom_build<true, uint32_t>();
        if (rc) return rc;
    }
    CTX_TRY(hipEventRecord(e1, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0;
    CTX_TRY(hipEventElapsedTime(&ms, e0, e1));
    ctx->info.ms_stage2 = ms;
    ctx->info.n_reads = ctx->n_reads;
    ctx->info.n_occ = ctx->n_occ;
    ctx->info.num_words = ctx->num_words;
    ctx->built = true;
    return KC_OK;
}

int kc_count(kc_ctx* ctx, kc_info* info) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "kc_count before kc_bloom_build");
    CTX_TRY(hipSetDevice(ctx->o.device));
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    Run run(ctx);
    int rc = KC_OK;
    if (ctx->n_occ == 0) {
        ctx->n_counts = 0;
        std::vector<uint32_t> hist(HIST_BINS, 0);
        uint32_t top = 0;
        ctx->info.cutoff = Run::cutoff_from_hist(hist, 0, ctx->o.min_coverage, &top);
        ctx->info.top_cov = top;
        ctx->info.n_counts = ctx->info.n_solid = 0;
        if (ctx->o.error_correct) {
            CTX_TRY(ctx->d_corr.ensure(ctx->d_data.bytes));
            CTX_TRY(hipMemcpyAsync(ctx->d_corr.p, ctx->d_data.p, ctx->d_data.bytes, hipMemcpyDeviceToDevice, ctx->st));
        }
    } else if (!ctx->wide) {
        rc = ctx->ev64 ? run.count<false, uint64_t>() : run.count<false, uint32_t>();
    } else {
        rc = ctx->ev64 ? run.count<true, uint64_t>() : run.count<true, uint32_t>();
    }
    if (rc) return rc;
    CTX_TRY(hipEventRecord(e1, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0;
    CTX_TRY(hipEventElapsedTime(&ms, e0, e1));
    ctx->info.ms_stage3 = ms;
    ctx->counted = true;
    if (ctx->timing()) ctx->collect_marks();
    if (info) *info = ctx->info;
    return KC_OK;
}

int kc_run(kc_ctx* ctx, kc_info* info) {
    int rc = kc_bloom_build(ctx);
    if (rc) return rc;
    return kc_count(ctx, info);
}

int kc_info_get(const kc_ctx* ctx, kc_info* info) {
    if (!ctx || !info) return KC_EINVAL;
    *info = ctx->info;
    return KC_OK;
}

int kc_bloom_export(kc_ctx* ctx, uint64_t* seen, uint64_t* solid, size_t num_words) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "kc_bloom_export before kc_bloom_build");
    if (num_words != ctx->num_words) return ctx->fail(KC_EINVAL, "num_words mismatch");
    CTX_TRY(hipSetDevice(ctx->o.device));
    if (seen) CTX_TRY(hipMemcpyAsync(seen, ctx->d_seen.p, num_words * 8, hipMemcpyDeviceToHost, ctx->st));
    if (solid) CTX_TRY(hipMemcpyAsync(solid, ctx->d_solid.p, num_words * 8, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_bloom_query(kc_ctx* ctx, const uint64_t* keys4, size_t n, uint8_t* out) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "kc_bloom_query before kc_bloom_build");
    if (n == 0) return KC_OK;
    if (!keys4 || !out) return ctx->fail(KC_EINVAL, "null keys/out");
    CTX_TRY(hipSetDevice(ctx->o.device));
    CTX_TRY(ctx->d_keys.ensure(n * 32));
    CTX_TRY(ctx->d_out8.ensure(n));
    CTX_TRY(hipMemcpyAsync(ctx->d_keys.p, keys4, n * 32, hipMemcpyHostToDevice, ctx->st));
    k_query<<<grid_for(n), BLOCK, 0, ctx->st>>>(ctx->d_keys.as<uint64_t>(), n, ctx->M, ctx->d_solid.as<uint64_t>(),
                                                ctx->d_out8.as<uint8_t>());
    CTX_TRY(hipGetLastError());
    CTX_TRY(hipMemcpyAsync(out, ctx->d_out8.p, n, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_counts_export(kc_ctx* ctx, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask, size_t n) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "kc_counts_export before kc_count");
    if (n != ctx->n_counts) return ctx->fail(KC_EINVAL, "n mismatch");
    if (n == 0) return KC_OK;
    CTX_TRY(hipSetDevice(ctx->o.device));
    if (counts) CTX_TRY(hipMemcpyAsync(counts, ctx->d_cnt.p, n * 4, hipMemcpyDeviceToHost, ctx->st));
    if (solid_mask) CTX_TRY(hipMemcpyAsync(solid_mask, ctx->d_sflags.p, n, hipMemcpyDeviceToHost, ctx->st));
    std::vector<uint64_t> a, b;
    if (keys4) {
        a.resize(n);
        CTX_TRY(hipMemcpyAsync(a.data(), ctx->d_a.p, n * 8, hipMemcpyDeviceToHost, ctx->st));
        if (ctx->wide) {
            b.resize(n);
            CTX_TRY(hipMemcpyAsync(b.data(), ctx->d_b.p, n * 8, hipMemcpyDeviceToHost, ctx->st));
        }
    }
    CTX_TRY(hipStreamSynchronize(ctx->st));
    if (keys4) {
        for (size_t i = 0; i < n; ++i) {
            uint64_t* w = keys4 + 4 * i;
            if (!ctx->wide) {
                w[0] = a[i];
                w[1] = 0;
            } else {  // V = A * 2^63 + B
                w[0] = b[i] | (a[i] << 63);
                w[1] = a[i] >> 1;
            }
            w[2] = w[3] = 0;
        }
    }
    return KC_OK;
}

int kc_reads_export(kc_ctx* ctx, uint8_t* data, size_t data_len) {
    if (!ctx) return KC_EINVAL;
    if (data_len != ctx->data_len) return ctx->fail(KC_EINVAL, "data_len mismatch");
    if (data_len == 0) return KC_OK;
    if (!data) return ctx->fail(KC_EINVAL, "null data");
    CTX_TRY(hipSetDevice(ctx->o.device));
    const void* src = (ctx->o.error_correct && ctx->counted) ? ctx->d_corr.p : ctx->d_data.p;
    CTX_TRY(hipMemcpyAsync(data, src, data_len, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_kernel_times(const kc_ctx* ctx, const char** names, double* ms, uint32_t* launches, int cap) {
    if (!ctx) return KC_EINVAL;
    int n = 0;
    for (int i = 0; i < KID_COUNT && n < cap; ++i) {
        if (!ctx->k_launches[i]) continue;
        if (names) names[n] = KERNEL_NAMES[i];
        if (ms) ms[n] = ctx->k_ms[i];
        if (launches) launches[n] = ctx->k_launches[i];
        ++n;
    }
    return n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// C ABI: standalone filter
// ---------------------------------------------------------------------------

struct kc_filter {
    int device = 0;
    hipStream_t st = nullptr;
    uint64_t num_words = 0;
    Modulus M{};
    DevBuf d_seen, d_solid, d_fs, d_keys, d_prom;
    std::string err;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

#define F_TRY(expr)                                                                              \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return f->fail(_e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,    \
                                             std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

extern "C" {

int kc_filter_create(kc_filter** out, int32_t device, uint64_t capacity_bits) {
    if (!out) return KC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_EHIP;
    if (device < 0 || device >= ndev) return KC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return KC_EHIP;
    kc_filter* f = new kc_filter();
    f->device = device;
    f->num_words = std::max<uint64_t>((capacity_bits + 63) / 64, 1);
    f->M = make_modulus(f->num_words * 64);
    const uint64_t nw = f->num_words;
    if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess || f->d_seen.ensure(nw * 8) ||
        f->d_solid.ensure(nw * 8) || f->d_fs.ensure(nw * 64 * 4)) {
        kc_filter_destroy(f);
        return KC_ENOMEM;
    }
    k_fill<uint64_t><<<grid_for(nw), BLOCK, 0, f->st>>>(f->d_seen.as<uint64_t>(), nw, 0ull);
    k_fill<uint64_t><<<grid_for(nw), BLOCK, 0, f->st>>>(f->d_solid.as<uint64_t>(), nw, 0ull);
    k_fill<uint32_t><<<grid_for(nw * 64), BLOCK, 0, f->st>>>(f->d_fs.as<uint32_t>(), nw * 64, ~0u);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(f->st) != hipSuccess) {
        kc_filter_destroy(f);
        return KC_EHIP;
    }
    *out = f;
    return KC_OK;
}

void kc_filter_destroy(kc_filter* f) {
    if (!f) return;
    (void)hipSetDevice(f->device);
    if (f->st) (void)hipStreamSynchronize(f->st);
    f->d_seen.release();
    f->d_solid.release();
    f->d_fs.release();
    f->d_keys.release();
    f->d_prom.release();
    if (f->st) (void)hipStreamDestroy(f->st);
    delete f;
}

const char* kc_filter_last_error(const kc_filter* f) { return f ? f->err.c_str() : "null filter"; }
uint64_t kc_filter_memory_usage_bytes(const kc_filter* f) { return f ? f->num_words * 8 * 2 : 0; }
uint64_t kc_filter_num_words(const kc_filter* f) { return f ? f->num_words : 0; }

int kc_filter_insert_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* promoted) {
    if (!f) return KC_EINVAL;
    if (n == 0) return KC_OK;
    if (!keys4) return f->fail(KC_EINVAL, "null keys");
    F_TRY(hipSetDevice(f->device));
    const size_t batch_max = (size_t)1 << 28;  // 4*t + 3 fits the 32-bit first-setter words
    for (size_t base = 0; base < n; base += batch_max) {
        const size_t m = std::min(batch_max, n - base);
        F_TRY(f->d_keys.ensure(m * 32));
        F_TRY(f->d_prom.ensure(m));
        F_TRY(hipMemcpyAsync(f->d_keys.p, keys4 + 4 * base, m * 32, hipMemcpyHostToDevice, f->st));
        const uint32_t g = grid_for(m);
        k_ev_first_setter<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_fs.as<uint32_t>());
        k_ev_promote<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_fs.as<uint32_t>(),
                                              f->d_seen.as<uint64_t>(), f->d_prom.as<uint8_t>());
        k_ev_apply<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_prom.as<uint8_t>(),
                                            f->d_seen.as<uint64_t>(), f->d_solid.as<uint64_t>(), f->d_fs.as<uint32_t>());
        F_TRY(hipGetLastError());
        if (promoted) F_TRY(hipMemcpyAsync(promoted + base, f->d_prom.p, m, hipMemcpyDeviceToHost, f->st));
        F_TRY(hipStreamSynchronize(f->st));
    }
    return KC_OK;
}

int kc_filter_is_solid_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* out) {
    if (!f) return KC_EINVAL;
    if (n == 0) return KC_OK;
    if (!keys4 || !out) return f->fail(KC_EINVAL, "null keys/out");
    F_TRY(hipSetDevice(f->device));
    F_TRY(f->d_keys.ensure(n * 32));
    F_TRY(f->d_prom.ensure(n));
    F_TRY(hipMemcpyAsync(f->d_keys.p, keys4, n * 32, hipMemcpyHostToDevice, f->st));
    k_query<<<grid_for(n), BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), n, f->M, f->d_solid.as<uint64_t>(),
                                              f->d_prom.as<uint8_t>());
    F_TRY(hipGetLastError());
    F_TRY(hipMemcpyAsync(out, f->d_prom.p, n, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    return KC_OK;
}

int kc_filter_correct_reads(kc_filter* f, uint32_t k, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                            const uint32_t* lengths, size_t n_reads, uint8_t* out, uint64_t* n_corrected) {
    if (!f) return KC_EINVAL;
    if (k < 1 || k > KC_MAX_K) return f->fail(KC_EINVAL, "k out of range");
    if (n_reads && (!offsets || !lengths)) return f->fail(KC_EINVAL, "null offsets/lengths");
    if (data_len && (!data || !out)) return f->fail(KC_EINVAL, "null data/out");
    if (n_reads > 0xFFFFFFFFull) return f->fail(KC_EUNSUPPORTED, "more than 2^32-1 reads");
    std::vector<uint32_t> active;
    for (size_t i = 0; i < n_reads; ++i) {
        const uint64_t nb = ((uint64_t)lengths[i] + 3) / 4;
        if (offsets[i] > data_len || nb > data_len - offsets[i]) return f->fail(KC_EINVAL, "read extends past data_len");
        if (lengths[i] > HAMMER_MAX_LEN) return f->fail(KC_EUNSUPPORTED, "reads of at most 1024 bases");
        if (lengths[i] >= k) active.push_back((uint32_t)i);
    }
    F_TRY(hipSetDevice(f->device));
    DevBuf d_data, d_corr, d_off, d_len, d_act, d_ctr;
    F_TRY(d_data.ensure(data_len + DATA_PAD));
    F_TRY(d_corr.ensure(data_len + DATA_PAD));
    F_TRY(d_off.ensure(n_reads * 8 + 8));
    F_TRY(d_len.ensure(n_reads * 4 + 4));
    F_TRY(d_act.ensure(active.size() * 4 + 4));
    F_TRY(d_ctr.ensure(64));
    F_TRY(hipMemsetAsync(d_data.p, 0, d_data.bytes, f->st));
    F_TRY(hipMemsetAsync(d_ctr.p, 0, 64, f->st));
    if (data_len) F_TRY(hipMemcpyAsync(d_data.p, data, data_len, hipMemcpyHostToDevice, f->st));
    if (n_reads) {
        F_TRY(hipMemcpyAsync(d_off.p, offsets, n_reads * 8, hipMemcpyHostToDevice, f->st));
        F_TRY(hipMemcpyAsync(d_len.p, lengths, n_reads * 4, hipMemcpyHostToDevice, f->st));
    }
    if (!active.empty()) F_TRY(hipMemcpyAsync(d_act.p, active.data(), active.size() * 4, hipMemcpyHostToDevice, f->st));
    F_TRY(hipMemcpyAsync(d_corr.p, d_data.p, d_data.bytes, hipMemcpyDeviceToDevice, f->st));
    if (!active.empty()) {
        const uint32_t g = grid_for(active.size(), 64, 65536);
        if (k > 32)
            k_hammer<true><<<g, 64, 0, f->st>>>(d_data.as<uint8_t>(), d_corr.as<uint8_t>(), d_act.as<uint32_t>(),
                                                 active.size(), d_off.as<uint64_t>(), d_len.as<uint32_t>(), k,
                                                 f->d_solid.as<uint64_t>(), f->M, d_ctr.as<unsigned long long>(),
                                                 d_ctr.as<uint32_t>() + 4);
        else
            k_hammer<false><<<g, 64, 0, f->st>>>(d_data.as<uint8_t>(), d_corr.as<uint8_t>(), d_act.as<uint32_t>(),
                                                  active.size(), d_off.as<uint64_t>(), d_len.as<uint32_t>(), k,
                                                  f->d_solid.as<uint64_t>(), f->M, d_ctr.as<unsigned long long>(),
                                                  d_ctr.as<uint32_t>() + 4);
        F_TRY(hipGetLastError());
    }
    uint64_t ctr[2] = {0, 0};
    F_TRY(hipMemcpyAsync(ctr, d_ctr.p, 16, hipMemcpyDeviceToHost, f->st));
    if (data_len) F_TRY(hipMemcpyAsync(out, d_corr.p, data_len, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    for (DevBuf* b : {&d_data, &d_corr, &d_off, &d_len, &d_act, &d_ctr}) b->release();
    if (n_corrected) *n_corrected = ctr[0];
    return KC_OK;
}

int kc_filter_export(kc_filter* f, uint64_t* seen, uint64_t* solid, size_t num_words) {
    if (!f) return KC_EINVAL;
    if (num_words != f->num_words) return f->fail(KC_EINVAL, "num_words mismatch");
    F_TRY(hipSetDevice(f->device));
    if (seen) F_TRY(hipMemcpyAsync(seen, f->d_seen.p, num_words * 8, hipMemcpyDeviceToHost, f->st));
    if (solid) F_TRY(hipMemcpyAsync(solid, f->d_solid.p, num_words * 8, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    return KC_OK;
}

}  // extern "C"
