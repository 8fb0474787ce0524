// kc_engine.hip — gfx950 k-mer spectrum engine behind include/kc.h.
//
// Stage 2 (assemble.rs:129-161) and Stage 3 (assemble.rs:166-292) of spades-rs on one
// MI355X, for one or many independent libraries (read sets) processed by the same
// launches.  The reference inserts every window into the two-tier filter in read order;
// this engine computes the same filter state from the distinct canonical k-mers:
//
//   * every occurrence is inserted into a device hash table that keeps, per canonical
//     k-mer, its occurrence count and its first event index t0 (window number within
//     its library, read order then window order);
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
//
// Libraries: every library has its own Bloom tiers, its own region of the table pool and
// its own event numbering; windows of all libraries form one global window space cut
// into tiles, and table scans find a slot's library from the region starts.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "../../include/kc.h"
#include "kc_device.h"
#include "kc_hammer.h"
#include "kc_mlp.h"
#include "kc_host.h"
#include "kc_superkmer.h"

using namespace kc;

namespace {

constexpr uint32_t TILE_WINDOWS = 1024;  // windows per work tile
constexpr uint32_t HLL_P = 12;
constexpr uint32_t HLL_M = 1u << HLL_P;
constexpr uint32_t HIST_BINS = 1024;  // count histogram bins; the last bin holds >= 1023
constexpr uint32_t SCAN_ITEMS = 16;   // table slots per thread per chunk in table scans
constexpr uint32_t LIB_LOCAL = 8;     // libraries a block aggregates in LDS per chunk

// Per-library state the kernels read.
struct LibDesc {
    uint64_t ev0;       // first global window number of the library
    uint64_t bucket0;   // first 64-byte bucket of its table region
    uint64_t nbuckets;  // buckets in its table region
    uint64_t word0;     // first word of its Bloom tiers in the seen/solid pools
    uint64_t* seen;
    uint64_t* solid;
    void* fs;           // num_words * 64 first-setter entries (valid while its group runs)
    Modulus M;
    uint64_t num_words;
    uint32_t boost;     // prior / trusted-contig boost per window (assemble.rs:270)
    uint32_t k;
};

// prior_boost of assemble.rs:270, (min_coverage as u32 * 4).max(50): the float-to-int
// cast saturates (NaN and negatives give 0) and the u32 product wraps.
uint32_t prior_boost(double min_cov) {
    uint32_t mc;
    if (!(min_cov > 0.0)) mc = 0;
    else if (min_cov >= 4294967295.0) mc = 0xFFFFFFFFu;
    else mc = (uint32_t)min_cov;
    const uint32_t b = mc * 4u;
    return b > 50u ? b : 50u;
}

// Per-library results written on the device.
struct LibOut {
    unsigned long long n_distinct, n_counts, n_solid, corrected;
    uint32_t overflow, cutoff, top, pad;
};

// How a count-table entry is laid out for the sort: a library index and the 2k-bit
// k-mer value V as two words (hi, lo) that sort as (library, V), or one word when
// everything fits in 64 bits.
struct SortLayout {
    uint32_t k, lib_bits, lo_bits, hi_bits;  // hi_bits == 0: one word
    uint32_t single;                         // 1: (lib, V) packed in the lo word
    uint32_t single_bits;
    uint32_t wide;                           // 1: (V, library) records (WideEntry)
};

// A count-table entry when (library, V) does not fit two words (k > 63, or k close to 63
// with many libraries): the k-mer value V (four words, lowest first) and its library;
// entries sort as (library, V).  The library is also written into the hi word array
// (layout: lib_bits == hi_bits, no value bits), which the count passes read.
struct WideEntry {
    uint64_t v[4];
    uint32_t lib, pad;
};
struct WideLess {
    __device__ __host__ bool operator()(const WideEntry& a, const WideEntry& b) const {
        if (a.lib != b.lib) return a.lib < b.lib;
        for (int w = 3; w >= 0; --w)
            if (a.v[w] != b.v[w]) return a.v[w] < b.v[w];
        return false;
    }
};

SortLayout make_layout(uint32_t k, uint32_t n_libs) {
    SortLayout s{};
    s.k = k;
    s.lib_bits = 0;
    while ((1u << s.lib_bits) < n_libs) ++s.lib_bits;
    if (k > 63 || (2 * k > 63 && 2 * k - 63 + s.lib_bits > 64)) {
        s.wide = 1;
        s.lo_bits = 0;
        s.hi_bits = s.lib_bits;
        s.single = 0;
        s.single_bits = s.lib_bits;
        return s;
    }
    if (2 * k <= 63) {
        s.lo_bits = 2 * k;
        s.hi_bits = s.lib_bits;
    } else {
        s.lo_bits = 63;
        s.hi_bits = 2 * k - 63 + s.lib_bits;
    }
    s.single = s.lo_bits + s.hi_bits <= 64;
    s.single_bits = s.lo_bits + s.hi_bits;
    return s;
}

__device__ __forceinline__ void encode_entry(const SortLayout& L, uint32_t lib, const Kv& v, uint64_t& hi,
                                             uint64_t& lo) {
    if (2 * L.k <= 63) {
        lo = v.lo;
        hi = lib;
    } else {
        lo = v.lo & 0x7FFFFFFFFFFFFFFFull;
        hi = ((uint64_t)lib << (2 * L.k - 63)) | (v.hi << 1) | (v.lo >> 63);
    }
    if (L.single) {
        lo = L.hi_bits ? ((hi << L.lo_bits) | lo) : lo;
        hi = 0;
    }
}

__device__ __host__ __forceinline__ uint32_t entry_lib(const SortLayout& L, uint64_t hi, uint64_t lo) {
    if (L.lib_bits == 0) return 0;
    if (L.single) return (uint32_t)(lo >> (L.lo_bits));
    return (uint32_t)(hi >> (L.hi_bits - L.lib_bits));
}

// Largest l with starts[l] <= x (starts ascending, starts[0] == 0).
__device__ __forceinline__ uint32_t find_region(const uint64_t* __restrict__ starts, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Block-level per-library accumulation for a chunk whose first library is lib0: counts
// for lib0..lib0+LIB_LOCAL-1 are summed in LDS and flushed with one atomic each.
struct LibAccum {
    unsigned long long* s;  // LDS [LIB_LOCAL]
    __device__ void reset() {
        if (threadIdx.x < LIB_LOCAL) s[threadIdx.x] = 0;
    }
    __device__ void add(uint32_t lib, uint32_t lib0, unsigned long long v, unsigned long long* global_field,
                        size_t stride_ull) {
        if (lib - lib0 < LIB_LOCAL) atomicAdd(&s[lib - lib0], v);
        else atomicAdd(global_field + (size_t)lib * stride_ull, v);
    }
    __device__ void flush(uint32_t lib0, uint32_t n_libs, unsigned long long* global_field, size_t stride_ull) {
        if (threadIdx.x < LIB_LOCAL && lib0 + threadIdx.x < n_libs && s[threadIdx.x])
            atomicAdd(global_field + (size_t)(lib0 + threadIdx.x) * stride_ull, s[threadIdx.x]);
    }
};

constexpr size_t LIBOUT_STRIDE = sizeof(LibOut) / sizeof(unsigned long long);
// Counter slot (of the context's 64) receiving the number of is_solid tests Hammer makes.
constexpr uint32_t HAMMER_QUERIES = 20;
// Counter slot a replayed graph sets when its data-dependent sizes leave the plan.
constexpr uint32_t PLAN_GUARD = 22;
// Counter slot holding the binned count's queue of bins (the next bin to take).
constexpr uint32_t BIN_QUEUE = 23;
// Counter slot receiving the number of bins that hold records.
constexpr uint32_t BIN_LIST_N = 24;
// Counter slot of the count-table sort's runs: listed long runs, and a run too long.
constexpr uint32_t TIE_CTR = 25;
static_assert(sizeof(LibOut) % sizeof(unsigned long long) == 0, "LibOut layout");

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

// For every active read: its first global window, byte offset and library, and the first
// active read of every tile whose first window it holds.  act_ev[n_active] is the total.
__global__ void k_active(const uint32_t* __restrict__ active, uint64_t n_active, const uint64_t* __restrict__ prefix,
                         const uint64_t* __restrict__ nwin, const uint64_t* __restrict__ offsets,
                         const uint32_t* __restrict__ read_lib, uint64_t* __restrict__ act_ev,
                         uint64_t* __restrict__ act_off, uint32_t* __restrict__ act_lib,
                         uint32_t* __restrict__ tile_first) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n_active;
         a += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = active[a];
        const uint64_t s = prefix[r], e = s + nwin[r];
        act_ev[a] = s;
        if (a + 1 == n_active) act_ev[n_active] = e;  // so that act_ev[a + 1] - act_ev[a] = windows of a
        act_off[a] = offsets[r];
        act_lib[a] = read_lib[r];
        for (uint64_t b = (s + TILE_WINDOWS - 1) / TILE_WINDOWS; b * TILE_WINDOWS < e; ++b) tile_first[b] = (uint32_t)a;
    }
}

// First global window of each library: prefix at its first read (or the total).
__global__ void k_lib_ev0(const uint64_t* __restrict__ prefix, const uint64_t* __restrict__ nwin, uint64_t n_reads,
                          const uint64_t* __restrict__ lib_read0, uint32_t n_libs, uint64_t* __restrict__ ev0) {
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l <= n_libs; l += gridDim.x * blockDim.x) {
        const uint64_t r = lib_read0[l];
        ev0[l] = r < n_reads ? prefix[r] : (n_reads ? prefix[n_reads - 1] + nwin[n_reads - 1] : 0);
    }
}

// Window w of tile: index of its active read within s_ev (largest i with s_ev[i] <= w).
__device__ __forceinline__ uint32_t tile_read(const uint64_t* s_ev, uint32_t na, uint64_t w) {
    uint32_t lo = 0, hi = na - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (s_ev[mid] <= w) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// distinct-count estimate per library (HyperLogLog, 2^12 registers) to size its table
// ---------------------------------------------------------------------------

template <int KW>
__global__ void __launch_bounds__(BLOCK) k_hll(const uint8_t* __restrict__ data, const uint64_t* __restrict__ act_ev,
                                               const uint64_t* __restrict__ act_off,
                                               const uint32_t* __restrict__ act_lib,
                                               const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                               uint64_t n_occ, uint64_t n_tiles, uint32_t k,
                                               uint32_t* __restrict__ regs) {
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    __shared__ uint32_t s_reg[HLL_M];
    constexpr int NW = nw_of(KW);
    uint32_t cur = ~0u;  // library whose registers s_reg holds (block-uniform)
    for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) s_reg[i] = 0;
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t a0 = tile_first[tile];
        const uint32_t lib0 = act_lib[a0];
        if (lib0 != cur) {
            __syncthreads();
            if (cur != ~0u)
                for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x)
                    if (s_reg[i]) atomicMax(&regs[(size_t)cur * HLL_M + i], s_reg[i]);
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x) s_reg[i] = 0;
            cur = lib0;
        }
        const uint64_t w0 = tile * TILE_WINDOWS;
        const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
        const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
        __syncthreads();
        for (uint64_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
            const uint32_t lo = tile_read(s_ev, na, w);
            const Kv v = load_window<KW>(data, act_off[a0 + lo], (uint32_t)(w - s_ev[lo]), k);
            uint64_t key[NW];
            pack_key<NW>(canonical<KW>(v, k), key);
            const uint64_t h = table_hash<NW>(key);
            const uint32_t idx = (uint32_t)(h >> (64 - HLL_P));
            const uint64_t hw = h << HLL_P;
            const uint32_t rho = hw ? (uint32_t)__clzll((long long)hw) + 1 : 64 - HLL_P + 1;
            const uint32_t lib = act_lib[a0 + lo];
            if (lib == cur) atomicMax(&s_reg[idx], rho);
            else atomicMax(&regs[(size_t)lib * HLL_M + idx], rho);
        }
    }
    __syncthreads();
    if (cur != ~0u)
        for (uint32_t i = threadIdx.x; i < HLL_M; i += blockDim.x)
            if (s_reg[i]) atomicMax(&regs[(size_t)cur * HLL_M + i], s_reg[i]);
}

// ---------------------------------------------------------------------------
// table passes
// ---------------------------------------------------------------------------

template <int NW, typename Ev>
__global__ void k_table_clear(Slot<NW, Ev>* __restrict__ tab, uint64_t n_slots) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_slots;
         i += (uint64_t)gridDim.x * blockDim.x) {
        Slot<NW, Ev> s;
#pragma unroll
        for (int w = 0; w < NW; ++w) s.k[w] = EMPTY_KEY;
        s.t0 = (Ev)~(Ev)0;
        s.cnt = 0;
        tab[i] = s;
    }
}

// Stage 2 insert (TRACK_T0, every window) or Stage 3 recount after Hammer (SOLID_ONLY:
// only windows whose canonical k-mer is Bloom-solid, assemble.rs:196).  Block b takes
// tiles b, b + gridDim.x, ... (one tile each when the grid covers every tile).  Within a
// tile each thread takes WPT windows at a time and issues their first bucket loads
// together, so a wave keeps WPT independent random table accesses in flight.
template <int KW, typename Ev, bool TRACK_T0, bool SOLID_ONLY, uint32_t WPT>
__global__ void __launch_bounds__(BLOCK) k_insert(const uint8_t* __restrict__ data, const uint64_t* __restrict__ act_ev,
                                                  const uint64_t* __restrict__ act_off,
                                                  const uint32_t* __restrict__ act_lib,
                                                  const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                                  uint64_t n_occ, uint64_t n_tiles, uint32_t k,
                                                  Slot<nw_of(KW), Ev>* __restrict__ tab,
                                                  const LibDesc* __restrict__ libs, uint32_t max_probe,
                                                  LibOut* __restrict__ out) {
    constexpr int NW = nw_of(KW);
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    constexpr uint32_t PASS = BLOCK * WPT;
    static_assert(TILE_WINDOWS % PASS == 0, "a tile is a whole number of passes");
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t a0 = tile_first[tile];
        const uint64_t w0 = tile * TILE_WINDOWS;
        const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
        const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
        __syncthreads();  // the previous tile's reads of s_ev are done
        for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
        __syncthreads();
        for (uint64_t wp = w0; wp < w1; wp += PASS) {
            uint64_t key[WPT][NW], bkt[WPT], pre[WPT][SPB];
            Ev pre_t0[WPT][SPB];
            uint32_t lib[WPT];
            bool valid[WPT];
#pragma unroll
            for (uint32_t j = 0; j < WPT; ++j) {
                const uint64_t w = wp + j * BLOCK + threadIdx.x;
                valid[j] = w < w1;
                if (!valid[j]) continue;
                const uint32_t lo = tile_read(s_ev, na, w);
                lib[j] = act_lib[a0 + lo];
                const Kv c =
                    canonical<KW>(load_window<KW>(data, act_off[a0 + lo], (uint32_t)(w - s_ev[lo]), k), k);
                if (SOLID_ONLY && !bloom_is_solid<KW>(c, k, libs[lib[j]].solid, libs[lib[j]].M)) {
                    valid[j] = false;
                    continue;
                }
                pack_key<NW>(c, key[j]);
                bkt[j] = libs[lib[j]].bucket0 + home_bucket<NW>(key[j], libs[lib[j]].nbuckets);
            }
#pragma unroll
            for (uint32_t j = 0; j < WPT; ++j)
                if (valid[j]) bucket_load<NW, Ev>(tab, bkt[j], pre[j], TRACK_T0 ? pre_t0[j] : nullptr);
#pragma unroll
            for (uint32_t j = 0; j < WPT; ++j) {
                if (!valid[j]) continue;
                const LibDesc& L = libs[lib[j]];
                const Ev t = (Ev)(wp + j * BLOCK + threadIdx.x - L.ev0);
                if (!table_insert_from<NW, Ev, TRACK_T0>(tab, L.bucket0, L.nbuckets, bkt[j], pre[j], pre_t0[j], key[j],
                                                         t, max_probe))
                    atomicOr(&out[lib[j]].overflow, 1u);
            }
        }
    }
}

// One 64-bit key word of an LDS slot: EMPTY -> v exactly once, given the word as loaded
// (`cur`; a non-EMPTY word is final).  `claimed` is set when this call made the change.
__device__ __forceinline__ bool lds_claim_loaded(uint64_t* w, uint64_t cur, uint64_t v, bool& claimed) {
    if (cur == EMPTY_KEY) {
        cur = atomicCAS(reinterpret_cast<unsigned long long*>(w), (unsigned long long)EMPTY_KEY,
                        (unsigned long long)v);
        if (cur == EMPTY_KEY) {
            claimed = true;
            cur = v;
        }
    }
    return cur == v;
}

// ---------------------------------------------------------------------------
// Minimizer-binned Stage 2 count (one library, k <= 63)
// ---------------------------------------------------------------------------
//
// Every window is routed to one of 2^BIN_BITS bins by the hash of its canonical
// minimizer: the smallest mix of the canonical m-mers it contains, which is the same for a
// k-mer and its reverse complement, so all occurrences of a canonical k-mer share a bin.
// A read is cut into super-k-mers, runs of consecutive windows in one bin, each an 8-byte
// record (active read, first window, window count); the records are radix-sorted by bin.
// A workgroup then counts a bin (or a range of a small library's bins) exactly in an LDS
// table (count and first event of every distinct k-mer) and writes the distinct k-mers
// densely to the table pool with plain stores: no global atomic per window, one per
// flushed pass.  A bin with too many distinct k-mers for the LDS table is split by key-hash
// bits into passes (2, 4, ...) that each count a disjoint subset.  The table then holds exactly the entries k_insert
// leaves (every distinct key with its count and first event), packed at the start of the
// library's region with no hash placement, which is all the Stage 2 / 3 scans need.

// Super-k-mers, one wave per active read (superkmers() in kc_superkmer.h is the
// definition).  Windows are taken 64 at a time: lane l holds the keys of m-mers c0 + l
// and c0 + 64 + l, a sparse-table sliding minimum over shuffles gives window c0 + l its
// minimizer key, and a ballot marks the windows whose bin differs from the previous
// window's.  The lane at such a start writes the record of the run that ends before it;
// the read's last run is written after the loop.  A read has at most as many records as
// windows, so its records go to scratch from its global window number act_ev[a] on, as
// bin | first window << 24 | windows << 48 (bins < 2^24, first window < 2^24), and its
// record count to cnt[a]; k_sk_copy then packs them in read order.  No lane diverges on
// the data.  (Reserving packed room chunk by chunk on one global counter was measured
// 40x slower: the same-address atomics serialize.  Counting first and cutting again, the
// second pass reading back every window's bin, was about 0.6 ms a k slower than this
// scratch on 5 Mb / 100x.)
static_assert(BIN_BITS_MAX <= 24, "scratch records hold 24-bit bins");
__device__ __forceinline__ uint64_t sk_scratch(uint32_t bin, uint32_t first, uint32_t nw) {
    return (uint64_t)bin | (uint64_t)first << 24 | (uint64_t)nw << 48;
}

__global__ void __launch_bounds__(BLOCK) k_sk_wave(const uint8_t* __restrict__ data,
                                                   const uint64_t* __restrict__ act_off,
                                                   const uint64_t* __restrict__ act_ev, uint64_t n_active, uint32_t k,
                                                   uint32_t m, uint32_t bits, uint32_t* __restrict__ cnt,
                                                   uint64_t* __restrict__ scratch) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * (BLOCK / 64);
    const uint32_t w = k - m + 1;
    uint32_t p = 1;  // largest power of two <= w
    while (2 * p <= w) p <<= 1;
    const uint64_t mmask = m >= 32 ? ~0ull : (1ull << (2 * m)) - 1;
    const uint32_t* dw = reinterpret_cast<const uint32_t*>(data);
    // The loop is software-pipelined over the wave's reads: the (offset, first window, end
    // window) of the read after the next are loaded while the current read is cut, so a
    // read's first packed words wait on one load, not on a chain of three.  (Holding a
    // read's packed dwords one per lane and reading an m-mer's words by shuffles was 8%
    // slower than these cached loads.)
    struct Meta {
        uint64_t off, ev, ev1;
    };
    auto load_meta = [&](uint64_t a) -> Meta {
        Meta r{0, 0, 0};
        if (a < n_active) r = Meta{act_off[a], act_ev[a], act_ev[a + 1]};
        return r;
    };
    uint64_t a = (uint64_t)blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6);
    Meta cur = load_meta(a);
    Meta nxt = load_meta(a + waves);
    for (; a < n_active; a += waves) {
        const uint64_t bit0 = cur.off * 8;
        const uint32_t nw = (uint32_t)(cur.ev1 - cur.ev);
        uint64_t* out = scratch + cur.ev;
        cur = nxt;
        nxt = load_meta(a + 2 * waves);
        const uint32_t nm = nw + w - 1;  // m-mers of the read
        auto key_at = [&](uint32_t i) -> uint32_t {
            if (i >= nm) return 0xFFFFFFFFu;  // past the read: no valid window reaches it
            const uint64_t bp = bit0 + 2ull * i;
            const uint64_t q = bp >> 5;
            const uint32_t sh = (uint32_t)(bp & 31);
            const uint64_t win = ((uint64_t)__builtin_bswap32(dw[q]) << 32) | __builtin_bswap32(dw[q + 1]);
            uint64_t f;
            if (m <= 17) {  // sh + 2m <= 64: the m-mer lies in two dwords
                f = (win >> (64 - sh - 2 * m)) & mmask;
            } else {  // a third dword (inside the read, or the buffer's padding past it)
                const uint64_t x = (win << sh) | ((uint64_t)__builtin_bswap32(dw[q + 2]) << sh >> 32);
                f = x >> (64 - 2 * m);
            }
            const uint64_t r = rev_groups(~f) >> (64 - 2 * m);
            return mmer_key(f < r ? f : r);
        };
        uint32_t open = 0, carry_bin = 0, total = 0;
        uint32_t K = key_at(lane);  // keys of positions c0 + lane
        for (uint32_t c0 = 0; c0 < nw; c0 += 64) {
            const bool valid = c0 + lane < nw;
            // A = positions c0 + lane, B = positions c0 + 64 + lane (the windows read B up to
            // lane w - 2; B is the next chunk's A)
            const uint32_t Kn = key_at(c0 + 64 + lane);
            uint32_t A = K, B = Kn;
            K = Kn;
            for (uint32_t d = 1; d < p; d <<= 1) {  // A, B: minimum over [pos, pos + 2d)
                const uint32_t src = (lane + d) & 63;
                const uint32_t tA = __shfl(A, src), tB = __shfl(B, src);
                const bool wrap = lane + d >= 64;
                A = min(A, wrap ? tB : tA);
                if (!wrap) B = min(B, tB);
            }
            const uint32_t e = lane + w - p;  // [lane, lane + w) = [lane, +p) and [e, e + p)
            const uint32_t tA = __shfl(A, e & 63), tB = __shfl(B, e & 63);
            const uint32_t bin = minimizer_bin(min(A, e >= 64 ? tB : tA), bits);
            const uint32_t prev = __shfl(bin, (lane + 63) & 63);
            const uint32_t pb = lane ? prev : carry_bin;
            const bool start = valid && c0 + lane > 0 && bin != pb;
            const uint64_t sm = __ballot(start);
            uint32_t nrec = 0, rs = 0, rl = 0;
            if (start) {
                const uint64_t below = sm & ((1ull << lane) - 1);
                rs = below ? c0 + (63 - __clzll(below)) : open;
                rl = c0 + lane - rs;
                nrec = (rl + SK_MAX_WINDOWS - 1) / SK_MAX_WINDOWS;
            }
            uint32_t o, nchunk;  // the lane's first output; the chunk's records
            if (__ballot(nrec > 1) == 0) {  // one record per start: rank the starts
                o = total + (uint32_t)__popcll(sm & ((1ull << lane) - 1));
                nchunk = (uint32_t)__popcll(sm);
            } else {  // a run longer than SK_MAX_WINDOWS: scan the record counts
                uint32_t incl = nrec;
                for (uint32_t d = 1; d < 64; d <<= 1) {
                    const uint32_t t = __shfl_up(incl, d);
                    if (lane >= d) incl += t;
                }
                o = total + incl - nrec;
                nchunk = __shfl(incl, 63);
            }
            for (uint32_t t = 0; t < nrec; ++t, ++o)
                out[o] = sk_scratch(pb, rs + t * SK_MAX_WINDOWS, min(SK_MAX_WINDOWS, rl - t * SK_MAX_WINDOWS));
            total += nchunk;
            if (sm) open = c0 + (63 - __clzll(sm));
            carry_bin = __shfl(bin, min(63u, nw - 1 - c0));
        }
        // the last run [open, nw)
        const uint32_t rl = nw - open, nlast = (rl + SK_MAX_WINDOWS - 1) / SK_MAX_WINDOWS;
        for (uint32_t t = lane; t < nlast; t += 64)
            out[total + t] = sk_scratch(carry_bin, open + t * SK_MAX_WINDOWS,
                                        min(SK_MAX_WINDOWS, rl - t * SK_MAX_WINDOWS));
        if (lane == 0) cnt[a] = total + nlast;
    }
}

// The records of each read from scratch to their packed place sk_off[a] ..: bin key
// (library << bits | bin), and a | first window << 32 | windows << 56.  Sixteen lanes per
// read (a read holds about 5-25 records), so a wave follows four reads' load chains at once.
// With pk != 0 (the pk key bits, the read index, the first window and the window count fit
// 64 bits: every read has at most 256 windows) a record is one word, a << (pk + 16) |
// first window << (pk + 8) | windows << pk | key, sorted on its low pk bits as a key alone,
// and keys is unused.
constexpr uint32_t SK_COPY_LANES = 16;
__host__ __device__ __forceinline__ uint32_t rec_read(uint64_t rec, uint32_t pk) {
    return pk ? (uint32_t)(rec >> (pk + 16)) : (uint32_t)rec;
}
__host__ __device__ __forceinline__ uint32_t rec_first(uint64_t rec, uint32_t pk) {
    return pk ? (uint32_t)(rec >> (pk + 8)) & 0xFFu : (uint32_t)(rec >> 32) & 0xFFFFFFu;
}
__host__ __device__ __forceinline__ uint32_t rec_windows(uint64_t rec, uint32_t pk) {
    return pk ? (uint32_t)(rec >> pk) & 0xFFu : (uint32_t)(rec >> 56);
}
__global__ void __launch_bounds__(BLOCK) k_sk_copy(const uint64_t* __restrict__ act_ev,
                                                   const uint32_t* __restrict__ act_lib, uint64_t n_active,
                                                   uint32_t bits, const uint32_t* __restrict__ cnt,
                                                   const uint64_t* __restrict__ sk_off,
                                                   const uint64_t* __restrict__ scratch, uint32_t* __restrict__ keys,
                                                   uint64_t* __restrict__ recs, uint64_t cap, uint32_t pk) {
    const uint32_t sub = threadIdx.x % SK_COPY_LANES;
    const uint64_t groups = (uint64_t)gridDim.x * (BLOCK / SK_COPY_LANES);
    for (uint64_t a = (uint64_t)blockIdx.x * (BLOCK / SK_COPY_LANES) + threadIdx.x / SK_COPY_LANES; a < n_active;
         a += groups) {
        const uint32_t n = cnt[a];
        const uint64_t* src = scratch + act_ev[a];
        const uint64_t o = sk_off[a];
        const uint32_t lkey = act_lib[a] << bits;
        // (cap: the record buffers' room; a replayed plan that no longer fits writes
        // nothing past it, and its guard sends the run back to planning)
        for (uint32_t t = sub; t < n && o + t < cap; t += SK_COPY_LANES) {
            const uint64_t r = src[t];
            const uint32_t key = lkey | (uint32_t)(r & 0xFFFFFFu);
            if (pk) {
                recs[o + t] = a << (pk + 16) | ((r >> 24) & 0xFFull) << (pk + 8) | (r >> 48) << pk | key;
            } else {
                keys[o + t] = key;
                recs[o + t] = a | ((r >> 24) & 0xFFFFFFull) << 32 | (r >> 48) << 56;
            }
        }
    }
}

// Guard of a replayed binned count: sets *guard when the replay's sizes leave the plan
// (record count, per-library distinct k-mers, a bin past its split limit, an output
// region overflow), so that the caller plans again.
__global__ void k_plan_guard(const uint64_t* __restrict__ sko, const uint32_t* __restrict__ skc, uint64_t na,
                             uint64_t nrec_plan, const unsigned long long* __restrict__ lcnt,
                             const uint64_t* __restrict__ lcnt_plan, uint32_t n_libs,
                             const unsigned long long* __restrict__ bin_ctr, unsigned long long* __restrict__ guard) {
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l <= n_libs; l += gridDim.x * blockDim.x) {
        bool bad = false;
        if (l < n_libs) bad = lcnt[l] != lcnt_plan[l];
        else bad = (na && sko[na - 1] + skc[na - 1] != nrec_plan) || (bin_ctr[1] & 0xFFFFFFFFull) ||
                   (bin_ctr[2] & 0xFFFFFFFFull);
        if (bad) atomicOr(guard, 1ull);
    }
}

// bounds[b] = first sorted record of bin b (b <= nbins): keys from the key array, or with
// pk != 0 from the low pk bits of the one-word records.
__global__ void k_bin_bounds(const uint32_t* __restrict__ keys, const uint64_t* __restrict__ recs, uint32_t pk,
                             uint64_t n, uint32_t nbins, uint64_t* __restrict__ bounds) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nbins; b += gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            const uint32_t key = pk ? (uint32_t)(recs[mid] & ((1ull << pk) - 1)) : keys[mid];
            if (key < b) lo = mid + 1;
            else hi = mid;
        }
        bounds[b] = lo;
    }
}

// The binned count's work items: library l's 2^bits bins in groups of 2^gsh[l]
// consecutive bins (gmeta = gsh[0 .. n_libs), then the first item of each library,
// gbase[0 .. n_libs]), item i = lo | hi << 32 for its bins [lo, hi), flagged when it holds
// records.
__global__ void k_bin_items(const uint64_t* __restrict__ bounds, uint32_t n_libs, uint32_t bits,
                            const uint32_t* __restrict__ gmeta, uint64_t n_items, uint64_t* __restrict__ items,
                            uint8_t* __restrict__ flags) {
    const uint32_t* gbase = gmeta + n_libs;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_items;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = n_libs - 1;  // the last library with gbase <= i
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (gbase[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t sh = gmeta[lo];
        const uint32_t b0 = (lo << bits) | ((uint32_t)(i - gbase[lo]) << sh), b1 = b0 + (1u << sh);
        items[i] = b0 | (uint64_t)b1 << 32;
        flags[i] = bounds[b1] > bounds[b0];
    }
}

// One work item (a range of bins of one library) per workgroup pass (see above and in the
// kernel).  The item's records are staged in LDS in chunks
// with the prefix sum of their window counts, and the chunk's windows are dealt to the
// threads in order: neighbouring lanes take neighbouring windows of one record (distinct
// k-mers), rather than the same window of reads from the same locus, which would send
// the whole wave's LDS atomics to one slot.  (Runs of 4-8 windows per thread, the first
// extracted and the others rolled in base by base, were measured slower: a wave then
// spans 8x more records, and copies of one super-k-mer collide on a slot.)  Library l's
// entries go to out[cap0[l] ..],
// reserved per pass with one atomicAdd on lcnt[l]; a pass that would pass capl[l] writes
// nothing and sets ctr[2] (lcnt still totals the library's distinct k-mers).  ctr[1] is set
// when an item needs more than BIN3_MAX_SPLIT passes (the caller then counts with the
// table insert instead).
// The bin sort's onesweep passes take 10 bits each, so the 15-20 bits of (library, bin)
// keys sort in two passes instead of three (and 16 items a thread: 1.60 -> 1.50 ms at
// k = 21 against 12).  The first-setter items (64-bit keys, 15-bit
// partition numbers) sort fastest with 1024-thread blocks of 8 items (5 Mb / 100x, k = 55:
// 3.71 -> 3.01 ms against rocPRIM's 512 x 12 for gfx950; 6, 10, 12 or 16 items were slower).
using BinSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>, 10,
                                        rocprim::block_radix_rank_algorithm::match>>;
using FsSortConfig = rocprim::radix_sort_config<
    rocprim::default_config, rocprim::default_config,
    rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 8>, 8,
                                        rocprim::block_radix_rank_algorithm::match>>;
constexpr uint32_t BIN3_RCHUNK = BLOCK;  // records staged per chunk (one per thread)
// A staged record: stream base number (2-bit position) of its first window, the event of
// its first window (library-relative), and its first window in the chunk (before the
// chunk's scan: its number of windows).  One LDS read gives a window all three.
template <typename Ev>
struct StagedRec {
    uint64_t pos;
    uint32_t woff;
    Ev ev;
};
constexpr uint32_t BIN3_MAP = 2048;      // windows per chunk (the window -> record map)
// Transient buffers above this size are released after use instead of kept for reuse.
constexpr uint64_t BIG_SCRATCH = 8ull << 30;
// Device bytes Stage 3 takes per count-table entry (two copies of key words and counts,
// sort indices and temporary storage).
constexpr uint64_t STAGE3_BYTES_PER_ENTRY = 64;

// A window's packed words as loaded (3 for k <= 32, 5 for k <= 63) and its bit shift;
// raw_extract gives the k-mer value load_window would.
template <int KW>
struct RawWindow {
    uint32_t w[KW == 1 ? 3 : 5];
    uint32_t sh;
};

// The window whose first base is base number `pos` of the packed stream (4 bases a byte).
template <int KW>
__device__ __forceinline__ void raw_fetch(const uint8_t* __restrict__ data, uint64_t pos, RawWindow<KW>& r) {
    static_assert(KW <= 2, "k <= 63");
    const uint64_t b0 = pos >> 2;
    const uint64_t a = b0 & ~3ull;
    r.sh = (uint32_t)(b0 - a) * 8 + 2 * (uint32_t)(pos & 3);
    const uint32_t* p = reinterpret_cast<const uint32_t*>(data + a);
#pragma unroll
    for (int i = 0; i < (KW == 1 ? 3 : 5); ++i) r.w[i] = p[i];
}

// Start slot (bits 0..9) and split-pass bits (20..31) of a key in a bin's LDS table: the
// key words folded to 32 bits and mixed by fmix32.  Only the probe start and the pass of a
// key depend on it, never the result, so a cheap hash serves (the keys are compared whole).
template <int NW>
__device__ __forceinline__ uint32_t bin_slot_hash(const uint64_t key[NW]) {
    uint32_t x = (uint32_t)key[0] ^ ((uint32_t)(key[0] >> 32) * 0x9E3779B1u);
    if (NW == 2) x ^= ((uint32_t)key[NW - 1] + 0x7F4A7C15u) * 0x85EBCA77u ^ (uint32_t)(key[NW - 1] >> 32);
    return mix32(x);
}
static_assert(BIN3_CAP <= (1u << 20) && BIN3_MAX_SPLIT <= (1u << 12), "slot and pass bits of bin_slot_hash");

template <int KW>
__device__ __forceinline__ Kv raw_extract(const RawWindow<KW>& r, uint32_t k) {
    return extract_be<KW>(r.w, r.sh, k);
}

// (byte offset, first global window) of every active read side by side, so that a
// staged record gathers both with one 16-byte load.
__global__ void k_pair_oe(const uint64_t* __restrict__ act_off, const uint64_t* __restrict__ act_ev, uint64_t n,
                          ulonglong2* __restrict__ oe) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n; a += (uint64_t)gridDim.x * blockDim.x)
        oe[a] = make_ulonglong2(act_off[a], act_ev[a]);
}

template <int KW, typename Ev>
// Waves a SIMD holds: LDS allows 5 workgroups of 4 waves a CU at k <= 63 (NW = 2) and 6 at
// k <= 32; registers are held to what that many waves leave.
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(KW == 1 ? 6 : 5, KW == 1 ? 6 : 5))) k_bin_count(const uint64_t* __restrict__ recs,
                                                     const uint64_t* __restrict__ bounds,
                                                     const uint8_t* __restrict__ data,
                                                     const ulonglong2* __restrict__ act_oe, uint32_t k,
                                                     const LibDesc* __restrict__ libs, uint32_t bits, uint32_t nbins,
                                                     Slot<nw_of(KW), Ev>* __restrict__ out,
                                                     const uint64_t* __restrict__ cap0,
                                                     const uint64_t* __restrict__ capl,
                                                     unsigned long long* __restrict__ lcnt,
                                                     unsigned long long* __restrict__ ctr,
                                                     unsigned int* __restrict__ queue,
                                                     const uint64_t* __restrict__ blist,
                                                     const unsigned int* __restrict__ n_blist, uint32_t pk) {
    constexpr int NW = nw_of(KW);
    static_assert(NW <= 2, "k <= 63");
    __shared__ uint64_t s_k0[BIN3_CAP];
    __shared__ uint64_t s_k1[NW == 2 ? BIN3_CAP : 1];
    __shared__ uint32_t s_cnt[BIN3_CAP];
    __shared__ Ev s_t0[BIN3_CAP];
    __shared__ StagedRec<Ev> s_rec[BIN3_RCHUNK];  // the chunk's staged records
    __shared__ uint8_t s_map[BIN3_MAP];       // the staged record of every window of the chunk
    __shared__ uint16_t s_list[BIN3_FILL];    // the pass's claimed slots, in claim order
    __shared__ uint32_t s_wsum[BLOCK / 64], s_wfit[BLOCK / 64];
    __shared__ uint32_t s_fill, s_over, s_np, s_nr, s_nwin;
    __shared__ uint32_t s_pass[2 * 32];  // pending passes: (parts, part) pairs
    __shared__ unsigned long long s_base;
    __shared__ uint64_t s_bin[2];
    static_assert(BIN3_CAP <= 65536, "slot numbers are 16-bit");
    static_assert(BIN3_RCHUNK <= 256 && BIN3_MAP >= SK_MAX_WINDOWS, "map entries are 8-bit, a record fits a chunk");
    // The table starts empty and every pass leaves it empty: its claimed slots are listed
    // and reset as they are written out; a pass that overflows (the list is then
    // incomplete) clears the whole table.
    auto clear_table = [&]() {
        for (uint32_t i = threadIdx.x; i < BIN3_CAP; i += blockDim.x) {
            s_k0[i] = EMPTY_KEY;
            if (NW == 2) s_k1[i] = EMPTY_KEY;
            s_cnt[i] = 0;
            s_t0[i] = (Ev)~(Ev)0;
        }
    };
    clear_table();
    // Workgroups stay resident and take work items from a queue over the list
    // blist[0 .. *n_blist) (so that uneven items balance): an item is a range of bins
    // [lo, hi) of one library, lo | hi << 32, that holds records (a library with far fewer
    // k-mers than the largest counts several of its bins at once).  And
    // the loads that start a bin (its record range, the first chunk's records and their
    // (offset, first event) pairs) are issued during the previous bin: the next bin's range
    // at the start of a bin, its records after the first chunk is staged, their pairs after
    // the first chunk's windows.  The record range is loaded by lanes 0 and 1 of each wave
    // as vector loads and read back with readlane.
    const uint32_t tid = threadIdx.x;
    const uint64_t none = nbins | (uint64_t)nbins << 32;
    auto load_range = [&](uint64_t it) -> uint64_t {
        return (uint32_t)it < nbins ? bounds[(tid & 1) ? (uint32_t)(it >> 32) : (uint32_t)it] : 0;
    };
    auto lane_u64 = [](uint64_t v, int l) -> uint64_t {
        return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l) |
               (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l) << 32;
    };
    const uint32_t nl = *n_blist;
    auto listed = [&](uint32_t q) -> uint64_t { return q < nl ? blist[q] : none; };
    if (tid == 0) {
        const uint32_t q = atomicAdd(queue, 2u);
        s_bin[0] = listed(q);
        s_bin[1] = listed(q + 1);
    }
    __syncthreads();
    uint64_t item = s_bin[0], nitem = s_bin[1];
    // A bin's first chunk of records is staged in LDS before the bin starts: pos and ev as
    // the staging below writes them, and each record's windows in woff.
    auto prestage = [&](uint64_t rec, ulonglong2 oe, uint64_t ev0) {
        const uint32_t first = rec_first(rec, pk);
        s_rec[tid].pos = oe.x * 4 + first;
        s_rec[tid].ev = (Ev)(oe.y + first - ev0);
        s_rec[tid].woff = rec_windows(rec, pk);
    };
    uint64_t c_rb, c_re;  // the current bin's records
    {
        const uint64_t v = load_range(item);
        c_rb = lane_u64(v, 0);
        c_re = lane_u64(v, 1);
        if (tid < min<uint64_t>(BIN3_RCHUNK, c_re - c_rb)) {
            const uint64_t rec = recs[c_rb + tid];
            prestage(rec, act_oe[rec_read(rec, pk)], libs[(uint32_t)item >> bits].ev0);
        }
    }
    uint64_t n_v = load_range(nitem);
    while ((uint32_t)item < nbins) {
        __syncthreads();  // every thread has read s_bin
        uint32_t q_next = 0;  // the item after nitem: its list index, then the item
        uint64_t nn_item = none;
        if (tid == 0) q_next = atomicAdd(queue, 1u);
        const uint64_t rb = c_rb, re = c_re;
        uint64_t n_rb = 0, n_re = 0, n_rec = 0;
        ulonglong2 n_oe = make_ulonglong2(0, 0);
        const uint64_t n_ev0 = (uint32_t)nitem < nbins ? libs[(uint32_t)nitem >> bits].ev0 : 0;
        bool pf_rec = false, pf_oe = false;  // (uniform)
        auto issue_next_records = [&]() {
            if (pf_rec) return;
            pf_rec = true;
            n_rb = lane_u64(n_v, 0);
            n_re = lane_u64(n_v, 1);
            if (tid < min<uint64_t>(BIN3_RCHUNK, n_re - n_rb)) n_rec = recs[n_rb + tid];
        };
        auto issue_next_pairs = [&]() {
            if (pf_oe) return;
            issue_next_records();
            pf_oe = true;
            if (tid == 0) nn_item = listed(q_next);
            if (tid < min<uint64_t>(BIN3_RCHUNK, n_re - n_rb)) n_oe = act_oe[rec_read(n_rec, pk)];
        };
        if (rb == re) goto next_bin;
        {
        const uint32_t lib = (uint32_t)item >> bits;
        const uint64_t ev0 = libs[lib].ev0;
        __syncthreads();
        if (threadIdx.x == 0) {
            s_np = 1;
            s_pass[0] = 1;
            s_pass[1] = 0;
        }
        __syncthreads();
        bool one_chunk = false;  // (uniform) the bin's records fit one staged chunk
        for (uint32_t pass_no = 0; s_np; ++pass_no) {
            uint32_t parts = s_pass[2 * (s_np - 1)];
            const uint32_t part = s_pass[2 * (s_np - 1) + 1];
            __syncthreads();
            if (threadIdx.x == 0) {
                s_fill = 0;
                s_over = 0;
                --s_np;
            }
            bool gave_up = false;  // (uniform) the bin needs more passes than allowed
            for (uint64_t c0 = rb; c0 < re && !gave_up;) {
                const uint32_t nr = (uint32_t)min<uint64_t>(BIN3_RCHUNK, re - c0);
                __syncthreads();  // the previous chunk's staging reads are done
                // the first chunk of the first pass is already in LDS; a bin of one chunk keeps
                // it staged for its later passes
                const bool first = c0 == rb;
                if (!(first && pass_no > 0 && one_chunk)) {
                    // stage one record per thread; scan of the window counts (waves, then
                    // across); the chunk takes the leading records whose windows fit the map
                    const uint32_t i = threadIdx.x, lane = i & 63, wave = i >> 6;
                    uint32_t n0 = 0;
                    if (i < nr) {
                        if (first && pass_no == 0) {
                            n0 = s_rec[i].woff;
                        } else {
                            const uint64_t rec = recs[c0 + i];
                            const uint32_t fw = rec_first(rec, pk);
                            n0 = rec_windows(rec, pk);
                            const ulonglong2 oe = act_oe[rec_read(rec, pk)];
                            s_rec[i].pos = oe.x * 4 + fw;
                            s_rec[i].ev = (Ev)(oe.y + fw - ev0);
                        }
                    }
                    uint32_t inc = n0;
#pragma unroll
                    for (uint32_t d = 1; d < 64; d <<= 1) {
                        const uint32_t v = (uint32_t)__shfl_up((int)inc, d, 64);
                        if (lane >= d) inc += v;
                    }
                    if (lane == 63) s_wsum[wave] = inc;
                    __syncthreads();
                    uint32_t before = 0;
#pragma unroll
                    for (uint32_t w = 0; w < BLOCK / 64; ++w) before += w < wave ? s_wsum[w] : 0u;
                    const uint32_t end = before + inc;  // the windows up to and including record i
                    const bool fits = i < nr && end <= BIN3_MAP;  // a prefix of the records
                    const uint64_t fit = __ballot(fits);
                    if (lane == 0) s_wfit[wave] = (uint32_t)__popcll(fit);
                    if (i < nr) s_rec[i].woff = end - n0;
                    __syncthreads();
                    uint32_t nfit = 0;
#pragma unroll
                    for (uint32_t w = 0; w < BLOCK / 64; ++w) nfit += s_wfit[w];
                    // nfit >= 1: a record has at most SK_MAX_WINDOWS <= BIN3_MAP windows
                    if (fits) {
                        for (uint32_t j = end - n0; j < end; ++j) s_map[j] = (uint8_t)i;
                        if (i + 1 == nfit) {
                            s_nwin = end;
                            s_nr = nfit;
                        }
                    }
                }
                __syncthreads();
                const uint32_t nwin = s_nwin;
                c0 += s_nr;
                if (first && pass_no == 0) one_chunk = c0 >= re;
                issue_next_records();
                // window x: its stream base number and its event (one read of its record)
                auto locate = [&](uint32_t x, uint64_t& pos, Ev& ev) {
                    const StagedRec<Ev> r = s_rec[s_map[x]];
                    pos = r.pos + (x - r.woff);
                    ev = (Ev)(r.ev + (x - r.woff));
                };
                // Each thread takes windows x = tid, tid + 256, ...; a full table stops every
                // thread at its first window not counted, and counting resumes there.
                uint32_t x = threadIdx.x;
                for (;;) {
                    // The packed words of the next window are fetched one iteration ahead, so
                    // their load is in flight while this window's LDS work runs.
                    RawWindow<KW> raw;
                    uint64_t pos = 0;
                    Ev ev = 0;
                    if (x < nwin) {
                        locate(x, pos, ev);
                        raw_fetch<KW>(data, pos, raw);
                    }
                    for (; x < nwin; x += blockDim.x) {
                        // the full-table flag's load completes under the key's arithmetic
                        // instead of stalling the iteration
                        const uint32_t over =
                            __hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        const RawWindow<KW> cur = raw;
                        const Ev t = ev;
                        if (x + blockDim.x < nwin) {
                            locate(x + blockDim.x, pos, ev);
                            raw_fetch<KW>(data, pos, raw);
                        }
                        uint64_t key[NW];
                        pack_key<NW>(canonical<KW>(raw_extract<KW>(cur, k), k), key);
                        const uint32_t h = bin_slot_hash<NW>(key);
                        if (over) break;  // window x is not counted
                        if (parts > 1 && ((h >> 20) & (parts - 1)) != part) continue;
                        uint32_t slot = h & (BIN3_CAP - 1);
                        bool done = false;
                        for (uint32_t probe = 0; probe < BIN3_CAP; ++probe) {
                            // both key words are loaded at once: an occupied slot's words are
                            // final once not EMPTY, and an EMPTY one is settled by the CAS
                            uint64_t w0 =
                                __hip_atomic_load(&s_k0[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                            uint64_t w1 = NW == 2 ? __hip_atomic_load(&s_k1[slot], __ATOMIC_RELAXED,
                                                                      __HIP_MEMORY_SCOPE_WORKGROUP)
                                                  : 0;
                            bool claimed = false;
                            bool hit = lds_claim_loaded(&s_k0[slot], w0, key[0], claimed);
                            if (claimed) {
                                const uint32_t f = atomicAdd(&s_fill, 1u);
                                if (f < BIN3_FILL) s_list[f] = (uint16_t)slot;
                                else atomicOr(&s_over, 1u);
                            }
                            if (hit && NW == 2) {
                                bool c2 = false;
                                hit = lds_claim_loaded(&s_k1[slot], w1, key[NW - 1], c2);
                            }
                            if (hit) {
                                // neither returns a value, so the window waits on neither
                                atomicAdd(&s_cnt[slot], 1u);
                                atomicMin(&s_t0[slot], t);
                                done = true;
                                break;
                            }
                            slot = (slot + 1) & (BIN3_CAP - 1);
                        }
                        if (!done) {  // no free slot: window x is not counted
                            atomicOr(&s_over, 1u);
                            break;
                        }
                    }
                    issue_next_pairs();
                    __syncthreads();
                    if (!s_over) break;
                    // The table is full.  The keys of this pass whose next key-hash bit is 0
                    // stay (with their counts and first events) and go on counting from where
                    // each thread stopped; those with the bit set get a pass of their own.
                    if (2 * parts > BIN3_MAX_SPLIT || s_np + 1 > 32) {
                        __syncthreads();
                        clear_table();
                        if (threadIdx.x == 0) {
                            atomicOr(reinterpret_cast<unsigned int*>(&ctr[1]), 1u);
                            s_np = 0;
                        }
                        gave_up = true;
                        __syncthreads();
                        break;
                    }
                    // each thread holds its BIN3_CAP / BLOCK slots in registers across the clear
                    constexpr uint32_t PER = BIN3_CAP / BLOCK;
                    static_assert(BIN3_CAP % BLOCK == 0, "slots a thread keeps");
                    Slot<NW, Ev> keep[PER];
#pragma unroll
                    for (uint32_t u = 0; u < PER; ++u) {
                        const uint32_t sl = threadIdx.x + u * BLOCK;
                        keep[u].k[0] = s_k0[sl];
                        if (NW == 2) keep[u].k[NW - 1] = s_k1[sl];
                        keep[u].t0 = s_t0[sl];
                        keep[u].cnt = s_cnt[sl];
                        if (keep[u].k[0] != EMPTY_KEY && ((bin_slot_hash<NW>(keep[u].k) >> 20) & parts))
                            keep[u].k[0] = EMPTY_KEY;
                    }
                    __syncthreads();
                    clear_table();
                    if (threadIdx.x == 0) {
                        s_fill = 0;
                        s_over = 0;
                        s_pass[2 * s_np] = 2 * parts;
                        s_pass[2 * s_np + 1] = part + parts;
                        ++s_np;
                    }
                    parts *= 2;
                    __syncthreads();
                    // put the kept keys back (distinct keys: a slot is taken by the CAS of its first word)
#pragma unroll
                    for (uint32_t u = 0; u < PER; ++u) {
                        const Slot<NW, Ev> e = keep[u];
                        if (e.k[0] == EMPTY_KEY) continue;
                        uint32_t slot = bin_slot_hash<NW>(e.k) & (BIN3_CAP - 1);
                        // (at most BIN3_CAP keys were kept, so one of the BIN3_CAP slots tried
                        // is free)
                        for (uint32_t probe = 0; probe < BIN3_CAP &&
                                                 atomicCAS(reinterpret_cast<unsigned long long*>(&s_k0[slot]),
                                                           (unsigned long long)EMPTY_KEY,
                                                           (unsigned long long)e.k[0]) != EMPTY_KEY;
                             ++probe)
                            slot = (slot + 1) & (BIN3_CAP - 1);
                        if (NW == 2) s_k1[slot] = e.k[NW - 1];
                        s_cnt[slot] = e.cnt;
                        s_t0[slot] = e.t0;
                        const uint32_t f = atomicAdd(&s_fill, 1u);
                        if (f < BIN3_FILL) s_list[f] = (uint16_t)slot;
                        else atomicOr(&s_over, 1u);
                    }
                    __syncthreads();
                }
            }
            __syncthreads();
            if (gave_up) continue;
            if (threadIdx.x == 0) {
                s_base = atomicAdd(&lcnt[lib], (unsigned long long)s_fill);
                if (s_base + s_fill > capl[lib]) {
                    atomicOr(reinterpret_cast<unsigned int*>(&ctr[2]), 1u);
                    s_base = ~0ull;
                } else {
                    s_base += cap0[lib];
                }
            }
            __syncthreads();
            const unsigned long long base = s_base;
            for (uint32_t f = threadIdx.x; f < s_fill; f += blockDim.x) {
                const uint32_t sl = s_list[f];
                if (base != ~0ull) {
                    Slot<NW, Ev> e;
                    e.k[0] = s_k0[sl];
                    if (NW == 2) e.k[NW - 1] = s_k1[sl];
                    e.t0 = s_t0[sl];
                    e.cnt = s_cnt[sl];
                    out[base + f] = e;
                }
                s_k0[sl] = EMPTY_KEY;
                if (NW == 2) s_k1[sl] = EMPTY_KEY;
                s_cnt[sl] = 0;
                s_t0[sl] = (Ev)~(Ev)0;
            }
            __syncthreads();
        }
        }
    next_bin:
        issue_next_pairs();  // (an empty bin, or one whose passes all ended early)
        // (every read of the staging arrays is behind the last barrier)
        if (tid < min<uint64_t>(BIN3_RCHUNK, n_re - n_rb)) prestage(n_rec, n_oe, n_ev0);
        if (tid == 0) s_bin[0] = nn_item;
        __syncthreads();
        item = nitem;
        nitem = s_bin[0];
        c_rb = n_rb;
        c_re = n_re;
        n_v = load_range(nitem);
    }
}

// Packed entries (binned count) of library `lib` re-inserted with hash placement into its region:
// each key is distinct, so its claimed slot has one writer for t0 and cnt.
template <int NW, typename Ev>
__global__ void k_rehash(const Slot<NW, Ev>* __restrict__ src, uint64_t n, Slot<NW, Ev>* __restrict__ tab,
                         const LibDesc* __restrict__ libs, uint32_t lib, unsigned long long* __restrict__ overflow) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    const LibDesc& L = libs[lib];
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const Slot<NW, Ev> e = src[i];
        if (e.k[0] == EMPTY_KEY) continue;
        uint64_t key[NW], w0[SPB];
        for (int w = 0; w < NW; ++w) key[w] = e.k[w];
        const uint64_t b = L.bucket0 + home_bucket<NW>(key, L.nbuckets);
        bucket_load<NW, Ev>(tab, b, w0);
        const uint64_t j = table_find_or_claim<NW, Ev>(tab, L.bucket0, L.nbuckets, b, w0, key, 1u << 20);
        if (j == ~0ull) {
            atomicOr(reinterpret_cast<unsigned int*>(overflow), 1u);
            continue;
        }
        tab[j].t0 = e.t0;
        tab[j].cnt = e.cnt;
    }
}

// Iterates the slots [s_begin, s_end) of the table pool in chunks of BLOCK * SCAN_ITEMS;
// f(slot index, library, chunk's first library) for every occupied slot.
template <int NW, typename Ev, class F, class G>
__device__ __forceinline__ void scan_slots(const Slot<NW, Ev>* __restrict__ tab, uint64_t s_begin, uint64_t s_end,
                                           const uint64_t* __restrict__ lib_b0, uint32_t n_libs, F&& f,
                                           G&& chunk_end) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    const uint64_t chunk = (uint64_t)BLOCK * SCAN_ITEMS;
    for (uint64_t c0 = s_begin + blockIdx.x * chunk; c0 < s_end; c0 += (uint64_t)gridDim.x * chunk) {
        const uint32_t lib0 = find_region(lib_b0, n_libs, c0 / SPB);
        uint64_t lib_end = (lib0 + 1 < n_libs) ? lib_b0[lib0 + 1] * SPB : ~0ull;
        if (lib_end >= min(c0 + chunk, s_end)) {
            // the chunk lies in one library: f sees a workgroup-uniform library, so its
            // descriptor loads are scalar and leave the per-slot dependency chain
            for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
                const uint64_t i = c0 + (uint64_t)j * BLOCK + threadIdx.x;
                if (i >= s_end) break;
                f(i, lib0, lib0);
            }
        } else {
            uint32_t lib = lib0;
            for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
                const uint64_t i = c0 + (uint64_t)j * BLOCK + threadIdx.x;
                if (i >= s_end) break;
                if (i >= lib_end) {
                    lib = find_region(lib_b0, n_libs, i / SPB);
                    lib_end = (lib + 1 < n_libs) ? lib_b0[lib + 1] * SPB : ~0ull;
                }
                f(i, lib, lib0);
            }
        }
        chunk_end(lib0);
    }
}

// fs[p] = min(4*t0 + j) over the distinct k-mers' bits; counts distinct k-mers per library.
// SEEN (the sparse mode): also sets the seen bits directly instead of deriving them from fs.
template <int NW, typename Ev, bool SEEN>
__global__ void __launch_bounds__(BLOCK) k_first_setter(const Slot<NW, Ev>* __restrict__ tab, uint64_t s_begin,
                                                        uint64_t s_end, const uint64_t* __restrict__ lib_b0,
                                                        uint32_t n_libs, const LibDesc* __restrict__ libs,
                                                        LibOut* __restrict__ out) {
    __shared__ unsigned long long s_acc[LIB_LOCAL];
    LibAccum acc{s_acc};
    acc.reset();
    __syncthreads();
    scan_slots<NW, Ev>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t lib0) {
            const Slot<NW, Ev> s = tab[i];
            if (s.k[0] == EMPTY_KEY) return;
            const LibDesc& L = libs[lib];
            const Kv v = unpack_key<NW>(s.k);
            uint64_t p[3];
            kv_positions<kw_of_nw(NW)>(v, L.k, L.M, p);
            Ev* fs = reinterpret_cast<Ev*>(L.fs);
            const Ev base = (Ev)(s.t0 * 4);
#pragma unroll
            for (int j = 0; j < 3; ++j) atomicMin(&fs[p[j]], (Ev)(base + j));
            if (SEEN) {
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    atomicOr(reinterpret_cast<unsigned long long*>(&L.seen[p[j] >> 6]), 1ull << (p[j] & 63));
            }
            acc.add(lib, lib0, 1, &out[0].n_distinct, LIBOUT_STRIDE);
        },
        [&](uint32_t lib0) {
            __syncthreads();
            acc.flush(lib0, n_libs, &out[0].n_distinct, LIBOUT_STRIDE);
            __syncthreads();
            acc.reset();
            __syncthreads();
        });
}

// seen word w (pool index) = the 64 bits (fs != inf) of its 64 first-setter entries, read
// as 16-byte vectors by one thread; clears the solid word.
template <typename Ev>
__global__ void __launch_bounds__(BLOCK) k_seen_from_fs(uint64_t w_begin, uint64_t w_end,
                                                        const uint64_t* __restrict__ lib_w0, uint32_t n_libs,
                                                        const LibDesc* __restrict__ libs) {
    constexpr int PER = 16 / sizeof(Ev);       // entries per 16-byte load
    constexpr int LOADS = 64 / PER;
    for (uint64_t w = w_begin + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < w_end;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lib = find_region(lib_w0, n_libs, w);
        const LibDesc& L = libs[lib];
        const uint64_t lw = w - L.word0;
        const uint4* src = reinterpret_cast<const uint4*>(reinterpret_cast<const Ev*>(L.fs) + lw * 64);
        uint4 v[LOADS];
#pragma unroll
        for (int q = 0; q < LOADS; ++q) v[q] = src[q];
        uint64_t m = 0;
#pragma unroll
        for (int q = 0; q < LOADS; ++q) {
            const uint32_t x[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
            if (sizeof(Ev) == 4) {
#pragma unroll
                for (int e = 0; e < 4; ++e) m |= (uint64_t)(x[e] != 0xFFFFFFFFu) << (q * 4 + e);
            } else {
#pragma unroll
                for (int e = 0; e < 2; ++e)
                    m |= (uint64_t)((x[2 * e] & x[2 * e + 1]) != 0xFFFFFFFFu) << (q * 2 + e);
            }
        }
        L.seen[lw] = m;
        L.solid[lw] = 0;
    }
}

// Sorted first setters (dense mode, 32-bit events, a group array of under 2^32 entries).
// Item of (distinct k-mer x, hash j) = e << 32 | (4*t0(x) + j), with e the bit's entry in the
// group's first-setter array; slot i of the scanned range writes items 3i .. 3i+2, and an
// empty slot writes ~0, whose partition key sorts past every partition.  Sorted by partition
// (e >> FS_PART_LOG2), each partition's minima are taken in LDS by one workgroup and its
// first setters and seen words written in order, in place of three random device-scope
// atomicMin per k-mer.  The minimum does not depend on the order of the items, so the
// result is the one of the atomics.  When the partition keys end below bit 63 (MULTI), bit
// 63 of an item marks a k-mer counted at least twice, which always promotes: k_fs_min
// then also writes those k-mers' solid bits, and k_promote skips them.
constexpr uint32_t FS_PART_LOG2 = 14;  // entries per partition: 64 KiB of LDS
constexpr uint32_t FS_PART = 1u << FS_PART_LOG2;
constexpr uint32_t FS_KEY_SHIFT = 32 + FS_PART_LOG2;
constexpr uint32_t FS_BLOCK = 1024;

template <int NW, bool MULTI>
__global__ void __launch_bounds__(BLOCK) k_fs_items(const Slot<NW, uint32_t>* __restrict__ tab, uint64_t s_begin,
                                                    uint64_t s_end, const uint64_t* __restrict__ lib_b0,
                                                    uint32_t n_libs, const LibDesc* __restrict__ libs,
                                                    LibOut* __restrict__ out, const uint8_t* __restrict__ fs_base,
                                                    uint64_t* __restrict__ items) {
    __shared__ unsigned long long s_acc[LIB_LOCAL];
    LibAccum acc{s_acc};
    acc.reset();
    __syncthreads();
    scan_slots<NW, uint32_t>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t lib0) {
            const Slot<NW, uint32_t> s = tab[i];
            uint64_t* o = items + 3 * (i - s_begin);
            if (s.k[0] == EMPTY_KEY) {
                o[0] = o[1] = o[2] = ~0ull;
                return;
            }
            const LibDesc& L = libs[lib];
            const Kv v = unpack_key<NW>(s.k);
            uint64_t p[3];
            kv_positions<kw_of_nw(NW)>(v, L.k, L.M, p);
            const uint64_t e0 = (uint64_t)(reinterpret_cast<const uint8_t*>(L.fs) - fs_base) / sizeof(uint32_t);
            const uint32_t base = s.t0 * 4u;
            const uint64_t multi = MULTI && s.cnt >= 2 ? 1ull << 63 : 0ull;
#pragma unroll
            for (int j = 0; j < 3; ++j) o[j] = multi | (e0 + p[j]) << 32 | (uint32_t)(base + j);
            acc.add(lib, lib0, 1, &out[0].n_distinct, LIBOUT_STRIDE);
        },
        [&](uint32_t lib0) {
            __syncthreads();
            acc.flush(lib0, n_libs, &out[0].n_distinct, LIBOUT_STRIDE);
            __syncthreads();
            acc.reset();
            __syncthreads();
        });
}

// bounds[b] = the first sorted item whose partition (key bits under kmask) is >= b, for
// b = 0 .. nparts.
__global__ void k_fs_bounds(const uint64_t* __restrict__ items, uint64_t n, uint32_t nparts, uint32_t shift,
                            uint32_t kmask, uint64_t* __restrict__ bounds) {
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nparts; b += gridDim.x * blockDim.x) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (((uint32_t)(items[mid] >> shift) & kmask) < b) lo = mid + 1;
            else hi = mid;
        }
        bounds[b] = lo;
    }
}

// One partition per workgroup pass: the minimum setter of each of its FS_PART entries in
// LDS (and, MULTI, the solid bits of the items marked counted twice), then, one wave per
// 64 entries, the entries to fs, their seen word (entry != ~0) by ballot, and the solid
// word.  fs_n is a multiple of 64.
template <bool MULTI>
__global__ void __launch_bounds__(FS_BLOCK) k_fs_min(const uint64_t* __restrict__ items,
                                                     const uint64_t* __restrict__ bounds, uint32_t nparts,
                                                     uint32_t* __restrict__ fs, uint64_t fs_n,
                                                     uint64_t* __restrict__ seen, uint64_t* __restrict__ solid) {
    __shared__ uint32_t m[FS_PART];
    __shared__ uint32_t sol[MULTI ? FS_PART / 32 : 1];
    __shared__ unsigned long long sw[FS_PART / 64];  // a sparse partition's seen words
    constexpr uint32_t U = 4;  // items in flight per thread
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t part = blockIdx.x; part < nparts; part += gridDim.x) {
        const uint64_t b = bounds[part], end = bounds[part + 1];
        const uint64_t e0 = (uint64_t)part * FS_PART;
        const uint32_t nw = (uint32_t)(min<uint64_t>(FS_PART, fs_n - e0) / 64);
        if (end == b) {  // no item: no entry is seen
            for (uint32_t w = threadIdx.x; w < nw; w += FS_BLOCK) {
                seen[e0 / 64 + w] = 0;
                solid[e0 / 64 + w] = 0;
            }
            continue;
        }
        if (end - b < FS_PART / 8) {
            // A sparse partition (small libraries on large tiers) touches only its items'
            // entries and builds its seen (and counted-twice solid) words from the items:
            // the promotion pass reads fs back only at its items' entries.
            for (uint64_t i = b + threadIdx.x; i < end; i += FS_BLOCK) m[(uint32_t)(items[i] >> 32) & (FS_PART - 1)] = ~0u;
            for (uint32_t w = threadIdx.x; w < FS_PART / 64; w += FS_BLOCK) sw[w] = 0;
            if (MULTI)
                for (uint32_t e = threadIdx.x; e < FS_PART / 32; e += FS_BLOCK) sol[e] = 0;
            __syncthreads();
            for (uint64_t i = b + threadIdx.x; i < end; i += FS_BLOCK) {
                const uint64_t it = items[i];
                const uint32_t e = (uint32_t)(it >> 32) & (FS_PART - 1);
                atomicMin(&m[e], (uint32_t)it);
                atomicOr(&sw[e >> 6], 1ull << (e & 63));
                if (MULTI && (it >> 63)) atomicOr(&sol[e >> 5], 1u << (e & 31));
            }
            __syncthreads();
            for (uint64_t i = b + threadIdx.x; i < end; i += FS_BLOCK) {
                const uint32_t e = (uint32_t)(items[i] >> 32) & (FS_PART - 1);
                fs[e0 + e] = m[e];  // (items of one entry write the same minimum)
            }
            for (uint32_t w = threadIdx.x; w < nw; w += FS_BLOCK) {
                seen[e0 / 64 + w] = sw[w];
                solid[e0 / 64 + w] = MULTI ? (uint64_t)sol[2 * w] | (uint64_t)sol[2 * w + 1] << 32 : 0ull;
            }
            __syncthreads();
            continue;
        }
        for (uint32_t e = threadIdx.x; e < FS_PART; e += FS_BLOCK) m[e] = ~0u;
        if (MULTI)
            for (uint32_t e = threadIdx.x; e < FS_PART / 32; e += FS_BLOCK) sol[e] = 0;
        __syncthreads();
        for (uint64_t i0 = b; i0 < end; i0 += (uint64_t)FS_BLOCK * U) {
            uint64_t it[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t i = i0 + u * FS_BLOCK + threadIdx.x;
                it[u] = i < end ? items[i] : ~0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                if (it[u] == ~0ull) continue;
                const uint32_t e = (uint32_t)(it[u] >> 32) & (FS_PART - 1);
                atomicMin(&m[e], (uint32_t)it[u]);
                if (MULTI && (it[u] >> 63)) atomicOr(&sol[e >> 5], 1u << (e & 31));
            }
        }
        __syncthreads();
        for (uint32_t w = wave; w < nw; w += FS_BLOCK / 64) {
            const uint32_t v = m[w * 64 + lane];
            fs[e0 + w * 64 + lane] = v;
            const uint64_t bits = __ballot(v != ~0u);
            if (lane == 0) {
                seen[e0 / 64 + w] = bits;
                solid[e0 / 64 + w] = MULTI ? (uint64_t)sol[2 * w] | (uint64_t)sol[2 * w + 1] << 32 : 0ull;
            }
        }
        __syncthreads();
    }
}

// Sorted first setters with 64-bit events (dense mode).  The setter 4*t0 + j needs up to S
// bits (S from 4 x the group's largest library's windows; every setter is below 2^S - 1),
// so an item is e << S | setter with the entry e in the top 64 - S bits.  Slot i of the
// scanned range writes items 3i .. 3i+2; an empty slot writes the group's last entry with
// setter 2^S - 1, which no real setter reaches, so the minimum of an entry that only such
// items reach is read back as "no setter".  Partitions are 8192 entries (their u64 minima
// fill 64 KiB of LDS).  The caller cuts the table's slots into chunks, each sorted and
// folded into fs in turn (MERGE: the partition's minima start from fs).
constexpr uint32_t FS64_PART_LOG2 = 13;
constexpr uint32_t FS64_PART = 1u << FS64_PART_LOG2;

template <int NW>
__global__ void __launch_bounds__(BLOCK) k_fs_items64(const Slot<NW, uint64_t>* __restrict__ tab, uint64_t s_begin,
                                                      uint64_t s_end, const uint64_t* __restrict__ lib_b0,
                                                      uint32_t n_libs, const LibDesc* __restrict__ libs,
                                                      LibOut* __restrict__ out, const uint8_t* __restrict__ fs_base,
                                                      uint32_t S, uint64_t last_entry, uint64_t* __restrict__ items) {
    __shared__ unsigned long long s_acc[LIB_LOCAL];
    LibAccum acc{s_acc};
    acc.reset();
    __syncthreads();
    scan_slots<NW, uint64_t>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t lib0) {
            const Slot<NW, uint64_t> s = tab[i];
            uint64_t* o = items + 3 * (i - s_begin);
            if (s.k[0] == EMPTY_KEY) {
                o[0] = o[1] = o[2] = last_entry << S | ((1ull << S) - 1);
                return;
            }
            const LibDesc& L = libs[lib];
            const Kv v = unpack_key<NW>(s.k);
            uint64_t p[3];
            kv_positions<kw_of_nw(NW)>(v, L.k, L.M, p);
            const uint64_t e0 = (uint64_t)(reinterpret_cast<const uint8_t*>(L.fs) - fs_base) / sizeof(uint64_t);
            const uint64_t setter = s.t0 * 4;
#pragma unroll
            for (int j = 0; j < 3; ++j) o[j] = (e0 + p[j]) << S | (setter + j);
            acc.add(lib, lib0, 1, &out[0].n_distinct, LIBOUT_STRIDE);
        },
        [&](uint32_t lib0) {
            __syncthreads();
            acc.flush(lib0, n_libs, &out[0].n_distinct, LIBOUT_STRIDE);
            __syncthreads();
            acc.reset();
            __syncthreads();
        });
}

template <bool MERGE>
__global__ void __launch_bounds__(FS_BLOCK) k_fs_min64(const uint64_t* __restrict__ items,
                                                       const uint64_t* __restrict__ bounds, uint32_t nparts,
                                                       uint32_t S, uint64_t* __restrict__ fs, uint64_t fs_n) {
    __shared__ unsigned long long m[FS64_PART];
    constexpr uint32_t U = 4;  // items in flight per thread
    const uint64_t smask = (1ull << S) - 1;
    for (uint32_t part = blockIdx.x; part < nparts; part += gridDim.x) {
        const uint64_t e0 = (uint64_t)part * FS64_PART;
        const uint32_t ne = (uint32_t)min<uint64_t>(FS64_PART, fs_n - e0);
        for (uint32_t e = threadIdx.x; e < FS64_PART; e += FS_BLOCK) m[e] = MERGE && e < ne ? fs[e0 + e] : ~0ull;
        __syncthreads();
        const uint64_t b = bounds[part], end = bounds[part + 1];
        for (uint64_t i0 = b; i0 < end; i0 += (uint64_t)FS_BLOCK * U) {
            uint64_t it[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t i = i0 + u * FS_BLOCK + threadIdx.x;
                it[u] = i < end ? items[i] : ~0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint64_t i = i0 + u * FS_BLOCK + threadIdx.x;
                if (i >= end) continue;
                atomicMin(&m[(uint32_t)(it[u] >> S) & (FS64_PART - 1)], (unsigned long long)(it[u] & smask));
            }
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < ne; e += FS_BLOCK) {
            const uint64_t v = m[e];
            fs[e0 + e] = v == smask ? ~0ull : v;  // an empty slot's item is no setter
        }
        __syncthreads();
    }
}

// Sparse mode: puts back the "no setter" value at every first-setter entry the table's
// k-mers touched, so the array is all ~0 again for the next run.
template <int NW, typename Ev>
__global__ void __launch_bounds__(BLOCK) k_fs_reset(const Slot<NW, Ev>* __restrict__ tab, uint64_t s_begin,
                                                    uint64_t s_end, const uint64_t* __restrict__ lib_b0,
                                                    uint32_t n_libs, const LibDesc* __restrict__ libs) {
    scan_slots<NW, Ev>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t) {
            const Slot<NW, Ev> s = tab[i];
            if (s.k[0] == EMPTY_KEY) return;
            const LibDesc& L = libs[lib];
            const Kv v = unpack_key<NW>(s.k);
            uint64_t p[3];
            kv_positions<kw_of_nw(NW)>(v, L.k, L.M, p);
            Ev* fs = reinterpret_cast<Ev*>(L.fs);
#pragma unroll
            for (int j = 0; j < 3; ++j) fs[p[j]] = (Ev)~(Ev)0;
        },
        [](uint32_t) {});
}

// Solid tier: OR of the bits of every promoted distinct k-mer (bloom.rs:99-103).
template <int NW, typename Ev>
__global__ void __launch_bounds__(BLOCK) k_promote(const Slot<NW, Ev>* __restrict__ tab, uint64_t s_begin,
                                                   uint64_t s_end, const uint64_t* __restrict__ lib_b0,
                                                   uint32_t n_libs, const LibDesc* __restrict__ libs,
                                                   bool skip_multi) {
    scan_slots<NW, Ev>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t) {
            const Slot<NW, Ev> s = tab[i];
            if (s.k[0] == EMPTY_KEY) return;
            if (skip_multi && s.cnt >= 2) return;  // already set by the sorted first-setter pass
            const LibDesc& L = libs[lib];
            const Kv v = unpack_key<NW>(s.k);
            uint64_t p[3];
            kv_positions<kw_of_nw(NW)>(v, L.k, L.M, p);
            bool promote = s.cnt >= 2;
            if (!promote) {  // each later first setter is read only while the earlier ones are others'
                const Ev* fs = reinterpret_cast<const Ev*>(L.fs);
                const Ev base = (Ev)(s.t0 * 4);
                promote = fs[p[0]] != base;
                if (promote) promote = fs[p[1]] != (Ev)(base + 1);
                if (promote) promote = fs[p[2]] != (Ev)(base + 2);
            }
            if (promote) {
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    atomicOr(reinterpret_cast<unsigned long long*>(&L.solid[p[j] >> 6]), 1ull << (p[j] & 63));
            }
        },
        [](uint32_t) {});
}

// The promotion pass of the sorted first setters (items with the counted-twice mark), from
// the items themselves: the three items of slot i, 3i .. 3i+2 of the unsorted array, hold
// its entries and setters, so no key is read or hashed.  A k-mer counted once promotes iff
// none of its items is its entry's first setter; a marked item's k-mer was promoted by the
// partition pass, an empty slot's items are all ones.
template <int NW>
__global__ void __launch_bounds__(BLOCK) k_promote_items(const Slot<NW, uint32_t>* __restrict__ tab,
                                                         uint64_t s_begin, uint64_t s_end,
                                                         const uint64_t* __restrict__ lib_b0, uint32_t n_libs,
                                                         const LibDesc* __restrict__ libs,
                                                         const uint8_t* __restrict__ fs_base,
                                                         const uint64_t* __restrict__ items) {
    const uint32_t* fs = reinterpret_cast<const uint32_t*>(fs_base);
    scan_slots<NW, uint32_t>(
        tab, s_begin, s_end, lib_b0, n_libs,
        [&](uint64_t i, uint32_t lib, uint32_t) {
            const uint64_t* it = items + 3 * (i - s_begin);
            const uint64_t a = it[0];
            if (a >> 63) return;  // empty slot (all ones) or counted twice
            const uint64_t b = it[1], c = it[2];
            bool promote = fs[(a >> 32) & 0x7FFFFFFFu] != (uint32_t)a;
            if (promote) promote = fs[(b >> 32) & 0x7FFFFFFFu] != (uint32_t)b;
            if (promote) promote = fs[(c >> 32) & 0x7FFFFFFFu] != (uint32_t)c;
            if (promote) {
                const LibDesc& L = libs[lib];
                const uint64_t e0 = (uint64_t)(reinterpret_cast<const uint8_t*>(L.fs) - fs_base) / sizeof(uint32_t);
                for (const uint64_t x : {a, b, c}) {
                    const uint64_t p = ((x >> 32) & 0x7FFFFFFFu) - e0;
                    atomicOr(reinterpret_cast<unsigned long long*>(&L.solid[p >> 6]), 1ull << (p & 63));
                }
            }
        },
        [](uint32_t) {});
}

// Gather the count table: occupied slots whose k-mer is Bloom-solid (or every occupied
// slot when the table was built from Bloom-solid windows only), as (hi, lo, count) sort
// entries.  Each block reserves its output range with one atomic per chunk.
// With priors (psum/pflag not null, assemble.rs:266-283), a flagged slot is in the table
// whatever its Bloom state, and its count is (counted occurrences, if Bloom-solid) + the
// boost sum, in wrapping u32 arithmetic.
template <int NW, typename Ev, bool CHECK_SOLID>
__global__ void __launch_bounds__(BLOCK) k_compact(const Slot<NW, Ev>* __restrict__ tab, uint64_t s_end,
                                                   const uint64_t* __restrict__ lib_b0, uint32_t n_libs,
                                                   const LibDesc* __restrict__ libs, SortLayout lay,
                                                   uint64_t* __restrict__ out_hi, uint64_t* __restrict__ out_lo,
                                                   uint32_t* __restrict__ out_cnt,
                                                   unsigned long long* __restrict__ counter,
                                                   const uint32_t* __restrict__ psum,
                                                   const uint8_t* __restrict__ pflag,
                                                   WideEntry* __restrict__ out_wide, bool twice_solid = true) {
    __shared__ uint32_t s_scan[BLOCK / 64];
    __shared__ unsigned long long s_base;
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    const uint64_t chunk = (uint64_t)BLOCK * SCAN_ITEMS;
    const uint32_t tid = threadIdx.x;
    for (uint64_t c0 = blockIdx.x * chunk; c0 < s_end; c0 += (uint64_t)gridDim.x * chunk) {
        uint32_t bits = 0, counted = 0, cnt = 0;  // bits: in the table; counted: occurrences count
        // a chunk inside one library reads its descriptor once (scalar loads)
        const uint32_t lib0 = find_region(lib_b0, n_libs, c0 / SPB);
        const bool single = lib0 + 1 >= n_libs || lib_b0[lib0 + 1] * SPB >= min(c0 + chunk, s_end);
        for (uint32_t j = 0; j < SCAN_ITEMS; ++j) {
            const uint64_t i = c0 + (uint64_t)j * BLOCK + tid;
            if (i >= s_end) break;
            const Slot<NW, Ev> sl = tab[i];
            const bool occ = sl.k[0] != EMPTY_KEY;
            bool base = occ;
            if (CHECK_SOLID && base) {
                // a slot claimed only by a prior holds no occurrence
                const Kv v = unpack_key<NW>(sl.k);
                // a k-mer inserted twice found its three bits set the second time and
                // promoted (bloom.rs:95-103), so it is solid without a look (in the table of
                // the reads the filter was built from, not in a recount of corrected reads)
                if (twice_solid && sl.cnt >= 2) {
                    base = true;
                } else if (single) {
                    const LibDesc& L = libs[lib0];
                    base = sl.cnt > 0 && bloom_is_solid_lazy<kw_of_nw(NW)>(v, L.k, L.solid, L.M);
                } else {
                    const LibDesc& L = libs[find_region(lib_b0, n_libs, i / SPB)];
                    base = sl.cnt > 0 && bloom_is_solid_lazy<kw_of_nw(NW)>(v, L.k, L.solid, L.M);
                }
            }
            const bool take = base || (occ && pflag && pflag[i]);
            bits |= (uint32_t)take << j;
            counted |= (uint32_t)base << j;
            cnt += take;
        }
        uint32_t inc = cnt;  // inclusive scan: in the wave, then across the waves
        const uint32_t lane = tid & 63, wave = tid >> 6;
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t v = (uint32_t)__shfl_up((int)inc, d, 64);
            if (lane >= d) inc += v;
        }
        if (lane == 63) s_scan[wave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < BLOCK / 64; ++w) {
            const uint32_t t = s_scan[w];
            before += w < wave ? t : 0u;
            total += t;
        }
        if (tid == 0) s_base = total ? atomicAdd(counter, (unsigned long long)total) : 0ull;
        __syncthreads();
        uint64_t o = s_base + before + inc - cnt;
        while (bits) {
            const uint32_t j = __ffs(bits) - 1;
            bits &= bits - 1;
            const uint64_t i = c0 + (uint64_t)j * BLOCK + tid;
            const Slot<NW, Ev> sl = tab[i];
            const uint32_t lib = single ? lib0 : find_region(lib_b0, n_libs, i / SPB);
            if (lay.wide) {
                const Kv v = unpack_key<NW>(sl.k);
                out_wide[o] = WideEntry{{v.lo, v.hi, v.w2, v.w3}, lib, 0u};
            } else {
                uint64_t hi, lo;
                encode_entry(lay, lib, unpack_key<NW>(sl.k), hi, lo);
                out_lo[o] = lo;
                if (!lay.single) out_hi[o] = hi;
            }
            out_cnt[o] = ((counted >> j) & 1u ? sl.cnt : 0u) + (psum ? psum[i] : 0u);
            ++o;
        }
        __syncthreads();
    }
}

// Prior / trusted contigs (assemble.rs:266-283): every window of a prior whose k bases
// are all ACGT/acgt (Kmer256::from_bytes, dna.rs:299-321) adds its library's boost to the
// slot of its canonical k-mer, claiming the slot when the k-mer is absent, and flags the
// slot as present in the count table.  Window w belongs to prior p = the last with
// pw0[p] <= w (every listed prior has at least one window); it starts at byte
// pstart[p] + (w - pw0[p]) of seq.
template <int KW, typename Ev>
__global__ void k_prior(const uint8_t* __restrict__ seq, const uint64_t* __restrict__ pstart,
                        const uint64_t* __restrict__ pw0, const uint32_t* __restrict__ plib, uint32_t n_priors,
                        uint64_t n_win, uint32_t k, Slot<nw_of(KW), Ev>* __restrict__ tab,
                        const LibDesc* __restrict__ libs, uint32_t* __restrict__ psum, uint8_t* __restrict__ pflag,
                        LibOut* __restrict__ out) {
    constexpr int NW = nw_of(KW);
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n_win;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = find_region(pw0, n_priors, w);
        const uint8_t* s = seq + pstart[p] + (w - pw0[p]);
        Kv v{0, 0, 0, 0};
        bool ok = true;
        for (uint32_t i = 0; i < k; ++i) {
            const uint32_t c = s[i] | 0x20u;  // 'A' and 'a' both map to 'a', and so on
            uint64_t code;
            if (c == 'a') code = 0;
            else if (c == 'c') code = 1;
            else if (c == 'g') code = 2;
            else if (c == 't') code = 3;
            else {
                ok = false;
                break;
            }
            v.w3 = (v.w3 << 2) | (v.w2 >> 62);
            v.w2 = (v.w2 << 2) | (v.hi >> 62);
            v.hi = (v.hi << 2) | (v.lo >> 62);
            v.lo = (v.lo << 2) | code;
        }
        if (!ok) continue;
        uint64_t key[NW];
        pack_key<NW>(canonical<KW>(v, k), key);
        const uint32_t l = plib[p];
        const LibDesc& L = libs[l];
        const uint64_t b = L.bucket0 + home_bucket<NW>(key, L.nbuckets);
        uint64_t w0[SPB];
        bucket_load<NW, Ev>(tab, b, w0);
        const uint64_t i = table_find_or_claim<NW, Ev>(tab, L.bucket0, L.nbuckets, b, w0, key, 1u << 20);
        if (i == ~0ull) {
            atomicOr(&out[l].overflow, 1u);
            continue;
        }
        atomicAdd(&psum[i], L.boost);
        pflag[i] = 1;
    }
}

// Byte copy (the graph's stand-in for a device-to-device memcpy): 8-byte words, then the tail.
__global__ void k_copy_bytes(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t n) {
    const uint64_t nw = n / 8;
    const uint64_t* s8 = reinterpret_cast<const uint64_t*>(src);
    uint64_t* d8 = reinterpret_cast<uint64_t*>(dst);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nw; i += stride) d8[i] = s8[i];
    for (uint64_t i = nw * 8 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// The library of every sorted wide entry into the hi word array.
__global__ void k_wide_lib(const WideEntry* __restrict__ in, uint64_t* __restrict__ hi, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        hi[i] = in[i].lib;
}

// ---------------------------------------------------------------------------
// learned first pass: features of the count-table entries
// ---------------------------------------------------------------------------

constexpr int N_FEAT = 16;

// Sum of the k base qualities of every window of library `lib`, added to its k-mer's slot
// of the (uncorrected) table.  aqoff[a] = quality offset of active read a, or ~0 when
// the read has none.
template <int KW, typename Ev>
__global__ void __launch_bounds__(BLOCK) k_qsum(const uint8_t* __restrict__ data, const uint64_t* __restrict__ act_ev,
                                                const uint64_t* __restrict__ act_off,
                                                const uint32_t* __restrict__ act_lib,
                                                const uint32_t* __restrict__ tile_first,
                                                const uint64_t* __restrict__ aqoff, const uint8_t* __restrict__ quals,
                                                uint64_t n_active, uint64_t n_occ, uint64_t tile0, uint64_t tile1,
                                                uint32_t k, uint32_t lib, const Slot<nw_of(KW), Ev>* __restrict__ tab,
                                                const LibDesc* __restrict__ libs,
                                                unsigned long long* __restrict__ qsum) {
    constexpr int NW = nw_of(KW);
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    const LibDesc& L = libs[lib];
    for (uint64_t tile = tile0 + blockIdx.x; tile < tile1; tile += gridDim.x) {
        const uint64_t a0 = tile_first[tile];
        const uint64_t w0 = tile * TILE_WINDOWS;
        const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
        const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
        __syncthreads();
        for (uint64_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
            const uint32_t lo = tile_read(s_ev, na, w);
            if (act_lib[a0 + lo] != lib) continue;
            const uint64_t qo = aqoff[a0 + lo];
            if (qo == ~0ull) continue;
            const uint32_t pos = (uint32_t)(w - s_ev[lo]);
            uint64_t key[NW];
            pack_key<NW>(canonical<KW>(load_window<KW>(data, act_off[a0 + lo], pos, k), k), key);
            const uint64_t i =
                table_find<NW, Ev>(tab, L.bucket0, L.nbuckets, L.bucket0 + home_bucket<NW>(key, L.nbuckets), key);
            if (i == ~0ull) continue;
            uint32_t s = 0;
            for (uint32_t j = 0; j < k; ++j) s += quals[qo + pos + j];
            atomicAdd(&qsum[i], (unsigned long long)s);
        }
    }
}

// Raw count of a k-mer value in the library's table (0 when absent).
template <int KW, typename Ev>
__device__ __forceinline__ uint32_t table_count(const Slot<nw_of(KW), Ev>* __restrict__ tab, const LibDesc& L,
                                                const Kv& v, uint32_t k, uint64_t* slot) {
    constexpr int NW = nw_of(KW);
    uint64_t key[NW];
    pack_key<NW>(canonical<KW>(v, k), key);
    const uint64_t i = table_find<NW, Ev>(tab, L.bucket0, L.nbuckets, L.bucket0 + home_bucket<NW>(key, L.nbuckets), key);
    if (slot) *slot = i;
    return i == ~0ull ? 0u : tab[i].cnt;
}

// The N_FEAT features of count-table entries [e0, e0 + n) (one library): its count c, the
// mean quality of its occurrences' bases, and statistics of the raw counts of its 3k
// Hamming-1 neighbours (canonical), relative to the library's cutoff and 95th-percentile
// count.
template <int KW, typename Ev>
__global__ void __launch_bounds__(BLOCK) k_features(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo,
                                                    const uint32_t* __restrict__ cnt, uint64_t e0, uint64_t n,
                                                    SortLayout lay, uint32_t k, uint32_t lib,
                                                    const Slot<nw_of(KW), Ev>* __restrict__ tab,
                                                    const LibDesc* __restrict__ libs, const LibOut* __restrict__ outs,
                                                    const unsigned long long* __restrict__ qsum,
                                                    float* __restrict__ feat, const WideEntry* __restrict__ wide) {
    const LibDesc& L = libs[lib];
    const float cutoff = (float)outs[lib].cutoff, top = (float)max(outs[lib].top, 1u);
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = e0 + t;
        // entry key -> k-mer value (the inverse of encode_entry)
        Kv v{0, 0, 0, 0};
        if (lay.wide) {
            v.lo = wide[e].v[0];
            v.hi = wide[e].v[1];
        } else if (lay.single) {
            v.lo = 2 * k >= 64 ? lo[e] : (lo[e] & ((1ull << (2 * k)) - 1));
            v.hi = 0;
        } else if (2 * k <= 63) {
            v.lo = lo[e];
            v.hi = 0;
        } else {
            const uint32_t vb = 2 * k - 63;
            const uint64_t a = hi[e] & ((1ull << vb) - 1);
            v.lo = lo[e] | (a << 63);
            v.hi = a >> 1;
        }
        const float c = (float)cnt[e];
        uint64_t slot = ~0ull;
        (void)table_count<KW, Ev>(tab, L, v, k, &slot);
        const float q = (slot != ~0ull && c > 0.f) ? (float)qsum[slot] / (c * (float)k) : 0.f;
        uint32_t mx = 0, mx2 = 0, present = 0, above_cut = 0, above_c = 0;
        float sum = 0.f;
        for (uint32_t pos = 0; pos < k; ++pos) {
            const uint32_t sh = 2 * (k - 1 - pos);  // bits of base `pos` (first base most significant)
            const uint32_t cur = sh < 64 ? (uint32_t)(v.lo >> sh) & 3u : (uint32_t)(v.hi >> (sh - 64)) & 3u;
            for (uint32_t b = 0; b < 4; ++b) {
                if (b == cur) continue;
                Kv y = v;
                if (sh < 64) y.lo = (y.lo & ~(3ull << sh)) | ((uint64_t)b << sh);
                else y.hi = (y.hi & ~(3ull << (sh - 64))) | ((uint64_t)b << (sh - 64));
                const uint32_t ny = table_count<KW, Ev>(tab, L, y, k, nullptr);
                if (ny > mx) {
                    mx2 = mx;
                    mx = ny;
                } else if (ny > mx2) {
                    mx2 = ny;
                }
                present += ny > 0;
                above_cut += (float)ny >= cutoff;
                above_c += (float)ny > c;
                sum += (float)ny;
            }
        }
        float* f = feat + t * N_FEAT;
        const float lc = log2f(1.f + c), lm = log2f(1.f + (float)mx);
        f[0] = lc;
        f[1] = q / 40.f;
        f[2] = lm;
        f[3] = log2f(1.f + sum);
        f[4] = (float)present / (3.f * (float)k);
        f[5] = c / (c + (float)mx);
        f[6] = lc - log2f(1.f + cutoff);
        f[7] = c >= cutoff ? 1.f : 0.f;
        f[8] = fminf((float)above_cut, 4.f) / 4.f;
        f[9] = lc - lm;
        f[10] = (q / 40.f) * (q / 40.f);
        f[11] = log2f(1.f + top) / 8.f;
        f[12] = c / top;
        f[13] = fminf((float)above_c, 4.f) / 4.f;
        f[14] = log2f(1.f + (float)mx2);
        f[15] = 1.f;
    }
}

// labels: the network's decision outside the margin, else the exact rule count >= cutoff.
__global__ void k_learned_labels(const float* __restrict__ p, const uint32_t* __restrict__ cnt, uint64_t e0,
                                 uint64_t n, const LibOut* __restrict__ outs, uint32_t lib, float margin,
                                 uint8_t* __restrict__ lab) {
    const uint32_t cutoff = outs[lib].cutoff;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const float pt = p[t];
        uint8_t v;
        if (pt >= 1.f - margin) v = 1;
        else if (pt <= margin) v = 0;
        else v = 2 | (cnt[e0 + t] >= cutoff ? 1 : 0);  // bit 1: the exact rule decided
        lab[t] = v;
    }
}

__global__ void k_aqoff(const uint32_t* __restrict__ active, uint64_t n_active, const uint64_t* __restrict__ qoff,
                        uint64_t* __restrict__ aqoff) {
    for (uint64_t a = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; a < n_active;
         a += (uint64_t)gridDim.x * blockDim.x)
        aqoff[a] = qoff[active[a]];
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

// Count-table order of two-word entries, most significant part first.  An entry's key is
// hi * 2^lo_bits + lo (lo_bits + hi_bits > 64 bits); its top 64 bits are
// hi << (64 - hi_bits) | lo >> r, r = lo_bits + hi_bits - 64, and the r bits below them are
// lo's low bits.  The entries are sorted by their top 64 bits (one 64-bit radix sort), then
// every run of equal top bits is put in order by its low bits.
__global__ void k_msd_keys(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo, uint64_t n,
                           uint32_t hi_bits, uint32_t r, uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        key[i] = hi[i] << (64 - hi_bits) | lo[i] >> r;
        idx[i] = (uint32_t)i;
    }
}

constexpr uint32_t TIE_SHORT = 16;    // runs a thread orders by insertion
constexpr uint32_t TIE_LONG = 1024;   // runs a workgroup orders in LDS; longer: the two-pass sort
constexpr uint32_t TIE_LONG_CAP = 1u << 16;

// Runs of equal top bits (keys sorted, idx the permutation): the run's head orders a short
// run by insertion on the low bits lo[idx] & low_mask (distinct entries: no two equal),
// lists a long one (start, length) for k_tie_long, and flags a run it cannot order.
__global__ void k_tie_short(const uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                            const uint64_t* __restrict__ lo, uint64_t n, uint64_t low_mask,
                            uint2* __restrict__ longs, unsigned int* __restrict__ ctr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i > 0 && key[i] == key[i - 1]) continue;
        if (i + 1 >= n || key[i + 1] != key[i]) continue;  // a run of one
        uint64_t j = i + 2;
        while (j < n && j - i <= TIE_LONG && key[j] == key[i]) ++j;
        const uint64_t len = j - i;
        if (len > TIE_LONG) {
            atomicOr(&ctr[1], 1u);  // too long: the caller sorts by both words instead
            continue;
        }
        if (len > TIE_SHORT) {
            const uint32_t at = atomicAdd(&ctr[0], 1u);
            if (at < TIE_LONG_CAP) longs[at] = make_uint2((uint32_t)i, (uint32_t)len);
            else atomicOr(&ctr[1], 1u);
            continue;
        }
        uint64_t lv[TIE_SHORT];
        uint32_t iv[TIE_SHORT];
        for (uint32_t a = 0; a < len; ++a) {
            const uint32_t x = idx[i + a];
            const uint64_t v = lo[x] & low_mask;
            uint32_t b = a;
            for (; b > 0 && lv[b - 1] > v; --b) {
                lv[b] = lv[b - 1];
                iv[b] = iv[b - 1];
            }
            lv[b] = v;
            iv[b] = x;
        }
        for (uint32_t a = 0; a < len; ++a) idx[i + a] = iv[a];
    }
}

// One listed run per workgroup pass: a bitonic sort of (low bits, index) pairs in LDS.
__global__ void __launch_bounds__(512) k_tie_long(uint32_t* __restrict__ idx, const uint64_t* __restrict__ lo,
                                                  uint64_t low_mask, const uint2* __restrict__ longs,
                                                  const unsigned int* __restrict__ ctr) {
    __shared__ uint64_t sv[TIE_LONG];
    __shared__ uint32_t si[TIE_LONG];
    const uint32_t nl = min(ctr[0], TIE_LONG_CAP);
    for (uint32_t r = blockIdx.x; r < nl; r += gridDim.x) {
        const uint2 run = longs[r];
        uint32_t np = 1;
        while (np < run.y) np <<= 1;
        for (uint32_t t = threadIdx.x; t < np; t += blockDim.x) {
            if (t < run.y) {
                const uint32_t x = idx[run.x + t];
                sv[t] = lo[x] & low_mask;
                si[t] = x;
            } else {
                sv[t] = ~0ull;  // pads sort last (low bits are below 2^63)
                si[t] = 0;
            }
        }
        __syncthreads();
        for (uint32_t size = 2; size <= np; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t t = threadIdx.x; t < np; t += blockDim.x) {
                    const uint32_t u = t ^ stride;
                    if (u > t) {
                        const bool up = (t & size) == 0;
                        if ((sv[t] > sv[u]) == up) {
                            const uint64_t tv = sv[t];
                            sv[t] = sv[u];
                            sv[u] = tv;
                            const uint32_t ti = si[t];
                            si[t] = si[u];
                            si[u] = ti;
                        }
                    }
                }
                __syncthreads();
            }
        for (uint32_t t = threadIdx.x; t < run.y; t += blockDim.x) idx[run.x + t] = si[t];
        __syncthreads();
    }
}

// Per-library histogram of min(count, HIST_BINS - 1) over the sorted entries, for
// determine_solid_kmer_cutoff (assemble.rs:692-735).  Entries are grouped by library, so
// a block keeps the histogram of its first entry's library in LDS.
__global__ void __launch_bounds__(BLOCK) k_count_hist(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo,
                                                      const uint32_t* __restrict__ cnt, uint64_t n, SortLayout lay,
                                                      uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[HIST_BINS];
    const uint64_t per_block = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = blockIdx.x * per_block, e = min(n, b + per_block);
    if (b >= e) return;
    const uint32_t lib0 = entry_lib(lay, lay.single ? 0 : hi[b], lo[b]);
    for (uint32_t i = threadIdx.x; i < HIST_BINS; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const uint32_t lib = entry_lib(lay, lay.single ? 0 : hi[i], lo[i]);
        const uint32_t bin = min(cnt[i], HIST_BINS - 1);
        if (lib == lib0) atomicAdd(&h[bin], 1u);
        else atomicAdd(&hist[(size_t)lib * HIST_BINS + bin], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < HIST_BINS; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[(size_t)lib0 * HIST_BINS + i], h[i]);
}

// determine_solid_kmer_cutoff (assemble.rs:692-735) per library from its histogram of
// min(count, HIST_BINS-1).  The 95th-percentile count `top` only enters through top >= 30
// and clamp(top/4, 10, 35), which the capped histogram decides exactly.
__device__ __host__ uint32_t cutoff_rule(const uint32_t* hist, uint64_t n, double min_cov, uint32_t* top_out) {
    *top_out = 0;
    if (min_cov > 5.0) {
        const double v = min_cov * 0.35 > 2.0 ? min_cov * 0.35 : 2.0;
        return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
    }
    if (n <= 100) return 2;
    const uint64_t rank = n * 95 / 100;
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
    uint32_t mvs = top / 4;
    mvs = mvs < 10 ? 10 : (mvs > 35 ? 35 : mvs);
    uint32_t valley = 2;
    uint64_t min_val = ~0ull;
    for (uint32_t c = 2; c <= mvs; ++c) {
        const uint64_t h = hist[c];
        if (h <= min_val) {
            min_val = h;
            valley = c;
        } else if (h > min_val * 2 && c > valley + 2) {
            return valley;
        }
    }
    return 2;
}

__global__ void k_cutoff(const uint32_t* __restrict__ hist, uint32_t n_libs, const double* __restrict__ min_cov,
                         LibOut* __restrict__ out) {
    for (uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; l < n_libs; l += gridDim.x * blockDim.x) {
        const uint32_t* h = hist + (size_t)l * HIST_BINS;
        uint64_t n = 0;
        for (uint32_t b = 0; b < HIST_BINS; ++b) n += h[b];
        uint32_t top = 0;
        out[l].cutoff = cutoff_rule(h, n, min_cov[l], &top);
        out[l].top = top;
        out[l].n_counts = n;
    }
}

__global__ void __launch_bounds__(BLOCK) k_solid_flags(const uint64_t* __restrict__ hi, const uint64_t* __restrict__ lo,
                                                       const uint32_t* __restrict__ cnt, uint64_t n, SortLayout lay,
                                                       uint32_t n_libs, const LibOut* __restrict__ cut,
                                                       uint8_t* __restrict__ f, LibOut* __restrict__ out) {
    __shared__ unsigned long long s_acc[LIB_LOCAL];
    LibAccum acc{s_acc};
    const uint64_t per_block = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t b = blockIdx.x * per_block, e = min(n, b + per_block);
    if (b >= e) return;
    acc.reset();
    __syncthreads();
    const uint32_t lib0 = entry_lib(lay, lay.single ? 0 : hi[b], lo[b]);
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const uint32_t lib = entry_lib(lay, lay.single ? 0 : hi[i], lo[i]);
        const bool s = cnt[i] >= cut[lib].cutoff;
        f[i] = s;
        if (s) acc.add(lib, lib0, 1, &out[0].n_solid, LIBOUT_STRIDE);
    }
    __syncthreads();
    acc.flush(lib0, n_libs, &out[0].n_solid, LIBOUT_STRIDE);
}

enum KernelId {
    KID_NWIN, KID_SCAN, KID_SELECT, KID_ACTIVE, KID_HLL, KID_CLEAR, KID_INSERT, KID_FS_CLEAR, KID_FIRST_SETTER,
    KID_SEEN, KID_PROMOTE, KID_HAMMER, KID_RECOUNT, KID_PRIOR, KID_COMPACT, KID_SORT, KID_HIST, KID_FLAGS,
    KID_SK, KID_BIN_SORT, KID_BIN_COUNT, KID_FS_ITEMS, KID_FS_SORT, KID_FS_MIN, KID_COUNT
};
const char* const KERNEL_NAMES[KID_COUNT] = {
    "k_nwin", "scan_windows", "select_active", "k_active", "k_hll", "k_table_clear", "k_insert", "k_fs_clear",
    "k_first_setter", "k_seen_from_fs", "k_promote", "k_hammer", "k_recount", "k_prior", "k_compact", "radix_sort",
    "k_count_hist", "k_solid_flags", "k_superkmers", "bin_sort", "k_bin_count",
    "k_fs_items", "fs_sort", "k_fs_min"};

// ---------------------------------------------------------------------------
// verification passes (kc_library_verify): closed forms recomputed window by window,
// independently of the first-setter construction and of the count's table layout
// ---------------------------------------------------------------------------

// Per library: every window's three seen bits ORed into `recount` (the seen tier is
// the OR over all occurrences, bloom.rs:95-97), and the windows of the counted reads
// (corrected when Hammer ran) whose canonical k-mer is Bloom-solid counted into
// solid_win (the count table holds exactly their multiplicities, assemble.rs:196).
template <int KW>
__global__ void __launch_bounds__(BLOCK) k_verify_windows(const uint8_t* __restrict__ data,
                                                          const uint8_t* __restrict__ counted,
                                                          const uint64_t* __restrict__ act_ev,
                                                          const uint64_t* __restrict__ act_off,
                                                          const uint32_t* __restrict__ act_lib,
                                                          const uint32_t* __restrict__ tile_first, uint64_t n_active,
                                                          uint64_t n_occ, uint64_t n_tiles, uint32_t k,
                                                          const LibDesc* __restrict__ libs,
                                                          unsigned long long* __restrict__ recount,
                                                          unsigned long long* __restrict__ solid_win) {
    __shared__ uint64_t s_ev[TILE_WINDOWS + 1];
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t a0 = tile_first[tile];
        const uint32_t lib0 = act_lib[a0];
        const uint64_t w0 = tile * TILE_WINDOWS;
        const uint64_t w1 = min(w0 + TILE_WINDOWS, n_occ);
        const uint32_t na = (uint32_t)min<uint64_t>(TILE_WINDOWS + 1, n_active - a0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < na; i += blockDim.x) s_ev[i] = act_ev[a0 + i];
        __syncthreads();
        uint32_t mine = 0;  // solid windows of lib0 seen by this thread
        for (uint64_t w = w0 + threadIdx.x; w < w1; w += blockDim.x) {
            const uint32_t lo = tile_read(s_ev, na, w);
            const uint32_t lib = act_lib[a0 + lo];
            const LibDesc& L = libs[lib];
            const uint32_t j = (uint32_t)(w - s_ev[lo]);
            const Kv v = canonical<KW>(load_window<KW>(data, act_off[a0 + lo], j, k), k);
            uint64_t p[3];
            kv_positions<KW>(v, k, L.M, p);
#pragma unroll
            for (int q = 0; q < 3; ++q) atomicOr(&recount[L.word0 + (p[q] >> 6)], 1ull << (p[q] & 63));
            const Kv u = counted == data ? v : canonical<KW>(load_window<KW>(counted, act_off[a0 + lo], j, k), k);
            if (bloom_is_solid<KW>(u, k, L.solid, L.M)) {
                if (lib == lib0) ++mine;
                else atomicAdd(&solid_win[lib], 1ull);
            }
        }
        for (uint32_t d = 32; d >= 1; d >>= 1) mine += __shfl_xor(mine, d);
        if ((threadIdx.x & 63) == 0 && mine) atomicAdd(&solid_win[lib0], (unsigned long long)mine);
    }
}

// Words [0, n) of the seen tier that differ from the recount.
__global__ void k_verify_seen(const uint64_t* __restrict__ seen, const unsigned long long* __restrict__ recount,
                              uint64_t n, unsigned long long* __restrict__ differing) {
    uint32_t mine = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        mine += seen[i] != recount[i];
    for (uint32_t d = 32; d >= 1; d >>= 1) mine += __shfl_xor(mine, d);
    if ((threadIdx.x & 63) == 0 && mine) atomicAdd(differing, (unsigned long long)mine);
}

// The k-mer value of sorted entry i (as the counts export decodes it).
__device__ __forceinline__ void entry_value(const SortLayout& lay, const uint64_t* __restrict__ hi,
                                            const uint64_t* __restrict__ lo, const WideEntry* __restrict__ wide,
                                            uint64_t i, uint64_t v[4]) {
    const uint32_t k = lay.k;
    v[2] = v[3] = 0;
    if (lay.wide) {
#pragma unroll
        for (int w = 0; w < 4; ++w) v[w] = wide[i].v[w];
    } else if (lay.single) {
        v[0] = 2 * k >= 64 ? lo[i] : (lo[i] & ((1ull << (2 * k)) - 1));
        v[1] = 0;
    } else if (2 * k <= 63) {
        v[0] = lo[i];
        v[1] = 0;
    } else {
        const uint32_t vb = 2 * k - 63;
        const uint64_t a = hi[i] & ((1ull << vb) - 1);
        v[0] = lo[i] | (a << 63);
        v[1] = a >> 1;
    }
}

// A 64-bit mix of (position, value, count); the table checksum is its sum over entries.
__device__ __host__ __forceinline__ uint64_t entry_mix(uint64_t i, const uint64_t v[4], uint32_t cnt) {
    auto f = [](uint64_t x) {
        x ^= x >> 33;
        x *= 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        x *= 0xc4ceb9fe1a85ec53ull;
        x ^= x >> 33;
        return x;
    };
    uint64_t x = f(i + 0x9E3779B97F4A7C15ull * (v[0] + 1));
    x = f(x ^ v[1]);
    x = f(x ^ v[2]);
    x = f(x ^ v[3]);
    return f(x ^ cnt);
}

// Over entries [e0, e0 + n): out[0] += counts, out[1] += pairs (i, i + 1) not strictly
// ascending, out[2] += solid flags that differ from count >= cutoff, out[3] += checksum.
__global__ void __launch_bounds__(BLOCK) k_verify_table(SortLayout lay, const uint64_t* __restrict__ hi,
                                                        const uint64_t* __restrict__ lo,
                                                        const WideEntry* __restrict__ wide,
                                                        const uint32_t* __restrict__ cnt,
                                                        const uint8_t* __restrict__ sflags, uint64_t e0, uint64_t n,
                                                        uint32_t cutoff, unsigned long long* __restrict__ out) {
    unsigned long long sum = 0, bad = 0, mism = 0, ck = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t v[4];
        entry_value(lay, hi, lo, wide, e0 + i, v);
        const uint32_t c = cnt[e0 + i];
        sum += c;
        mism += (sflags[e0 + i] != 0) != (c >= cutoff);
        ck += entry_mix(i, v, c);
        if (i + 1 < n) {
            uint64_t u[4];
            entry_value(lay, hi, lo, wide, e0 + i + 1, u);
            bool less = false, eq = true;
            for (int w = 3; w >= 0 && eq; --w)
                if (v[w] != u[w]) {
                    less = v[w] < u[w];
                    eq = false;
                }
            bad += !less;
        }
    }
    for (uint32_t d = 32; d >= 1; d >>= 1) {
        sum += __shfl_xor(sum, d);
        bad += __shfl_xor(bad, d);
        mism += __shfl_xor(mism, d);
        ck += __shfl_xor(ck, d);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], sum);
        if (bad) atomicAdd(&out[1], bad);
        if (mism) atomicAdd(&out[2], mism);
        atomicAdd(&out[3], ck);
    }
}

// A library as uploaded: its reads (host copy, PackedReads layout) and options.
struct HostLib {
    std::vector<uint8_t> data;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    uint64_t bloom_bits = 0;
    double min_cov = 5.0;
    // prior / trusted contigs (ASCII), applied to the count table (assemble.rs:244-284)
    std::vector<uint8_t> prior;
    std::vector<uint64_t> prior_start, prior_len;
    // per-base Phred qualities (features of the learned first pass): read i's bases at
    // quals[qoff[i] ..]
    std::vector<uint8_t> quals;
    std::vector<uint64_t> qoff;
    uint64_t prior_windows(uint32_t k) const {
        uint64_t n = 0;
        for (uint64_t len : prior_len) n += len >= k ? len - k + 1 : 0;
        return n;
    }
    // placement in the device pools (set by the planner)
    uint64_t read0 = 0, data0 = 0, word0 = 0, num_words = 0, bucket0 = 0, nbuckets = 0, fs_off = 0, est = 0;
    uint64_t ev0 = 0;  // first global window number (set by the reads preparation)
    uint32_t max_len = 0;  // longest read
    uint32_t est_k = 0;    // k of an exact `est` (binned count)
    uint32_t group = 0;
    kc_info info{};
    uint64_t entry0 = 0;  // first sorted entry
};

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
    std::vector<HostLib> libs;
    bool uploaded = false;  // device read buffers match `libs`
    // reads of all libraries, concatenated
    DevBuf d_data, d_off, d_len, d_rlib;
    uint64_t data_len = 0, n_reads = 0;
    // read bookkeeping
    DevBuf d_nwin, d_prefix, d_flags, d_iota, d_active, d_act_ev, d_act_off, d_act_lib, d_tile_first;
    DevBuf d_lib_read0, d_lib_ev0;
    uint64_t n_occ = 0, n_active = 0, n_tiles = 0;
    // filter pools, first setters, table pool
    DevBuf d_seen, d_solid, d_fs, d_tab, d_libs, d_lib_b0, d_lib_w0, d_blooms, d_out, d_mincov;
    uint64_t total_words = 0, total_buckets = 0, fs_bytes = 0;
    uint64_t fs_clean_bytes = 0;  // leading bytes of d_fs known to hold ~0 (sparse mode)
    bool bloom_sparse = false;    // first-setter mode of the last Bloom build
    std::vector<std::pair<uint32_t, uint32_t>> groups;  // [lib begin, lib end) sharing d_fs
    // count entries (sorted) + scratch
    DevBuf d_hi, d_lo, d_cnt, d_hi2, d_lo2, d_cnt2, d_idx, d_idx2, d_sflags;
    DevBuf d_wide, d_wide2;  // k > 63: count-table entries as (V, library) records
    // minimizer-binned Stage 2 count: super-k-mer bin keys / records (sorted copies), bin
    // bounds, per-read counts and offsets
    DevBuf d_bk, d_bk2, d_br, d_br2, d_bbounds, d_skc, d_sko, d_pack, d_wbin, d_bcap, d_bcnt;  // d_wbin: cut records
    DevBuf d_bflag, d_bitems, d_blist, d_bgmeta;  // binned count: its work items (ranges of bins)
    DevBuf d_ties;  // count-table sort: the long runs of equal top bits
    DevBuf d_fsi, d_fsi2, d_fsb;  // sorted first setters: items, their sort, partition bounds
    bool last_binned = false;     // the last Stage 2 left a packed (unhashed) table
    bool graph_planning = false;  // the eager run that plans a graph capture
    DevBuf d_corr, d_hll, d_hist, d_counters, d_temp, d_keys, d_out8;
    DevBuf d_vrec, d_vout;  // kc_library_verify: seen-tier recount, per-library results
    DevBuf d_act_oe;        // binned count: (byte offset, first window) per active read
    DevBuf d_pseq, d_pstart, d_pw0, d_plib, d_psum, d_pflag;  // prior contigs
    DevBuf d_quals, d_rqoff, d_aqoff, d_qsum, d_feat, d_prob, d_lab;  // learned first pass
    DevBuf d_gscratch[GRAPH_SCRATCH];  // unitig builder (kc_graph.hip), reused across calls
    uint64_t n_entries = 0;
    SortLayout lay{};
    kc_tune tune{1, 0, 0.0f, 0, 0, 0, 0, 0.0};  // launch shapes, table load, paths, bin size, m (kc_tune_set / kc_autotune)
    uint32_t n_cu = 0;
    uint32_t kw = 1;  // value words of the current k (1, 2 or 4)
    bool ev64 = false, built = false, counted = false;
    uint64_t* h_stage = nullptr;  // small pinned staging
    std::vector<LibOut> h_out;
    // timing
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    size_t mark_idx = 0;
    double k_ms[KID_COUNT] = {0};
    uint32_t k_launches[KID_COUNT] = {0};

    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
    // Table slots per 64-byte bucket for the current k and event width.
    uint64_t slots_per_bucket() const {
        if (kw == 4) return 1;
        return (kw == 2 || ev64) ? 2 : 4;
    }
    bool timing() const { return (o.flags & KC_FLAG_TIMING) != 0 && !capturing; }
    // Whether `need` more device bytes are free now, with a 4 GiB margin.  Large transient
    // buffers are kept between runs (reallocating tens of GB costs hundreds of ms) unless
    // this says the next stage would lack the room.
    bool fits(uint64_t need) const {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return true;
        return free_b > need + (4ull << 30);
    }
    // graph replay (kc_graph_run)
    bool capturing = false, graph_ok = false;
    hipGraphExec_t gexec = nullptr;
    uint64_t g_entries = 0;         // count-table entries of the planned run
    // the planned run's binned count (a capture replays it): bins, records, per-library room
    // and distinct k-mers (also on the device for the replay guard)
    bool g_binned = false;        // the last Stage 2 took the binned count
    uint32_t g_memset_nodes = 0;  // memset nodes the last capture recorded (replaced by fills)
    uint32_t g_bin_bits = 0;
    uint64_t g_bin_nrec = 0, g_bin_items = 0;
    std::vector<uint32_t> h_bgmeta;  // binned count: the work items' group sizes, as uploaded
    std::vector<uint64_t> g_bin_cap, g_bin_lcnt;
    DevBuf d_glcnt;
    uint32_t g_np = 0;              // prior sequences / windows uploaded by the last count
    uint64_t g_nw = 0;
    void graph_release() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        graph_ok = false;
    }
    hipEvent_t new_event() {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        ev_pool.push_back(e);
        return e;
    }
    void mark_begin(int, hipEvent_t* b) {
        *b = nullptr;
        if (!timing()) return;
        while (mark_idx * 2 + 1 >= ev_pool.size()) new_event();
        *b = ev_pool[mark_idx * 2];
        (void)hipEventRecord(*b, st);
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
    uint32_t L() const { return (uint32_t)c->libs.size(); }

    // Concatenate every library's reads into the device read buffers.
    int upload() {
        hipStream_t st = c->st;
        uint64_t data_len = 0, n = 0;
        for (auto& lib : c->libs) {
            lib.data0 = data_len;
            lib.read0 = n;
            data_len += lib.data.size();
            n += lib.len.size();
        }
        if (n > 0xFFFFFFFFull) return fail(KC_EUNSUPPORTED, "more than 2^32-1 reads in one context");
        KC_TRY(c->d_data.ensure(data_len + DATA_PAD));
        KC_TRY(c->d_off.ensure(n * 8 + 8));
        KC_TRY(c->d_len.ensure(n * 4 + 4));
        KC_TRY(c->d_rlib.ensure(n * 4 + 4));
        k_fill<uint8_t><<<grid_for(data_len + DATA_PAD), BLOCK, 0, st>>>(c->d_data.as<uint8_t>(), data_len + DATA_PAD,
                                                                         (uint8_t)0);
        KC_TRY(hipGetLastError());
        std::vector<uint64_t> off;
        std::vector<uint32_t> rl;
        for (uint32_t l = 0; l < L(); ++l) {
            auto& lib = c->libs[l];
            const uint64_t m = lib.len.size();
            if (!lib.data.empty())
                KC_TRY(hipMemcpyAsync(c->d_data.as<uint8_t>() + lib.data0, lib.data.data(), lib.data.size(),
                                      hipMemcpyHostToDevice, st));
            if (m) {
                off.resize(m);
                for (uint64_t i = 0; i < m; ++i) off[i] = lib.off[i] + lib.data0;
                rl.assign(m, l);
                KC_TRY(hipMemcpyAsync(c->d_off.as<uint64_t>() + lib.read0, off.data(), m * 8, hipMemcpyHostToDevice, st));
                KC_TRY(hipMemcpyAsync(c->d_len.as<uint32_t>() + lib.read0, lib.len.data(), m * 4, hipMemcpyHostToDevice,
                                      st));
                KC_TRY(hipMemcpyAsync(c->d_rlib.as<uint32_t>() + lib.read0, rl.data(), m * 4, hipMemcpyHostToDevice, st));
                KC_TRY(hipStreamSynchronize(st));  // host vectors are reused
            }
        }
        KC_TRY(hipStreamSynchronize(st));
        c->data_len = data_len;
        c->n_reads = n;
        c->uploaded = true;
        return KC_OK;
    }

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
        KC_TRY(c->d_lib_read0.ensure((L() + 1) * 8));
        KC_TRY(c->d_lib_ev0.ensure((L() + 1) * 8));
        c->n_occ = c->n_active = c->n_tiles = 0;
        std::vector<uint64_t> r0(L() + 1);
        for (uint32_t l = 0; l < L(); ++l) r0[l] = c->libs[l].read0;
        r0[L()] = n;
        KC_TRY(hipMemcpyAsync(c->d_lib_read0.p, r0.data(), r0.size() * 8, hipMemcpyHostToDevice, st));
        if (n == 0) {
            for (auto& lib : c->libs) {
                lib.info.n_occ = 0;
                lib.ev0 = 0;
            }
            lib_ev0_.assign(L() + 1, 0);
            KC_TRY(hipStreamSynchronize(st));
            return KC_OK;
        }
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
        k_lib_ev0<<<1, 256, 0, st>>>(c->d_prefix.as<uint64_t>(), c->d_nwin.as<uint64_t>(), n,
                                     c->d_lib_read0.as<uint64_t>(), L(), c->d_lib_ev0.as<uint64_t>());
        KC_TRY(hipGetLastError());
        std::vector<uint64_t> ev0(L() + 1);
        KC_TRY(hipMemcpyAsync(ev0.data(), c->d_lib_ev0.p, (L() + 1) * 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipMemcpyAsync(c->h_stage + 2, c->d_counters.as<uint64_t>() + 8, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        c->n_occ = ev0[L()];
        c->n_active = c->h_stage[2];
        c->n_tiles = (c->n_occ + TILE_WINDOWS - 1) / TILE_WINDOWS;
        for (uint32_t l = 0; l < L(); ++l) {
            c->libs[l].info.n_occ = ev0[l + 1] - ev0[l];
            c->libs[l].ev0 = ev0[l];
        }
        lib_ev0_ = ev0;
        if (c->n_occ == 0) return KC_OK;
        if (c->n_tiles > 0xFFFFFFFFull) return fail(KC_EUNSUPPORTED, "too many windows for one launch");
        KC_TRY(c->d_act_ev.ensure(c->n_active * 8 + 8));
        KC_TRY(c->d_act_off.ensure(c->n_active * 8 + 8));
        KC_TRY(c->d_act_lib.ensure(c->n_active * 4 + 4));
        KC_TRY(c->d_tile_first.ensure(c->n_tiles * 4 + 4));
        TIMED(KID_ACTIVE, (k_active<<<grid_for(c->n_active), BLOCK, 0, st>>>(
                              c->d_active.as<uint32_t>(), c->n_active, c->d_prefix.as<uint64_t>(),
                              c->d_nwin.as<uint64_t>(), c->d_off.as<uint64_t>(), c->d_rlib.as<uint32_t>(),
                              c->d_act_ev.as<uint64_t>(), c->d_act_off.as<uint64_t>(), c->d_act_lib.as<uint32_t>(),
                              c->d_tile_first.as<uint32_t>())));
        return KC_OK;
    }
    std::vector<uint64_t> lib_ev0_;

    // k_insert over every tile with the context's tuning (windows in flight per thread,
    // workgroups per CU; 0 = one workgroup per tile).
    template <int KW, typename Ev, bool TRACK, bool SOLID>
    int launch_insert(const uint8_t* data, int kid) {
        constexpr int NW = nw_of(KW);
        const uint64_t tiles = c->n_tiles;
        if (!tiles) return KC_OK;
        uint64_t g = tiles;
        if (c->tune.blocks_per_cu) g = std::min<uint64_t>(tiles, (uint64_t)c->n_cu * c->tune.blocks_per_cu);
        if (g > 0xFFFFFFFFull) return fail(KC_EUNSUPPORTED, "too many tiles for one launch");
        hipStream_t st = c->st;
#define KC_INSERT_LAUNCH(W)                                                                                       \
    TIMED(kid, (k_insert<KW, Ev, TRACK, SOLID, W><<<(uint32_t)g, BLOCK, 0, st>>>(                              \
                   data, c->d_act_ev.as<uint64_t>(), c->d_act_off.as<uint64_t>(), c->d_act_lib.as<uint32_t>(), \
                   c->d_tile_first.as<uint32_t>(), c->n_active, c->n_occ, tiles, c->o.k,                         \
                   c->d_tab.as<Slot<NW, Ev>>(), c->d_libs.as<LibDesc>(), 1u << 20, c->d_out.as<LibOut>())))
        if (c->tune.windows_per_thread == 2) KC_INSERT_LAUNCH(2);
        else if (c->tune.windows_per_thread == 1) KC_INSERT_LAUNCH(1);
        else KC_INSERT_LAUNCH(4);
#undef KC_INSERT_LAUNCH
        return KC_OK;
    }

    template <int KW>
    int estimate_distinct() {
        hipStream_t st = c->st;
        const size_t nreg = (size_t)L() * HLL_M;
        KC_TRY(c->d_hll.ensure(nreg * 4));
        k_fill<uint32_t><<<grid_for(nreg), BLOCK, 0, st>>>(c->d_hll.as<uint32_t>(), nreg, 0u);
        KC_TRY(hipGetLastError());
        const uint32_t g = (uint32_t)std::min<uint64_t>(c->n_tiles, 2048);
        TIMED(KID_HLL, (k_hll<KW><<<g, BLOCK, 0, st>>>(c->d_data.as<uint8_t>(), c->d_act_ev.as<uint64_t>(),
                                                         c->d_act_off.as<uint64_t>(), c->d_act_lib.as<uint32_t>(),
                                                         c->d_tile_first.as<uint32_t>(), c->n_active, c->n_occ,
                                                         c->n_tiles, c->o.k, c->d_hll.as<uint32_t>())));
        return hll_read();
    }

    // Distinct-count estimate of every library from the HyperLogLog registers in d_hll.
    int hll_read() {
        const size_t nreg = (size_t)L() * HLL_M;
        std::vector<uint32_t> regs(nreg);
        KC_TRY(hipMemcpyAsync(regs.data(), c->d_hll.p, nreg * 4, hipMemcpyDeviceToHost, c->st));
        KC_TRY(hipStreamSynchronize(c->st));
        for (uint32_t l = 0; l < L(); ++l) {
            double sum = 0;
            uint32_t zeros = 0;
            for (uint32_t i = 0; i < HLL_M; ++i) {
                const uint32_t r = regs[(size_t)l * HLL_M + i];
                sum += std::ldexp(1.0, -(int)r);
                zeros += r == 0;
            }
            const double m = HLL_M;
            double E = (0.7213 / (1.0 + 1.079 / m)) * m * m / sum;
            if (E <= 2.5 * m && zeros) E = m * std::log(m / zeros);
            c->libs[l].est = (uint64_t)std::ceil(E);
        }
        return KC_OK;
    }

    // Whether Stage 2 takes the minimizer-binned count: the default (insert_path 0) and
    // insert_path 3, k <= 63, reads shorter than 2^24; in a graph capture (and the eager
    // run that plans one) only without Hammer and priors.  Hammer's recount and the priors
    // hash into the table, so Stage 3 lays it out for hashing first (hashed_layout, rehash).
    bool use_binned() const {
        const uint32_t ip = c->tune.insert_path;
        if (c->kw == 4 || (ip != 0 && ip != 3)) return false;
        if (c->n_active == 0) return false;
        // a graph replays the binned count without Hammer's recount or priors (both re-lay
        // the table out for hashing, which a capture cannot size in advance)
        if (c->capturing || c->graph_planning) {
            if (c->o.error_correct) return false;
            for (const HostLib& H : c->libs)
                if (H.prior_windows(c->o.k)) return false;
        }
        for (const HostLib& H : c->libs)
            if (H.max_len >= SK_MAX_LEN) return false;
        return true;
    }

    // Every library's region sized for hash placement of its distinct keys (est) and
    // prior windows, as plan() sizes one, grown 2^grow times; the descriptors follow.
    template <int NW, typename Ev>
    int hashed_layout(uint32_t grow) {
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        const float tl = c->tune.table_load > 0.0f ? c->tune.table_load : c->o.table_load;
        const double load = (tl > 0.05f && tl < 0.95f) ? tl : 0.5;
        uint64_t b0 = 0;
        for (HostLib& H : c->libs) {
            const uint64_t keys = H.est + H.prior_windows(c->o.k);
            const uint64_t cap = ((uint64_t)std::ceil((double)keys * 1.10 / load) + 64) << grow;
            H.nbuckets = std::max<uint64_t>((cap + SPB - 1) / SPB, 1);
            H.bucket0 = b0;
            b0 += H.nbuckets;
        }
        c->total_buckets = b0;
        KC_TRY(c->d_tab.ensure(b0 * 64));
        return upload_desc();
    }

    // The binned count of every library into the table pool.  On success each library's
    // region holds exactly its entries, packed, regions in library order.  BIN_FALLBACK: a
    // bin needed more passes than allowed (or the keys need more than 31 bits), so the
    // caller counts with the table insert instead.
    static constexpr int BIN_FALLBACK = 1;
    // planned (a graph capture): the bins, record count and per-library sizes are those
    // the planning run recorded (g_bin_*), nothing is read back, allocated or copied from
    // the host, and k_plan_guard flags a replay whose sizes differ.  recount: Stage 3's
    // count of the reads Hammer corrected (src = the corrected copy): every window is
    // counted (the compaction then keeps the Bloom-solid k-mers), and the libraries'
    // distinct counts and the capture plan, which size Stage 2's bins, are left as they are.
    template <int KW, typename Ev>
    int binned_count(bool planned = false, const uint8_t* src = nullptr, bool recount = false) {
        constexpr int NW = nw_of(KW);
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        using S = Slot<NW, Ev>;
        hipStream_t st = c->st;
        const uint8_t* data = src ? src : c->d_data.as<uint8_t>();
        const uint64_t na = c->n_active;
        const uint32_t k = c->o.k, nL = L();
        const uint32_t gw = grid_for(na, BLOCK / 64);
        // bins for about BIN_TARGET distinct k-mers each in the largest library, and output
        // room per library: from the last exact count at this k (plus 1/8), else a third of
        // the library's windows; a run that overflows is counted again with the exact sizes
        uint32_t lib_bits = 0;
        while ((1u << lib_bits) < nL) ++lib_bits;
        std::vector<uint64_t> cap(nL), cap0(nL + 1, 0), lcnt(nL, 0);
        uint64_t est_max = 1;
        for (uint32_t l = 0; l < nL; ++l) {
            const HostLib& H = c->libs[l];
            const bool exact = H.est_k == k && H.est;
            const uint64_t e = exact ? H.est : H.info.n_occ / 3;
            est_max = std::max(est_max, e);
            cap[l] = (exact ? e + e / 8 : e) + 1024;
        }
        const uint32_t bits = planned ? c->g_bin_bits
                                      : bin_bits_for(est_max, c->tune.kmers_per_bin ? c->tune.kmers_per_bin
                                                                                   : BIN_TARGET);
        if (lib_bits + bits > 31) return BIN_FALLBACK;
        const uint32_t nbins = nL << bits;
        unsigned long long* ctr = c->d_counters.as<unsigned long long>() + 12;
        const uint32_t key_bits = lib_bits + bits;
        // one-word records (sorted as keys alone: 8 bytes a record through the sort instead
        // of 12) when the key, the read index and 8-bit first window and window count fit
        uint32_t pk = 0;
        {
            uint32_t abits = 1, maxw = 0;
            while ((1ull << abits) < na) ++abits;
            for (uint32_t l = 0; l < nL; ++l)
                if (c->libs[l].max_len >= k) maxw = std::max(maxw, c->libs[l].max_len - k + 1);
            if (maxw <= 256 && key_bits + 16 + abits <= 64) pk = key_bits;
        }
        uint64_t nrec = 0;
        {
            // records per read, their offsets, then the records in read order
            KC_TRY(c->d_skc.ensure(na * 4 + 4));
            KC_TRY(c->d_sko.ensure(na * 8 + 8));
            KC_TRY(c->d_wbin.ensure(c->n_occ * 8 + 8));
            TIMED(KID_SK, (k_sk_wave<<<gw, BLOCK, 0, st>>>(data, c->d_act_off.as<uint64_t>(),
                                                          c->d_act_ev.as<uint64_t>(), na, k,
                                                          c->tune.minimizer_len && c->tune.minimizer_len <= k
                                                              ? c->tune.minimizer_len
                                                              : minimizer_len(k),
                                                          bits,
                                                          c->d_skc.as<uint32_t>(), c->d_wbin.as<uint64_t>())));
            size_t tb0 = 0;
            KC_TRY(rocprim::exclusive_scan(nullptr, tb0, c->d_skc.as<uint32_t>(), c->d_sko.as<uint64_t>(),
                                           (uint64_t)0, (size_t)na, rocprim::plus<uint64_t>(), st));
            KC_TRY(c->d_temp.ensure(tb0));
            tb0 = c->d_temp.bytes;
            KC_TRY(rocprim::exclusive_scan(c->d_temp.p, tb0, c->d_skc.as<uint32_t>(), c->d_sko.as<uint64_t>(),
                                           (uint64_t)0, (size_t)na, rocprim::plus<uint64_t>(), st));
            if (planned) {
                nrec = c->g_bin_nrec;
            } else {
                KC_TRY(hipMemcpyAsync(c->h_stage, c->d_sko.as<uint64_t>() + na - 1, 8, hipMemcpyDeviceToHost, st));
                KC_TRY(hipMemcpyAsync(c->h_stage + 1, c->d_skc.as<uint32_t>() + na - 1, 4, hipMemcpyDeviceToHost,
                                      st));
                KC_TRY(hipStreamSynchronize(st));
                nrec = c->h_stage[0] + (uint32_t)c->h_stage[1];
            }
            if (!pk) {
                KC_TRY(c->d_bk.ensure(nrec * 4 + 4));
                KC_TRY(c->d_bk2.ensure(nrec * 4 + 4));
            }
            KC_TRY(c->d_br.ensure(nrec * 8 + 8));
            KC_TRY(c->d_br2.ensure(nrec * 8 + 8));
            TIMED(KID_SK, (k_sk_copy<<<gw, BLOCK, 0, st>>>(c->d_act_ev.as<uint64_t>(), c->d_act_lib.as<uint32_t>(), na,
                                                          bits, c->d_skc.as<uint32_t>(), c->d_sko.as<uint64_t>(),
                                                          c->d_wbin.as<uint64_t>(), c->d_bk.as<uint32_t>(),
                                                          c->d_br.as<uint64_t>(), nrec, pk)));
            // the cut's per-window scratch is kept for the next run unless the device is
            // short of room for what follows (metagenome-scale inputs)
            if (!planned && !c->graph_planning && c->d_wbin.bytes > BIG_SCRATCH && !c->fits(c->d_wbin.bytes)) {
                KC_TRY(hipStreamSynchronize(st));
                c->d_wbin.release();
            }
        }
        KC_TRY(c->d_bbounds.ensure(((uint64_t)nbins + 1) * 8));
        size_t tb = 0;
        if (pk) {
            KC_TRY(rocprim::radix_sort_keys<BinSortConfig>(nullptr, tb, c->d_br.as<uint64_t>(), c->d_br2.as<uint64_t>(),
                                                           (size_t)nrec, 0, pk, st));
            KC_TRY(c->d_temp.ensure(tb));
            tb = c->d_temp.bytes;
            TIMED(KID_BIN_SORT, KC_TRY(rocprim::radix_sort_keys<BinSortConfig>(c->d_temp.p, tb, c->d_br.as<uint64_t>(),
                                                                               c->d_br2.as<uint64_t>(), (size_t)nrec,
                                                                               0, pk, st)));
        } else {
            KC_TRY(rocprim::radix_sort_pairs<BinSortConfig>(nullptr, tb, c->d_bk.as<uint32_t>(), c->d_bk2.as<uint32_t>(),
                                                            c->d_br.as<uint64_t>(), c->d_br2.as<uint64_t>(),
                                                            (size_t)nrec, 0, key_bits, st));
            KC_TRY(c->d_temp.ensure(tb));
            tb = c->d_temp.bytes;
            TIMED(KID_BIN_SORT, KC_TRY(rocprim::radix_sort_pairs<BinSortConfig>(
                                    c->d_temp.p, tb, c->d_bk.as<uint32_t>(), c->d_bk2.as<uint32_t>(),
                                    c->d_br.as<uint64_t>(), c->d_br2.as<uint64_t>(), (size_t)nrec, 0, key_bits, st)));
        }
        k_bin_bounds<<<grid_for((uint64_t)nbins + 1), BLOCK, 0, st>>>(c->d_bk2.as<uint32_t>(), c->d_br2.as<uint64_t>(),
                                                                     pk, nrec, nbins, c->d_bbounds.as<uint64_t>());
        KC_TRY(hipGetLastError());
        // Work items: the bins are sized for the largest library, so a library with far
        // fewer distinct k-mers counts 2^gsh of its bins at once (about as many k-mers as a
        // bin of the largest); only items that hold records are listed.
        uint64_t n_items = 0;
        if (planned) {
            n_items = c->g_bin_items;
        } else {
            const uint32_t target = c->tune.kmers_per_bin ? c->tune.kmers_per_bin : BIN_TARGET;
            std::vector<uint32_t>& gmeta = c->h_bgmeta;  // (kept until the copy is done)
            gmeta.assign(2 * nL + 1, 0);
            for (uint32_t l = 0; l < nL; ++l) {
                const HostLib& H = c->libs[l];
                const double e = (double)(H.est_k == k && H.est ? H.est : H.info.n_occ / 3);
                uint32_t sh = 0;
                while (sh < bits && e / (double)(1ull << bits) * (double)(2ull << sh) <= target) ++sh;
                gmeta[l] = sh;
                gmeta[nL + l] = (uint32_t)n_items;
                n_items += (1ull << bits) >> sh;
            }
            gmeta[2 * nL] = (uint32_t)n_items;
            KC_TRY(c->d_bgmeta.ensure(gmeta.size() * 4));
            KC_TRY(hipMemcpyAsync(c->d_bgmeta.p, gmeta.data(), gmeta.size() * 4, hipMemcpyHostToDevice, st));
            if (!recount) c->g_bin_items = n_items;
        }
        KC_TRY(c->d_bflag.ensure(n_items + 1));
        KC_TRY(c->d_bitems.ensure((n_items + 1) * 8));
        KC_TRY(c->d_blist.ensure((n_items + 1) * 8));
        k_bin_items<<<grid_for(n_items), BLOCK, 0, st>>>(c->d_bbounds.as<uint64_t>(), nL, bits,
                                                         c->d_bgmeta.as<uint32_t>(), n_items,
                                                         c->d_bitems.as<uint64_t>(), c->d_bflag.as<uint8_t>());
        KC_TRY(hipGetLastError());
        unsigned int* n_blist = reinterpret_cast<unsigned int*>(c->d_counters.as<unsigned long long>() + BIN_LIST_N);
        {
            size_t tbl = 0;
            KC_TRY(rocprim::select(nullptr, tbl, c->d_bitems.as<uint64_t>(), c->d_bflag.as<uint8_t>(),
                                   c->d_blist.as<uint64_t>(), n_blist, (size_t)n_items, st));
            KC_TRY(c->d_temp.ensure(tbl));
            tbl = c->d_temp.bytes;
            KC_TRY(rocprim::select(c->d_temp.p, tbl, c->d_bitems.as<uint64_t>(), c->d_bflag.as<uint8_t>(),
                                   c->d_blist.as<uint64_t>(), n_blist, (size_t)n_items, st));
        }
        KC_TRY(c->d_act_oe.ensure(na * 16 + 16));
        k_pair_oe<<<grid_for(na), BLOCK, 0, st>>>(c->d_act_off.as<uint64_t>(), c->d_act_ev.as<uint64_t>(), na,
                                                 c->d_act_oe.as<ulonglong2>());
        // count: one library writes straight into the table pool, several into staging
        const bool direct = nL == 1;
        // workgroups: as many as the device holds at once (they take bins from a queue);
        // kc_tune.blocks_per_cu overrides
        static const int resident = [] {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_bin_count<KW, Ev>, BLOCK, 0) != hipSuccess || n < 1)
                n = 1;
            return n;
        }();
        const uint64_t g_max = (uint64_t)c->n_cu * (c->tune.blocks_per_cu ? c->tune.blocks_per_cu : (uint32_t)resident);
        const uint32_t g = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({nbins, g_max, (uint64_t)1 << 23}));
        unsigned int* queue = reinterpret_cast<unsigned int*>(c->d_counters.as<unsigned long long>() + BIN_QUEUE);
        if (planned) {
            cap = c->g_bin_cap;
            lcnt = c->g_bin_lcnt;
        }
        for (int attempt = 0;; ++attempt) {
            for (uint32_t l = 0; l < nL; ++l) cap0[l + 1] = cap0[l] + cap[l];
            if (direct) KC_TRY(c->d_tab.ensure((cap0[nL] + SPB - 1) / SPB * 64));
            else KC_TRY(c->d_pack.ensure(std::max<uint64_t>(cap0[nL], 1) * sizeof(S)));
            KC_TRY(c->d_bcap.ensure(nL * 16 + 16));
            KC_TRY(c->d_bcnt.ensure(nL * 8 + 8));
            if (!planned) {  // a replay finds the plan's room in d_bcap
                KC_TRY(hipMemcpyAsync(c->d_bcap.p, cap0.data(), nL * 8, hipMemcpyHostToDevice, st));
                KC_TRY(hipMemcpyAsync(c->d_bcap.as<uint64_t>() + nL, cap.data(), nL * 8, hipMemcpyHostToDevice, st));
            }
            k_fill<unsigned long long><<<grid_for(nL), BLOCK, 0, st>>>(c->d_bcnt.as<unsigned long long>(), nL, 0ull);
            k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr, 3, 0ull);
            k_fill<unsigned int><<<1, 64, 0, st>>>(queue, 1, 0u);
            TIMED(KID_BIN_COUNT, (k_bin_count<KW, Ev><<<g, BLOCK, 0, st>>>(
                                     c->d_br2.as<uint64_t>(), c->d_bbounds.as<uint64_t>(), data,
                                     c->d_act_oe.as<ulonglong2>(), k,
                                     c->d_libs.as<LibDesc>(), bits, nbins,
                                     direct ? c->d_tab.as<S>() : c->d_pack.as<S>(), c->d_bcap.as<uint64_t>(),
                                     c->d_bcap.as<uint64_t>() + nL, c->d_bcnt.as<unsigned long long>(), ctr,
                                     queue, c->d_blist.as<uint64_t>(), n_blist, pk)));
            if (planned) {
                k_plan_guard<<<grid_for((uint64_t)nL + 1), BLOCK, 0, st>>>(
                    c->d_sko.as<uint64_t>(), c->d_skc.as<uint32_t>(), na, nrec, c->d_bcnt.as<unsigned long long>(),
                    c->d_glcnt.as<uint64_t>(), nL, ctr, c->d_counters.as<unsigned long long>() + PLAN_GUARD);
                KC_TRY(hipGetLastError());
                break;
            }
            KC_TRY(hipMemcpyAsync(lcnt.data(), c->d_bcnt.p, nL * 8, hipMemcpyDeviceToHost, st));
            KC_TRY(hipMemcpyAsync(c->h_stage, ctr, 24, hipMemcpyDeviceToHost, st));
            KC_TRY(hipStreamSynchronize(st));
            if (c->h_stage[1]) return BIN_FALLBACK;
            if (!c->h_stage[2]) break;
            if (attempt) return fail(KC_EHIP, "binned count passed its exact output sizes");
            cap = lcnt;  // the counters totalled every library's distinct k-mers
        }
        if (!planned && !c->graph_planning) {
            // the sorted records are dead now: dropped only when Stage 3 would lack the room
            uint64_t tot = 0;
            for (uint32_t l = 0; l < nL; ++l) tot += lcnt[l];
            const uint64_t rec_bytes = c->d_bk.bytes + c->d_bk2.bytes + c->d_br.bytes + c->d_br2.bytes;
            if (rec_bytes > BIG_SCRATCH && !c->fits(tot * STAGE3_BYTES_PER_ENTRY))
                for (DevBuf* b : {&c->d_bk, &c->d_bk2, &c->d_br, &c->d_br2}) b->release();
        }
        if (!planned && !recount) {  // what a capture of this run replays
            c->g_bin_bits = bits;
            c->g_bin_nrec = nrec;
            c->g_bin_cap = cap;
            c->g_bin_lcnt = lcnt;
            KC_TRY(c->d_glcnt.ensure(nL * 8 + 8));
            KC_TRY(hipMemcpyAsync(c->d_glcnt.p, lcnt.data(), nL * 8, hipMemcpyHostToDevice, st));
            KC_TRY(hipStreamSynchronize(st));
        }
        // pack: library l's entries from bucket0_l, the rest of its last bucket empty
        uint64_t b0 = 0;
        for (uint32_t l = 0; l < nL; ++l) {
            HostLib& H = c->libs[l];
            if (!recount) {
                H.est = lcnt[l];  // exact: the first-setter mode decision and the next run's bins read it
                H.est_k = k;
            }
            H.bucket0 = b0;
            H.nbuckets = (lcnt[l] + SPB - 1) / SPB;
            b0 += H.nbuckets;
        }
        c->total_buckets = b0;
        if (!direct) KC_TRY(c->d_tab.ensure(std::max<uint64_t>(b0, 1) * 64));
        for (uint32_t l = 0; l < nL; ++l) {
            const HostLib& H = c->libs[l];
            S* dst = c->d_tab.as<S>() + H.bucket0 * SPB;
            if (!direct && lcnt[l]) {
                const uint64_t nb = lcnt[l] * sizeof(S);
                k_copy_bytes<<<grid_for(nb / 8 + 1), BLOCK, 0, st>>>(
                    reinterpret_cast<const uint8_t*>(c->d_pack.as<S>() + cap0[l]), reinterpret_cast<uint8_t*>(dst), nb);
            }
            if (H.nbuckets * SPB > lcnt[l])
                k_table_clear<NW, Ev><<<1, BLOCK, 0, st>>>(dst + lcnt[l], H.nbuckets * SPB - lcnt[l]);
        }
        KC_TRY(hipGetLastError());
        return planned ? KC_OK : upload_desc();
    }

    // A packed table (left by the binned count) back to hash placement, for the passes
    // that look k-mers up (priors, features): the entries are copied aside and re-inserted
    // into a region sized as plan() sizes one for that many distinct keys and the prior
    // windows.
    template <int NW, typename Ev>
    int rehash() {
        if (!c->last_binned) return KC_OK;
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        hipStream_t st = c->st;
        const uint64_t packed = c->total_buckets * SPB;
        std::vector<uint64_t> pb0(L()), pn(L());  // packed regions, in slots
        for (uint32_t l = 0; l < L(); ++l) {
            pb0[l] = c->libs[l].bucket0 * SPB;
            pn[l] = c->libs[l].nbuckets * SPB;
        }
        KC_TRY(c->d_pack.ensure(std::max<uint64_t>(packed, 1) * sizeof(Slot<NW, Ev>)));
        if (packed)
            KC_TRY(hipMemcpyAsync(c->d_pack.p, c->d_tab.p, packed * sizeof(Slot<NW, Ev>), hipMemcpyDeviceToDevice, st));
        unsigned long long* flag = c->d_counters.as<unsigned long long>() + 14;
        for (uint32_t grow = 0;; ++grow) {
            if (int rc = hashed_layout<NW, Ev>(grow)) return rc;
            k_fill<unsigned long long><<<1, 64, 0, st>>>(flag, 1, 0ull);
            const uint64_t slots = c->total_buckets * SPB;
            k_table_clear<NW, Ev><<<grid_for(slots), BLOCK, 0, st>>>(c->d_tab.as<Slot<NW, Ev>>(), slots);
            for (uint32_t l = 0; l < L(); ++l)
                if (pn[l])
                    k_rehash<NW, Ev><<<grid_for(pn[l]), BLOCK, 0, st>>>(c->d_pack.as<Slot<NW, Ev>>() + pb0[l], pn[l],
                                                                        c->d_tab.as<Slot<NW, Ev>>(),
                                                                        c->d_libs.as<LibDesc>(), l, flag);
            KC_TRY(hipGetLastError());
            KC_TRY(hipMemcpyAsync(c->h_stage, flag, 8, hipMemcpyDeviceToHost, st));
            KC_TRY(hipStreamSynchronize(st));
            if (!c->h_stage[0]) break;
            if (grow >= 4) return fail(KC_ENOMEM, "device hash table overflow");
        }
        c->last_binned = false;
        return KC_OK;
    }

    // Lay out the table pool, the Bloom pools and the first-setter groups.  dense: the
    // binned count's layout, whose table regions binned_count() sets once it has counted.
    template <int NW, typename Ev>
    int plan(const std::vector<uint32_t>& grow, bool dense = false) {
        const float tl = c->tune.table_load > 0.0f ? c->tune.table_load : c->o.table_load;
        const double load = (tl > 0.05f && tl < 0.95f) ? tl : 0.5;
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        uint64_t b0 = 0, w0 = 0;
        for (uint32_t l = 0; l < L(); ++l) {
            auto& lib = c->libs[l];
            const uint64_t pw = lib.prior_windows(c->o.k);
            const uint64_t occ = lib.info.n_occ + pw;
            uint64_t cap = (uint64_t)std::ceil((double)(lib.est + pw) * 1.10 / load) + 64;
            cap = std::min<uint64_t>(cap, occ * 2 + 64) << grow[l];
            lib.nbuckets = occ && !dense ? std::max<uint64_t>((cap + SPB - 1) / SPB, 1) : 0;
            lib.bucket0 = b0;
            b0 += lib.nbuckets;
            lib.num_words = std::max<uint64_t>((lib.bloom_bits + 63) / 64, 1);
            lib.word0 = w0;
            w0 += lib.num_words;
        }
        c->total_buckets = b0;
        c->total_words = w0;
        if (!dense) KC_TRY(c->d_tab.ensure(std::max<uint64_t>(b0, 1) * 64));
        KC_TRY(c->d_seen.ensure(w0 * 8));
        KC_TRY(c->d_solid.ensure(w0 * 8));
        // first-setter groups: consecutive libraries whose fs arrays fit the budget together
        // (with 32-bit events and sorted first setters, at most 2^31 entries, so that an
        // item's entry leaves its top bit to the counted-twice mark: four libraries on
        // 2^29-bit tiers)
        const uint64_t budget = sizeof(Ev) == 4 && c->tune.setter_path == 0 ? (1ull << 31) * sizeof(Ev)
                                                                             : 64ull << 30;
        c->groups.clear();
        uint64_t used = 0, maxg = 0;
        uint32_t gb = 0;
        for (uint32_t l = 0; l < L(); ++l) {
            const uint64_t need = c->libs[l].num_words * 64 * sizeof(Ev);
            if (l > gb && used + need > budget) {
                c->groups.push_back({gb, l});
                maxg = std::max(maxg, used);
                gb = l;
                used = 0;
            }
            c->libs[l].fs_off = used;
            c->libs[l].group = (uint32_t)c->groups.size();
            used += need;
        }
        c->groups.push_back({gb, L()});
        maxg = std::max(maxg, used);
        void* fs_before = c->d_fs.p;
        KC_TRY(c->d_fs.ensure(maxg));
        if (c->d_fs.p != fs_before) c->fs_clean_bytes = 0;
        c->fs_bytes = maxg;
        return upload_desc();
    }

    // The libraries' descriptors (table regions, Bloom pools, first-setter slices) to the
    // device, from the host state.
    int upload_desc() {
        std::vector<LibDesc> desc(L());
        std::vector<LibBloom> blooms(L());
        std::vector<uint64_t> lb0(L()), lw0(L());
        std::vector<double> mc(L());
        for (uint32_t l = 0; l < L(); ++l) {
            auto& lib = c->libs[l];
            LibDesc& d = desc[l];
            d.ev0 = lib.ev0;  // kept per library: a Run made after the reads preparation has no lib_ev0_
            d.bucket0 = lib.bucket0;
            d.nbuckets = lib.nbuckets;
            d.word0 = lib.word0;
            d.seen = c->d_seen.as<uint64_t>() + lib.word0;
            d.solid = c->d_solid.as<uint64_t>() + lib.word0;
            d.fs = c->d_fs.as<uint8_t>() + lib.fs_off;
            d.M = make_modulus(lib.num_words * 64);
            d.num_words = lib.num_words;
            d.boost = prior_boost(lib.min_cov);
            d.k = c->o.k;
            blooms[l] = LibBloom{d.solid, d.M};
            lb0[l] = lib.bucket0;
            lw0[l] = lib.word0;
            mc[l] = lib.min_cov;
        }
        KC_TRY(c->d_libs.ensure(L() * sizeof(LibDesc)));
        KC_TRY(c->d_blooms.ensure(L() * sizeof(LibBloom)));
        KC_TRY(c->d_lib_b0.ensure(L() * 8));
        KC_TRY(c->d_lib_w0.ensure(L() * 8));
        KC_TRY(c->d_mincov.ensure(L() * 8));
        KC_TRY(c->d_out.ensure(L() * sizeof(LibOut)));
        hipStream_t st = c->st;
        KC_TRY(hipMemcpyAsync(c->d_libs.p, desc.data(), L() * sizeof(LibDesc), hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_blooms.p, blooms.data(), L() * sizeof(LibBloom), hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_lib_b0.p, lb0.data(), L() * 8, hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_lib_w0.p, lw0.data(), L() * 8, hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_mincov.p, mc.data(), L() * 8, hipMemcpyHostToDevice, st));
        KC_TRY(hipStreamSynchronize(st));
        return KC_OK;
    }

    int clear_out() {
        k_fill<uint8_t><<<grid_for(L() * sizeof(LibOut)), BLOCK, 0, c->st>>>(c->d_out.as<uint8_t>(),
                                                                           L() * sizeof(LibOut), (uint8_t)0);
        KC_TRY(hipGetLastError());
        return KC_OK;
    }

    int read_out() {
        c->h_out.resize(L());
        KC_TRY(hipMemcpyAsync(c->h_out.data(), c->d_out.p, L() * sizeof(LibOut), hipMemcpyDeviceToHost, c->st));
        KC_TRY(hipStreamSynchronize(c->st));
        return KC_OK;
    }

    template <int KW, typename Ev>
    int bloom_build() {
        constexpr int NW = nw_of(KW);
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        hipStream_t st = c->st;
        c->last_binned = false;
        c->g_binned = false;
        if constexpr (KW <= 2) {
            if (use_binned()) {
                std::vector<uint32_t> none(L(), 0);
                int rc = plan<NW, Ev>(none, true);
                if (rc) return rc;
                rc = clear_out();
                if (rc) return rc;
                rc = binned_count<KW, Ev>();
                if (rc == KC_OK) {
                    c->last_binned = true;
                    c->g_binned = true;
                    rc = bloom_groups<NW, Ev>(true);
                    if (rc) return rc;
                    return read_out();
                }
                if (rc != BIN_FALLBACK) return rc;
            }
        }
        int rc = estimate_distinct<KW>();
        if (rc) return rc;
        std::vector<uint32_t> grow(L(), 0);
        for (int attempt = 0;; ++attempt) {
            rc = plan<NW, Ev>(grow);
            if (rc) return rc;
            rc = clear_out();
            if (rc) return rc;
            const uint64_t n_slots = c->total_buckets * SPB;
            TIMED(KID_CLEAR, (k_table_clear<NW, Ev><<<grid_for(n_slots), BLOCK, 0, st>>>(
                                 c->d_tab.as<Slot<NW, Ev>>(), n_slots)));
            rc = launch_insert<KW, Ev, true, false>(c->d_data.as<uint8_t>(), KID_INSERT);
            if (rc) return rc;
            rc = read_out();
            if (rc) return rc;
            bool any = false;
            for (uint32_t l = 0; l < L(); ++l)
                if (c->h_out[l].overflow) {
                    any = true;
                    ++grow[l];
                }
            if (!any) break;
            if (attempt >= 4) return fail(KC_ENOMEM, "device hash table overflow");
        }
        rc = bloom_groups<NW, Ev>(true);
        if (rc) return rc;
        return read_out();
    }

    // Whether a dense group's first setters are taken by sorting (k_fs_items, fs_sort,
    // k_fs_min) rather than by device atomics: 32-bit events, item entries under 2^32 - FS_PART
    // (so no partition key is all ones), the setting allows it, and not for a graph capture
    // or the eager run that plans one (rocPRIM's radix sort records memset nodes, which are
    // kept out of graphs).
    template <typename Ev>
    bool sorted_setters(uint64_t fs_n) const {
        return sizeof(Ev) == 4 && c->tune.setter_path == 0 && fs_n + FS_PART <= (1ull << 32);
    }

    // With 64-bit events: the bits S a setter 4*t0 + j of the group needs, when an item
    // e << S | setter and its partition key fit 64 bits; 0 when the group takes the atomics
    // (the setting, a graph capture or its planning run, whose sizes must not be read back).
    template <typename Ev>
    uint32_t sorted_setters64(const std::pair<uint32_t, uint32_t>& g, uint64_t fs_n) const {
        if (sizeof(Ev) != 8 || c->tune.setter_path != 0 || c->capturing || c->graph_planning) return 0;
        uint64_t wmax = 1;
        for (uint32_t l = g.first; l < g.second; ++l) wmax = std::max<uint64_t>(wmax, c->libs[l].info.n_occ);
        uint32_t S = 2;
        while (S < 64 && (1ull << S) < 4 * wmax) ++S;
        const uint32_t nparts = (uint32_t)std::min<uint64_t>((fs_n + FS64_PART - 1) / FS64_PART, 1ull << 31);
        uint32_t key_bits = 1;
        while ((1ull << key_bits) < nparts) ++key_bits;
        return S + FS64_PART_LOG2 + key_bits <= 64 ? S : 0;
    }

    // First setters -> seen tier -> solid tier, one group of libraries at a time.
    // Dense mode fills each group's first-setter array and derives the seen tier from it.
    // Sparse mode (every group's k-mers touch under 1/32 of its entries, by the distinct
    // estimate) keeps the array all ~0 between runs: the first setters also set the seen
    // bits, and afterwards only the touched entries are reset.  `decide` (eager runs)
    // picks the mode and restores a clean array when needed; a captured graph replays
    // the mode of the run it was planned from.
    template <int NW, typename Ev>
    int bloom_groups(bool decide) {
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        hipStream_t st = c->st;
        if (decide) {
            bool sparse = true;
            for (auto& g : c->groups) {
                uint64_t est = 0;
                for (uint32_t l = g.first; l < g.second; ++l) est += c->libs[l].est;
                const auto& le = c->libs[g.second - 1];
                const uint64_t fs_n = (le.fs_off + le.num_words * 64 * sizeof(Ev)) / sizeof(Ev);
                if (3 * est * 32 >= fs_n) sparse = false;
            }
            c->bloom_sparse = sparse;
            if (sparse && c->fs_clean_bytes < c->fs_bytes) {
                TIMED(KID_FS_CLEAR, (k_fill<uint8_t><<<grid_for(c->fs_bytes), BLOCK, 0, st>>>(
                                        c->d_fs.as<uint8_t>(), c->fs_bytes, (uint8_t)0xFF)));
                c->fs_clean_bytes = c->fs_bytes;
            }
        }
        const bool sparse = c->bloom_sparse;
        const uint64_t clean = c->fs_clean_bytes;
        c->fs_clean_bytes = 0;  // until this run's resets are enqueued
        for (auto& g : c->groups) {
            const auto& lb = c->libs[g.first];
            const auto& le = c->libs[g.second - 1];
            const uint64_t s_begin = lb.bucket0 * SPB, s_end = (le.bucket0 + le.nbuckets) * SPB;
            const uint64_t w_begin = lb.word0, w_end = le.word0 + le.num_words;
            const uint64_t fs_n = (le.fs_off + le.num_words * 64 * sizeof(Ev)) / sizeof(Ev);
            const uint32_t gs = grid_for(s_end - s_begin, BLOCK * SCAN_ITEMS);
            if (sparse) {
                k_fill<uint64_t><<<grid_for(w_end - w_begin), BLOCK, 0, st>>>(c->d_seen.as<uint64_t>() + w_begin,
                                                                             w_end - w_begin, 0ull);
                k_fill<uint64_t><<<grid_for(w_end - w_begin), BLOCK, 0, st>>>(c->d_solid.as<uint64_t>() + w_begin,
                                                                             w_end - w_begin, 0ull);
                KC_TRY(hipGetLastError());
                if (s_end > s_begin) {
                    TIMED(KID_FIRST_SETTER, (k_first_setter<NW, Ev, true><<<gs, BLOCK, 0, st>>>(
                                                c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end,
                                                c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                                                c->d_out.as<LibOut>())));
                    TIMED(KID_PROMOTE, (k_promote<NW, Ev><<<gs, BLOCK, 0, st>>>(
                                           c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end, c->d_lib_b0.as<uint64_t>(),
                                           L(), c->d_libs.as<LibDesc>(), false)));
                    TIMED(KID_FS_CLEAR, (k_fs_reset<NW, Ev><<<gs, BLOCK, 0, st>>>(
                                            c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end, c->d_lib_b0.as<uint64_t>(),
                                            L(), c->d_libs.as<LibDesc>())));
                }
            } else if (sorted_setters<Ev>(fs_n)) {
                const uint64_t n_items = 3 * (s_end - s_begin);
                const uint32_t nparts = (uint32_t)((fs_n + FS_PART - 1) / FS_PART);
                uint32_t key_bits = 1;  // partition keys 0 .. nparts - 1; empty slots' items all ones
                while ((1ull << key_bits) <= nparts) ++key_bits;
                const bool multi = FS_KEY_SHIFT + key_bits <= 63;  // bit 63 free for the counted-twice mark
                KC_TRY(c->d_fsi.ensure(std::max<uint64_t>(n_items, 1) * 8));
                KC_TRY(c->d_fsi2.ensure(std::max<uint64_t>(n_items, 1) * 8));
                KC_TRY(c->d_fsb.ensure(((uint64_t)nparts + 1) * 8));
                if (n_items) {
                    auto items = multi ? k_fs_items<NW, true> : k_fs_items<NW, false>;
                    TIMED(KID_FS_ITEMS, (items<<<gs, BLOCK, 0, st>>>(
                                            c->d_tab.as<Slot<NW, uint32_t>>(), s_begin, s_end,
                                            c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                                            c->d_out.as<LibOut>(), c->d_fs.as<uint8_t>(), c->d_fsi.as<uint64_t>())));
                    size_t tb = 0;
                    KC_TRY(rocprim::radix_sort_keys<FsSortConfig>(nullptr, tb, c->d_fsi.as<uint64_t>(),
                                                                  c->d_fsi2.as<uint64_t>(), (size_t)n_items,
                                                                  FS_KEY_SHIFT, FS_KEY_SHIFT + key_bits, st));
                    KC_TRY(c->d_temp.ensure(tb));
                    tb = c->d_temp.bytes;
                    TIMED(KID_FS_SORT,
                          KC_TRY(rocprim::radix_sort_keys<FsSortConfig>(c->d_temp.p, tb, c->d_fsi.as<uint64_t>(),
                                                          c->d_fsi2.as<uint64_t>(), (size_t)n_items, FS_KEY_SHIFT,
                                                          FS_KEY_SHIFT + key_bits, st)));
                }
                k_fs_bounds<<<grid_for((uint64_t)nparts + 1), BLOCK, 0, st>>>(
                    c->d_fsi2.as<uint64_t>(), n_items, nparts, FS_KEY_SHIFT, (uint32_t)((1ull << key_bits) - 1),
                    c->d_fsb.as<uint64_t>());
                KC_TRY(hipGetLastError());
                auto fsmin = multi ? k_fs_min<true> : k_fs_min<false>;
                TIMED(KID_FS_MIN, (fsmin<<<nparts, FS_BLOCK, 0, st>>>(
                                      c->d_fsi2.as<uint64_t>(), c->d_fsb.as<uint64_t>(), nparts,
                                      c->d_fs.as<uint32_t>(), fs_n, c->d_seen.as<uint64_t>() + w_begin,
                                      c->d_solid.as<uint64_t>() + w_begin)));
                if (s_end > s_begin && multi && n_items)
                    TIMED(KID_PROMOTE, (k_promote_items<NW><<<gs, BLOCK, 0, st>>>(
                                           c->d_tab.as<Slot<NW, uint32_t>>(), s_begin, s_end,
                                           c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                                           c->d_fs.as<uint8_t>(), c->d_fsi.as<uint64_t>())));
                else if (s_end > s_begin)
                    TIMED(KID_PROMOTE, (k_promote<NW, Ev><<<gs, BLOCK, 0, st>>>(
                                           c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end, c->d_lib_b0.as<uint64_t>(),
                                           L(), c->d_libs.as<LibDesc>(), multi)));
            } else if (const uint32_t S = sorted_setters64<Ev>(g, fs_n)) {
                // items e << S | setter, compacted; the slots in chunks whose items and their
                // sort fit the device, each chunk folded into fs (the first writes every entry)
                const uint32_t nparts = (uint32_t)((fs_n + FS64_PART - 1) / FS64_PART);
                uint32_t key_bits = 1;
                while ((1ull << key_bits) < nparts) ++key_bits;
                const uint32_t shift = S + FS64_PART_LOG2;
                // chunks of at most BIG_SCRATCH bytes of items and their sort
                const uint64_t per = std::max<uint64_t>(1, std::min<uint64_t>(s_end - s_begin, BIG_SCRATCH / 48));
                KC_TRY(c->d_fsi.ensure(per * 24));
                KC_TRY(c->d_fsi2.ensure(per * 24));
                KC_TRY(c->d_fsb.ensure(((uint64_t)nparts + 1) * 8));
                for (uint64_t c0 = s_begin;; c0 += per) {
                    const uint64_t c1 = std::min(s_end, c0 + per);
                    const uint64_t n = 3 * (c1 - c0);
                    if (c1 > c0)
                        TIMED(KID_FS_ITEMS, (k_fs_items64<NW><<<grid_for(c1 - c0, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                                                reinterpret_cast<const Slot<NW, uint64_t>*>(c->d_tab.p), c0, c1,
                                                c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                                                c->d_out.as<LibOut>(), c->d_fs.as<uint8_t>(), S, fs_n - 1,
                                                c->d_fsi.as<uint64_t>())));
                    if (n) {
                        size_t tb = 0;
                        KC_TRY(rocprim::radix_sort_keys<FsSortConfig>(nullptr, tb, c->d_fsi.as<uint64_t>(),
                                                                      c->d_fsi2.as<uint64_t>(), (size_t)n, shift,
                                                                      shift + key_bits, st));
                        KC_TRY(c->d_temp.ensure(tb));
                        tb = c->d_temp.bytes;
                        TIMED(KID_FS_SORT,
                              KC_TRY(rocprim::radix_sort_keys<FsSortConfig>(c->d_temp.p, tb, c->d_fsi.as<uint64_t>(),
                                                              c->d_fsi2.as<uint64_t>(), (size_t)n, shift,
                                                              shift + key_bits, st)));
                    }
                    k_fs_bounds<<<grid_for((uint64_t)nparts + 1), BLOCK, 0, st>>>(
                        c->d_fsi2.as<uint64_t>(), n, nparts, shift, (uint32_t)((1ull << key_bits) - 1),
                        c->d_fsb.as<uint64_t>());
                    KC_TRY(hipGetLastError());
                    auto fsmin = c0 == s_begin ? k_fs_min64<false> : k_fs_min64<true>;
                    TIMED(KID_FS_MIN, (fsmin<<<nparts, FS_BLOCK, 0, st>>>(c->d_fsi2.as<uint64_t>(),
                                                                         c->d_fsb.as<uint64_t>(), nparts, S,
                                                                         c->d_fs.as<uint64_t>(), fs_n)));
                    if (c1 >= s_end) break;
                }
                TIMED(KID_SEEN, (k_seen_from_fs<Ev><<<grid_for(w_end - w_begin), BLOCK, 0, st>>>(
                                    w_begin, w_end, c->d_lib_w0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>())));
                if (s_end > s_begin)
                    TIMED(KID_PROMOTE, (k_promote<NW, Ev><<<gs, BLOCK, 0, st>>>(
                                           c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end, c->d_lib_b0.as<uint64_t>(),
                                           L(), c->d_libs.as<LibDesc>(), false)));
            } else {
                TIMED(KID_FS_CLEAR,
                      (k_fill<Ev><<<grid_for(fs_n), BLOCK, 0, st>>>(c->d_fs.as<Ev>(), fs_n, (Ev)~(Ev)0)));
                if (s_end > s_begin)
                    TIMED(KID_FIRST_SETTER, (k_first_setter<NW, Ev, false><<<gs, BLOCK, 0, st>>>(
                                                c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end,
                                                c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                                                c->d_out.as<LibOut>())));
                TIMED(KID_SEEN, (k_seen_from_fs<Ev><<<grid_for(w_end - w_begin), BLOCK, 0, st>>>(
                                    w_begin, w_end, c->d_lib_w0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>())));
                if (s_end > s_begin)
                    TIMED(KID_PROMOTE, (k_promote<NW, Ev><<<gs, BLOCK, 0, st>>>(
                                           c->d_tab.as<Slot<NW, Ev>>(), s_begin, s_end, c->d_lib_b0.as<uint64_t>(),
                                           L(), c->d_libs.as<LibDesc>(), false)));
            }
        }
        if (sparse) c->fs_clean_bytes = clean;
        return KC_OK;
    }

    // Sort entries ascending by (library, k-mer).
    int sort_entries(uint64_t n) {
        hipStream_t st = c->st;
        if (n == 0) return KC_OK;
        const SortLayout& lay = c->lay;
        size_t tb = 0;
        if (lay.wide) {  // (library, V) records: one merge sort, then the library words
            KC_TRY(rocprim::merge_sort(nullptr, tb, c->d_wide.as<WideEntry>(), c->d_wide2.as<WideEntry>(),
                                       c->d_cnt.as<uint32_t>(), c->d_cnt2.as<uint32_t>(), (size_t)n, WideLess(), st));
            KC_TRY(c->d_temp.ensure(tb));
            tb = c->d_temp.bytes;
            TIMED(KID_SORT, KC_TRY(rocprim::merge_sort(c->d_temp.p, tb, c->d_wide.as<WideEntry>(),
                                                        c->d_wide2.as<WideEntry>(), c->d_cnt.as<uint32_t>(),
                                                        c->d_cnt2.as<uint32_t>(), (size_t)n, WideLess(), st)));
            std::swap(c->d_wide, c->d_wide2);
            std::swap(c->d_cnt, c->d_cnt2);
            k_wide_lib<<<grid_for(n), BLOCK, 0, st>>>(c->d_wide.as<WideEntry>(), c->d_hi.as<uint64_t>(), n);
            KC_TRY(hipGetLastError());
            return KC_OK;
        }
        if (lay.single) {
            const uint32_t bits = lay.single_bits;
            KC_TRY(rocprim::radix_sort_pairs(nullptr, tb, c->d_lo.as<uint64_t>(), c->d_lo2.as<uint64_t>(),
                                             c->d_cnt.as<uint32_t>(), c->d_cnt2.as<uint32_t>(), (size_t)n, 0u, bits, st));
            KC_TRY(c->d_temp.ensure(tb));
            tb = c->d_temp.bytes;
            TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_lo.as<uint64_t>(),
                                                              c->d_lo2.as<uint64_t>(), c->d_cnt.as<uint32_t>(),
                                                              c->d_cnt2.as<uint32_t>(), (size_t)n, 0u, bits, st)));
            std::swap(c->d_lo, c->d_lo2);
            std::swap(c->d_cnt, c->d_cnt2);
            return KC_OK;
        }
        // Top 64 bits first, then the runs of equal top bits by their low bits (one 64-bit
        // sort instead of two), when no run is too long to order that way; a graph capture
        // (no read-back of that verdict) and such runs take the two-word sort below, and so
        // do tables over 2^27 entries (on the 1.38e9-entry metagenome table, with its many
        // long runs of error k-mers, the first pass cost more than it saved: 261 -> 504 ms).
        if (!c->capturing && !c->graph_planning && n <= (1ull << 27)) {
            const uint32_t r = lay.lo_bits + lay.hi_bits - 64;
            const uint64_t low_mask = (1ull << r) - 1;
            KC_TRY(c->d_idx.ensure(n * 4));
            KC_TRY(c->d_idx2.ensure(n * 4));
            KC_TRY(c->d_ties.ensure((uint64_t)TIE_LONG_CAP * 8));
            k_msd_keys<<<grid_for(n), BLOCK, 0, st>>>(c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), n, lay.hi_bits,
                                                       r, c->d_lo2.as<uint64_t>(), c->d_idx.as<uint32_t>());
            KC_TRY(hipGetLastError());
            size_t need = 0;
            KC_TRY(rocprim::radix_sort_pairs(nullptr, need, c->d_lo2.as<uint64_t>(), c->d_hi2.as<uint64_t>(),
                                             c->d_idx.as<uint32_t>(), c->d_idx2.as<uint32_t>(), (size_t)n, 0u, 64u,
                                             st));
            KC_TRY(c->d_temp.ensure(need));
            tb = c->d_temp.bytes;
            TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_lo2.as<uint64_t>(),
                                                              c->d_hi2.as<uint64_t>(), c->d_idx.as<uint32_t>(),
                                                              c->d_idx2.as<uint32_t>(), (size_t)n, 0u, 64u, st)));
            unsigned int* tctr = reinterpret_cast<unsigned int*>(c->d_counters.as<unsigned long long>() + TIE_CTR);
            k_fill<unsigned int><<<1, 64, 0, st>>>(tctr, 2, 0u);
            TIMED(KID_SORT, (k_tie_short<<<grid_for(n), BLOCK, 0, st>>>(
                                c->d_hi2.as<uint64_t>(), c->d_idx2.as<uint32_t>(), c->d_lo.as<uint64_t>(), n, low_mask,
                                c->d_ties.as<uint2>(), tctr)));
            TIMED(KID_SORT, (k_tie_long<<<(uint32_t)c->n_cu * 2, 512, 0, st>>>(
                                c->d_idx2.as<uint32_t>(), c->d_lo.as<uint64_t>(), low_mask, c->d_ties.as<uint2>(),
                                tctr)));
            KC_TRY(hipMemcpyAsync(c->h_stage, tctr, 8, hipMemcpyDeviceToHost, st));
            KC_TRY(hipStreamSynchronize(st));
            const uint32_t too_long = (uint32_t)(c->h_stage[0] >> 32);
            if (!too_long) {
                TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(
                                    c->d_hi.as<uint64_t>(), c->d_idx2.as<uint32_t>(), c->d_lo2.as<uint64_t>(), n)));
                TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(
                                    c->d_lo.as<uint64_t>(), c->d_idx2.as<uint32_t>(), c->d_hi2.as<uint64_t>(), n)));
                TIMED(KID_SORT, (k_gather<uint32_t><<<grid_for(n), BLOCK, 0, st>>>(
                                    c->d_cnt.as<uint32_t>(), c->d_idx2.as<uint32_t>(), c->d_cnt2.as<uint32_t>(), n)));
                std::swap(c->d_hi, c->d_lo2);
                std::swap(c->d_lo, c->d_hi2);
                std::swap(c->d_cnt, c->d_cnt2);
                return KC_OK;
            }
        }
        // LSD over the two words; hi, lo, cnt stay intact until the final gathers.
        //   1. (lo, iota) sorted by lo        -> perm p1 in d_idx2 (sorted lo discarded)
        //   2. h1 = hi[p1]                    -> d_hi2
        //   3. (h1, p1) stably sorted by h1   -> sorted hi in d_lo2, perm p2 in d_idx
        //   4. lo[p2] -> d_hi2, cnt[p2] -> d_cnt2
        KC_TRY(c->d_idx.ensure(n * 4));
        KC_TRY(c->d_idx2.ensure(n * 4));
        k_iota<<<grid_for(n), BLOCK, 0, st>>>(c->d_idx.as<uint32_t>(), n);
        KC_TRY(hipGetLastError());
        size_t need1 = 0, need2 = 0;
        KC_TRY(rocprim::radix_sort_pairs(nullptr, need1, c->d_lo.as<uint64_t>(), c->d_lo2.as<uint64_t>(),
                                         c->d_idx.as<uint32_t>(), c->d_idx2.as<uint32_t>(), (size_t)n, 0u, lay.lo_bits,
                                         st));
        KC_TRY(rocprim::radix_sort_pairs(nullptr, need2, c->d_hi2.as<uint64_t>(), c->d_lo2.as<uint64_t>(),
                                         c->d_idx2.as<uint32_t>(), c->d_idx.as<uint32_t>(), (size_t)n, 0u, lay.hi_bits,
                                         st));
        KC_TRY(c->d_temp.ensure(std::max(need1, need2)));
        tb = c->d_temp.bytes;
        TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_lo.as<uint64_t>(),
                                                          c->d_lo2.as<uint64_t>(), c->d_idx.as<uint32_t>(),
                                                          c->d_idx2.as<uint32_t>(), (size_t)n, 0u, lay.lo_bits, st)));
        TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_hi.as<uint64_t>(),
                                                                           c->d_idx2.as<uint32_t>(),
                                                                           c->d_hi2.as<uint64_t>(), n)));
        tb = c->d_temp.bytes;
        TIMED(KID_SORT, KC_TRY(rocprim::radix_sort_pairs(c->d_temp.p, tb, c->d_hi2.as<uint64_t>(),
                                                          c->d_lo2.as<uint64_t>(), c->d_idx2.as<uint32_t>(),
                                                          c->d_idx.as<uint32_t>(), (size_t)n, 0u, lay.hi_bits, st)));
        TIMED(KID_SORT, (k_gather<uint64_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_lo.as<uint64_t>(),
                                                                           c->d_idx.as<uint32_t>(),
                                                                           c->d_hi2.as<uint64_t>(), n)));
        TIMED(KID_SORT, (k_gather<uint32_t><<<grid_for(n), BLOCK, 0, st>>>(c->d_cnt.as<uint32_t>(),
                                                                           c->d_idx.as<uint32_t>(),
                                                                           c->d_cnt2.as<uint32_t>(), n)));
        std::swap(c->d_hi, c->d_lo2);
        std::swap(c->d_lo, c->d_hi2);
        std::swap(c->d_cnt, c->d_cnt2);
        return KC_OK;
    }

    // Upload every library's prior windows and add them to the (final) count table.
    // Sets `any` when at least one prior window exists.
    template <int KW, typename Ev>
    int apply_priors(uint64_t slots, bool& any) {
        constexpr int NW = nw_of(KW);
        hipStream_t st = c->st;
        const uint32_t k = c->o.k;
        std::vector<uint8_t> seq;
        std::vector<uint64_t> pstart, pw0;
        std::vector<uint32_t> plib;
        uint64_t nw = 0;
        for (uint32_t l = 0; l < L(); ++l) {
            const auto& lib = c->libs[l];
            for (size_t p = 0; p < lib.prior_len.size(); ++p) {
                const uint64_t len = lib.prior_len[p];
                if (len < k) continue;
                pstart.push_back(seq.size());
                pw0.push_back(nw);
                plib.push_back(l);
                seq.insert(seq.end(), lib.prior.begin() + lib.prior_start[p],
                           lib.prior.begin() + lib.prior_start[p] + len);
                nw += len - k + 1;
            }
        }
        any = nw > 0;
        c->g_np = 0;
        c->g_nw = 0;
        if (!any) return KC_OK;
        if (pstart.size() > 0xFFFFFFFFull) return fail(KC_EUNSUPPORTED, "too many prior sequences");
        const uint32_t np = (uint32_t)pstart.size();
        KC_TRY(c->d_pseq.ensure(seq.size()));
        KC_TRY(c->d_pstart.ensure(np * 8));
        KC_TRY(c->d_pw0.ensure(np * 8));
        KC_TRY(c->d_plib.ensure(np * 4));
        KC_TRY(c->d_psum.ensure(slots * 4));
        KC_TRY(c->d_pflag.ensure(slots));
        KC_TRY(hipMemcpyAsync(c->d_pseq.p, seq.data(), seq.size(), hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_pstart.p, pstart.data(), np * 8, hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_pw0.p, pw0.data(), np * 8, hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(c->d_plib.p, plib.data(), np * 4, hipMemcpyHostToDevice, st));
        k_fill<uint32_t><<<grid_for(slots), BLOCK, 0, st>>>(c->d_psum.as<uint32_t>(), slots, 0u);
        k_fill<uint8_t><<<grid_for(slots), BLOCK, 0, st>>>(c->d_pflag.as<uint8_t>(), slots, (uint8_t)0);
        KC_TRY(hipGetLastError());
        TIMED(KID_PRIOR, (k_prior<KW, Ev><<<grid_for(nw), BLOCK, 0, st>>>(
                             c->d_pseq.as<uint8_t>(), c->d_pstart.as<uint64_t>(), c->d_pw0.as<uint64_t>(),
                             c->d_plib.as<uint32_t>(), np, nw, k, c->d_tab.as<Slot<NW, Ev>>(), c->d_libs.as<LibDesc>(),
                             c->d_psum.as<uint32_t>(), c->d_pflag.as<uint8_t>(), c->d_out.as<LibOut>())));
        int rc = read_out();  // also keeps the host vectors alive until the copies are done
        if (rc) return rc;
        c->g_np = np;
        c->g_nw = nw;
        for (uint32_t l = 0; l < L(); ++l)
            if (c->h_out[l].overflow) return fail(KC_ENOMEM, "device hash table overflow (priors)");
        return KC_OK;
    }

    // Stage 3 after the Bloom build: optional Hammer + recount, compaction, sort, cutoff.
    template <int KW, typename Ev>
    int count() {
        constexpr int NW = nw_of(KW);
        hipStream_t st = c->st;
        unsigned long long* ctr = c->d_counters.as<unsigned long long>();
        k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr, 16, 0ull);
        KC_TRY(hipGetLastError());
        int rc = clear_out();
        if (rc) return rc;
        const bool ec = c->o.error_correct != 0;
        bool recount_binned = false;  // Hammer's recount counted every window (binned)
        c->lay = make_layout(c->o.k, L());
        if (c->lay.hi_bits > 64)
            return fail(KC_EUNSUPPORTED, "too many libraries for this k in one context (need 2k - 63 + log2(libraries) <= 64)");
        if (ec) {
            // corr = data; Hammer writes the reads it modifies
            KC_TRY(c->d_corr.ensure(c->d_data.bytes));
            KC_TRY(hipMemcpyAsync(c->d_corr.p, c->d_data.p, c->d_data.bytes, hipMemcpyDeviceToDevice, st));
            k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr + HAMMER_QUERIES, 1, 0ull);
            if (c->n_active) {
                TIMED(KID_HAMMER, (k_hammer<KW><<<hammer_grid(c->n_active), 64, 0, st>>>(
                                      c->d_data.as<uint8_t>(), c->d_corr.as<uint8_t>(), c->d_active.as<uint32_t>(),
                                      c->n_active, c->d_off.as<uint64_t>(), c->d_len.as<uint32_t>(),
                                      c->d_rlib.as<uint32_t>(), c->d_blooms.as<LibBloom>(), c->o.k,
                                      &c->d_out.as<LibOut>()[0].corrected, LIBOUT_STRIDE, ctr + HAMMER_QUERIES, nullptr)));
            }
            // The recount: by default the binned count over the corrected reads (every
            // window; the compaction keeps the Bloom-solid k-mers), else the table insert of
            // the Bloom-solid windows (assemble.rs:188-205), as a graph capture needs.
            if constexpr (KW <= 2 && sizeof(Ev) == 4) {
                if (use_binned() && !c->capturing && !c->graph_planning) {
                    rc = binned_count<KW, uint32_t>(false, c->d_corr.as<uint8_t>(), true);
                    if (rc == KC_OK) {
                        recount_binned = true;
                        c->last_binned = true;
                    } else if (rc != BIN_FALLBACK) {
                        return rc;
                    }
                }
            }
            if (!recount_binned) {
                if (c->last_binned) {  // the recount hashes: regions sized for the first count's keys
                    rc = hashed_layout<NW, uint32_t>(0);
                    if (rc) return rc;
                    c->last_binned = false;
                }
                const uint64_t SPB = slots_per_bucket<NW, uint32_t>();
                const uint64_t n_slots = c->total_buckets * SPB;
                TIMED(KID_CLEAR, (k_table_clear<NW, uint32_t><<<grid_for(n_slots), BLOCK, 0, st>>>(
                                     c->d_tab.as<Slot<NW, uint32_t>>(), n_slots)));
                if (c->n_tiles) {
                    rc = launch_insert<KW, uint32_t, false, true>(c->d_corr.as<uint8_t>(), KID_RECOUNT);
                    if (rc) return rc;
                }
                rc = read_out();
                if (rc) return rc;
                for (uint32_t l = 0; l < L(); ++l)
                    if (c->h_out[l].overflow)  // the recount table holds a subset of the first table's keys
                        return fail(KC_ENOMEM, "device hash table overflow (recount)");
            }
        }
        bool any_prior = false;
        for (const HostLib& H : c->libs) any_prior = any_prior || H.prior_windows(c->o.k) > 0;
        if (c->last_binned && any_prior) {
            // the priors find or claim their k-mers by hash (a binned recount's table has
            // 32-bit events, Ev here)
            rc = ec ? rehash<NW, uint32_t>() : rehash<NW, Ev>();
            if (rc) return rc;
        }
        const uint64_t slots = c->total_buckets * (ec ? slots_per_bucket<NW, uint32_t>() : slots_per_bucket<NW, Ev>());
        bool priors = false;
        rc = ec ? apply_priors<KW, uint32_t>(slots, priors) : apply_priors<KW, Ev>(slots, priors);
        if (rc) return rc;
        const uint32_t* psum = priors ? c->d_psum.as<uint32_t>() : nullptr;
        const uint8_t* pflag = priors ? c->d_pflag.as<uint8_t>() : nullptr;
        const uint64_t cap = std::max<uint64_t>(slots, 1);
        if (c->d_hi.bytes < cap * 8 && !c->graph_planning && !c->fits(cap * STAGE3_BYTES_PER_ENTRY)) {
            // Stage 2's transients are dead here
            KC_TRY(hipStreamSynchronize(st));
            for (DevBuf* b : {&c->d_wbin, &c->d_bk, &c->d_bk2, &c->d_br, &c->d_br2, &c->d_pack, &c->d_fsi, &c->d_fsi2})
                b->release();
        }
        KC_TRY(c->d_hi.ensure(cap * 8));
        KC_TRY(c->d_lo.ensure(cap * 8));
        KC_TRY(c->d_cnt.ensure(cap * 4));
        KC_TRY(c->d_hi2.ensure(cap * 8));
        KC_TRY(c->d_lo2.ensure(cap * 8));
        KC_TRY(c->d_cnt2.ensure(cap * 4));
        if (c->lay.wide) {
            KC_TRY(c->d_wide.ensure(cap * sizeof(WideEntry)));
            KC_TRY(c->d_wide2.ensure(cap * sizeof(WideEntry)));
        }
        if (ec && recount_binned) {  // every window was counted: the full Bloom-solid test
            TIMED(KID_COMPACT, (k_compact<NW, uint32_t, true><<<grid_for(slots, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                                   c->d_tab.as<Slot<NW, uint32_t>>(), slots, c->d_lib_b0.as<uint64_t>(), L(),
                                   c->d_libs.as<LibDesc>(), c->lay, c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(),
                                   c->d_cnt.as<uint32_t>(), ctr + 4, psum, pflag, c->d_wide.as<WideEntry>(), false)));
        } else if (ec) {
            TIMED(KID_COMPACT, (k_compact<NW, uint32_t, false><<<grid_for(slots, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                                   c->d_tab.as<Slot<NW, uint32_t>>(), slots, c->d_lib_b0.as<uint64_t>(), L(),
                                   c->d_libs.as<LibDesc>(), c->lay, c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(),
                                   c->d_cnt.as<uint32_t>(), ctr + 4, psum, pflag, c->d_wide.as<WideEntry>())));
        } else {
            TIMED(KID_COMPACT, (k_compact<NW, Ev, true><<<grid_for(slots, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                                   c->d_tab.as<Slot<NW, Ev>>(), slots, c->d_lib_b0.as<uint64_t>(), L(),
                                   c->d_libs.as<LibDesc>(), c->lay, c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(),
                                   c->d_cnt.as<uint32_t>(), ctr + 4, psum, pflag, c->d_wide.as<WideEntry>())));
        }
        KC_TRY(hipMemcpyAsync(c->h_stage, ctr + 4, 8, hipMemcpyDeviceToHost, st));
        KC_TRY(hipStreamSynchronize(st));
        const uint64_t n = c->h_stage[0];
        c->n_entries = n;
        rc = sort_entries(n);
        if (rc) return rc;
        // cutoff (assemble.rs:692-735) per library from its count histogram
        const size_t nh = (size_t)L() * HIST_BINS;
        KC_TRY(c->d_hist.ensure(nh * 4));
        k_fill<uint32_t><<<grid_for(nh), BLOCK, 0, st>>>(c->d_hist.as<uint32_t>(), nh, 0u);
        KC_TRY(hipGetLastError());
        if (n) TIMED(KID_HIST, (k_count_hist<<<grid_for(n, BLOCK * 16, 4096), BLOCK, 0, st>>>(
                                   c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), n, c->lay,
                                   c->d_hist.as<uint32_t>())));
        k_cutoff<<<grid_for(L(), 64), 64, 0, st>>>(c->d_hist.as<uint32_t>(), L(), c->d_mincov.as<double>(),
                                                   c->d_out.as<LibOut>());
        KC_TRY(hipGetLastError());
        KC_TRY(c->d_sflags.ensure(n + 1));
        if (n) TIMED(KID_FLAGS, (k_solid_flags<<<grid_for(n, BLOCK * 16, 4096), BLOCK, 0, st>>>(
                                    c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), n, c->lay,
                                    L(), c->d_out.as<LibOut>(), c->d_sflags.as<uint8_t>(), c->d_out.as<LibOut>())));
        return read_out();
    }

    // The device work of one Stage 2 + Stage 3 run, for capture into a HIP graph: the
    // sizes are the last eager run's (reads prep, table plan, first-setter groups, prior
    // windows, count-table entries), nothing is read back and nothing is allocated.  The
    // results land where the eager run's did (d_hi/d_lo/d_cnt, d_sflags, d_out); d_out
    // and the counters are cleared once, so the caller can check the overflow flags,
    // the read-length flag and the number of compacted entries after a replay.
    template <int KW, typename Ev>
    int enqueue_planned() {
        constexpr int NW = nw_of(KW);
        constexpr uint64_t SPB = slots_per_bucket<NW, Ev>();
        hipStream_t st = c->st;
        const bool ec = c->o.error_correct != 0;
        const uint32_t multi = L() > 1;
        unsigned long long* ctr = c->d_counters.as<unsigned long long>();
        (void)multi;
        // ---- Stage 2 ----
        k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr, 32, 0ull);
        KC_TRY(hipGetLastError());
        int rc = clear_out();
        if (rc) return rc;
        if constexpr (KW <= 2) {
            if (c->g_binned) {
                rc = binned_count<KW, Ev>(true);
                if (rc) return rc == BIN_FALLBACK ? fail(KC_EHIP, "planned binned count fell back") : rc;
            }
        }
        if (!c->g_binned) {
            const uint64_t n_slots = c->total_buckets * SPB;
            k_table_clear<NW, Ev><<<grid_for(n_slots), BLOCK, 0, st>>>(c->d_tab.as<Slot<NW, Ev>>(), n_slots);
            KC_TRY(hipGetLastError());
            rc = launch_insert<KW, Ev, true, false>(c->d_data.as<uint8_t>(), KID_INSERT);
            if (rc) return rc;
        }
        rc = bloom_groups<NW, Ev>(false);
        if (rc) return rc;
        // ---- Stage 3 ----
        if (ec) {
            k_copy_bytes<<<grid_for(c->d_data.bytes / 8 + 1), BLOCK, 0, st>>>(c->d_data.as<uint8_t>(),
                                                                             c->d_corr.as<uint8_t>(), c->d_data.bytes);
            k_fill<unsigned long long><<<1, 64, 0, st>>>(ctr + HAMMER_QUERIES, 1, 0ull);
            if (c->n_active)
                k_hammer<KW><<<hammer_grid(c->n_active), 64, 0, st>>>(
                    c->d_data.as<uint8_t>(), c->d_corr.as<uint8_t>(), c->d_active.as<uint32_t>(), c->n_active,
                    c->d_off.as<uint64_t>(), c->d_len.as<uint32_t>(), c->d_rlib.as<uint32_t>(),
                    c->d_blooms.as<LibBloom>(), c->o.k, &c->d_out.as<LibOut>()[0].corrected, LIBOUT_STRIDE,
                    ctr + HAMMER_QUERIES, nullptr);
            const uint64_t n_slots32 = c->total_buckets * slots_per_bucket<NW, uint32_t>();
            k_table_clear<NW, uint32_t><<<grid_for(n_slots32), BLOCK, 0, st>>>(c->d_tab.as<Slot<NW, uint32_t>>(),
                                                                              n_slots32);
            KC_TRY(hipGetLastError());
            rc = launch_insert<KW, uint32_t, false, true>(c->d_corr.as<uint8_t>(), KID_RECOUNT);
            if (rc) return rc;
        }
        const uint64_t slots = c->total_buckets * (ec ? slots_per_bucket<NW, uint32_t>() : SPB);
        const bool priors = c->g_nw > 0;
        if (priors) {
            k_fill<uint32_t><<<grid_for(slots), BLOCK, 0, st>>>(c->d_psum.as<uint32_t>(), slots, 0u);
            k_fill<uint8_t><<<grid_for(slots), BLOCK, 0, st>>>(c->d_pflag.as<uint8_t>(), slots, (uint8_t)0);
            if (ec)
                k_prior<KW, uint32_t><<<grid_for(c->g_nw), BLOCK, 0, st>>>(
                    c->d_pseq.as<uint8_t>(), c->d_pstart.as<uint64_t>(), c->d_pw0.as<uint64_t>(),
                    c->d_plib.as<uint32_t>(), c->g_np, c->g_nw, c->o.k, c->d_tab.as<Slot<NW, uint32_t>>(),
                    c->d_libs.as<LibDesc>(), c->d_psum.as<uint32_t>(), c->d_pflag.as<uint8_t>(), c->d_out.as<LibOut>());
            else
                k_prior<KW, Ev><<<grid_for(c->g_nw), BLOCK, 0, st>>>(
                    c->d_pseq.as<uint8_t>(), c->d_pstart.as<uint64_t>(), c->d_pw0.as<uint64_t>(),
                    c->d_plib.as<uint32_t>(), c->g_np, c->g_nw, c->o.k, c->d_tab.as<Slot<NW, Ev>>(),
                    c->d_libs.as<LibDesc>(), c->d_psum.as<uint32_t>(), c->d_pflag.as<uint8_t>(), c->d_out.as<LibOut>());
            KC_TRY(hipGetLastError());
        }
        const uint32_t* psum = priors ? c->d_psum.as<uint32_t>() : nullptr;
        const uint8_t* pflag = priors ? c->d_pflag.as<uint8_t>() : nullptr;
        // compaction, then the radix sort of the eager run (its memset nodes become fill
        // kernels after the capture)
        if (ec)
            k_compact<NW, uint32_t, false><<<grid_for(slots, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                c->d_tab.as<Slot<NW, uint32_t>>(), slots, c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(),
                c->lay, c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), ctr + 4, psum, pflag,
                c->d_wide.as<WideEntry>());
        else
            k_compact<NW, Ev, true><<<grid_for(slots, BLOCK * SCAN_ITEMS), BLOCK, 0, st>>>(
                c->d_tab.as<Slot<NW, Ev>>(), slots, c->d_lib_b0.as<uint64_t>(), L(), c->d_libs.as<LibDesc>(), c->lay,
                c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), ctr + 4, psum, pflag,
                c->d_wide.as<WideEntry>());
        KC_TRY(hipGetLastError());
        const uint64_t n = c->g_entries;
        rc = sort_entries(n);
        if (rc) return rc;
        const size_t nh = (size_t)L() * HIST_BINS;
        k_fill<uint32_t><<<grid_for(nh), BLOCK, 0, st>>>(c->d_hist.as<uint32_t>(), nh, 0u);
        if (n)
            k_count_hist<<<grid_for(n, BLOCK * 16, 4096), BLOCK, 0, st>>>(c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(),
                                                                         c->d_cnt.as<uint32_t>(), n, c->lay,
                                                                         c->d_hist.as<uint32_t>());
        k_cutoff<<<grid_for(L(), 64), 64, 0, st>>>(c->d_hist.as<uint32_t>(), L(), c->d_mincov.as<double>(),
                                                   c->d_out.as<LibOut>());
        if (n)
            k_solid_flags<<<grid_for(n, BLOCK * 16, 4096), BLOCK, 0, st>>>(
                c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), n, c->lay, L(),
                c->d_out.as<LibOut>(), c->d_sflags.as<uint8_t>(), c->d_out.as<LibOut>());
        KC_TRY(hipGetLastError());
        return KC_OK;
    }

    // The learned first pass's features of library lib's count-table entries, into d_feat
    // (entry order, N_FEAT floats each).  The table must hold the uncorrected counts.
    template <int KW, typename Ev>
    int features(uint32_t lib) {
        constexpr int NW = nw_of(KW);
        hipStream_t st = c->st;
        if (int rc = rehash<NW, Ev>()) return rc;
        const HostLib& H = c->libs[lib];
        const uint64_t n = H.info.n_counts, e0 = H.entry0;
        const uint64_t slots = c->total_buckets * slots_per_bucket<NW, Ev>();
        KC_TRY(c->d_qsum.ensure(std::max<uint64_t>(slots, 1) * 8));
        k_fill<unsigned long long><<<grid_for(slots), BLOCK, 0, st>>>(c->d_qsum.as<unsigned long long>(), slots, 0ull);
        KC_TRY(hipGetLastError());
        if (!H.quals.empty() && c->n_active && H.info.n_occ) {
            std::vector<uint64_t> qo(c->n_reads, ~0ull);
            for (size_t i = 0; i < H.len.size(); ++i) qo[H.read0 + i] = H.qoff[i];
            KC_TRY(c->d_quals.ensure(H.quals.size() + 64));
            KC_TRY(c->d_rqoff.ensure(c->n_reads * 8 + 8));
            KC_TRY(c->d_aqoff.ensure(c->n_active * 8 + 8));
            KC_TRY(hipMemcpyAsync(c->d_quals.p, H.quals.data(), H.quals.size(), hipMemcpyHostToDevice, st));
            KC_TRY(hipMemcpyAsync(c->d_rqoff.p, qo.data(), c->n_reads * 8, hipMemcpyHostToDevice, st));
            k_aqoff<<<grid_for(c->n_active), BLOCK, 0, st>>>(c->d_active.as<uint32_t>(), c->n_active,
                                                           c->d_rqoff.as<uint64_t>(), c->d_aqoff.as<uint64_t>());
            KC_TRY(hipGetLastError());
            const uint64_t t0 = H.ev0 / TILE_WINDOWS;
            const uint64_t t1 = (H.ev0 + H.info.n_occ + TILE_WINDOWS - 1) / TILE_WINDOWS;
            const uint32_t g = (uint32_t)std::min<uint64_t>(t1 - t0, 4096);
            k_qsum<KW, Ev><<<g, BLOCK, 0, st>>>(c->d_data.as<uint8_t>(), c->d_act_ev.as<uint64_t>(),
                                                  c->d_act_off.as<uint64_t>(), c->d_act_lib.as<uint32_t>(),
                                                  c->d_tile_first.as<uint32_t>(), c->d_aqoff.as<uint64_t>(),
                                                  c->d_quals.as<uint8_t>(), c->n_active, c->n_occ, t0, t1, c->o.k, lib,
                                                  c->d_tab.as<Slot<NW, Ev>>(), c->d_libs.as<LibDesc>(),
                                                  c->d_qsum.as<unsigned long long>());
            KC_TRY(hipGetLastError());
            KC_TRY(hipStreamSynchronize(st));  // the host vector qo is released on return
        }
        KC_TRY(c->d_feat.ensure(std::max<uint64_t>(n, 1) * N_FEAT * 4));
        if (n) {
            k_features<KW, Ev><<<grid_for(n), BLOCK, 0, st>>>(
                c->d_hi.as<uint64_t>(), c->d_lo.as<uint64_t>(), c->d_cnt.as<uint32_t>(), e0, n, c->lay, c->o.k, lib,
                c->d_tab.as<Slot<NW, Ev>>(), c->d_libs.as<LibDesc>(), c->d_out.as<LibOut>(),
                c->d_qsum.as<unsigned long long>(), c->d_feat.as<float>(), c->d_wide.as<WideEntry>());
            KC_TRY(hipGetLastError());
        }
        return KC_OK;
    }

    int features_any(uint32_t lib) {
        if (c->kw == 4) return fail(KC_EUNSUPPORTED, "learned-pass features support k <= 63");
        if (c->kw == 1) return c->ev64 ? features<1, uint64_t>(lib) : features<1, uint32_t>(lib);
        return c->ev64 ? features<2, uint64_t>(lib) : features<2, uint32_t>(lib);
    }

    // Buffers the planned run needs beyond the eager run's (called before capture).
    int prepare_planned() {
        // the planning run sized the sort buffers and temporary storage for these entries
        const uint64_t n = c->g_entries;
        KC_TRY(c->d_hist.ensure((size_t)L() * HIST_BINS * 4));
        KC_TRY(c->d_sflags.ensure(n + 1));
        if (c->o.error_correct) KC_TRY(c->d_corr.ensure(c->d_data.bytes));
        return KC_OK;
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

static int add_library(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                       const uint32_t* lengths, size_t n_reads, uint64_t bloom_bits, double min_cov) {
    if (n_reads && (!offsets || !lengths)) return ctx->fail(KC_EINVAL, "null offsets/lengths");
    if (data_len && !data) return ctx->fail(KC_EINVAL, "null data");
    for (size_t i = 0; i < n_reads; ++i) {
        const uint64_t nb = ((uint64_t)lengths[i] + 3) / 4;
        if (offsets[i] > data_len || nb > data_len - offsets[i])
            return ctx->fail(KC_EINVAL, "read " + std::to_string(i) + " extends past data_len");
    }
    HostLib lib;
    lib.data.assign(data, data + data_len);
    lib.off.assign(offsets, offsets + n_reads);
    lib.len.assign(lengths, lengths + n_reads);
    lib.max_len = n_reads ? *std::max_element(lengths, lengths + n_reads) : 0;
    lib.bloom_bits = bloom_bits;
    lib.min_cov = min_cov;
    ctx->libs.push_back(std::move(lib));
    ctx->uploaded = ctx->built = ctx->counted = ctx->graph_ok = false;
    return KC_OK;
}

extern "C" {

int kc_create(kc_ctx** out, const kc_opts* opts) {
    if (!out || !opts) return KC_EINVAL;
    *out = nullptr;
    if (opts->k < 1 || opts->k > 128) return KC_EINVAL;  // Kmer256 holds k <= 128 (dna.rs:134)
    if (opts->k > KC_MAX_K) return KC_EUNSUPPORTED;
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
    ctx->kw = kw_of_k(opts->k);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, opts->device) != hipSuccess || ncu <= 0)
        ncu = 256;
    ctx->n_cu = (uint32_t)ncu;
    *out = ctx;
    return KC_OK;
}

void kc_destroy(kc_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->o.device);
    (void)hipStreamSynchronize(ctx->st);
    DevBuf* bufs[] = {&ctx->d_data,   &ctx->d_off,     &ctx->d_len,     &ctx->d_rlib,     &ctx->d_nwin,
                      &ctx->d_prefix, &ctx->d_flags,   &ctx->d_iota,    &ctx->d_active,   &ctx->d_act_ev,
                      &ctx->d_act_off, &ctx->d_act_lib, &ctx->d_tile_first, &ctx->d_lib_read0, &ctx->d_lib_ev0,
                      &ctx->d_seen,   &ctx->d_solid,   &ctx->d_fs,      &ctx->d_tab,      &ctx->d_libs,
                      &ctx->d_lib_b0, &ctx->d_lib_w0,  &ctx->d_blooms,  &ctx->d_out,      &ctx->d_mincov,
                      &ctx->d_hi,     &ctx->d_lo,      &ctx->d_cnt,     &ctx->d_hi2,      &ctx->d_lo2,
                      &ctx->d_cnt2,   &ctx->d_idx,     &ctx->d_idx2,    &ctx->d_sflags,   &ctx->d_corr,
                      &ctx->d_hll,    &ctx->d_hist,    &ctx->d_counters, &ctx->d_temp,    &ctx->d_keys,
                      &ctx->d_out8,   &ctx->d_vrec,    &ctx->d_vout,    &ctx->d_act_oe};
    for (DevBuf* b : bufs) b->release();
    ctx->graph_release();
    DevBuf* pbufs[] = {&ctx->d_pseq,  &ctx->d_pstart, &ctx->d_pw0,  &ctx->d_plib, &ctx->d_psum,
                       &ctx->d_pflag, &ctx->d_quals,  &ctx->d_rqoff, &ctx->d_aqoff, &ctx->d_qsum,
                       &ctx->d_feat,  &ctx->d_prob,   &ctx->d_lab,   &ctx->d_wide, &ctx->d_wide2};
    for (DevBuf* b : pbufs) b->release();
    for (DevBuf& b : ctx->d_gscratch) b.release();
    for (DevBuf* b : {&ctx->d_fsi, &ctx->d_fsi2, &ctx->d_fsb, &ctx->d_bk, &ctx->d_bk2, &ctx->d_br, &ctx->d_br2,
                      &ctx->d_bbounds, &ctx->d_skc, &ctx->d_sko,
                      &ctx->d_pack, &ctx->d_wbin, &ctx->d_bcap, &ctx->d_bcnt, &ctx->d_bflag,
                      &ctx->d_bitems, &ctx->d_blist, &ctx->d_bgmeta, &ctx->d_ties})
        b->release();
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

const char* kc_last_error(const kc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int kc_reads_upload(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                    const uint32_t* lengths, size_t n_reads) {
    if (!ctx) return KC_EINVAL;
    ctx->libs.clear();
    return add_library(ctx, data, data_len, offsets, lengths, n_reads, ctx->o.bloom_bits, ctx->o.min_coverage);
}

int kc_library_add(kc_ctx* ctx, const uint8_t* data, size_t data_len, const uint64_t* offsets, const uint32_t* lengths,
                   size_t n_reads, uint64_t bloom_bits, double min_coverage, uint32_t* lib) {
    if (!ctx) return KC_EINVAL;
    if (ctx->libs.size() >= 65535) return ctx->fail(KC_EUNSUPPORTED, "at most 65535 libraries per context");
    const int rc = add_library(ctx, data, data_len, offsets, lengths, n_reads, bloom_bits, min_coverage);
    if (rc == KC_OK && lib) *lib = (uint32_t)ctx->libs.size() - 1;
    return rc;
}

int kc_library_priors_set(kc_ctx* ctx, uint32_t lib, const uint8_t* seqs, size_t seqs_len, const uint64_t* starts,
                          const uint64_t* lengths, size_t n_seqs) {
    if (!ctx) return KC_EINVAL;
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    if (n_seqs && (!starts || !lengths)) return ctx->fail(KC_EINVAL, "null starts/lengths");
    if (seqs_len && !seqs) return ctx->fail(KC_EINVAL, "null seqs");
    for (size_t i = 0; i < n_seqs; ++i)
        if (starts[i] > seqs_len || lengths[i] > seqs_len - starts[i])
            return ctx->fail(KC_EINVAL, "prior " + std::to_string(i) + " extends past seqs_len");
    HostLib& L = ctx->libs[lib];
    L.prior.assign(seqs, seqs + seqs_len);
    L.prior_start.assign(starts, starts + n_seqs);
    L.prior_len.assign(lengths, lengths + n_seqs);
    ctx->built = ctx->counted = ctx->graph_ok = false;
    return KC_OK;
}

int kc_priors_set(kc_ctx* ctx, const uint8_t* seqs, size_t seqs_len, const uint64_t* starts, const uint64_t* lengths,
                  size_t n_seqs) {
    return kc_library_priors_set(ctx, 0, seqs, seqs_len, starts, lengths, n_seqs);
}

int kc_set_k(kc_ctx* ctx, uint32_t k) {
    if (!ctx) return KC_EINVAL;
    if (k < 1 || k > KC_MAX_K)
        return ctx->fail(k == 128 ? KC_EUNSUPPORTED : KC_EINVAL, "k outside 1..127");
    ctx->o.k = k;
    ctx->kw = kw_of_k(k);
    ctx->built = ctx->counted = false;
    ctx->graph_release();
    return KC_OK;
}

int kc_library_count(const kc_ctx* ctx, uint32_t* n) {
    if (!ctx || !n) return KC_EINVAL;
    *n = (uint32_t)ctx->libs.size();
    return KC_OK;
}

int kc_bloom_build(kc_ctx* ctx) {
    if (!ctx) return KC_EINVAL;
    if (ctx->libs.empty()) return ctx->fail(KC_ESTATE, "no reads uploaded");
    ctx->graph_ok = false;  // an eager run may move buffers a captured graph points at
    CTX_TRY(hipSetDevice(ctx->o.device));
    ctx->reset_marks();
    ctx->built = ctx->counted = false;
    Run run(ctx);
    int rc = KC_OK;
    if (!ctx->uploaded) {
        rc = run.upload();
        if (rc) return rc;
    }
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    rc = run.prep();
    if (rc) return rc;
    for (auto& lib : ctx->libs) {
        const uint64_t occ = lib.info.n_occ;
        lib.info = kc_info{};
        lib.info.n_occ = occ;
        lib.info.n_reads = lib.len.size();
    }
    // 32-bit events hold 4 t + j for t < 2^30 - 1; larger libraries (or KC_FLAG_EVENTS64)
    // take 64-bit events
    ctx->ev64 = (ctx->o.flags & KC_FLAG_EVENTS64) != 0;
    for (auto& lib : ctx->libs) ctx->ev64 = ctx->ev64 || lib.info.n_occ >= (1ull << 30) - 1;
    if (ctx->n_occ == 0) {
        std::vector<uint32_t> grow(ctx->libs.size(), 0);
        for (auto& lib : ctx->libs) lib.est = 0;
        rc = ctx->kw == 1 ? run.plan<1, uint32_t>(grow)
             : ctx->kw == 2 ? run.plan<2, uint32_t>(grow)
                            : run.plan<5, uint32_t>(grow);
        if (rc) return rc;
        // an empty table for prior windows (sized by the plan)
        const uint64_t n_slots = ctx->total_buckets * ctx->slots_per_bucket();
        if (n_slots) {
            if (ctx->kw == 1)
                k_table_clear<1, uint32_t><<<grid_for(n_slots), BLOCK, 0, ctx->st>>>(
                    ctx->d_tab.as<Slot<1, uint32_t>>(), n_slots);
            else if (ctx->kw == 2)
                k_table_clear<2, uint32_t><<<grid_for(n_slots), BLOCK, 0, ctx->st>>>(
                    ctx->d_tab.as<Slot<2, uint32_t>>(), n_slots);
            else
                k_table_clear<5, uint32_t><<<grid_for(n_slots), BLOCK, 0, ctx->st>>>(
                    ctx->d_tab.as<Slot<5, uint32_t>>(), n_slots);
            CTX_TRY(hipGetLastError());
        }
        k_fill<uint64_t><<<grid_for(ctx->total_words), BLOCK, 0, ctx->st>>>(ctx->d_seen.as<uint64_t>(),
                                                                           ctx->total_words, 0ull);
        k_fill<uint64_t><<<grid_for(ctx->total_words), BLOCK, 0, ctx->st>>>(ctx->d_solid.as<uint64_t>(),
                                                                           ctx->total_words, 0ull);
        CTX_TRY(hipGetLastError());
        rc = run.clear_out();
        if (rc) return rc;
        rc = run.read_out();
    } else if (ctx->kw == 1) {
        rc = ctx->ev64 ? run.bloom_build<1, uint64_t>() : run.bloom_build<1, uint32_t>();
    } else if (ctx->kw == 2) {
        rc = ctx->ev64 ? run.bloom_build<2, uint64_t>() : run.bloom_build<2, uint32_t>();
    } else {
        rc = ctx->ev64 ? run.bloom_build<4, uint64_t>() : run.bloom_build<4, uint32_t>();
    }
    if (rc) return rc;
    CTX_TRY(hipEventRecord(e1, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0;
    CTX_TRY(hipEventElapsedTime(&ms, e0, e1));
    for (uint32_t l = 0; l < ctx->libs.size(); ++l) {
        auto& lib = ctx->libs[l];
        lib.info.ms_stage2 = ms;
        lib.info.num_words = lib.num_words;
        lib.info.n_distinct = ctx->h_out[l].n_distinct;
        lib.info.table_capacity = lib.nbuckets * ctx->slots_per_bucket();
    }
    ctx->built = true;
    return KC_OK;
}

int kc_count(kc_ctx* ctx, kc_info* info) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "kc_count before kc_bloom_build");
    ctx->graph_ok = false;
    CTX_TRY(hipSetDevice(ctx->o.device));
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    Run run(ctx);
    int rc = KC_OK;
    uint64_t prior_windows = 0;
    for (auto& lib : ctx->libs) prior_windows += lib.prior_windows(ctx->o.k);
    if (ctx->n_occ == 0 && prior_windows == 0) {
        ctx->n_entries = 0;
        ctx->lay = make_layout(ctx->o.k, (uint32_t)ctx->libs.size());
        ctx->h_out.assign(ctx->libs.size(), LibOut{});
        for (uint32_t l = 0; l < ctx->libs.size(); ++l) {
            std::vector<uint32_t> hist(HIST_BINS, 0);
            uint32_t top = 0;
            ctx->h_out[l].cutoff = cutoff_rule(hist.data(), 0, ctx->libs[l].min_cov, &top);
        }
        if (ctx->o.error_correct) {
            CTX_TRY(ctx->d_corr.ensure(ctx->d_data.bytes));
            CTX_TRY(hipMemcpyAsync(ctx->d_corr.p, ctx->d_data.p, ctx->d_data.bytes, hipMemcpyDeviceToDevice, ctx->st));
        }
    } else if (ctx->kw == 1) {
        rc = ctx->ev64 ? run.count<1, uint64_t>() : run.count<1, uint32_t>();
    } else if (ctx->kw == 2) {
        rc = ctx->ev64 ? run.count<2, uint64_t>() : run.count<2, uint32_t>();
    } else {
        rc = ctx->ev64 ? run.count<4, uint64_t>() : run.count<4, uint32_t>();
    }
    if (rc) return rc;
    CTX_TRY(hipEventRecord(e1, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    float ms = 0;
    CTX_TRY(hipEventElapsedTime(&ms, e0, e1));
    uint64_t hq = 0;
    if (ctx->o.error_correct)
        CTX_TRY(hipMemcpy(&hq, ctx->d_counters.as<unsigned long long>() + HAMMER_QUERIES, 8, hipMemcpyDeviceToHost));
    uint64_t e = 0;
    for (uint32_t l = 0; l < ctx->libs.size(); ++l) {
        auto& lib = ctx->libs[l];
        const LibOut& o = ctx->h_out[l];
        lib.info.ms_stage3 = ms;
        lib.info.hammer_queries = hq;
        lib.info.n_counts = o.n_counts;
        lib.info.n_solid = o.n_solid;
        lib.info.cutoff = o.cutoff;
        lib.info.top_cov = o.top;
        lib.info.corrected_reads = o.corrected;
        lib.entry0 = e;
        e += o.n_counts;
    }
    if (e != ctx->n_entries) return ctx->fail(KC_EHIP, "per-library counts do not add up to the entries");
    ctx->counted = true;
    if (ctx->timing()) ctx->collect_marks();
    if (info) *info = ctx->libs[0].info;
    return KC_OK;
}

int kc_run(kc_ctx* ctx, kc_info* info) {
    int rc = kc_bloom_build(ctx);
    if (rc) return rc;
    return kc_count(ctx, info);
}

int kc_run_many(kc_ctx* const* ctxs, uint32_t n, kc_info* infos) {
    if (!ctxs && n) return KC_EINVAL;
    for (uint32_t i = 0; i < n; ++i) {
        if (!ctxs[i]) return KC_EINVAL;
        for (uint32_t j = 0; j < i; ++j)
            if (ctxs[j] == ctxs[i]) return ctxs[i]->fail(KC_EINVAL, "kc_run_many: a context is listed twice");
    }
    std::vector<int> rc(n, KC_OK);
    std::vector<std::thread> th;
    th.reserve(n ? n - 1 : 0);
    for (uint32_t i = 1; i < n; ++i)
        th.emplace_back([&, i]() { rc[i] = kc_run(ctxs[i], infos ? &infos[i] : nullptr); });
    if (n) rc[0] = kc_run(ctxs[0], infos ? &infos[0] : nullptr);
    for (std::thread& t : th) t.join();
    for (uint32_t i = 0; i < n; ++i)
        if (rc[i]) return rc[i];
    return KC_OK;
}

// Eager run, then capture of the planned device work into ctx->gexec.
// A memset node of a captured graph (library calls such as rocPRIM's radix sort record
// them) as a fill kernel: width elements of elementSize bytes in each of height rows.
__global__ void k_graph_memset(uint8_t* __restrict__ dst, uint64_t pitch, uint32_t value, uint32_t esize,
                               uint64_t width, uint64_t height) {
    const uint64_t n = width * height;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = i / width, col = i - r * width;
        uint8_t* p = dst + r * pitch + col * esize;
        if (esize == 4) *reinterpret_cast<uint32_t*>(p) = value;
        else if (esize == 2) *reinterpret_cast<uint16_t*>(p) = (uint16_t)value;
        else *p = (uint8_t)value;
    }
}

// Every memset node of g replaced by a k_graph_memset kernel node with the same
// dependencies: memset nodes have given wrong results on replays here, fill kernels not.
static hipError_t replace_memset_nodes(hipGraph_t g, uint32_t* replaced) {
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    if (e != hipSuccess) return e;
    std::vector<hipGraphNode_t> nodes(n);
    if (n && (e = hipGraphGetNodes(g, nodes.data(), &n)) != hipSuccess) return e;
    *replaced = 0;
    for (hipGraphNode_t node : nodes) {
        hipGraphNodeType t;
        if ((e = hipGraphNodeGetType(node, &t)) != hipSuccess) return e;
        if (t != hipGraphNodeTypeMemset) continue;
        hipMemsetParams mp{};
        if ((e = hipGraphMemsetNodeGetParams(node, &mp)) != hipSuccess) return e;
        size_t nd = 0, no = 0;
        if ((e = hipGraphNodeGetDependencies(node, nullptr, &nd)) != hipSuccess) return e;
        std::vector<hipGraphNode_t> deps(nd);
        if (nd && (e = hipGraphNodeGetDependencies(node, deps.data(), &nd)) != hipSuccess) return e;
        if ((e = hipGraphNodeGetDependentNodes(node, nullptr, &no)) != hipSuccess) return e;
        std::vector<hipGraphNode_t> outs(no);
        if (no && (e = hipGraphNodeGetDependentNodes(node, outs.data(), &no)) != hipSuccess) return e;
        uint8_t* dst = static_cast<uint8_t*>(mp.dst);
        uint32_t value = mp.value, esize = mp.elementSize ? mp.elementSize : 1;
        uint64_t width = mp.width, height = mp.height ? mp.height : 1;
        uint64_t pitch = mp.pitch ? mp.pitch : width * esize;
        void* args[] = {&dst, &pitch, &value, &esize, &width, &height};
        hipKernelNodeParams kp{};
        kp.func = reinterpret_cast<void*>(k_graph_memset);
        kp.gridDim = dim3(grid_for(width * height));
        kp.blockDim = dim3(BLOCK);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t kn;
        if ((e = hipGraphAddKernelNode(&kn, g, nd ? deps.data() : nullptr, nd, &kp)) != hipSuccess) return e;
        for (hipGraphNode_t o : outs)
            if ((e = hipGraphAddDependencies(g, &kn, &o, 1)) != hipSuccess) return e;
        if ((e = hipGraphDestroyNode(node)) != hipSuccess) return e;
        ++*replaced;
    }
    return hipSuccess;
}

static int graph_plan_and_capture(kc_ctx* ctx, kc_info* info) {
    // The binned count's bins are sized from each library's distinct k-mers at this k; a
    // context that has not counted them yet runs once to learn them, so that the plan the
    // graph replays is the one an eager run at this size would take.
    bool exact = true;
    for (const HostLib& H : ctx->libs) exact = exact && H.est_k == ctx->o.k && H.est;
    const bool binned = ctx->o.k <= 63 && (ctx->tune.insert_path == 0 || ctx->tune.insert_path == 3);
    if (binned && !exact) {
        const int rc0 = kc_run(ctx, info);
        if (rc0) return rc0;
    }
    ctx->graph_planning = true;
    int rc = kc_run(ctx, info);
    ctx->graph_planning = false;
    if (rc) return rc;
    ctx->graph_release();
    if (ctx->n_occ == 0) return KC_OK;  // nothing to replay: kc_graph_run runs eagerly
    ctx->g_entries = ctx->n_entries;
    Run run(ctx);
    rc = run.prepare_planned();
    if (rc) return rc;
    CTX_TRY(hipStreamSynchronize(ctx->st));
    CTX_TRY(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    if (ctx->kw == 1) rc = ctx->ev64 ? run.enqueue_planned<1, uint64_t>() : run.enqueue_planned<1, uint32_t>();
    else if (ctx->kw == 2) rc = ctx->ev64 ? run.enqueue_planned<2, uint64_t>() : run.enqueue_planned<2, uint32_t>();
    else rc = ctx->ev64 ? run.enqueue_planned<4, uint64_t>() : run.enqueue_planned<4, uint32_t>();
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->st, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return ctx->fail(KC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    uint32_t n_memset = 0;
    const hipError_t em = replace_memset_nodes(g, &n_memset);
    if (em != hipSuccess) {
        (void)hipGraphDestroy(g);
        return ctx->fail(KC_EHIP, std::string("replacing memset nodes: ") + hipGetErrorString(em));
    }
    ctx->g_memset_nodes = n_memset;
    const hipError_t ei = hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->gexec = nullptr;
        return ctx->fail(KC_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    // The capture moved the sort's buffer roles on the host (it sorts into the spare
    // buffers and swaps them in), so the planning run's results are no longer where the
    // exports look: one replay puts them there.
    CTX_TRY(hipGraphLaunch(ctx->gexec, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    ctx->graph_ok = true;
    return KC_OK;
}

int kc_graph_run(kc_ctx* ctx, kc_info* info) {
    if (!ctx) return KC_EINVAL;
    if (ctx->libs.empty()) return ctx->fail(KC_ESTATE, "no reads uploaded");
    CTX_TRY(hipSetDevice(ctx->o.device));
    if (!ctx->graph_ok || !ctx->gexec) return graph_plan_and_capture(ctx, info);
    ctx->reset_marks();
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    CTX_TRY(hipGraphLaunch(ctx->gexec, ctx->st));
    CTX_TRY(hipEventRecord(e1, ctx->st));
    const uint32_t L = (uint32_t)ctx->libs.size();
    ctx->h_out.resize(L);
    CTX_TRY(hipMemcpyAsync(ctx->h_out.data(), ctx->d_out.p, L * sizeof(LibOut), hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipMemcpyAsync(ctx->h_stage, ctx->d_counters.p, 32 * 8, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    // the plan holds when no table overflowed, the binned count kept the planned sizes and
    // the compaction produced exactly the planned number of entries; otherwise plan again
    bool ok = ctx->h_stage[4] == ctx->g_entries && ctx->h_stage[PLAN_GUARD] == 0;
    for (uint32_t l = 0; l < L; ++l) ok = ok && !ctx->h_out[l].overflow;
    if (!ok) return graph_plan_and_capture(ctx, info);
    float ms = 0;
    CTX_TRY(hipEventElapsedTime(&ms, e0, e1));
    uint64_t e = 0;
    for (uint32_t l = 0; l < L; ++l) {
        auto& lib = ctx->libs[l];
        const LibOut& o = ctx->h_out[l];
        lib.info.ms_stage2 = ms;  // the whole replay
        lib.info.ms_stage3 = 0;
        lib.info.n_distinct = o.n_distinct;
        lib.info.n_counts = o.n_counts;
        lib.info.n_solid = o.n_solid;
        lib.info.cutoff = o.cutoff;
        lib.info.top_cov = o.top;
        lib.info.corrected_reads = o.corrected;
        lib.entry0 = e;
        e += o.n_counts;
    }
    if (e != ctx->n_entries) return ctx->fail(KC_EHIP, "per-library counts do not add up to the entries");
    ctx->built = ctx->counted = true;
    if (info) *info = ctx->libs[0].info;
    return KC_OK;
}

int kc_graph_release(kc_ctx* ctx) {
    if (!ctx) return KC_EINVAL;
    ctx->graph_release();
    return KC_OK;
}

int kc_library_verify(kc_ctx* ctx, uint32_t lib, kc_verify* out) {
    if (!ctx || !out) return KC_EINVAL;
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "verify before kc_count");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    CTX_TRY(hipSetDevice(ctx->o.device));
    hipStream_t st = ctx->st;
    const uint32_t L = (uint32_t)ctx->libs.size();
    const HostLib& H = ctx->libs[lib];
    const uint64_t words = ctx->total_words;
    CTX_TRY(ctx->d_vrec.ensure(std::max<uint64_t>(words, 1) * 8));
    CTX_TRY(ctx->d_vout.ensure(((size_t)L + 8) * 8));
    unsigned long long* vo = ctx->d_vout.as<unsigned long long>();
    k_fill<unsigned long long><<<grid_for(words), BLOCK, 0, st>>>(ctx->d_vrec.as<unsigned long long>(), words, 0ull);
    k_fill<unsigned long long><<<1, 64, 0, st>>>(vo, (uint64_t)L + 8, 0ull);
    CTX_TRY(hipGetLastError());
    const uint8_t* counted = ctx->o.error_correct ? ctx->d_corr.as<uint8_t>() : ctx->d_data.as<uint8_t>();
    if (ctx->n_tiles) {
        const uint32_t g = (uint32_t)std::min<uint64_t>(ctx->n_tiles, 8192);
#define KC_VERIFY_LAUNCH(KW)                                                                                       \
    k_verify_windows<KW><<<g, BLOCK, 0, st>>>(ctx->d_data.as<uint8_t>(), counted, ctx->d_act_ev.as<uint64_t>(),   \
                                              ctx->d_act_off.as<uint64_t>(), ctx->d_act_lib.as<uint32_t>(),      \
                                              ctx->d_tile_first.as<uint32_t>(), ctx->n_active, ctx->n_occ,       \
                                              ctx->n_tiles, ctx->o.k, ctx->d_libs.as<LibDesc>(),                 \
                                              ctx->d_vrec.as<unsigned long long>(), vo + 8)
        if (ctx->kw == 1) KC_VERIFY_LAUNCH(1);
        else if (ctx->kw == 2) KC_VERIFY_LAUNCH(2);
        else KC_VERIFY_LAUNCH(4);
#undef KC_VERIFY_LAUNCH
    }
    k_verify_seen<<<grid_for(H.num_words), BLOCK, 0, st>>>(ctx->d_seen.as<uint64_t>() + H.word0,
                                                           ctx->d_vrec.as<unsigned long long>() + H.word0,
                                                           H.num_words, vo + 4);
    const uint64_t n = H.info.n_counts;
    if (n)
        k_verify_table<<<grid_for(n), BLOCK, 0, st>>>(ctx->lay, ctx->d_hi.as<uint64_t>(), ctx->d_lo.as<uint64_t>(),
                                                     ctx->d_wide.as<WideEntry>(), ctx->d_cnt.as<uint32_t>(),
                                                     ctx->d_sflags.as<uint8_t>(), H.entry0, n, H.info.cutoff, vo);
    CTX_TRY(hipGetLastError());
    std::vector<unsigned long long> h((size_t)L + 8);
    CTX_TRY(hipMemcpyAsync(h.data(), vo, h.size() * 8, hipMemcpyDeviceToHost, st));
    CTX_TRY(hipStreamSynchronize(st));
    out->windows = H.info.n_occ;
    out->solid_windows = h[8 + lib];
    out->sum_counts = h[0];
    out->seen_words_differing = h[4];
    out->unordered_pairs = h[1];
    out->mask_mismatches = h[2];
    out->checksum = h[3];
    return KC_OK;
}

int kc_library_info(const kc_ctx* ctx, uint32_t lib, kc_info* info) {
    if (!ctx || !info || lib >= ctx->libs.size()) return KC_EINVAL;
    *info = ctx->libs[lib].info;
    return KC_OK;
}

int kc_info_get(const kc_ctx* ctx, kc_info* info) {
    if (!ctx || !info) return KC_EINVAL;
    if (ctx->libs.empty()) {
        *info = kc_info{};
        return KC_OK;
    }
    return kc_library_info(ctx, 0, info);
}

int kc_library_bloom_export(kc_ctx* ctx, uint32_t lib, uint64_t* seen, uint64_t* solid, size_t num_words) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "bloom export before kc_bloom_build");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    const HostLib& L = ctx->libs[lib];
    if (num_words != L.num_words) return ctx->fail(KC_EINVAL, "num_words mismatch");
    CTX_TRY(hipSetDevice(ctx->o.device));
    if (seen)
        CTX_TRY(hipMemcpyAsync(seen, ctx->d_seen.as<uint64_t>() + L.word0, num_words * 8, hipMemcpyDeviceToHost, ctx->st));
    if (solid)
        CTX_TRY(hipMemcpyAsync(solid, ctx->d_solid.as<uint64_t>() + L.word0, num_words * 8, hipMemcpyDeviceToHost,
                               ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_bloom_export(kc_ctx* ctx, uint64_t* seen, uint64_t* solid, size_t num_words) {
    return kc_library_bloom_export(ctx, 0, seen, solid, num_words);
}

int kc_library_bloom_query(kc_ctx* ctx, uint32_t lib, const uint64_t* keys4, size_t n, uint8_t* out) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->built) return ctx->fail(KC_ESTATE, "bloom query before kc_bloom_build");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    if (n == 0) return KC_OK;
    if (!keys4 || !out) return ctx->fail(KC_EINVAL, "null keys/out");
    CTX_TRY(hipSetDevice(ctx->o.device));
    CTX_TRY(ctx->d_keys.ensure(n * 32));
    CTX_TRY(ctx->d_out8.ensure(n));
    CTX_TRY(hipMemcpyAsync(ctx->d_keys.p, keys4, n * 32, hipMemcpyHostToDevice, ctx->st));
    const HostLib& L = ctx->libs[lib];
    CTX_TRY(query_solid(ctx->d_keys.as<uint64_t>(), n, make_modulus(L.num_words * 64),
                        ctx->d_solid.as<uint64_t>() + L.word0, ctx->d_out8.as<uint8_t>(), ctx->st));
    CTX_TRY(hipMemcpyAsync(out, ctx->d_out8.p, n, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_bloom_query(kc_ctx* ctx, const uint64_t* keys4, size_t n, uint8_t* out) {
    return kc_library_bloom_query(ctx, 0, keys4, n, out);
}

// The Kmer256 limbs (lo u128 = limbs 0-1, hi u128 = limbs 2-3) of the k-mer whose 2k-bit
// value is v[0..3] (lowest word first): for k <= 64, Kmer256(V, 0); for k > 64,
// Kmer256(V >> 2(k-64), V mod 2^(2(k-64))) (dna.rs:134-138, 299-321).
static void kmer256_limbs_host(const uint64_t v[4], uint32_t k, uint64_t* w) {
    if (k <= 64) {
        w[0] = v[0];
        w[1] = v[1];
        w[2] = w[3] = 0;
        return;
    }
    const uint32_t s = 2 * k - 128;  // 2..126
    uint64_t t[4];
    const uint32_t q = s >> 6, r = s & 63;
    for (uint32_t i = 0; i < 4; ++i) {
        const uint64_t a = i + q < 4 ? v[i + q] : 0, b = i + q + 1 < 4 ? v[i + q + 1] : 0;
        t[i] = r ? (a >> r) | (b << (64 - r)) : a;
    }
    w[0] = t[0];
    w[1] = t[1];
    w[2] = s >= 64 ? v[0] : (v[0] & ((1ull << s) - 1));
    w[3] = s > 64 ? (v[1] & ((1ull << (s - 64)) - 1)) : 0;
}

int kc_library_counts_export(kc_ctx* ctx, uint32_t lib, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask,
                             size_t n) {
    if (!ctx) return KC_EINVAL;
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "counts export before kc_count");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    const HostLib& L = ctx->libs[lib];
    if (n != L.info.n_counts) return ctx->fail(KC_EINVAL, "n mismatch");
    if (n == 0) return KC_OK;
    CTX_TRY(hipSetDevice(ctx->o.device));
    const uint64_t e0 = L.entry0;
    if (counts) CTX_TRY(hipMemcpyAsync(counts, ctx->d_cnt.as<uint32_t>() + e0, n * 4, hipMemcpyDeviceToHost, ctx->st));
    if (solid_mask)
        CTX_TRY(hipMemcpyAsync(solid_mask, ctx->d_sflags.as<uint8_t>() + e0, n, hipMemcpyDeviceToHost, ctx->st));
    std::vector<uint64_t> hi, lo;
    const SortLayout& lay = ctx->lay;
    if (keys4 && lay.wide) {
        std::vector<WideEntry> we(n);
        CTX_TRY(hipMemcpyAsync(we.data(), ctx->d_wide.as<WideEntry>() + e0, n * sizeof(WideEntry),
                               hipMemcpyDeviceToHost, ctx->st));
        CTX_TRY(hipStreamSynchronize(ctx->st));
        for (size_t i = 0; i < n; ++i) kmer256_limbs_host(we[i].v, lay.k, keys4 + 4 * i);
        return KC_OK;
    }
    if (keys4) {
        lo.resize(n);
        CTX_TRY(hipMemcpyAsync(lo.data(), ctx->d_lo.as<uint64_t>() + e0, n * 8, hipMemcpyDeviceToHost, ctx->st));
        if (!lay.single) {
            hi.resize(n);
            CTX_TRY(hipMemcpyAsync(hi.data(), ctx->d_hi.as<uint64_t>() + e0, n * 8, hipMemcpyDeviceToHost, ctx->st));
        }
    }
    CTX_TRY(hipStreamSynchronize(ctx->st));
    if (keys4) {
        const uint32_t k = lay.k;
        for (size_t i = 0; i < n; ++i) {
            uint64_t* w = keys4 + 4 * i;
            uint64_t vlo, vhi;
            if (lay.single) {  // (lib << 2k) | V, or V itself when 2k == 64
                vlo = 2 * k >= 64 ? lo[i] : (lo[i] & ((1ull << (2 * k)) - 1));
                vhi = 0;
            } else if (2 * k <= 63) {
                vlo = lo[i];
                vhi = 0;
            } else {  // hi = (lib << (2k-63)) | (V >> 63), lo = V mod 2^63
                const uint32_t vb = 2 * k - 63;
                const uint64_t a = hi[i] & ((1ull << vb) - 1);
                vlo = lo[i] | (a << 63);
                vhi = a >> 1;
            }
            w[0] = vlo;
            w[1] = vhi;
            w[2] = w[3] = 0;
        }
    }
    return KC_OK;
}

int kc_counts_export(kc_ctx* ctx, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask, size_t n) {
    return kc_library_counts_export(ctx, 0, keys4, counts, solid_mask, n);
}

}  // extern "C"

// ---- learned first pass: the library-level margin -------------------------------
//
// Hammer correction changes the count histogram, and where determine_solid_kmer_cutoff's
// valley search (assemble.rs:692-735) lands then decides a whole band of labels, which no
// per-k-mer feature sees.  A library is fragile, and takes the exact selection, when the
// search gives the default cutoff 2 without finding a valley (the labels then depend on
// which low-count k-mers survive correction), or when its found / not-found outcome
// changes on any of `resamples` Poisson resamplings of the histogram.  The resamplings
// draw from SplitMix64 streams keyed by (seed, resampling), so the decision is
// reproducible.

namespace {

bool valley_found(const uint64_t* h, uint32_t mvs) {
    uint32_t valley = 2;
    uint64_t min_val = ~0ull;
    for (uint32_t c = 2; c <= mvs; ++c) {
        if (h[c] <= min_val) {
            min_val = h[c];
            valley = c;
        } else if (h[c] > min_val * 2 && c > valley + 2) {
            return true;
        }
    }
    return false;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// Poisson(lam): the product of uniforms (Knuth) below lam = 10, Hormann's transformed
// rejection with squeeze (PTRS, 1993) above.
uint64_t poisson(double lam, uint64_t& s) {
    if (lam <= 0.0) return 0;
    if (lam < 10.0) {
        const double L = std::exp(-lam);
        uint64_t k = 0;
        double p = uniform01(s);
        while (p > L) {
            ++k;
            p *= uniform01(s);
        }
        return k;
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = uniform01(s) - 0.5, V = uniform01(s);
        const double us = 0.5 - std::fabs(U);
        const double k = std::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return (uint64_t)k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <= -lam + k * loglam - std::lgamma(k + 1.0))
            return (uint64_t)k;
    }
}

}  // namespace

// hist[c] = entries with count c for c < n_bins - 1, hist[n_bins - 1] those with more.
static int cutoff_fragile(const uint64_t* hist, uint32_t n_bins, uint64_t n, double min_cov, uint32_t resamples,
                          uint64_t seed) {
    if (min_cov > 5.0) return 0;  // a fixed cutoff, max(0.35 min_cov, 2): no valley search
    if (n <= 100) return 1;
    const uint64_t rank = n * 95 / 100;
    uint64_t cum = 0;
    uint32_t top = n_bins - 1;
    for (uint32_t c = 0; c < n_bins; ++c) {
        cum += hist[c];
        if (cum > rank) {
            top = c;
            break;
        }
    }
    if (top < 30) return 1;
    const uint32_t mvs = std::min<uint32_t>(std::max<uint32_t>(top / 4, 10), 35);
    if (mvs >= n_bins - 1) return 1;  // the histogram does not resolve the search
    if (!valley_found(hist, mvs)) return 1;
    std::vector<uint64_t> h(mvs + 1, 0);
    for (uint32_t r = 0; r < resamples; ++r) {
        uint64_t st = seed ^ (0xD1B54A32D192ED03ull * (r + 1));
        for (uint32_t c = 2; c <= mvs; ++c) h[c] = poisson((double)hist[c], st);
        if (!valley_found(h.data(), mvs)) return 1;
    }
    return 0;
}

extern "C" {

int kc_cutoff_is_fragile(const uint64_t* hist, uint32_t n_bins, uint64_t n_counts, double min_coverage,
                         uint32_t resamples, uint64_t seed, int32_t* fragile) {
    if (!fragile || (n_bins && !hist) || n_bins < 2) return KC_EINVAL;
    *fragile = cutoff_fragile(hist, n_bins, n_counts, min_coverage, resamples, seed);
    return KC_OK;
}

int kc_library_cutoff_fragile(kc_ctx* ctx, uint32_t lib, uint32_t resamples, uint64_t seed, int32_t* fragile) {
    if (!ctx || !fragile) return KC_EINVAL;
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "cutoff margin before kc_count");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    CTX_TRY(hipSetDevice(ctx->o.device));
    std::vector<uint32_t> h32(HIST_BINS);
    CTX_TRY(hipMemcpyAsync(h32.data(), ctx->d_hist.as<uint32_t>() + (size_t)lib * HIST_BINS, HIST_BINS * 4,
                           hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    std::vector<uint64_t> h(h32.begin(), h32.end());
    const HostLib& H = ctx->libs[lib];
    *fragile = cutoff_fragile(h.data(), HIST_BINS, H.info.n_counts, H.min_cov, resamples, seed);
    return KC_OK;
}

}  // extern "C"

// ---- A library's solid set on the device, for the unitig builder (kc_graph.hip) -------

// Solid entry i (entry e0 + idx[i] of the sorted count table) as its value V in vw words,
// lowest first, and its count.
__global__ void k_solid_values(SortLayout lay, const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                               const WideEntry* __restrict__ wide, const uint32_t* __restrict__ cnt,
                               const uint64_t* __restrict__ idx, uint64_t n, uint64_t e0, uint32_t vw,
                               uint64_t* __restrict__ vals, uint32_t* __restrict__ counts) {
    const uint32_t k = lay.k;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t e = e0 + idx[i];
        uint64_t v[4] = {0, 0, 0, 0};
        if (lay.wide) {
            for (int w = 0; w < 4; ++w) v[w] = wide[e].v[w];
        } else if (lay.single) {  // (lib << 2k) | V, or V itself when 2k == 64
            v[0] = 2 * k >= 64 ? lo[e] : (lo[e] & ((1ull << (2 * k)) - 1));
        } else if (2 * k <= 63) {
            v[0] = lo[e];
        } else {  // hi = (lib << (2k-63)) | (V >> 63), lo = V mod 2^63
            const uint64_t a = hi[e] & ((1ull << (2 * k - 63)) - 1);
            v[0] = lo[e] | (a << 63);
            v[1] = a >> 1;
        }
        for (uint32_t w = 0; w < vw; ++w) vals[i * vw + w] = v[w];
        counts[i] = cnt[e];
    }
}

namespace kc {

int ctx_fail(kc_ctx* ctx, int code, const std::string& msg) { return ctx->fail(code, msg); }

DevBuf* graph_scratch(kc_ctx* ctx) { return ctx->d_gscratch; }

int solid_set(kc_ctx* ctx, uint32_t lib, SolidSet* s) {
    if (!ctx || !s) return KC_EINVAL;
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "solid set before kc_count");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    s->device = ctx->o.device;
    s->stream = ctx->st;
    s->k = ctx->lay.k;
    s->n = ctx->libs[lib].info.n_solid;
    return KC_OK;
}

int solid_values(kc_ctx* ctx, uint32_t lib, uint32_t vw, uint64_t* d_vals, uint32_t* d_counts, DevBuf& idx,
                 DevBuf& tmp) {
    SolidSet s{};
    const int rc = solid_set(ctx, lib, &s);
    if (rc != KC_OK) return rc;
    const HostLib& L = ctx->libs[lib];
    const uint64_t n = L.info.n_counts, e0 = L.entry0;
    if (s.n == 0 || n == 0) return KC_OK;
    CTX_TRY(hipSetDevice(ctx->o.device));
    CTX_TRY(idx.ensure((n + 1) * 8));
    uint64_t* d_nsel = idx.as<uint64_t>() + n;
    rocprim::counting_iterator<uint64_t> iota(0);
    size_t tb = 0;
    CTX_TRY(rocprim::select(nullptr, tb, iota, ctx->d_sflags.as<uint8_t>() + e0, idx.as<uint64_t>(), d_nsel,
                            (size_t)n, ctx->st));
    CTX_TRY(tmp.ensure(tb));
    tb = tmp.bytes;
    CTX_TRY(rocprim::select(tmp.p, tb, iota, ctx->d_sflags.as<uint8_t>() + e0, idx.as<uint64_t>(), d_nsel,
                            (size_t)n, ctx->st));
    uint64_t got = 0;
    CTX_TRY(hipMemcpyAsync(&got, d_nsel, 8, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    if (got != s.n) return ctx->fail(KC_EHIP, "solid flags disagree with the solid count");
    const bool wide = ctx->lay.wide != 0;
    k_solid_values<<<grid_for(s.n), BLOCK, 0, ctx->st>>>(
        ctx->lay, ctx->d_lo.as<uint64_t>(), wide ? nullptr : ctx->d_hi.as<uint64_t>(),
        wide ? ctx->d_wide.as<WideEntry>() : nullptr, ctx->d_cnt.as<uint32_t>(), idx.as<uint64_t>(), s.n, e0, vw,
        d_vals, d_counts);
    CTX_TRY(hipGetLastError());
    return KC_OK;
}

}  // namespace kc

extern "C" {

int kc_library_quals_set(kc_ctx* ctx, uint32_t lib, const uint8_t* quals, size_t n_quals, const uint64_t* qoffsets) {
    if (!ctx) return KC_EINVAL;
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    HostLib& H = ctx->libs[lib];
    if (n_quals && !quals) return ctx->fail(KC_EINVAL, "null quals");
    if (!H.len.empty() && n_quals && !qoffsets) return ctx->fail(KC_EINVAL, "null qoffsets");
    if (n_quals == 0) {
        H.quals.clear();
        H.qoff.clear();
        return KC_OK;
    }
    for (size_t i = 0; i < H.len.size(); ++i)
        if (qoffsets[i] > n_quals || H.len[i] > n_quals - qoffsets[i])
            return ctx->fail(KC_EINVAL, "qualities of read " + std::to_string(i) + " extend past n_quals");
    H.quals.assign(quals, quals + n_quals);
    H.qoff.assign(qoffsets, qoffsets + H.len.size());
    return KC_OK;
}

static int learned_ready(kc_ctx* ctx, uint32_t lib, size_t n) {
    if (!ctx->counted) return ctx->fail(KC_ESTATE, "features before kc_count");
    if (ctx->o.error_correct) return ctx->fail(KC_ESTATE, "features need the uncorrected table (error_correct off)");
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    if (n != ctx->libs[lib].info.n_counts) return ctx->fail(KC_EINVAL, "n must equal the library's n_counts");
    return KC_OK;
}

int kc_library_features(kc_ctx* ctx, uint32_t lib, float* out, size_t n) {
    if (!ctx) return KC_EINVAL;
    int rc = learned_ready(ctx, lib, n);
    if (rc) return rc;
    if (n && !out) return ctx->fail(KC_EINVAL, "null out");
    CTX_TRY(hipSetDevice(ctx->o.device));
    Run run(ctx);
    rc = run.features_any(lib);
    if (rc) return rc;
    if (n) CTX_TRY(hipMemcpyAsync(out, ctx->d_feat.p, n * N_FEAT * 4, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_library_learned_select(kc_ctx* ctx, uint32_t lib, kc_mlp* mlp, float margin, uint8_t* labels,
                              uint8_t* fallback, size_t n, double* ms) {
    if (!ctx || !mlp) return KC_EINVAL;
    int rc = learned_ready(ctx, lib, n);
    if (rc) return rc;
    if (mlp->device != ctx->o.device) return ctx->fail(KC_EINVAL, "network and context on different devices");
    if (!(margin >= 0.f && margin < 0.5f)) return ctx->fail(KC_EINVAL, "margin must be in [0, 0.5)");
    CTX_TRY(hipSetDevice(ctx->o.device));
    hipEvent_t e0 = ctx->new_event(), e1 = ctx->new_event();
    CTX_TRY(hipEventRecord(e0, ctx->st));
    Run run(ctx);
    rc = run.features_any(lib);
    if (rc) return rc;
    CTX_TRY(ctx->d_prob.ensure(std::max<size_t>(n, 1) * 4));
    CTX_TRY(ctx->d_lab.ensure(std::max<size_t>(n, 1)));
    CTX_TRY(mlp_infer_device(mlp, ctx->d_feat.as<float>(), n, ctx->d_prob.as<float>(), ctx->st));
    if (n) {
        k_learned_labels<<<grid_for(n), BLOCK, 0, ctx->st>>>(ctx->d_prob.as<float>(), ctx->d_cnt.as<uint32_t>(),
                                                             ctx->libs[lib].entry0, n, ctx->d_out.as<LibOut>(), lib,
                                                             margin, ctx->d_lab.as<uint8_t>());
        CTX_TRY(hipGetLastError());
    }
    CTX_TRY(hipEventRecord(e1, ctx->st));
    std::vector<uint8_t> lab(n);
    if (n) CTX_TRY(hipMemcpyAsync(lab.data(), ctx->d_lab.p, n, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    float t = 0;
    CTX_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    for (size_t i = 0; i < n; ++i) {
        if (labels) labels[i] = lab[i] & 1;
        if (fallback) fallback[i] = (lab[i] >> 1) & 1;
    }
    return KC_OK;
}

int kc_library_reads_export(kc_ctx* ctx, uint32_t lib, uint8_t* data, size_t data_len) {
    if (!ctx) return KC_EINVAL;
    if (lib >= ctx->libs.size()) return ctx->fail(KC_EINVAL, "no such library");
    const HostLib& L = ctx->libs[lib];
    if (data_len != L.data.size()) return ctx->fail(KC_EINVAL, "data_len mismatch");
    if (data_len == 0) return KC_OK;
    if (!data) return ctx->fail(KC_EINVAL, "null data");
    if (!(ctx->o.error_correct && ctx->counted) || !ctx->uploaded) {
        std::memcpy(data, L.data.data(), data_len);
        return KC_OK;
    }
    CTX_TRY(hipSetDevice(ctx->o.device));
    CTX_TRY(hipMemcpyAsync(data, ctx->d_corr.as<uint8_t>() + L.data0, data_len, hipMemcpyDeviceToHost, ctx->st));
    CTX_TRY(hipStreamSynchronize(ctx->st));
    return KC_OK;
}

int kc_reads_export(kc_ctx* ctx, uint8_t* data, size_t data_len) {
    return kc_library_reads_export(ctx, 0, data, data_len);
}

int kc_tune_get(const kc_ctx* ctx, kc_tune* tune) {
    if (!ctx || !tune) return KC_EINVAL;
    *tune = ctx->tune;
    return KC_OK;
}

int kc_tune_set(kc_ctx* ctx, const kc_tune* tune) {
    if (!ctx || !tune) return KC_EINVAL;
    const uint32_t w = tune->windows_per_thread;
    if (w != 1 && w != 2 && w != 4) return ctx->fail(KC_EINVAL, "windows_per_thread must be 1, 2 or 4");
    if (tune->blocks_per_cu > 1024) return ctx->fail(KC_EINVAL, "blocks_per_cu must be at most 1024");
    if (tune->table_load != 0.0f && !(tune->table_load > 0.05f && tune->table_load < 0.95f))
        return ctx->fail(KC_EINVAL, "table_load must be 0 or in (0.05, 0.95)");
    if (tune->insert_path > 3 || tune->insert_path == 2)
        return ctx->fail(KC_EINVAL, "insert_path must be 0, 1 or 3 (the partitioned count, 2, was removed)");
    if (tune->setter_path > 1) return ctx->fail(KC_EINVAL, "setter_path must be 0 or 1");
    if (tune->kmers_per_bin && (tune->kmers_per_bin < 64 || tune->kmers_per_bin > 4096))
        return ctx->fail(KC_EINVAL, "kmers_per_bin must be 0 or in 64..4096");
    if (tune->minimizer_len > 32 || (tune->minimizer_len && tune->minimizer_len > ctx->o.k))
        return ctx->fail(KC_EINVAL, "minimizer_len must be 0 or in 1..32 and at most k");
    ctx->tune = *tune;
    ctx->built = ctx->counted = ctx->graph_ok = false;  // the table layout depends on the load factor
    return KC_OK;
}

int kc_autotune(kc_ctx* ctx, uint32_t reps, kc_tune* best) {
    if (!ctx) return KC_EINVAL;
    if (ctx->libs.empty()) return ctx->fail(KC_ESTATE, "no reads uploaded");
    reps = std::max<uint32_t>(reps, 1);
    std::vector<kc_info> ref;  // per-library summary every setting must reproduce
    auto same = [&](const kc_info& a, const kc_info& b) {
        return a.n_occ == b.n_occ && a.n_distinct == b.n_distinct && a.n_counts == b.n_counts &&
               a.n_solid == b.n_solid && a.cutoff == b.cutoff && a.corrected_reads == b.corrected_reads;
    };
    auto trial = [&](const kc_tune& t, double* ms) -> int {
        int rc = kc_tune_set(ctx, &t);
        if (rc) return rc;
        double m = 1e300;
        for (uint32_t r = 0; r <= reps; ++r) {  // run 0 is the warm-up (and sizes the bins)
            kc_info info{};
            rc = kc_run(ctx, &info);
            if (rc) return rc;
            if (r) m = std::min(m, info.ms_stage2 + info.ms_stage3);
        }
        for (size_t l = 0; l < ctx->libs.size(); ++l) {
            if (ref.size() < ctx->libs.size()) ref.push_back(ctx->libs[l].info);
            else if (!same(ref[l], ctx->libs[l].info))
                return ctx->fail(KC_EHIP, "autotune: a setting changed the results of library " + std::to_string(l));
        }
        *ms = m;
        return KC_OK;
    };
    kc_tune cur = ctx->tune;
    double cur_ms = 0;
    int rc = trial(cur, &cur_ms);
    if (rc) return rc;
    // The knobs of the Stage 2 path this input takes.  Binned count: distinct k-mers a bin
    // is sized for (passes over the LDS table against workgroups), then workgroups per CU
    // (0: one per bin).  Table insert: load factor, windows in flight, workgroups per CU.
    // Then the first-setter path for either.
    const bool binned = ctx->g_binned;  // the last Stage 2 took the binned count
    if (!binned && cur.table_load == 0.0f)
        cur.table_load = (ctx->o.table_load > 0.05f && ctx->o.table_load < 0.95f) ? ctx->o.table_load : 0.5f;
    struct Dim {
        int field;
        std::vector<double> values;
    };
    enum { F_LOAD, F_WPT, F_BPC, F_KPB, F_M, F_SETTER };
    std::vector<Dim> dims;
    if (binned) {
        dims.push_back({F_KPB, {192, 256, 352, 448, 640}});
        // minimizer lengths around the default for this k (0 = the default itself)
        std::vector<double> ms{0};
        for (uint32_t m : {11u, 13u, 16u, 20u, 24u, 28u, 32u})
            if (m <= ctx->o.k && m != minimizer_len(ctx->o.k)) ms.push_back(m);
        dims.push_back({F_M, ms});
        dims.push_back({F_BPC, {0, 4, 8, 16}});
    } else {
        dims.push_back({F_LOAD, {0.35, 0.5, 0.65, 0.8}});
        dims.push_back({F_WPT, {1, 2, 4}});
        dims.push_back({F_BPC, {0, 4, 8, 16, 32}});
    }
    dims.push_back({F_SETTER, {0, 1}});
    auto get = [](const kc_tune& t, int f) -> double {
        switch (f) {
            case F_LOAD: return t.table_load;
            case F_WPT: return t.windows_per_thread;
            case F_BPC: return t.blocks_per_cu;
            case F_KPB: return t.kmers_per_bin ? t.kmers_per_bin : BIN_TARGET;
            case F_M: return t.minimizer_len;
            default: return t.setter_path;
        }
    };
    auto set = [](kc_tune& t, int f, double v) {
        switch (f) {
            case F_LOAD: t.table_load = (float)v; break;
            case F_WPT: t.windows_per_thread = (uint32_t)v; break;
            case F_BPC: t.blocks_per_cu = (uint32_t)v; break;
            case F_KPB: t.kmers_per_bin = (uint32_t)v; break;
            case F_M: t.minimizer_len = (uint32_t)v; break;
            default: t.setter_path = (uint32_t)v; break;
        }
    };
    for (const Dim& d : dims) {
        const kc_tune base = cur;
        for (double v : d.values) {
            if (v == get(base, d.field)) continue;
            kc_tune t = base;
            set(t, d.field, v);
            double ms = 0;
            rc = trial(t, &ms);
            if (rc) return rc;
            if (ms < cur_ms) {
                cur = t;
                cur_ms = ms;
            }
        }
    }
    cur.ms = cur_ms;
    rc = kc_tune_set(ctx, &cur);
    if (rc) return rc;
    ctx->tune.ms = cur_ms;
    if (best) *best = ctx->tune;
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
