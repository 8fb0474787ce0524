// kc_device.h — device primitives of the gfx950 k-mer spectrum engine.
//
// A k-mer (k <= 127) is held as its 2k-bit value V, first base most significant (the
// order of Kmer256's derived Ord, dna.rs:137-138, 299-321), in four 64-bit words: lo =
// bits 0..63, hi = 64..127, w2 = 128..191, w3 = 192..255.  A kernel is built for KW
// value words: KW = 1 for k <= 32, 2 for 33 <= k <= 63, 4 for 64 <= k <= 127; the words
// above KW are zero.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kc {

struct Kv {
    uint64_t lo, hi, w2, w3;
};

// Table key words per value width: 63-bit words, 1, 2 or 5 of them (section below).
__host__ __device__ constexpr int nw_of(int KW) { return KW == 4 ? 5 : KW; }
__host__ __device__ constexpr int kw_of_nw(int NW) { return NW == 5 ? 4 : NW; }
__host__ __device__ constexpr int kw_of_k(uint32_t k) { return k <= 32 ? 1 : k <= 63 ? 2 : 4; }

constexpr uint64_t EMPTY_KEY = ~0ull;

__device__ __forceinline__ bool kv_le(const Kv& a, const Kv& b) {
    if (a.w3 != b.w3) return a.w3 < b.w3;
    if (a.w2 != b.w2) return a.w2 < b.w2;
    return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo);
}

// Word j (0..3) of a value, 0 past the top.
__device__ __forceinline__ uint64_t kv_word(const Kv& x, uint32_t j) {
    return j == 0 ? x.lo : j == 1 ? x.hi : j == 2 ? x.w2 : j == 3 ? x.w3 : 0ull;
}

// x >> s as a 256-bit value, 0 <= s < 256.
__device__ __forceinline__ Kv kv_shr(const Kv& x, uint32_t s) {
    const uint32_t q = s >> 6, r = s & 63;
    uint64_t o[4];
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const uint64_t a = kv_word(x, i + q), b = kv_word(x, i + q + 1);
        o[i] = r ? (a >> r) | (b << (64 - r)) : a;
    }
    return Kv{o[0], o[1], o[2], o[3]};
}

// Reverse the order of the 16 2-bit groups of a 32-bit word, keeping the two bits of a
// group in order: a full bit reversal (v_bfrev_b32) puts the groups in place with their
// two bits swapped, and one masked merge of the word shifted each way swaps them back.
__device__ __forceinline__ uint32_t rev_groups32(uint32_t x) {
    const uint32_t y = __builtin_bitreverse32(x);
    return ((y >> 1) & 0x55555555u) | ((y << 1) & 0xAAAAAAAAu);
}

// Reverse the order of the 32 2-bit groups of x, keeping the two bits of a group in order.
__device__ __forceinline__ uint64_t rev_groups(uint64_t x) {
    return ((uint64_t)rev_groups32((uint32_t)x) << 32) | rev_groups32((uint32_t)(x >> 32));
}

// Reverse complement of a k-mer value (complement = code ^ 3, dna.rs:1-8): the same
// result as Kmer256::revcomp (dna.rs:152-180).  For KW = 4 the 128 groups of the
// complemented 256-bit value are reversed (words swapped end for end, each word's groups
// reversed), which leaves the k bases at the top, and a shift by 256 - 2k brings them down.
template <int KW>
__device__ __forceinline__ Kv revcomp(const Kv& v, uint32_t k) {
    if (KW == 1) return Kv{rev_groups(~v.lo) >> (64 - 2 * k), 0};
    if (KW == 2) {
        const uint64_t rhi = rev_groups(~v.lo), rlo = rev_groups(~v.hi);
        const uint32_t s2 = 128 - 2 * k;  // 2..62
        return Kv{(rlo >> s2) | (rhi << (64 - s2)), rhi >> s2};
    }
    const Kv r{rev_groups(~v.w3), rev_groups(~v.w2), rev_groups(~v.hi), rev_groups(~v.lo)};
    return kv_shr(r, 256 - 2 * k);  // 2..128 for 64 <= k <= 127
}

// Kmer256::canonical (dna.rs:185-192): min(fwd, rc), ties to the forward strand.
template <int KW>
__device__ __forceinline__ Kv canonical(const Kv& v, uint32_t k) {
    const Kv rc = revcomp<KW>(v, k);
    return kv_le(v, rc) ? v : rc;
}

// Bits [0, 2k) of x kept, the rest cleared.
template <int KW>
__device__ __forceinline__ Kv kv_mask(Kv x, uint32_t k) {
    const uint32_t b = 2 * k;
    uint64_t w[4] = {x.lo, x.hi, x.w2, x.w3};
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) {
        const uint32_t lo = 64 * i;
        if (b <= lo) w[i] = 0;
        else if (b < lo + 64) w[i] &= (1ull << (b - lo)) - 1;
    }
    return Kv{w[0], w[1], w[2], w[3]};
}

// Kmer256::extend_right (dna.rs): drop the first base, append b.
template <int KW>
__device__ __forceinline__ Kv extend_right(const Kv& v, uint32_t b, uint32_t k) {
    Kv r{(v.lo << 2) | b, 0, 0, 0};
    if (KW >= 2) r.hi = (v.hi << 2) | (v.lo >> 62);
    if (KW == 4) {
        r.w2 = (v.w2 << 2) | (v.hi >> 62);
        r.w3 = (v.w3 << 2) | (v.w2 >> 62);
    }
    return kv_mask<KW>(r, k);
}

// Kmer256::prepend_left (dna.rs): drop the last base, put a first.
template <int KW>
__device__ __forceinline__ Kv prepend_left(const Kv& v, uint32_t a, uint32_t k) {
    Kv r = kv_shr(v, 2);
    const uint32_t bit = 2 * k - 2, w = bit >> 6;
    const uint64_t m = (uint64_t)a << (bit & 63);
    if (w == 0) r.lo |= m;
    else if (w == 1) r.hi |= m;
    else if (w == 2) r.w2 |= m;
    else r.w3 |= m;
    return r;
}

// The 2k bits starting at bit `sh` (0..31) of a big-endian bit stream held in 32-bit
// words p[0..] (byte 0 = bits 31..24 of the first word after bswap).  p must be 4-byte
// aligned with 3 (KW = 1), 5 (KW = 2) or 9 (KW = 4) readable words.
template <int KW>
__device__ __forceinline__ Kv extract_be(const uint32_t* p, uint32_t sh, uint32_t k) {
    if (KW == 1) {
        const uint64_t w0 = __builtin_bswap32(p[0]), w1 = __builtin_bswap32(p[1]);
        const uint32_t w2 = __builtin_bswap32(p[2]);
        uint64_t top = (w0 << 32) | w1;
        top = sh ? ((top << sh) | (uint64_t)(w2 >> (32 - sh))) : top;
        return Kv{top >> (64 - 2 * k), 0};
    }
    if (KW == 2) {
        const uint64_t a = ((uint64_t)__builtin_bswap32(p[0]) << 32) | __builtin_bswap32(p[1]);
        const uint64_t b = ((uint64_t)__builtin_bswap32(p[2]) << 32) | __builtin_bswap32(p[3]);
        const uint64_t c = (uint64_t)__builtin_bswap32(p[4]) << 32;
        const uint64_t hi = sh ? ((a << sh) | (b >> (64 - sh))) : a;
        const uint64_t lo = sh ? ((b << sh) | (c >> (64 - sh))) : b;
        const uint32_t s2 = 128 - 2 * k;  // 2..62 for 33 <= k <= 63
        return Kv{(lo >> s2) | (hi << (64 - s2)), hi >> s2};
    }
    uint64_t a[5];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        a[i] = ((uint64_t)__builtin_bswap32(p[2 * i]) << 32) | __builtin_bswap32(p[2 * i + 1]);
    a[4] = (uint64_t)__builtin_bswap32(p[8]) << 32;
    uint64_t t[4];  // t[0] = the most significant word of the window's 256 leading bits
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = sh ? ((a[i] << sh) | (a[i + 1] >> (64 - sh))) : a[i];
    return kv_shr(Kv{t[3], t[2], t[1], t[0]}, 256 - 2 * k);
}

// The four u64 limbs hash_kmer256 reads (bloom.rs:61-66) of the Kmer256 holding V:
// for k <= 64, Kmer256(V, 0); for k > 64, Kmer256(first 64 bases, the other k - 64
// bases) (dna.rs:134-138, 299-321).
template <int KW>
__device__ __forceinline__ void kmer_limbs(const Kv& v, uint32_t k, uint64_t w[4]) {
    if (KW <= 2) {
        w[0] = v.lo;
        w[1] = v.hi;
        w[2] = w[3] = 0;
        return;
    }
    const uint32_t s = 2 * k - 128;  // bits of the second u128, 0..126
    const Kv top = kv_shr(v, s);
    w[0] = top.lo;
    w[1] = top.hi;
    w[2] = s >= 64 ? v.lo : (v.lo & ((1ull << s) - 1));
    w[3] = s > 64 ? (v.hi & ((1ull << (s - 64)) - 1)) : 0ull;
}

// Window starting at base s of a read whose packed bytes start at byte `off` of `data`
// (PackedReads layout, packed_reads.rs:547-578: base j in byte j/4, bits 6-2*(j%4)).
// `data` must be readable 24 bytes past the last read byte.
template <int KW>
__device__ __forceinline__ Kv load_window(const uint8_t* __restrict__ data, uint64_t off, uint32_t s, uint32_t k) {
    const uint64_t b0 = off + (s >> 2);
    const uint64_t a = b0 & ~3ull;
    const uint32_t sh = (uint32_t)(b0 - a) * 8 + 2 * (s & 3);
    return extract_be<KW>(reinterpret_cast<const uint32_t*>(data + a), sh, k);
}

// ---- bloom.rs:61-78 hash_kmer256 -------------------------------------------------

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

// h % d for a runtime d = 64 * num_words (bloom.rs:84-86).  Power-of-two sizes use a mask;
// otherwise magic = floor(2^64 / d) gives q' = mulhi(h, magic) in {q-1, q}, so one
// conditional subtraction yields the exact remainder.
struct Modulus {
    uint64_t d, magic, mask;
    uint32_t pow2;
    __device__ __forceinline__ uint64_t mod(uint64_t h) const {
        if (pow2) return h & mask;
        const uint64_t q = __umul64hi(h, magic);
        const uint64_t r = h - q * d;
        return r >= d ? r - d : r;
    }
};

constexpr uint64_t SEED1 = 0x517cc1b727220a95ull;
constexpr uint64_t SEED2 = 0x9e3779b97f4a7c15ull;
constexpr uint64_t SEED3 = 0x6c62272e07bb0142ull;

// The three Bloom bit positions h_j % size_bits of a Kmer256 given as limbs w0..w3.
__device__ __forceinline__ void bloom_positions(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3,
                                                const Modulus& M, uint64_t pos[3]) {
    const uint64_t m1 = rotl64(w1, 21) * 0x9e3779b97f4a7c15ull;
    const uint64_t m2 = rotl64(w2, 43) * 0xbf58476d1ce4e5b9ull;
    const uint64_t m3 = rotl64(w3, 11) * 0x94d049bb133111ebull;
    const uint64_t seeds[3] = {SEED1, SEED2, SEED3};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        uint64_t x = w0 ^ seeds[j];
        x = (x ^ m1) * 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        x = (x ^ m2) * 0xc4ceb9fe1a85ec53ull;
        x ^= x >> 33;
        x = (x ^ m3) * 0xff51afd7ed558ccdull;
        x ^= x >> 33;
        pos[j] = M.mod(x);
    }
}

// The Bloom bit positions of a k-mer value.
template <int KW>
__device__ __forceinline__ void kv_positions(const Kv& v, uint32_t k, const Modulus& M, uint64_t pos[3]) {
    uint64_t w[4];
    kmer_limbs<KW>(v, k, w);
    bloom_positions(w[0], w[1], w[2], w[3], M, pos);
}

// The same test reading the next bit only while every earlier one is set: a sparse
// solid tier rejects most k-mers on their first bit, so most lanes issue one load, not
// three (scans over many distinct k-mers are bound by these random reads).
template <int KW>
__device__ __forceinline__ bool bloom_is_solid_lazy(const Kv& v, uint32_t k, const uint64_t* __restrict__ solid,
                                                    const Modulus& M) {
    uint64_t p[3];
    kv_positions<KW>(v, k, M, p);
    if (!((solid[p[0] >> 6] >> (p[0] & 63)) & 1ull)) return false;
    if (!((solid[p[1] >> 6] >> (p[1] & 63)) & 1ull)) return false;
    return ((solid[p[2] >> 6] >> (p[2] & 63)) & 1ull) != 0;
}

// bloom.rs:111-126 is_solid_kmer256 for a k-mer value.
template <int KW>
__device__ __forceinline__ bool bloom_is_solid(const Kv& v, uint32_t k, const uint64_t* __restrict__ solid,
                                               const Modulus& M) {
    uint64_t p[3];
    kv_positions<KW>(v, k, M, p);
    const uint64_t a = solid[p[0] >> 6], b = solid[p[1] >> 6], c = solid[p[2] >> 6];
    return ((a >> (p[0] & 63)) & (b >> (p[1] & 63)) & (c >> (p[2] & 63)) & 1ull) != 0;
}

// ---- device hash table -----------------------------------------------------------
//
// Open addressing, linear probing, one slot = key words + first-occurrence event + count.
// NW = 1: the key is V itself (k <= 32; a canonical 32-mer is never all ones because
// revcomp(T^32) = A^32 < T^32, so ~0 can mark an empty word).  NW = 2 (33 <= k <= 63):
// V (< 2^126) is split into two 63-bit words A = V >> 63, B = V & (2^63 - 1), neither of
// which can be ~0.  NW = 5 (64 <= k <= 127): V (< 2^254) as five words of at most 63
// bits, lowest bits first.  A slot is claimed word by word with 64-bit CAS (word 0
// first): every word goes from EMPTY to its final value exactly once, so all inserters of
// one key take the same decision at every slot and the key lands in exactly one slot
// without locks.

// A slot is 16 bytes (one key word, 32-bit events), 32 bytes (NW <= 2) or 64 bytes
// (NW = 5), so a 64-byte line holds a bucket of SPB = 4, 2 or 1 whole slots.  Probing
// scans a bucket in slot order, then moves to the next bucket.
template <int NW, typename Ev>
struct alignas((NW == 1 && sizeof(Ev) == 4) ? 16 : NW <= 2 ? 32 : 64) Slot {
    uint64_t k[NW];
    Ev t0;         // smallest event index (global window number) of this key
    uint32_t cnt;  // occurrences
};

template <int NW, typename Ev>
constexpr uint32_t slots_per_bucket() {
    return 64 / sizeof(Slot<NW, Ev>);
}

constexpr uint64_t M63 = 0x7FFFFFFFFFFFFFFFull;

template <int NW>
__device__ __forceinline__ void pack_key(const Kv& v, uint64_t key[NW]) {
    if (NW == 1) {
        key[0] = v.lo;
    } else if (NW == 2) {
        key[0] = (v.hi << 1) | (v.lo >> 63);
        key[NW - 1] = v.lo & M63;
    } else {  // V bits 0..62, 63..125, 126..188, 189..251, 252..253
        key[0] = v.lo & M63;
        key[1] = ((v.lo >> 63) | (v.hi << 1)) & M63;
        key[2] = ((v.hi >> 62) | (v.w2 << 2)) & M63;
        key[3] = ((v.w2 >> 61) | (v.w3 << 3)) & M63;
        key[NW - 1] = v.w3 >> 60;
    }
}

template <int NW>
__device__ __forceinline__ Kv unpack_key(const uint64_t* key) {
    if (NW == 1) return Kv{key[0], 0};
    if (NW == 2) return Kv{key[NW - 1] | (key[0] << 63), key[0] >> 1};
    return Kv{key[0] | (key[1] << 63), (key[1] >> 1) | (key[2] << 62), (key[2] >> 2) | (key[3] << 61),
              (key[3] >> 3) | (key[NW - 1] << 60)};
}

__device__ __forceinline__ uint64_t fmix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

template <int NW>
__device__ __forceinline__ uint64_t table_hash(const uint64_t key[NW]) {
    uint64_t h = key[0] * 0x9E3779B97F4A7C15ull;
    if (NW == 2) h ^= key[NW - 1] + 0x632BE59BD9B4E019ull;
    if (NW == 5) {
#pragma unroll
        for (int w = 1; w < NW; ++w) h = fmix64(h ^ (key[w] + 0x632BE59BD9B4E019ull * (uint64_t)w));
    }
    return fmix64(h);
}

// Probe loads of table words.  Every key word goes from EMPTY to its final value once and
// t0 only decreases, so any value a load returns is either current or an older one that
// the protocol tolerates (an EMPTY key word is settled by the CAS, a larger t0 only costs
// an extra atomicMin).  Workgroup scope lets the per-XCD L2 serve the line.
template <typename T>
__device__ __forceinline__ T ld_relaxed(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// First key word (and, when T0 is given, the first event) of every slot of bucket b (one
// 64-byte line).
template <int NW, typename Ev>
__device__ __forceinline__ void bucket_load(const Slot<NW, Ev>* __restrict__ tab, uint64_t b,
                                            uint64_t w0[slots_per_bucket<NW, Ev>()],
                                            Ev* t0 = nullptr) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
#pragma unroll
    for (uint32_t s = 0; s < SPB; ++s) {
        w0[s] = ld_relaxed(&tab[b * SPB + s].k[0]);
        if (t0) t0[s] = ld_relaxed(&tab[b * SPB + s].t0);
    }
}

// Try to find or claim key in bucket b given the bucket's first key words w0 (possibly
// stale: a stale word is EMPTY, and the CAS then returns the current value).
// Returns the slot index, or ~0 when the bucket holds SPB other keys.
template <int NW, typename Ev>
__device__ __forceinline__ uint64_t bucket_find(Slot<NW, Ev>* __restrict__ tab, uint64_t b,
                                                const uint64_t w0[slots_per_bucket<NW, Ev>()],
                                                const uint64_t key[NW]) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
#pragma unroll
    for (uint32_t s = 0; s < SPB; ++s) {
        Slot<NW, Ev>* sl = tab + b * SPB + s;
        bool match = true;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            uint64_t cur = w == 0 ? w0[s] : ld_relaxed(&sl->k[w]);
            if (cur == EMPTY_KEY) {
                cur = atomicCAS(reinterpret_cast<unsigned long long*>(&sl->k[w]), (unsigned long long)EMPTY_KEY,
                                (unsigned long long)key[w]);
                if (cur == EMPTY_KEY) cur = key[w];
            }
            if (cur != key[w]) {
                match = false;
                break;
            }
        }
        if (match) return b * SPB + s;
    }
    return ~0ull;
}

// Record one occurrence (event t) in slot i.  seen_t0 is a value of the slot's t0 read
// earlier in this kernel: t0 only decreases, so t >= seen_t0 means t cannot be the
// minimum and the atomicMin is skipped.
template <int NW, typename Ev, bool TRACK_T0>
__device__ __forceinline__ void slot_hit(Slot<NW, Ev>* __restrict__ tab, uint64_t i, Ev t, Ev seen_t0) {
    // Neither result is needed, so both are issued without return and the wave does not wait.
    (void)__hip_atomic_fetch_add(&tab[i].cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (TRACK_T0 && t < seen_t0)
        (void)__hip_atomic_fetch_min(&tab[i].t0, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Insert one occurrence of key, starting at bucket b whose first key words are w0 and
// whose first events are t0 (when TRACK_T0).  Returns false when max_probe buckets were
// inspected without finding or claiming it.  The table region is buckets
// [b0, b0 + nbuckets); probing wraps inside it.
template <int NW, typename Ev, bool TRACK_T0>
__device__ __forceinline__ bool table_insert_from(Slot<NW, Ev>* __restrict__ tab, uint64_t b0, uint64_t nbuckets,
                                                  uint64_t b, uint64_t w0[slots_per_bucket<NW, Ev>()],
                                                  Ev t0[slots_per_bucket<NW, Ev>()], const uint64_t key[NW], Ev t,
                                                  uint32_t max_probe) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    const uint64_t probes = nbuckets < max_probe ? nbuckets : max_probe;
    for (uint64_t probe = 0; probe < probes; ++probe) {
        const uint64_t i = bucket_find<NW, Ev>(tab, b, w0, key);
        if (i != ~0ull) {
            Ev seen = (Ev)~(Ev)0;
            if (TRACK_T0) {
#pragma unroll
                for (uint32_t s = 0; s < SPB; ++s)
                    if (i == b * SPB + s) seen = t0[s];
            }
            slot_hit<NW, Ev, TRACK_T0>(tab, i, t, seen);
            return true;
        }
        b = (b + 1 == b0 + nbuckets) ? b0 : b + 1;
        bucket_load<NW, Ev>(tab, b, w0, TRACK_T0 ? t0 : nullptr);
    }
    return false;
}

// Find or claim key starting at bucket b (first key words w0); returns its slot, or ~0
// when max_probe buckets were inspected without finding or claiming it.
template <int NW, typename Ev>
__device__ __forceinline__ uint64_t table_find_or_claim(Slot<NW, Ev>* __restrict__ tab, uint64_t b0,
                                                        uint64_t nbuckets, uint64_t b,
                                                        uint64_t w0[slots_per_bucket<NW, Ev>()],
                                                        const uint64_t key[NW], uint32_t max_probe) {
    const uint64_t probes = nbuckets < max_probe ? nbuckets : max_probe;
    for (uint64_t probe = 0; probe < probes; ++probe) {
        const uint64_t i = bucket_find<NW, Ev>(tab, b, w0, key);
        if (i != ~0ull) return i;
        b = (b + 1 == b0 + nbuckets) ? b0 : b + 1;
        bucket_load<NW, Ev>(tab, b, w0);
    }
    return ~0ull;
}

// Read-only lookup: the slot holding key, or ~0 when absent.  A key lives in the first
// bucket from its home that had a free slot when it was inserted, and slots are never
// freed, so a bucket with an EMPTY slot ends the search.
template <int NW, typename Ev>
__device__ __forceinline__ uint64_t table_find(const Slot<NW, Ev>* __restrict__ tab, uint64_t b0, uint64_t nbuckets,
                                               uint64_t b, const uint64_t key[NW]) {
    constexpr uint32_t SPB = slots_per_bucket<NW, Ev>();
    for (uint64_t probe = 0; probe < nbuckets; ++probe) {
#pragma unroll
        for (uint32_t s = 0; s < SPB; ++s) {
            const Slot<NW, Ev>* sl = tab + b * SPB + s;
            const uint64_t w0 = sl->k[0];
            if (w0 == EMPTY_KEY) return ~0ull;
            bool match = w0 == key[0];
#pragma unroll
            for (int w = 1; w < NW; ++w) match = match && sl->k[w] == key[w];
            if (match) return b * SPB + s;
        }
        b = (b + 1 == b0 + nbuckets) ? b0 : b + 1;
    }
    return ~0ull;
}

template <int NW>
__device__ __forceinline__ uint64_t home_bucket(const uint64_t key[NW], uint64_t nbuckets) {
    return __umul64hi(table_hash<NW>(key), nbuckets);
}

// Wave-aggregated append: returns a unique index for every lane with pred set.
__device__ __forceinline__ uint64_t wave_append(bool pred, unsigned long long* counter) {
    const uint64_t mask = __ballot(pred);
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (mask != 0 && (int)lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader < 0 ? 0 : leader);
    const uint64_t below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
    return base + __popcll(below);
}

}  // namespace kc
