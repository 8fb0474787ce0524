// kc_superkmer.h — cutting a packed read into super-k-mers by canonical-minimizer bin
// (the minimizer-binned Stage 2 count in kc_engine.hip).  Host and device: the host
// build is what the CPU tests of the cut run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kc {

// Bins: 2^bits, bits chosen per run (bin_bits_for) for at most BIN_TARGET distinct k-mers a
// bin on average.  Measured on 5 Mb / 100x: an average of 190-330 counts fastest; twice as
// many bins pay their per-bin staging and output (k = 33: 3.9 -> 6.5 ms), half as many
// split more bins into passes (k = 33: 5.3 ms).
constexpr uint32_t BIN_BITS_MIN = 14, BIN_BITS_MAX = 24;
constexpr uint32_t BIN_TARGET = 352;
constexpr uint32_t BIN3_CAP = 1024;  // LDS slots per pass
constexpr uint32_t BIN3_FILL = BIN3_CAP * 3 / 4;
constexpr uint32_t BIN3_MAX_SPLIT = 1u << 12;  // passes per bin before giving up
constexpr uint32_t SK_MAX_WINDOWS = 255;       // windows per record
constexpr uint32_t SK_MAX_LEN = 1u << 24;      // read length the record can address

// Minimizer length: the whole k-mer up to 11, then 11 (k <= 32), then k - 17 (windows of
// 18 m-mers) between 16 and 32.  A longer m for the long k-mers shortens the super-k-mers
// and moves more of an error's k-mers out of the genomic super-k-mer's bin, so fewer bins
// pass the LDS table's fill and are counted in split passes.  Measured on 5 Mb / 100x
// (binned count, 2^18 bins): k = 55 at m = 16, 24, 28 and 32 took 5.95, 4.86, 4.46 and 4.16
// ms (the bin sort 0.44 -> 0.62 ms for the shorter super-k-mers); k = 63 at 16 and 32, 6.53
// and 4.47 ms; k = 45 at 16, 21 and 26, 5.01, 4.45 and 4.13 ms; k = 33 at 12, 14, 16, 18 and
// 20, 5.10, 4.25, 4.11, 4.00 and 4.03 ms with the cut and bin sort slower from 18 on.  An
// m-mer key takes 64 bits, so m stops at 32.
__host__ __device__ constexpr uint32_t minimizer_len(uint32_t k) {
    return k <= 11 ? k : k <= 32 ? 11 : k - 17 < 16 ? 16 : k - 17 > 32 ? 32 : k - 17;
}

// MurmurHash3's 32-bit finalizer (fmix32): a bijection of 32-bit values, two 32-bit
// multiplies.
__host__ __device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x85ebca6bu;
    x ^= x >> 13;
    x *= 0xc2b2ae35u;
    x ^= x >> 16;
    return x;
}

// Base j of a packed read at byte offset off (first base in bits 7..6).
__host__ __device__ __forceinline__ uint32_t packed_base(const uint8_t* __restrict__ data, uint64_t off, uint32_t j) {
    return (uint32_t)(data[off + (j >> 2)] >> (6 - 2 * (j & 3))) & 3u;
}

// Sequential reader of the 2-bit bases of a packed read (first base in bits 7..6 of its
// first byte): keeps the current 32-bit word, so a scan loads one word per 16 bases.
// The packed data must be 4-byte aligned and readable 3 bytes past the read's end.
struct BaseReader {
    const uint32_t* w;
    uint64_t bit0;
    uint64_t idx = ~0ull;
    uint32_t word = 0;
    __host__ __device__ BaseReader(const uint8_t* data, uint64_t off)
        : w(reinterpret_cast<const uint32_t*>(data)), bit0(off * 8) {}
    __host__ __device__ __forceinline__ uint32_t operator()(uint32_t j) {
        const uint64_t bp = bit0 + 2ull * j;
        if ((bp >> 5) != idx) {
            idx = bp >> 5;
            word = __builtin_bswap32(w[idx]);
        }
        return (word >> (30 - (uint32_t)(bp & 31))) & 3u;
    }
};

// The 32-bit order key of a canonical m-mer: its mix, a bijection for m <= 16 (distinct
// m-mers get distinct keys); for m > 16 the bits above the low 32 are folded in first (a
// key shared by two m-mers only ties their order, which any key function may).
__host__ __device__ __forceinline__ uint32_t mmer_key(uint64_t canonical_mmer) {
    return mix32((uint32_t)canonical_mmer ^ (uint32_t)(canonical_mmer >> 32) * 0x9E3779B1u);
}

// The bin of a window from its minimizer's key (the smallest key of its m-mers).  The
// minimum of w keys sits near 0, so its top bits would crowd the low bins: the key is
// spread by one multiplication by the golden-ratio constant (Fibonacci hashing), whose top
// bits depend on all of the key's bits.  (Against a 64-bit mix, on the 5 Mb / 100x reads:
// the windows counted with split passes 1.196 / 1.002 / 1.077 times the windows at
// k = 55 / 33 / 21, against 1.218 / 1.005 / 1.069.)
__host__ __device__ __forceinline__ uint32_t minimizer_bin(uint32_t minkey, uint32_t bits) {
    return (uint32_t)((minkey ^ 0x7f4a7c15u) * 0x9E3779B1u) >> (32 - bits);
}

// Bin bits for an expected number of distinct k-mers, about `target` a bin.
inline uint32_t bin_bits_for(uint64_t distinct, uint32_t target = BIN_TARGET) {
    uint32_t b = BIN_BITS_MIN;
    while (b < BIN_BITS_MAX && ((uint64_t)target << b) < distinct) ++b;
    return b;
}

// Calls emit(first window, windows, bin) for every super-k-mer of a read of len >= k
// bases at byte offset off: the maximal runs of consecutive windows in one bin, cut after
// SK_MAX_WINDOWS windows.  A window's bin is minimizer_bin (2^bits bins) of the smallest mmer_key over
// its w = k - m + 1 canonical m-mers, which a k-mer and its reverse complement share.
// This sequential form is the definition (and the host reference); the device cuts reads
// one wave per read (k_sk_wave).
template <class F>
__host__ __device__ __forceinline__ void superkmers(const uint8_t* __restrict__ data, uint64_t off, uint32_t len, uint32_t k,
                                           uint32_t bits, F&& emit) {
    const uint32_t m = minimizer_len(k), w = k - m + 1, nwin = len - k + 1;
    const uint64_t mask = m >= 32 ? ~0ull : (1ull << (2 * m)) - 1;
    const uint32_t top = 2 * (m - 1);
    BaseReader base(data, off);
    uint64_t f = 0, r = 0, minh = ~0ull;  // minh: a key, or ~0 before the first
    uint32_t minp = 0;
    // minimum over the m-mers [i0, i0 + w) of the window starting at i0; leaves f, r at
    // the last of them
    auto rescan = [&](uint32_t i0) {
        f = r = 0;
        for (uint32_t j = i0; j < i0 + m - 1; ++j) {
            const uint32_t b = base(j);
            f = ((f << 2) | b) & mask;
            r = (r >> 2) | ((uint64_t)(3 - b) << top);
        }
        minh = ~0ull;
        for (uint32_t i = i0; i < i0 + w; ++i) {
            const uint32_t b = base(i + m - 1);
            f = ((f << 2) | b) & mask;
            r = (r >> 2) | ((uint64_t)(3 - b) << top);
            const uint64_t h = mmer_key(f < r ? f : r);
            if (h < minh) {
                minh = h;
                minp = i;
            }
        }
    };
    rescan(0);
    uint32_t bin = minimizer_bin((uint32_t)minh, bits), first = 0;
    for (uint32_t j = 1; j < nwin; ++j) {
        if (minp < j) {
            rescan(j);
        } else {
            const uint32_t b = base(j + k - 1);
            f = ((f << 2) | b) & mask;
            r = (r >> 2) | ((uint64_t)(3 - b) << top);
            const uint64_t h = mmer_key(f < r ? f : r);
            if (h < minh) {
                minh = h;
                minp = j + w - 1;
            }
        }
        const uint32_t nb = minimizer_bin((uint32_t)minh, bits);
        if (nb != bin || j - first == SK_MAX_WINDOWS) {
            emit(first, j - first, bin);
            first = j;
            bin = nb;
        }
    }
    emit(first, nwin - first, bin);
}


}  // namespace kc
