// Host build of kc::superkmers (spades-rs_amd/csrc/kc_superkmer.h) checked on random,
// low-complexity and long reads for k = 1..63: the records tile each read's windows in
// order with 1..SK_MAX_WINDOWS windows each, every window's bin equals a direct
// recomputation from its k bases (smallest m-mer key over its canonical m-mers), and a read
// and its reverse complement put each canonical k-mer in the same bin -- the property that
// makes the binned count exact.  Exit status 0 when every check holds.
#include <cstdio>
#include <random>
#include <vector>

#include "../spades-rs_amd/csrc/kc_superkmer.h"

using namespace kc;

static uint32_t direct_bin(const std::vector<uint8_t>& b, uint32_t j, uint32_t k, uint32_t bits) {
    const uint32_t m = minimizer_len(k), w = k - m + 1;
    uint64_t best = ~0ull;
    for (uint32_t i = j; i < j + w; ++i) {
        uint64_t f = 0, r = 0;
        for (uint32_t t = 0; t < m; ++t) {
            f = (f << 2) | b[i + t];
            r = (r << 2) | (3 - b[i + m - 1 - t]);
        }
        const uint64_t h = mmer_key(f < r ? f : r);
        if (h < best) best = h;
    }
    return minimizer_bin((uint32_t)best, bits);
}

static std::vector<uint8_t> pack(const std::vector<uint8_t>& b, uint64_t off) {
    std::vector<uint8_t> p(off + (b.size() + 3) / 4 + 64, 0);
    for (size_t j = 0; j < b.size(); ++j) p[off + j / 4] |= b[j] << (6 - 2 * (j % 4));
    return p;
}

int main() {
    std::mt19937_64 rng(7);
    long bad = 0, checked = 0;
    for (uint32_t k : {1u, 3u, 5u, 11u, 12u, 21u, 31u, 32u, 33u, 47u, 55u, 63u}) {
        for (int it = 0; it < 200; ++it) {
            const uint32_t bits = BIN_BITS_MIN + it % (BIN_BITS_MAX - BIN_BITS_MIN + 1);
            uint32_t len = k + rng() % 300;
            if (it % 50 == 0) len = k + 2000 + rng() % 2000;  // runs cut at SK_MAX_WINDOWS
            std::vector<uint8_t> b(len);
            const int mode = it % 4;
            for (auto& x : b) x = mode == 3 ? (rng() % 8 == 0 ? rng() % 4 : 0) : rng() % 4;
            std::vector<uint8_t> rc(len);
            for (uint32_t j = 0; j < len; ++j) rc[j] = 3 - b[len - 1 - j];
            const uint64_t off = rng() % 5;
            const auto p = pack(b, off), prc = pack(rc, off);
            const uint32_t nwin = len - k + 1;
            std::vector<uint32_t> bins(nwin, ~0u), bins_rc(nwin, ~0u);
            uint32_t next = 0;
            superkmers(p.data(), off, len, k, bits, [&](uint32_t first, uint32_t nw, uint32_t bin) {
                if (first != next || nw == 0 || nw > SK_MAX_WINDOWS) ++bad;
                for (uint32_t j = first; j < first + nw && j < nwin; ++j) bins[j] = bin;
                next = first + nw;
            });
            if (next != nwin) ++bad;
            superkmers(prc.data(), off, len, k, bits, [&](uint32_t first, uint32_t nw, uint32_t bin) {
                for (uint32_t j = first; j < first + nw && j < nwin; ++j) bins_rc[j] = bin;
            });
            for (uint32_t j = 0; j < nwin; ++j) {
                ++checked;
                if (bins[j] != direct_bin(b, j, k, bits)) ++bad;
                if (bins[j] != bins_rc[nwin - 1 - j]) ++bad;
                if (bins[j] >> bits) ++bad;
            }
        }
    }
    printf("windows checked %ld, failures %ld\n", checked, bad);
    return bad != 0;
}
