// oracle.cpp — CPU restatement of the spades-rs k-mer spectrum path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the MI355X engine:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// the library built from it.  Nothing in the product links it.
//
// It restates, in plain C++17, the reference Rust code listed below.  Each
// function cites the reference file:line it follows.  Run with threads == 1
// it is the sequential order of the reference at `-t 1` (read index 0..n,
// window 0..L-k, hashes h1 -> h2 -> h3), which is the only well-defined
// bit-exact target because the reference's solid Bloom tier depends on
// insertion order under rayon (bloom.rs:95-103).  Run with threads > 1 it
// follows the reference's rayon chunking and mutex-sharded count maps
// (assemble.rs:136-242) and serves as the timed CPU baseline.
//
// Pinning: tests/test_oracle_*.py check every function here against the
// reference's own known-answer tests (dna.rs:329-432, bloom.rs:138-156,
// hammer.rs:150-192, tests/test_dna_kmer.rs, tests/test_hammer_and_bloom.rs,
// tests/test_packed_reads.rs), against an independent NumPy transcription of
// bloom.rs:61-126 (oracle/hash_np.py), and against closed forms / brute force
// (seen tier = OR over all occurrences; counts = brute-force multiplicities;
// cutoff on hand-built histograms).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

typedef unsigned __int128 u128;

// ---------------------------------------------------------------------------
// dna.rs
// ---------------------------------------------------------------------------

// dna.rs:11-22 — ASCII -> 2-bit code, 0xFF for anything that is not ACGTacgt.
static uint8_t ascii_to_2bit(uint8_t b) {
    switch (b) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return 0xFF;
    }
}

// dna.rs:25,41-43
static uint8_t bit2_to_base(uint8_t v) { return "ACGT"[v & 3]; }

static uint64_t reverse_bits64(uint64_t x) {
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

static u128 reverse_bits128(u128 x) {
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    return ((u128)reverse_bits64(lo) << 64) | (u128)reverse_bits64(hi);
}

// dna.rs:65-78 revcomp_kmer_u64
static uint64_t revcomp_kmer_u64(uint64_t kmer, unsigned k) {
    if (k == 0 || k > 32) return 0;
    kmer = reverse_bits64(~kmer);
    kmer = ((kmer >> 1) & 0x5555555555555555ULL) | ((kmer & 0x5555555555555555ULL) << 1);
    unsigned sh = 2 * (32 - k);
    return sh >= 64 ? 0 : (kmer >> sh);
}

// dna.rs:82-89 canonical_kmer_u64 (ties go to the forward strand)
static uint64_t canonical_kmer_u64(uint64_t kmer, unsigned k, int* is_rc) {
    uint64_t rc = revcomp_kmer_u64(kmer, k);
    if (kmer <= rc) { *is_rc = 0; return kmer; }
    *is_rc = 1;
    return rc;
}

// dna.rs:118-132 revcomp_kmer_u128
static u128 revcomp_kmer_u128(u128 kmer, unsigned k) {
    if (k == 0 || k > 64) return 0;
    const u128 m55 = ((u128)0x5555555555555555ULL << 64) | (u128)0x5555555555555555ULL;
    kmer = reverse_bits128(~kmer);
    kmer = ((kmer >> 1) & m55) | ((kmer & m55) << 1);
    if (k == 64) return kmer;
    return kmer >> (2 * (64 - k));
}

// dna.rs:137-138 — Kmer256(lo, hi); derived Ord is (lo, hi) lexicographic.
struct Kmer256 {
    u128 lo, hi;
    bool operator==(const Kmer256& o) const { return lo == o.lo && hi == o.hi; }
    bool operator<(const Kmer256& o) const { return lo < o.lo || (lo == o.lo && hi < o.hi); }
    bool operator<=(const Kmer256& o) const { return !(o < *this); }
};

struct Kmer256Hash {
    size_t operator()(const Kmer256& x) const {
        uint64_t a = (uint64_t)x.lo, b = (uint64_t)(x.lo >> 64);
        uint64_t c = (uint64_t)x.hi, d = (uint64_t)(x.hi >> 64);
        uint64_t h = a * 0x9E3779B97F4A7C15ULL;
        h ^= (b + 0x632BE59BD9B4E019ULL) * 0xC2B2AE3D27D4EB4FULL;
        h ^= (c ^ (d << 1)) * 0x165667B19E3779F9ULL;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ULL;
        h ^= h >> 29;
        return (size_t)h;
    }
};

// dna.rs:152-180 Kmer256::revcomp
static Kmer256 kmer_revcomp(const Kmer256& x, unsigned k) {
    if (k == 0 || k > 128) return Kmer256{0, 0};
    if (k <= 64) return Kmer256{revcomp_kmer_u128(x.lo, k), 0};
    unsigned k2 = k - 64;
    u128 lo_rc = revcomp_kmer_u128(x.lo, 64);
    u128 hi_rc = revcomp_kmer_u128(x.hi, k2);
    unsigned shift_lo = 2 * (64 - k2);
    u128 rc_lo = (shift_lo >= 128 ? (u128)0 : (hi_rc << shift_lo)) |
                 (k2 >= 64 ? (u128)0 : (lo_rc >> (2 * k2)));
    u128 mask_hi = k2 >= 64 ? ~(u128)0 : (((u128)1 << (2 * k2)) - 1);
    u128 rc_hi = lo_rc & mask_hi;
    return Kmer256{rc_lo, rc_hi};
}

// dna.rs:185-192 Kmer256::canonical
static Kmer256 kmer_canonical(const Kmer256& x, unsigned k, int* is_rc) {
    Kmer256 rc = kmer_revcomp(x, k);
    if (x <= rc) { if (is_rc) *is_rc = 0; return x; }
    if (is_rc) *is_rc = 1;
    return rc;
}

// dna.rs:196-221 Kmer256::extend_right
static Kmer256 kmer_extend_right(const Kmer256& x, uint8_t b, unsigned k) {
    if (k == 0 || k > 128) return Kmer256{0, 0};
    u128 b2 = (u128)(b & 3);
    if (k <= 64) {
        u128 mask = k == 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1);
        return Kmer256{((x.lo << 2) | b2) & mask, 0};
    }
    unsigned k2 = k - 64;
    u128 mask_hi = k2 == 64 ? ~(u128)0 : (((u128)1 << (2 * k2)) - 1);
    u128 s64 = (x.hi >> (2 * (k2 - 1))) & 3;
    return Kmer256{(x.lo << 2) | s64, ((x.hi << 2) | b2) & mask_hi};
}

// dna.rs:225-245 Kmer256::prepend_left
static Kmer256 kmer_prepend_left(const Kmer256& x, uint8_t a, unsigned k) {
    if (k == 0 || k > 128) return Kmer256{0, 0};
    u128 a2 = (u128)(a & 3);
    if (k <= 64) return Kmer256{(x.lo >> 2) | (a2 << (2 * (k - 1))), 0};
    unsigned k2 = k - 64;
    u128 mask_hi = k2 == 64 ? ~(u128)0 : (((u128)1 << (2 * k2)) - 1);
    u128 s63 = x.lo & 3;
    return Kmer256{(x.lo >> 2) | (a2 << 126), ((x.hi >> 2) | (s63 << (2 * (k2 - 1)))) & mask_hi};
}

// dna.rs:299-321 Kmer256::from_bytes — O(k) re-encode of an ASCII window.
static bool kmer_from_bytes(const uint8_t* s, size_t len, unsigned k, Kmer256* out) {
    if (k == 0 || len < k || k > 128) return false;
    u128 lo = 0, hi = 0;
    if (k <= 64) {
        for (unsigned i = 0; i < k; ++i) {
            uint8_t c = ascii_to_2bit(s[i]);
            if (c == 0xFF) return false;
            lo = (lo << 2) | (u128)c;
        }
    } else {
        for (unsigned i = 0; i < 64; ++i) {
            uint8_t c = ascii_to_2bit(s[i]);
            if (c == 0xFF) return false;
            lo = (lo << 2) | (u128)c;
        }
        for (unsigned i = 64; i < k; ++i) {
            uint8_t c = ascii_to_2bit(s[i]);
            if (c == 0xFF) return false;
            hi = (hi << 2) | (u128)c;
        }
    }
    *out = Kmer256{lo, hi};
    return true;
}

// dna.rs:274-295 Kmer256::to_string
static void kmer_to_string(const Kmer256& x, unsigned k, uint8_t* out) {
    if (k <= 64) {
        u128 v = x.lo;
        for (int i = (int)k - 1; i >= 0; --i) { out[i] = bit2_to_base((uint8_t)(v & 3)); v >>= 2; }
    } else {
        u128 v = x.lo;
        for (int i = 63; i >= 0; --i) { out[i] = bit2_to_base((uint8_t)(v & 3)); v >>= 2; }
        u128 w = x.hi;
        for (int i = (int)k - 1; i >= 64; --i) { out[i] = bit2_to_base((uint8_t)(w & 3)); w >>= 2; }
    }
}

static void kmer_to_words(const Kmer256& x, uint64_t* w) {
    w[0] = (uint64_t)x.lo; w[1] = (uint64_t)(x.lo >> 64);
    w[2] = (uint64_t)x.hi; w[3] = (uint64_t)(x.hi >> 64);
}

static Kmer256 kmer_from_words(const uint64_t* w) {
    return Kmer256{((u128)w[1] << 64) | w[0], ((u128)w[3] << 64) | w[2]};
}

// ---------------------------------------------------------------------------
// bloom.rs
// ---------------------------------------------------------------------------

static inline uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64 - r)); }

// bloom.rs:61-78 hash_kmer256 — three rounds over the four 64-bit limbs.
static uint64_t hash_kmer256(const Kmer256& km, uint64_t seed) {
    uint64_t w0 = (uint64_t)km.lo, w1 = (uint64_t)(km.lo >> 64);
    uint64_t w2 = (uint64_t)km.hi, w3 = (uint64_t)(km.hi >> 64);
    uint64_t x = w0 ^ seed;
    x = (x ^ (rotl64(w1, 21) * 0x9e3779b97f4a7c15ULL)) * 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x = (x ^ (rotl64(w2, 43) * 0xbf58476d1ce4e5b9ULL)) * 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    x = (x ^ (rotl64(w3, 11) * 0x94d049bb133111ebULL)) * 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return x;
}

// bloom.rs:84-86 seeds, in h1, h2, h3 order.
static const uint64_t SEEDS[3] = {0x517cc1b727220a95ULL, 0x9e3779b97f4a7c15ULL, 0x6c62272e07bb0142ULL};

// bloom.rs:13-43 TwoTierFilter
struct TwoTierFilter {
    uint64_t size_bits = 0, num_words = 0;
    std::vector<std::atomic<uint64_t>> seen, solid;
    explicit TwoTierFilter(uint64_t capacity_bits) {
        num_words = std::max<uint64_t>((capacity_bits + 63) / 64, 1);  // div_ceil(64).max(1)
        size_bits = num_words * 64;
        seen = std::vector<std::atomic<uint64_t>>(num_words);
        solid = std::vector<std::atomic<uint64_t>>(num_words);
        for (uint64_t i = 0; i < num_words; ++i) { seen[i].store(0); solid[i].store(0); }
    }
    // bloom.rs:83-107 insert_kmer256: seen fetch_or h1,h2,h3; promote iff all three were set.
    bool insert(const Kmer256& km) {
        uint64_t h[3];
        for (int j = 0; j < 3; ++j) h[j] = hash_kmer256(km, SEEDS[j]) % size_bits;
        uint64_t prev[3], bit[3];
        for (int j = 0; j < 3; ++j) {
            bit[j] = 1ULL << (h[j] % 64);
            prev[j] = seen[h[j] / 64].fetch_or(bit[j], std::memory_order_relaxed);
        }
        if ((prev[0] & bit[0]) && (prev[1] & bit[1]) && (prev[2] & bit[2])) {
            for (int j = 0; j < 3; ++j) solid[h[j] / 64].fetch_or(bit[j], std::memory_order_relaxed);
            return true;
        }
        return false;
    }
    // bloom.rs:111-126 is_solid_kmer256
    bool is_solid(const Kmer256& km) const {
        for (int j = 0; j < 3; ++j) {
            uint64_t h = hash_kmer256(km, SEEDS[j]) % size_bits;
            if (!(solid[h / 64].load(std::memory_order_relaxed) & (1ULL << (h % 64)))) return false;
        }
        return true;
    }
};

// ---------------------------------------------------------------------------
// hammer.rs
// ---------------------------------------------------------------------------

// hammer.rs:13-18
static uint32_t hamming_distance_2bit(uint64_t a, uint64_t b) {
    uint64_t d = a ^ b;
    return (uint32_t)__builtin_popcountll((d | (d >> 1)) & 0x5555555555555555ULL);
}

static bool window_solid(const TwoTierFilter& f, const uint8_t* s, unsigned k) {
    Kmer256 km;
    if (!kmer_from_bytes(s, k, k, &km)) return false;
    return f.is_solid(kmer_canonical(km, k, nullptr));
}

static uint8_t to_ascii_upper(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

// hammer.rs:34-120 ErrorCorrector::correct_read_in_place.
static bool correct_read_in_place(const TwoTierFilter& f, unsigned k, uint8_t* seq, size_t len,
                                  std::vector<uint8_t>& scratch) {
    if (k == 0 || k > 128 || len < k) return false;
    scratch.assign(seq, seq + len);
    bool modified = false;
    for (size_t i = 0; i + k <= len; ++i) {
        Kmer256 km;
        if (!kmer_from_bytes(&scratch[i], k, k, &km)) continue;
        if (f.is_solid(kmer_canonical(km, k, nullptr))) continue;
        bool have_best = false;
        size_t best_off = 0;
        uint8_t best_base = 0;
        size_t max_gain = 0;
        for (size_t offset = 0; offset < k; ++offset) {
            size_t pos = i + offset;
            uint8_t orig = scratch[pos];
            uint8_t orig_upper = to_ascii_upper(orig);
            size_t start_scan = pos >= (k - 1) ? pos - (k - 1) : 0;  // saturating_sub
            size_t end_scan = std::min(pos, len - k);
            size_t base_solid = 0;
            for (size_t s = start_scan; s <= end_scan; ++s) {
                if (s == i) continue;
                if (window_solid(f, &scratch[s], k)) ++base_solid;
            }
            static const uint8_t CAND[4] = {'A', 'C', 'G', 'T'};
            for (int c = 0; c < 4; ++c) {
                uint8_t cand = CAND[c];
                if (cand == orig_upper) continue;
                scratch[pos] = cand;
                if (window_solid(f, &scratch[i], k)) {
                    size_t solid_count = 1;  // window i is solid with this candidate
                    for (size_t s = start_scan; s <= end_scan; ++s) {
                        if (s == i) continue;
                        if (window_solid(f, &scratch[s], k)) ++solid_count;
                    }
                    size_t gain = solid_count > base_solid ? solid_count - base_solid : 0;  // saturating_sub
                    if (gain > max_gain) {
                        max_gain = gain;
                        have_best = true;
                        best_off = offset;
                        best_base = cand;
                    }
                }
                scratch[pos] = orig;
            }
        }
        if (have_best) {
            seq[i + best_off] = best_base;
            scratch[i + best_off] = best_base;
            modified = true;
        }
    }
    return modified;
}

// ---------------------------------------------------------------------------
// packed_reads.rs
// ---------------------------------------------------------------------------

// packed_reads.rs:547-578 pack_clean: MSB-first, 4 bases per byte, each read byte-aligned;
// non-ACGT bytes are skipped (release behaviour).
static void pack_clean(const uint8_t* read, size_t len, std::vector<uint8_t>& data,
                       std::vector<uint64_t>& offsets, std::vector<uint32_t>& lengths) {
    offsets.push_back(data.size());
    uint32_t clean_len = 0;
    uint8_t byte = 0;
    int shift = 6;
    for (size_t i = 0; i < len; ++i) {
        uint8_t b2 = ascii_to_2bit(read[i]);
        if (b2 == 0xFF) continue;
        ++clean_len;
        byte |= (uint8_t)(b2 << shift);
        if (shift == 0) { data.push_back(byte); byte = 0; shift = 6; }
        else shift -= 2;
    }
    if (shift != 6) data.push_back(byte);
    lengths.push_back(clean_len);
}

// packed_reads.rs:592-613 get_read
static void get_read(const uint8_t* data, uint64_t off, uint32_t len, std::vector<uint8_t>& buf) {
    buf.clear();
    if (len == 0) return;
    size_t nbytes = (len + 3) / 4;
    size_t rem = len;
    for (size_t j = 0; j < nbytes; ++j) {
        uint8_t byte = data[off + j];
        size_t take = std::min<size_t>(rem, 4);
        for (int s = 3; s >= (int)(4 - take); --s) buf.push_back(bit2_to_base((byte >> (s * 2)) & 3));
        rem -= take;
    }
}

// ---------------------------------------------------------------------------
// packed_reads.rs:60-235 FastxReader over an in-memory file; 436-460, 513-578, 630-690
// three-gate ingestion of paired files into the three-block layout
// ---------------------------------------------------------------------------

struct MemFastx {
    const uint8_t* p;
    size_t n, pos = 0;
    bool is_fastq = true, eof = false;
    std::string err;
    // read_until(b'\n'): appends through the newline; false when nothing is left
    bool read_line(std::vector<uint8_t>& out) {
        out.clear();
        if (pos >= n) return false;
        while (pos < n) {
            const uint8_t b = p[pos++];
            out.push_back(b);
            if (b == '\n') break;
        }
        return true;
    }
    static void strip_eol(std::vector<uint8_t>& v) {
        while (!v.empty() && (v.back() == '\n' || v.back() == '\r')) v.pop_back();
    }
    MemFastx(const uint8_t* p_, size_t n_) : p(p_), n(n_) {
        std::vector<uint8_t> line;
        for (;;) {  // FastxReader::open: skip blank / whitespace-only leading lines
            if (pos >= n) {
                eof = true;
                return;
            }
            if (p[pos] == '\n' || p[pos] == '\r') {
                read_line(line);
                continue;
            }
            size_t e = pos;
            while (e < n && p[e] != '\n') ++e;
            bool ws = true;
            for (size_t i = pos; i < e; ++i) ws = ws && (p[i] == ' ' || p[i] == '\t' || p[i] == '\r' || p[i] == '\f' || p[i] == 0x0B);
            if (ws) {
                read_line(line);
                continue;
            }
            if (p[pos] == '@') is_fastq = true;
            else if (p[pos] == '>') is_fastq = false;
            else {
                err = "expected '@' or '>' header";
                eof = true;
                return;
            }
            break;
        }
        if (!is_fastq) read_line(line);  // the first '>' header
    }
    // 1: a sequence in seq, 0: end of file, -1: malformed (err).  qual (FASTQ, optional)
    // receives the record's quality line without its line end.
    int next_seq(std::vector<uint8_t>& seq, std::vector<uint8_t>* qual = nullptr) {
        seq.clear();
        if (qual) qual->clear();
        if (eof) return 0;
        std::vector<uint8_t> line;
        if (is_fastq) {
            for (;;) {
                if (!read_line(line)) {
                    eof = true;
                    return 0;
                }
                strip_eol(line);
                if (line.empty()) continue;
                if (line[0] != '@') {
                    err = "expected '@' header";
                    return -1;
                }
                if (!read_line(line)) {
                    eof = true;
                    err = "unexpected EOF waiting for sequence line";
                    return -1;
                }
                strip_eol(line);
                seq = line;
                if (!read_line(line)) {
                    eof = true;
                    err = "unexpected EOF waiting for '+' separator";
                    return -1;
                }
                if (line.empty() || line[0] != '+') {
                    err = "expected '+' separator line";
                    return -1;
                }
                if (!read_line(line)) {
                    eof = true;
                    err = "unexpected EOF waiting for quality scores";
                    return -1;
                }
                strip_eol(line);
                if (line.size() != seq.size()) {
                    err = "sequence length != quality length";
                    return -1;
                }
                if (qual) *qual = line;
                return 1;
            }
        }
        for (;;) {  // FASTA
            if (!read_line(line)) {
                eof = true;
                return seq.empty() ? 0 : 1;
            }
            strip_eol(line);
            if (!line.empty() && line[0] == '>') return 1;
            seq.insert(seq.end(), line.begin(), line.end());
        }
    }
};

// The packed store, and (not part of the reference, which drops qualities) the Phred
// values of every packed base: read i's bases at quals[qoff[i] ..], each the quality
// character of the same base in the source record minus 33 (0 below '!').
struct PackedStore {
    std::vector<uint8_t> data;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    uint64_t ambiguous = 0;
    std::vector<uint8_t> quals;
    std::vector<uint64_t> qoff;
    void pack(const uint8_t* r, size_t n, const uint8_t* q = nullptr) {
        pack_clean(r, n, data, off, len);
        if (q) {
            qoff.push_back(quals.size());
            for (size_t i = 0; i < n; ++i) quals.push_back(q[i] >= 33 ? (uint8_t)(q[i] - 33) : 0);
        }
    }
    void append(const PackedStore& o) {
        const uint64_t base = data.size(), qbase = quals.size();
        data.insert(data.end(), o.data.begin(), o.data.end());
        for (uint64_t x : o.off) off.push_back(x + base);
        len.insert(len.end(), o.len.begin(), o.len.end());
        ambiguous += o.ambiguous;
        quals.insert(quals.end(), o.quals.begin(), o.quals.end());
        for (uint64_t x : o.qoff) qoff.push_back(x + qbase);
    }
};

static bool ambig(uint8_t b) { return ascii_to_2bit(b) == 0xFF; }

// trim_n_ends (packed_reads.rs:634-648): [s, e)
static void trim_n_ends(const std::vector<uint8_t>& r, size_t& s, size_t& e) {
    s = 0;
    while (s < r.size() && ambig(r[s])) ++s;
    e = r.size();
    while (e > 0 && ambig(r[e - 1])) --e;
    if (s >= e) s = e = 0;
}

// ingest_trimmed_unpaired (packed_reads.rs:513-535): split on N, pack segments >= MIN_K,
// else the ambiguous sidecar
static void ingest_trimmed(const uint8_t* t, size_t n, PackedStore& b3, const uint8_t* tq = nullptr) {
    const size_t MIN_K = 21;
    if (n == 0) return;
    bool any = false;
    size_t i = 0;
    while (i < n) {
        while (i < n && ambig(t[i])) ++i;
        size_t j = i;
        while (j < n && !ambig(t[j])) ++j;
        if (j > i && j - i >= MIN_K) {
            b3.pack(t + i, j - i, tq ? tq + i : nullptr);
            any = true;
        }
        i = j;
    }
    if (!any) b3.ambiguous += 1;
}

static void ingest_read(const std::vector<uint8_t>& r, PackedStore& b3, const std::vector<uint8_t>* q = nullptr) {
    size_t s, e;
    trim_n_ends(r, s, e);
    ingest_trimmed(r.data() + s, e - s, b3, q ? q->data() + s : nullptr);
}

// partition_pair_streaming (packed_reads.rs:654-680)
static void partition_pair(const std::vector<uint8_t>& r1, const std::vector<uint8_t>& r2, PackedStore& b1,
                           PackedStore& b2, PackedStore& b3, const std::vector<uint8_t>* q1 = nullptr,
                           const std::vector<uint8_t>* q2 = nullptr) {
    const size_t MIN_K = 21;
    size_t s1, e1, s2, e2;
    trim_n_ends(r1, s1, e1);
    trim_n_ends(r2, s2, e2);
    bool n1 = false, n2 = false;
    for (size_t i = s1; i < e1; ++i) n1 = n1 || ambig(r1[i]);
    for (size_t i = s2; i < e2; ++i) n2 = n2 || ambig(r2[i]);
    if (!n1 && !n2 && e1 - s1 >= MIN_K && e2 - s2 >= MIN_K) {
        b1.pack(r1.data() + s1, e1 - s1, q1 ? q1->data() + s1 : nullptr);
        b2.pack(r2.data() + s2, e2 - s2, q2 ? q2->data() + s2 : nullptr);
    } else {
        ingest_trimmed(r1.data() + s1, e1 - s1, b3, q1 ? q1->data() + s1 : nullptr);
        ingest_trimmed(r2.data() + s2, e2 - s2, b3, q2 ? q2->data() + s2 : nullptr);
    }
}

struct PackResult {
    PackedStore st;
    uint64_t pe_count = 0, rec1 = 0, rec2 = 0;
    std::string err;
};

// from_paired_fastq (packed_reads.rs:325-380) over two in-memory files; with r2 == NULL,
// append_from_file (packed_reads.rs:490-501) of one file (every read through ingest_read).
static PackResult* pack_fastq(const uint8_t* r1, size_t n1, const uint8_t* r2, size_t n2, bool keep_q = false) {
    auto* res = new PackResult();
    MemFastx a(r1, n1);
    if (!a.err.empty()) {
        res->err = a.err;
        return res;
    }
    std::vector<uint8_t> s1, s2, q1, q2;
    std::vector<uint8_t>* qa = keep_q && a.is_fastq ? &q1 : nullptr;
    if (!r2) {
        for (;;) {
            const int h = a.next_seq(s1, qa);
            if (h < 0) {
                res->err = a.err;
                return res;
            }
            if (h == 0) break;
            ++res->rec1;
            ingest_read(s1, res->st, qa);
        }
        return res;
    }
    MemFastx b(r2, n2);
    if (!b.err.empty()) {
        res->err = b.err;
        return res;
    }
    PackedStore b1, b2, b3;
    std::vector<uint8_t>* qb = keep_q && b.is_fastq ? &q2 : nullptr;
    for (;;) {
        const int h1 = a.next_seq(s1, qa);
        const int h2 = b.next_seq(s2, qb);
        if (h1 < 0 || h2 < 0) {
            res->err = h1 < 0 ? a.err : b.err;
            return res;
        }
        res->rec1 += h1;
        res->rec2 += h2;
        if (h1 && h2) partition_pair(s1, s2, b1, b2, b3, qa, qb);
        else if (h1) ingest_read(s1, b3, qa);
        else if (h2) ingest_read(s2, b3, qb);
        else break;
    }
    res->pe_count = b1.len.size();
    res->st = b1;
    res->st.append(b2);
    res->st.append(b3);
    return res;
}

// ---------------------------------------------------------------------------
// assemble.rs:692-735 determine_solid_kmer_cutoff
// ---------------------------------------------------------------------------
// Rust's `f as u32`: truncation toward zero, saturating at both ends, NaN -> 0.
static uint32_t f64_as_u32(double v) {
    if (!(v > 0.0)) return 0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

static uint32_t determine_solid_kmer_cutoff(std::vector<uint32_t> covs, double min_coverage, uint32_t* top_out) {
    if (top_out) *top_out = 0;
    if (min_coverage > 5.0) return f64_as_u32(std::max(min_coverage * 0.35, 2.0));
    if (covs.size() <= 100) return 2;
    std::sort(covs.begin(), covs.end());
    uint32_t top = covs[covs.size() * 95 / 100];
    if (top_out) *top_out = top;
    if (top < 30) return 2;
    size_t mvs = std::min<uint32_t>(std::max<uint32_t>(top / 4, 10), 35);
    std::vector<size_t> hist(mvs + 1, 0);
    for (uint32_t c : covs)
        if ((size_t)c <= mvs) hist[c] += 1;
    uint32_t valley = 2;
    size_t min_val = SIZE_MAX;
    bool found = false;
    for (size_t c = 2; c <= mvs; ++c) {
        size_t h = hist[c];
        if (h <= min_val) { min_val = h; valley = (uint32_t)c; }
        else if (h > min_val * 2 && c > (size_t)valley + 2) { found = true; break; }
    }
    return found ? valley : 2;
}

// ---------------------------------------------------------------------------
// assemble.rs:122-298 Stage 2 (Bloom build) + Stage 3 (optional Hammer, count, cutoff, solid set)
// ---------------------------------------------------------------------------

struct Spectrum {
    unsigned k = 0;
    uint64_t num_words = 0;
    std::vector<uint64_t> seen, solid;
    std::vector<Kmer256> keys;      // sorted ascending (Kmer256 Ord)
    std::vector<uint32_t> counts;
    std::vector<uint8_t> solid_mask;
    std::vector<uint8_t> corrected_data;  // packed reads after Hammer (same offsets/lengths)
    uint32_t cutoff = 0, top_cov = 0;
    uint64_t corrected_reads = 0, n_solid = 0, n_occ = 0;
    double stage2_s = 0, stage3_s = 0;
};

template <class F>
static void run_chunks(size_t num_chunks, int threads, F&& fn) {
    if (threads <= 1) {
        for (size_t c = 0; c < num_chunks; ++c) fn(c);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                size_t c = next.fetch_add(1);
                if (c >= num_chunks) break;
                fn(c);
            }
        });
    for (auto& th : pool) th.join();
}

static size_t next_pow2(size_t x) {
    size_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// `config.min_coverage as u32` (Rust's saturating float-to-int cast), then `* 4` and
// `.max(50)` as in assemble.rs:270 (u32 arithmetic wraps in a release build).
static uint32_t prior_boost_of(double min_coverage) {
    return std::max<uint32_t>(f64_as_u32(min_coverage) * 4u, 50u);
}

static Spectrum* spectrum_run(const uint8_t* data, const uint64_t* offsets, const uint32_t* lengths,
                              uint64_t n_reads, unsigned k, uint64_t bloom_bits, int error_correct,
                              double min_coverage, int threads, const uint8_t* prior = nullptr,
                              const uint64_t* prior_starts = nullptr, const uint64_t* prior_lens = nullptr,
                              uint64_t n_priors = 0) {
    auto* out = new Spectrum();
    out->k = k;
    auto t0 = std::chrono::steady_clock::now();
    TwoTierFilter filter(bloom_bits);  // assemble.rs:130
    out->num_words = filter.num_words;

    // assemble.rs:136-139 rayon chunking
    size_t num_threads = (size_t)std::max(threads, 1);
    size_t chunk_size = std::max<size_t>(n_reads / num_threads, 1000);
    size_t num_chunks = n_reads == 0 ? 0 : (n_reads + chunk_size - 1) / chunk_size;

    std::atomic<uint64_t> n_occ{0};
    // assemble.rs:141-157 Stage 2
    run_chunks(num_chunks, threads, [&](size_t c) {
        size_t start = c * chunk_size, end = std::min<size_t>(start + chunk_size, n_reads);
        std::vector<uint8_t> buf;
        buf.reserve(512);
        uint64_t local = 0;
        for (size_t idx = start; idx < end; ++idx) {
            get_read(data, offsets[idx], lengths[idx], buf);
            if (buf.size() < k) continue;
            for (size_t i = 0; i + k <= buf.size(); ++i) {
                Kmer256 km;
                if (kmer_from_bytes(&buf[i], k, k, &km)) {
                    filter.insert(kmer_canonical(km, k, nullptr));
                    ++local;
                }
            }
        }
        n_occ.fetch_add(local);
    });
    out->n_occ = n_occ.load();
    auto t1 = std::chrono::steady_clock::now();

    // assemble.rs:167-219 Stage 3
    size_t num_shards = std::min<size_t>(std::max<size_t>(next_pow2(num_threads * 4), 8), 64);
    size_t shard_mask = num_shards - 1;
    std::vector<std::mutex> locks(num_shards);
    std::vector<std::unordered_map<Kmer256, uint32_t, Kmer256Hash>> shards(num_shards);
    for (auto& m : shards) m.reserve(75000);
    std::atomic<uint64_t> corrected{0};
    std::vector<uint8_t> corrected_data;
    if (error_correct) {
        uint64_t total = 0;
        for (uint64_t r = 0; r < n_reads; ++r) total = std::max<uint64_t>(total, offsets[r] + (lengths[r] + 3) / 4);
        corrected_data.assign(data, data + total);
    }

    run_chunks(num_chunks, threads, [&](size_t c) {
        size_t start = c * chunk_size, end = std::min<size_t>(start + chunk_size, n_reads);
        std::vector<std::vector<Kmer256>> tb(num_shards);
        for (auto& v : tb) v.reserve(256);
        std::vector<uint8_t> buf, scratch;
        buf.reserve(512);
        scratch.reserve(512);
        for (size_t idx = start; idx < end; ++idx) {
            get_read(data, offsets[idx], lengths[idx], buf);
            if (buf.size() < k) continue;
            if (error_correct && correct_read_in_place(filter, k, buf.data(), buf.size(), scratch)) {
                corrected.fetch_add(1);
                // record the corrected read in packed form at its own offset (pack_clean layout)
                uint64_t off = offsets[idx];
                for (size_t j = 0; j < buf.size(); j += 4) {
                    uint8_t byte = 0;
                    for (size_t q = 0; q < 4 && j + q < buf.size(); ++q)
                        byte |= (uint8_t)(ascii_to_2bit(buf[j + q]) << (6 - 2 * q));
                    corrected_data[off + j / 4] = byte;
                }
            }
            for (size_t i = 0; i + k <= buf.size(); ++i) {
                Kmer256 km;
                if (!kmer_from_bytes(&buf[i], k, k, &km)) continue;
                Kmer256 can = kmer_canonical(km, k, nullptr);
                if (filter.is_solid(can)) {
                    size_t si = (size_t)(uint64_t)can.lo & shard_mask;
                    tb[si].push_back(can);
                    if (tb[si].size() >= 256) {
                        std::lock_guard<std::mutex> g(locks[si]);
                        for (auto& x : tb[si]) shards[si][x] += 1;
                        tb[si].clear();
                    }
                }
            }
        }
        for (size_t si = 0; si < num_shards; ++si) {
            if (tb[si].empty()) continue;
            std::lock_guard<std::mutex> g(locks[si]);
            for (auto& x : tb[si]) shards[si][x] += 1;
            tb[si].clear();
        }
    });
    out->corrected_reads = corrected.load();

    // Filter state is exported (the reference drops it here, assemble.rs:229).
    out->seen.resize(filter.num_words);
    out->solid.resize(filter.num_words);
    for (uint64_t w = 0; w < filter.num_words; ++w) {
        out->seen[w] = filter.seen[w].load();
        out->solid[w] = filter.solid[w].load();
    }

    // assemble.rs:238-242 merge; outputs are sorted here so they can be compared.
    std::unordered_map<Kmer256, uint32_t, Kmer256Hash> global;
    size_t total = 0;
    for (auto& m : shards) total += m.size();
    global.reserve(total);
    for (auto& m : shards) global.insert(m.begin(), m.end());
    // assemble.rs:266-283 prior / trusted contigs: every valid window of every prior adds
    // prior_boost to its canonical k-mer, inserting it when absent (u32 wraps).
    if (n_priors) {
        const uint32_t boost = prior_boost_of(min_coverage);
        for (uint64_t p = 0; p < n_priors; ++p) {
            const uint8_t* seq = prior + prior_starts[p];
            const uint64_t len = prior_lens[p];
            if (len < k) continue;
            for (uint64_t i = 0; i + k <= len; ++i) {
                Kmer256 km;
                if (!kmer_from_bytes(seq + i, k, k, &km)) continue;
                uint32_t& c = global[kmer_canonical(km, k, nullptr)];
                c += boost;
            }
        }
    }
    std::vector<std::pair<Kmer256, uint32_t>> all(global.begin(), global.end());
    std::sort(all.begin(), all.end(), [](const std::pair<Kmer256, uint32_t>& a, const std::pair<Kmer256, uint32_t>& b) {
        return a.first < b.first;
    });
    out->keys.resize(all.size());
    out->counts.resize(all.size());
    for (size_t i = 0; i < all.size(); ++i) { out->keys[i] = all[i].first; out->counts[i] = all[i].second; }

    // assemble.rs:286-292 cutoff + solid set
    out->cutoff = determine_solid_kmer_cutoff(out->counts, min_coverage, &out->top_cov);
    out->solid_mask.resize(all.size());
    uint64_t ns = 0;
    for (size_t i = 0; i < all.size(); ++i) {
        out->solid_mask[i] = out->counts[i] >= out->cutoff;
        ns += out->solid_mask[i];
    }
    out->n_solid = ns;
    if (error_correct) out->corrected_data.swap(corrected_data);
    auto t2 = std::chrono::steady_clock::now();
    out->stage2_s = std::chrono::duration<double>(t1 - t0).count();
    out->stage3_s = std::chrono::duration<double>(t2 - t1).count();
    return out;
}

// ---------------------------------------------------------------------------
// C ABI for the ctypes test wrapper (oracle/oracle.py)
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Unique-anchor k-mer indexes over contigs / unitigs (ASCII sequences)
// ---------------------------------------------------------------------------
// The value of one anchor: sequence index, window position, and whether the canonical
// form of that window is its reverse complement.
struct Anchor {
    uint32_t seq, pos;
    bool is_rc;
};
using AnchorMap = std::unordered_map<Kmer256, Anchor, Kmer256Hash>;
using KmerSet = std::unordered_set<Kmer256, Kmer256Hash>;

// polisher.rs:33-60 Polisher::build_kmer_index: a second sighting removes the k-mer (for
// k < 11 only when it is in another contig); removed k-mers are never re-added.
static AnchorMap polisher_index(unsigned k, const std::vector<std::vector<uint8_t>>& contigs) {
    AnchorMap pos;
    KmerSet repeats;
    for (size_t c = 0; c < contigs.size(); ++c) {
        const auto& sq = contigs[c];
        if (sq.size() < k) continue;
        for (size_t i = 0; i + k <= sq.size(); ++i) {
            Kmer256 km;
            if (!kmer_from_bytes(sq.data() + i, k, k, &km)) continue;
            int rc = 0;
            const Kmer256 can = kmer_canonical(km, k, &rc);
            if (repeats.count(can)) continue;
            auto it = pos.find(can);
            if (it != pos.end()) {
                if (k >= 11 || it->second.seq != c) {
                    pos.erase(it);
                    repeats.insert(can);
                }
            } else {
                pos[can] = Anchor{(uint32_t)c, (uint32_t)i, rc != 0};
            }
        }
    }
    return pos;
}

// paired_info.rs:34-70 PairedInfoIndex::build_kmer_index: any second sighting removes.
static AnchorMap paired_info_index(unsigned k, const std::vector<std::vector<uint8_t>>& unitigs) {
    AnchorMap map;
    if (k == 0 || k > 128) return map;
    KmerSet repeats;
    for (size_t u = 0; u < unitigs.size(); ++u) {
        const auto& sq = unitigs[u];
        if (sq.size() < k) continue;
        for (size_t i = 0; i + k <= sq.size(); ++i) {
            Kmer256 km;
            if (!kmer_from_bytes(sq.data() + i, k, k, &km)) continue;
            int rc = 0;
            const Kmer256 can = kmer_canonical(km, k, &rc);
            if (repeats.count(can)) continue;
            if (map.count(can)) {
                map.erase(can);
                repeats.insert(can);
            } else {
                map[can] = Anchor{(uint32_t)u, (uint32_t)i, rc != 0};
            }
        }
    }
    return map;
}

// splitter.rs:57-70: a sighting in another contig removes; the value is the contig.
static AnchorMap splitter_index(unsigned k, const std::vector<std::vector<uint8_t>>& contigs) {
    AnchorMap map;
    KmerSet ambiguous;
    for (size_t c = 0; c < contigs.size(); ++c) {
        const auto& sq = contigs[c];
        if (sq.size() < k) continue;
        for (size_t i = 0; i + k <= sq.size(); ++i) {
            Kmer256 km;
            if (!kmer_from_bytes(sq.data() + i, k, k, &km)) continue;
            int rc = 0;
            const Kmer256 can = kmer_canonical(km, k, &rc);
            if (ambiguous.count(can)) continue;
            auto it = map.find(can);
            if (it != map.end()) {
                if (it->second.seq != c) {
                    map.erase(it);
                    ambiguous.insert(can);
                }
            } else {
                map[can] = Anchor{(uint32_t)c, (uint32_t)i, rc != 0};
            }
        }
    }
    return map;
}


// ---------------------------------------------------------------------------
// graph.rs:65-234 CompactedGraph::build
// ---------------------------------------------------------------------------
// The reference iterates a hash set; this restatement iterates the solid k-mers in
// ascending order, which is one of the orders the reference can take.
struct OKmer {
    Kmer256 kmer;
    bool is_rc;
};
struct Unitig {
    std::string seq;
    double mean_cov;
    uint64_t kmers;
};

static std::vector<Unitig> compacted_graph(unsigned k, const std::vector<Kmer256>& solid_sorted,
                                           const std::unordered_map<Kmer256, uint32_t, Kmer256Hash>& cov) {
    std::vector<Unitig> out;
    if (k == 0 || solid_sorted.empty()) return out;
    KmerSet solid(solid_sorted.begin(), solid_sorted.end());
    auto explicit_seq = [&](const OKmer& o) { return o.is_rc ? kmer_revcomp(o.kmer, k) : o.kmer; };
    auto succs = [&](const OKmer& o) {
        std::vector<OKmer> r;
        const Kmer256 s = explicit_seq(o);
        for (uint8_t b = 0; b < 4; ++b) {
            int rc = 0;
            const Kmer256 c = kmer_canonical(kmer_extend_right(s, b, k), k, &rc);
            if (solid.count(c)) r.push_back(OKmer{c, rc != 0});
        }
        return r;
    };
    auto preds = [&](const OKmer& o) {
        std::vector<OKmer> r;
        const Kmer256 s = explicit_seq(o);
        for (uint8_t a = 0; a < 4; ++a) {
            int rc = 0;
            const Kmer256 c = kmer_canonical(kmer_prepend_left(s, a, k), k, &rc);
            if (solid.count(c)) r.push_back(OKmer{c, rc != 0});
        }
        return r;
    };
    KmerSet visited;
    auto extract = [&](OKmer start) {
        std::vector<OKmer> path;
        OKmer cur = start;
        while (!visited.count(cur.kmer)) {
            visited.insert(cur.kmer);
            path.push_back(cur);
            const auto sc = succs(cur);
            if (sc.size() != 1) break;
            const OKmer nx = sc[0];
            if (visited.count(nx.kmer)) break;
            if (preds(nx).size() != 1) break;
            cur = nx;
        }
        if (path.empty()) return path;
        std::vector<OKmer> back;
        OKmer bc = path[0];
        for (;;) {
            const auto pr = preds(bc);
            if (pr.size() != 1) break;
            const OKmer pv = pr[0];
            if (visited.count(pv.kmer)) break;
            if (succs(pv).size() != 1) break;
            visited.insert(pv.kmer);
            back.push_back(pv);
            bc = pv;
        }
        std::reverse(back.begin(), back.end());
        back.insert(back.end(), path.begin(), path.end());
        return back;
    };
    auto build = [&](const std::vector<OKmer>& path) {
        std::vector<uint8_t> buf(k);
        kmer_to_string(explicit_seq(path[0]), k, buf.data());
        Unitig u;
        u.seq.assign(buf.begin(), buf.end());
        auto covof = [&](const Kmer256& x) {
            auto it = cov.find(x);
            return it == cov.end() ? 2.0 : (double)it->second;
        };
        double total = covof(path[0].kmer);
        for (size_t i = 1; i < path.size(); ++i) {
            kmer_to_string(explicit_seq(path[i]), k, buf.data());
            u.seq.push_back((char)buf[k - 1]);  // last_base
            total += covof(path[i].kmer);
        }
        u.mean_cov = total / (double)path.size();
        u.kmers = path.size();
        return u;
    };
    for (const Kmer256& x : solid_sorted) {  // pass 1: paths from boundary nodes
        if (visited.count(x)) continue;
        OKmer start{x, false};
        const auto pf = preds(start);
        const auto prv = preds(OKmer{x, true});
        if (pf.size() == 1) {
            const OKmer p = pf[0];
            if (succs(p).size() == 1 && !visited.count(p.kmer)) {
                if (prv.size() == 1) {
                    const OKmer q = prv[0];
                    if (succs(q).size() == 1 && !visited.count(q.kmer)) continue;  // interior node
                    start.is_rc = true;
                } else {
                    start.is_rc = true;
                }
            }
        }
        const auto path = extract(start);
        if (!path.empty()) out.push_back(build(path));
    }
    for (const Kmer256& x : solid_sorted) {  // pass 2: isolated cycles and the rest
        if (visited.count(x)) continue;
        const auto path = extract(OKmer{x, false});
        if (!path.empty()) out.push_back(build(path));
    }
    return out;
}

struct GraphResult {
    std::vector<Unitig> u;
};

struct AnchorResult {
    std::vector<Kmer256> keys;  // sorted
    std::vector<Anchor> vals;
};

static AnchorResult* anchor_run(int which, unsigned k, const uint8_t* seqs, const uint64_t* starts,
                                const uint64_t* lengths, uint64_t n) {
    std::vector<std::vector<uint8_t>> v(n);
    for (uint64_t i = 0; i < n; ++i) v[i].assign(seqs + starts[i], seqs + starts[i] + lengths[i]);
    const AnchorMap m = which == 0 ? polisher_index(k, v) : which == 1 ? paired_info_index(k, v) : splitter_index(k, v);
    auto* r = new AnchorResult();
    std::vector<std::pair<Kmer256, Anchor>> e(m.begin(), m.end());
    std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto& x : e) {
        r->keys.push_back(x.first);
        r->vals.push_back(x.second);
    }
    return r;
}

extern "C" {

uint8_t or_ascii_to_2bit(uint8_t b) { return ascii_to_2bit(b); }
uint8_t or_bit2_to_base(uint8_t v) { return bit2_to_base(v); }
uint64_t or_revcomp_u64(uint64_t x, uint32_t k) { return revcomp_kmer_u64(x, k); }
uint64_t or_canonical_u64(uint64_t x, uint32_t k, int* is_rc) { return canonical_kmer_u64(x, k, is_rc); }
uint32_t or_hamming_2bit(uint64_t a, uint64_t b) { return hamming_distance_2bit(a, b); }

int or_kmer_from_bytes(const uint8_t* s, uint64_t len, uint32_t k, uint64_t* w4) {
    Kmer256 km;
    if (!kmer_from_bytes(s, len, k, &km)) return 0;
    kmer_to_words(km, w4);
    return 1;
}
void or_kmer_to_string(const uint64_t* w4, uint32_t k, uint8_t* out) { kmer_to_string(kmer_from_words(w4), k, out); }
void or_kmer_revcomp(const uint64_t* w4, uint32_t k, uint64_t* out) { kmer_to_words(kmer_revcomp(kmer_from_words(w4), k), out); }
int or_kmer_canonical(const uint64_t* w4, uint32_t k, uint64_t* out) {
    int rc = 0;
    kmer_to_words(kmer_canonical(kmer_from_words(w4), k, &rc), out);
    return rc;
}
void or_kmer_extend_right(const uint64_t* w4, uint8_t b, uint32_t k, uint64_t* out) {
    kmer_to_words(kmer_extend_right(kmer_from_words(w4), b, k), out);
}
void or_kmer_prepend_left(const uint64_t* w4, uint8_t a, uint32_t k, uint64_t* out) {
    kmer_to_words(kmer_prepend_left(kmer_from_words(w4), a, k), out);
}
uint64_t or_hash_kmer256(const uint64_t* w4, uint64_t seed) { return hash_kmer256(kmer_from_words(w4), seed); }
void or_hash_many(const uint64_t* w4, uint64_t n, uint64_t seed, uint64_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = hash_kmer256(kmer_from_words(w4 + 4 * i), seed);
}

void* or_filter_new(uint64_t capacity_bits) { return new TwoTierFilter(capacity_bits); }
void or_filter_free(void* f) { delete (TwoTierFilter*)f; }
uint64_t or_filter_num_words(void* f) { return ((TwoTierFilter*)f)->num_words; }
int or_filter_insert(void* f, const uint64_t* w4) { return ((TwoTierFilter*)f)->insert(kmer_from_words(w4)); }
int or_filter_is_solid(void* f, const uint64_t* w4) { return ((TwoTierFilter*)f)->is_solid(kmer_from_words(w4)); }
void or_filter_insert_many(void* f, const uint64_t* w4, uint64_t n, uint8_t* ret) {
    for (uint64_t i = 0; i < n; ++i) {
        int r = ((TwoTierFilter*)f)->insert(kmer_from_words(w4 + 4 * i));
        if (ret) ret[i] = (uint8_t)r;
    }
}
void or_filter_is_solid_many(void* f, const uint64_t* w4, uint64_t n, uint8_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = (uint8_t)((TwoTierFilter*)f)->is_solid(kmer_from_words(w4 + 4 * i));
}
void or_filter_export(void* f, uint64_t* seen, uint64_t* solid) {
    auto* F = (TwoTierFilter*)f;
    for (uint64_t w = 0; w < F->num_words; ++w) { seen[w] = F->seen[w].load(); solid[w] = F->solid[w].load(); }
}
int or_correct_read(void* f, uint32_t k, uint8_t* seq, uint64_t len) {
    std::vector<uint8_t> scratch;
    return correct_read_in_place(*(TwoTierFilter*)f, k, seq, len, scratch);
}

uint32_t or_cutoff(const uint32_t* counts, uint64_t n, double min_cov, uint32_t* top) {
    return determine_solid_kmer_cutoff(std::vector<uint32_t>(counts, counts + n), min_cov, top);
}

// Packs reads given as a concatenated ASCII buffer with per-read starts/lengths.
// Returns the packed byte length; data_out must hold sum(ceil(len/4)) bytes.
uint64_t or_pack(const uint8_t* ascii, const uint64_t* starts, const uint64_t* lens, uint64_t n,
                 uint8_t* data_out, uint64_t* offsets_out, uint32_t* lengths_out) {
    std::vector<uint8_t> data;
    std::vector<uint64_t> offs;
    std::vector<uint32_t> lns;
    for (uint64_t i = 0; i < n; ++i) pack_clean(ascii + starts[i], lens[i], data, offs, lns);
    std::memcpy(data_out, data.data(), data.size());
    std::memcpy(offsets_out, offs.data(), n * 8);
    std::memcpy(lengths_out, lns.data(), n * 4);
    return data.size();
}
void or_get_read(const uint8_t* data, uint64_t off, uint32_t len, uint8_t* out) {
    std::vector<uint8_t> buf;
    get_read(data, off, len, buf);
    if (!buf.empty()) std::memcpy(out, buf.data(), buf.size());
}

void* or_spectrum_run(const uint8_t* data, const uint64_t* offsets, const uint32_t* lengths, uint64_t n_reads,
                      uint32_t k, uint64_t bloom_bits, int error_correct, double min_cov, int threads) {
    return spectrum_run(data, offsets, lengths, n_reads, k, bloom_bits, error_correct, min_cov, threads);
}
// Same, with prior / trusted contigs given as ASCII sequences (starts/lens into `prior`).
void* or_spectrum_run_priors(const uint8_t* data, const uint64_t* offsets, const uint32_t* lengths, uint64_t n_reads,
                             uint32_t k, uint64_t bloom_bits, int error_correct, double min_cov, int threads,
                             const uint8_t* prior, const uint64_t* prior_starts, const uint64_t* prior_lens,
                             uint64_t n_priors) {
    return spectrum_run(data, offsets, lengths, n_reads, k, bloom_bits, error_correct, min_cov, threads, prior,
                        prior_starts, prior_lens, n_priors);
}
uint32_t or_prior_boost(double min_cov) { return prior_boost_of(min_cov); }

void* or_pack_fastq(const uint8_t* r1, uint64_t n1, const uint8_t* r2, uint64_t n2) { return pack_fastq(r1, n1, r2, n2); }
void* or_pack_fastq_q(const uint8_t* r1, uint64_t n1, const uint8_t* r2, uint64_t n2) {
    return pack_fastq(r1, n1, r2, n2, true);
}
uint64_t or_pack_n_quals(void* h) { return ((PackResult*)h)->st.quals.size(); }
void or_pack_export_quals(void* h, uint8_t* quals, uint64_t* qoff) {
    auto* r = (PackResult*)h;
    if (!r->st.quals.empty()) std::memcpy(quals, r->st.quals.data(), r->st.quals.size());
    if (!r->st.qoff.empty()) std::memcpy(qoff, r->st.qoff.data(), r->st.qoff.size() * 8);
}
void or_pack_free(void* h) { delete (PackResult*)h; }
// sizes[0..5] = n_reads, data_len, pe_count, ambiguous, r1 records, r2 records; returns 0 or -1 (malformed)
int or_pack_sizes(void* h, uint64_t* sizes) {
    auto* r = (PackResult*)h;
    sizes[0] = r->st.len.size();
    sizes[1] = r->st.data.size();
    sizes[2] = r->pe_count;
    sizes[3] = r->st.ambiguous;
    sizes[4] = r->rec1;
    sizes[5] = r->rec2;
    return r->err.empty() ? 0 : -1;
}
void or_pack_export(void* h, uint8_t* data, uint64_t* offsets, uint32_t* lengths) {
    auto* r = (PackResult*)h;
    if (!r->st.data.empty()) std::memcpy(data, r->st.data.data(), r->st.data.size());
    if (!r->st.off.empty()) std::memcpy(offsets, r->st.off.data(), r->st.off.size() * 8);
    if (!r->st.len.empty()) std::memcpy(lengths, r->st.len.data(), r->st.len.size() * 4);
}
// which: 0 = polisher.rs, 1 = paired_info.rs, 2 = splitter.rs
void* or_anchor_run(int which, uint32_t k, const uint8_t* seqs, const uint64_t* starts, const uint64_t* lengths,
                    uint64_t n) {
    return anchor_run(which, k, seqs, starts, lengths, n);
}
uint64_t or_anchor_size(void* h) { return ((AnchorResult*)h)->keys.size(); }
void or_anchor_export(void* h, uint64_t* keys4, uint32_t* seq, uint32_t* pos, uint8_t* is_rc) {
    auto* r = (AnchorResult*)h;
    for (size_t i = 0; i < r->keys.size(); ++i) {
        kmer_to_words(r->keys[i], keys4 + 4 * i);
        seq[i] = r->vals[i].seq;
        pos[i] = r->vals[i].pos;
        is_rc[i] = r->vals[i].is_rc;
    }
}
void or_anchor_free(void* h) { delete (AnchorResult*)h; }
// solid keys (any order; sorted here), their counts (cov: keys4c/counts over all count-
// table entries)
void* or_graph_build(uint32_t k, const uint64_t* solid4, uint64_t n_solid, const uint64_t* cov4,
                     const uint32_t* cov_counts, uint64_t n_cov) {
    std::vector<Kmer256> s(n_solid);
    for (uint64_t i = 0; i < n_solid; ++i) s[i] = kmer_from_words(solid4 + 4 * i);
    std::sort(s.begin(), s.end());
    std::unordered_map<Kmer256, uint32_t, Kmer256Hash> cov;
    for (uint64_t i = 0; i < n_cov; ++i) cov[kmer_from_words(cov4 + 4 * i)] = cov_counts[i];
    auto* r = new GraphResult();
    r->u = compacted_graph(k, s, cov);
    return r;
}
uint64_t or_graph_n(void* h) { return ((GraphResult*)h)->u.size(); }
uint64_t or_graph_bases(void* h) {
    uint64_t n = 0;
    for (auto& u : ((GraphResult*)h)->u) n += u.seq.size();
    return n;
}
void or_graph_export(void* h, uint8_t* seq, uint64_t* off, uint64_t* kmers, double* mean_cov) {
    auto* r = (GraphResult*)h;
    uint64_t o = 0;
    for (size_t i = 0; i < r->u.size(); ++i) {
        off[i] = o;
        std::memcpy(seq + o, r->u[i].seq.data(), r->u[i].seq.size());
        o += r->u[i].seq.size();
        kmers[i] = r->u[i].kmers;
        mean_cov[i] = r->u[i].mean_cov;
    }
    off[r->u.size()] = o;
}
void or_graph_free(void* h) { delete (GraphResult*)h; }
void or_spectrum_free(void* h) { delete (Spectrum*)h; }
uint64_t or_spectrum_num_words(void* h) { return ((Spectrum*)h)->num_words; }
uint64_t or_spectrum_num_counts(void* h) { return ((Spectrum*)h)->keys.size(); }
uint64_t or_spectrum_n_occ(void* h) { return ((Spectrum*)h)->n_occ; }
uint32_t or_spectrum_cutoff(void* h) { return ((Spectrum*)h)->cutoff; }
uint32_t or_spectrum_top_cov(void* h) { return ((Spectrum*)h)->top_cov; }
uint64_t or_spectrum_n_solid(void* h) { return ((Spectrum*)h)->n_solid; }
uint64_t or_spectrum_corrected_reads(void* h) { return ((Spectrum*)h)->corrected_reads; }
double or_spectrum_stage2_secs(void* h) { return ((Spectrum*)h)->stage2_s; }
double or_spectrum_stage3_secs(void* h) { return ((Spectrum*)h)->stage3_s; }
void or_spectrum_export_bloom(void* h, uint64_t* seen, uint64_t* solid) {
    auto* S = (Spectrum*)h;
    std::memcpy(seen, S->seen.data(), S->num_words * 8);
    std::memcpy(solid, S->solid.data(), S->num_words * 8);
}
void or_spectrum_export_counts(void* h, uint64_t* keys4, uint32_t* counts, uint8_t* solid_mask) {
    auto* S = (Spectrum*)h;
    for (size_t i = 0; i < S->keys.size(); ++i) kmer_to_words(S->keys[i], keys4 + 4 * i);
    std::memcpy(counts, S->counts.data(), S->counts.size() * 4);
    std::memcpy(solid_mask, S->solid_mask.data(), S->solid_mask.size());
}
uint64_t or_spectrum_corrected_bytes(void* h) { return ((Spectrum*)h)->corrected_data.size(); }
void or_spectrum_export_corrected(void* h, uint8_t* out) {
    auto* S = (Spectrum*)h;
    std::memcpy(out, S->corrected_data.data(), S->corrected_data.size());
}

}  // extern "C"
