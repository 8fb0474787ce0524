// kc_ingest.hip — FASTQ ingestion on the device: paired (or single) file contents ->
// the reference's three-block PackedReads layout (packed_reads.rs:325-380, 490-501), with
// FastxReader's record checks (packed_reads.rs:141-217) and the three N gates
// (packed_reads.rs:513-535, 630-690).
//
// Pipeline (one stream, buffers kept in a kc_ingest context between calls):
//   1. newline index: per 4 KiB tile a newline count (16 bytes per thread, SWAR byte
//      compare), an exclusive scan over tiles, then each tile writes its newline
//      positions in order (block scan of the per-thread counts);
//   2. per line (block-reduced): first line that is not whitespace only, last non-empty
//      line; per record (4 lines): the '@' / '+' / length checks;
//   3. per read (one wavefront, 64 bases per step, run boundaries from the ballot mask):
//      gate 1 (terminal non-ACGT trim), the internal-N flag and the number and packed size
//      of the >= 21-base ACGT segments, kept in a gate array;
//   4. per pair: clean flag (gate 2), scan -> clean rank; per read: its output unit (block
//      1 = clean pairs' R1, block 2 = their R2, block 3 = the other pairs' segments, then
//      the longer file's extra reads) and that unit's reads and bytes; exclusive scans give
//      each unit's first read and first byte;
//   5. one wavefront per input read writes its output reads' offsets, lengths and packed
//      bytes (lane q packs output bytes q, q + 64, ...).
// The device handles FASTQ (blank or whitespace-only lines before the first record; blank
// lines between records, skipped where a header is due, found by a pointer-doubling chain
// of headers) and FASTA (single- or multi-line: each record's sequence lines are copied
// together first).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/kc.h"
#include "kc_host.h"

using namespace kc;

struct kc_packed {
    std::vector<uint8_t> data;
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    std::vector<uint8_t> quals;  // when kept: Phred value of every packed base
    std::vector<uint64_t> qoff;
    kc_packed_info info{};
};

namespace {

thread_local std::string g_ingest_err;

// A device buffer released when it goes out of scope.
struct Buf : DevBuf {
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
};

constexpr uint32_t MIN_K = 21;     // packed_reads.rs:38
constexpr uint32_t TILE = BLOCK * 16;  // bytes per newline-index tile
constexpr uint64_t FILE_PAD = 64;  // zero bytes after each uploaded file
constexpr uint32_t WAVE = 64;

__device__ __forceinline__ uint32_t code2(uint8_t b) {  // dna.rs:11-30, 4 = not ACGT
    const uint32_t c = b | 0x20u;
    return c == 'a' ? 0u : c == 'c' ? 1u : c == 'g' ? 2u : c == 't' ? 3u : 4u;
}

__device__ __forceinline__ bool is_ws(uint8_t b) {
    return b == ' ' || b == '\t' || b == '\n' || b == '\r' || b == '\f' || b == 0x0B;
}

// High bit of each byte of w set where that byte is '\n' (exact, no borrow between bytes).
__device__ __forceinline__ uint32_t nl_bits(uint32_t w) {
    const uint32_t x = w ^ 0x0A0A0A0Au;
    const uint32_t t = (x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu;
    return ~(t | x | 0x7F7F7F7Fu);
}

__device__ __forceinline__ uint32_t nl_count16(const uint4 v) {
    return __popc(nl_bits(v.x)) + __popc(nl_bits(v.y)) + __popc(nl_bits(v.z)) + __popc(nl_bits(v.w));
}

// Exclusive block scan of one value per thread (BLOCK threads); returns the block total.
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl, uint32_t* sh) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t d = 1; d < BLOCK; d <<= 1) {
        const uint32_t a = t >= d ? sh[t - d] : 0u;
        __syncthreads();
        sh[t] += a;
        __syncthreads();
    }
    excl = sh[t] - v;
    const uint32_t total = sh[BLOCK - 1];
    __syncthreads();
    return total;
}

// Per tile of TILE bytes: its newline count.  The file is zero-padded to a multiple of 16.
__global__ __launch_bounds__(BLOCK) void k_nl_count(const uint8_t* __restrict__ b, uint64_t n,
                                                    uint32_t* __restrict__ tile_cnt) {
    __shared__ uint32_t sh[BLOCK];
    const uint64_t at = (uint64_t)blockIdx.x * TILE + threadIdx.x * 16ull;
    uint32_t c = 0;
    if (at < n) c = nl_count16(*reinterpret_cast<const uint4*>(b + at));
    uint32_t excl;
    const uint32_t total = block_exclusive_scan(c, excl, sh);
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = total;
}

// Per tile: writes the positions of its newlines at nl[tile_base[tile] ..] in file order.
__global__ __launch_bounds__(BLOCK) void k_nl_write(const uint8_t* __restrict__ b, uint64_t n,
                                                    const unsigned long long* __restrict__ tile_base,
                                                    unsigned long long* __restrict__ nl) {
    __shared__ uint32_t sh[BLOCK];
    const uint64_t at = (uint64_t)blockIdx.x * TILE + threadIdx.x * 16ull;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (at < n) v = *reinterpret_cast<const uint4*>(b + at);
    uint32_t excl;
    block_exclusive_scan(nl_count16(v), excl, sh);
    uint64_t o = tile_base[blockIdx.x] + excl;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (int q = 0; q < 4; ++q) {
        uint32_t m = nl_bits(w[q]);
        while (m) {
            const uint32_t byte = (uint32_t)(__ffs(m) - 1) >> 3;
            nl[o++] = at + 4u * q + byte;
            m &= m - 1;
        }
    }
}

__device__ __forceinline__ void line_span(const unsigned long long* nl, uint64_t m, uint64_t n, uint64_t j,
                                          uint64_t& s, uint64_t& e) {
    s = j ? nl[j - 1] + 1 : 0;
    e = j < m ? nl[j] : n;
}

// red[0] = first line that is not whitespace only (FastxReader::open), red[1] = last
// non-empty line (after stripping CR) + 1.  One atomic per block.
__global__ __launch_bounds__(BLOCK) void k_lines(const uint8_t* __restrict__ b, uint64_t n,
                                                 const unsigned long long* __restrict__ nl, uint64_t m,
                                                 uint64_t lines, unsigned long long* __restrict__ red) {
    __shared__ unsigned long long lo_sh[BLOCK], hi_sh[BLOCK];
    unsigned long long lo = ~0ull, hi = 0;
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < lines; j += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s, e;
        line_span(nl, m, n, j, s, e);
        uint64_t t = e;
        while (t > s && (b[t - 1] == '\r' || b[t - 1] == '\n')) --t;
        bool ws = true;
        for (uint64_t i = s; i < e && ws; ++i) ws = is_ws(b[i]);
        if (!ws && j < lo) lo = j;
        if (t > s && j + 1 > hi) hi = j + 1;
    }
    lo_sh[threadIdx.x] = lo;
    hi_sh[threadIdx.x] = hi;
    __syncthreads();
    for (uint32_t d = BLOCK / 2; d; d >>= 1) {
        if (threadIdx.x < d) {
            lo_sh[threadIdx.x] = min(lo_sh[threadIdx.x], lo_sh[threadIdx.x + d]);
            hi_sh[threadIdx.x] = max(hi_sh[threadIdx.x], hi_sh[threadIdx.x + d]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (lo_sh[0] != ~0ull) atomicMin(&red[0], lo_sh[0]);
        if (hi_sh[0]) atomicMax(&red[1], hi_sh[0]);
    }
}

// Empty lines (after stripping CR/LF) among lines [j0, j1).
__global__ void k_count_empty(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                              uint64_t m, uint64_t j0, uint64_t j1, unsigned long long* __restrict__ count) {
    for (uint64_t j = j0 + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < j1; j += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s, e;
        line_span(nl, m, n, j, s, e);
        while (e > s && (b[e - 1] == '\r' || b[e - 1] == '\n')) --e;
        if (e == s) atomicAdd(count, 1ull);
    }
}

// Record r = lines h .. h + 3, h = hdr[r] (or first + 4r without hdr).  Writes the sequence
// span (CR/LF stripped) and sets err bits: 1 = bad '@' header, 2 = bad '+' line, 4 =
// quality length, 8 = an empty header line (blank lines between records), 16 = the file
// ends inside the record; bad[0] = first bad record.
__global__ void k_records(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                          uint64_t m, uint64_t first, uint64_t R, const unsigned long long* __restrict__ hdr,
                          uint64_t lines, uint64_t* __restrict__ seq_s, uint32_t* __restrict__ seq_l,
                          uint64_t* __restrict__ qual_s, uint32_t* __restrict__ err,
                          unsigned long long* __restrict__ bad) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = hdr ? hdr[r] : first + 4 * r;
        if (h + 3 >= lines) {  // FastxReader: unexpected EOF inside the record
            atomicOr(err, 16u);
            atomicMin(bad, (unsigned long long)r);
            continue;
        }
        uint64_t s[4], e[4];
        for (int q = 0; q < 4; ++q) {
            line_span(nl, m, n, h + q, s[q], e[q]);
            while (e[q] > s[q] && (b[e[q] - 1] == '\r' || b[e[q] - 1] == '\n')) --e[q];
        }
        uint32_t f = 0;
        if (e[0] == s[0]) f |= 8u;
        else if (b[s[0]] != '@') f |= 1u;
        if (s[2] >= n || b[s[2]] != '+') f |= 2u;  // checked on the raw line
        if (e[3] - s[3] != e[1] - s[1]) f |= 4u;
        if (e[1] - s[1] > 0xFFFFFFFFull) f |= 4u;
        seq_s[r] = s[1];
        qual_s[r] = s[3];
        seq_l[r] = (uint32_t)(e[1] - s[1]);
        if (f) {
            atomicOr(err, f);
            atomicMin(bad, (unsigned long long)r);
        }
    }
}

// FASTQ with blank lines between records (FastxReader::next_seq, packed_reads.rs:154-166:
// empty lines are skipped where a header is expected, and a record then takes the next four
// lines whatever they hold).  The headers are the chain h0 = first, h(t+1) = nb(h(t) + 4),
// nb(i) = the first non-empty line at or after i; it is found by pointer doubling.
// rev[t] = line lines - 1 - t if it is not empty (after stripping CR/LF), else `lines`; a
// minimum scan of rev gives nb.
__global__ void k_nb_init(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                          uint64_t m, uint64_t lines, unsigned long long* __restrict__ rev) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < lines; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = lines - 1 - t;
        uint64_t s, e;
        line_span(nl, m, n, i, s, e);
        while (e > s && (b[e - 1] == '\r' || b[e - 1] == '\n')) --e;
        rev[t] = e > s ? i : lines;
    }
}

// jump[i] = nb(i + 4) (`lines` past the end), on[i] = (i == first).
__global__ void k_chain_init(const unsigned long long* __restrict__ nbrev, uint64_t lines, uint64_t first,
                             unsigned long long* __restrict__ jump, uint32_t* __restrict__ on) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lines; i += (uint64_t)gridDim.x * blockDim.x) {
        jump[i] = i + 4 < lines ? nbrev[lines - 1 - (i + 4)] : lines;
        on[i] = i == first;
    }
}

// One doubling round: every marked line marks its jump target (any mark is a header: it
// is reached from a marked header), then jump2 = jump o jump.  *grew = a mark was added.
__global__ void k_chain_mark(const unsigned long long* __restrict__ jump, uint64_t lines, uint32_t* __restrict__ on,
                             uint32_t* __restrict__ grew) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lines; i += (uint64_t)gridDim.x * blockDim.x) {
        if (!on[i]) continue;
        const uint64_t j = jump[i];
        if (j < lines && !on[j]) {
            on[j] = 1;
            *grew = 1;
        }
    }
}

__global__ void k_chain_jump(const unsigned long long* __restrict__ jump, uint64_t lines,
                             unsigned long long* __restrict__ jump2) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lines; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = jump[i];
        jump2[i] = j < lines ? jump[j] : lines;
    }
}

// hdr[pre[i]] = i for every marked line.
__global__ void k_chain_list(const uint32_t* __restrict__ on, const unsigned long long* __restrict__ pre,
                             uint64_t lines, unsigned long long* __restrict__ hdr) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < lines; i += (uint64_t)gridDim.x * blockDim.x)
        if (on[i]) hdr[pre[i]] = i;
}

// FASTA (FastxReader::next_seq, packed_reads.rs:217-235): after the first '>' header, a
// record's sequence is the concatenation of its lines (CR/LF stripped) up to the next line
// that starts with '>' (after stripping) or the end of the file; a last record with no
// bases is no record.  Line j = first + i: hdr[i] = it is a header, len[i] = the bases it
// contributes (0 for a header); err bit 1 = a line of 2^32 bytes or more.
__global__ void k_fa_lines(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                           uint64_t m, uint64_t first, uint64_t nlines, uint32_t* __restrict__ hdr,
                           uint32_t* __restrict__ len, uint32_t* __restrict__ err) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nlines; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s, e;
        line_span(nl, m, n, first + i, s, e);
        while (e > s && (b[e - 1] == '\r' || b[e - 1] == '\n')) --e;
        const bool h = e > s && b[s] == '>';
        const uint64_t l = h ? 0 : e - s;
        if (l > 0xFFFFFFFFull) atomicOr(err, 1u);
        hdr[i] = h;
        len[i] = (uint32_t)l;
    }
}

// Sixteen lanes per sequence line (four lines a wave: FASTA lines are mostly 60-80 bytes):
// its bases to out[off[i] ..].  Record r (the r-th header, rec[i] = headers before line i)
// starts at its header's offset.  Lines longer than FA_LONG_LINE (unwrapped records:
// assemblies, long reads) are left to k_fa_copy_long, a workgroup a line.
constexpr uint32_t FA_LONG_LINE = 4096;
__global__ __launch_bounds__(BLOCK) void k_fa_copy(const uint8_t* __restrict__ b, uint64_t n,
                                                   const unsigned long long* __restrict__ nl, uint64_t m,
                                                   uint64_t first, uint64_t nlines, const uint32_t* __restrict__ hdr,
                                                   const uint32_t* __restrict__ len,
                                                   const unsigned long long* __restrict__ rec,
                                                   const unsigned long long* __restrict__ off,
                                                   uint8_t* __restrict__ out, uint64_t* __restrict__ seq_s) {
    constexpr uint32_t LANES = 16;
    const uint32_t sub = threadIdx.x & (LANES - 1);
    const uint64_t groups = (uint64_t)gridDim.x * (BLOCK / LANES);
    for (uint64_t i = blockIdx.x * (uint64_t)(BLOCK / LANES) + threadIdx.x / LANES; i < nlines; i += groups) {
        if (hdr[i]) {
            if (sub == 0) seq_s[rec[i]] = off[i];
            continue;
        }
        if (len[i] > FA_LONG_LINE) continue;
        uint64_t s, e;
        line_span(nl, m, n, first + i, s, e);
        const uint64_t o = off[i];
        for (uint64_t q = sub; q < len[i]; q += LANES) out[o + q] = b[s + q];
    }
}

// The sequence lines longer than FA_LONG_LINE, one workgroup each (a workgroup passes over
// the short lines of its stride at one scalar test a line).
__global__ __launch_bounds__(BLOCK) void k_fa_copy_long(const uint8_t* __restrict__ b, uint64_t n,
                                                        const unsigned long long* __restrict__ nl, uint64_t m,
                                                        uint64_t first, uint64_t nlines,
                                                        const uint32_t* __restrict__ hdr,
                                                        const uint32_t* __restrict__ len,
                                                        const unsigned long long* __restrict__ off,
                                                        uint8_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x; i < nlines; i += gridDim.x) {
        if (hdr[i] || len[i] <= FA_LONG_LINE) continue;
        uint64_t s, e;
        line_span(nl, m, n, first + i, s, e);
        const uint64_t o = off[i];
        for (uint64_t q = threadIdx.x; q < len[i]; q += BLOCK) out[o + q] = b[s + q];
    }
}

// seq_l[r] = the next record's start (or the total) - seq_s[r]; err bit 1 = a record of
// 2^32 bases or more.
__global__ void k_fa_lens(const uint64_t* __restrict__ seq_s, uint64_t R, uint64_t total,
                          uint32_t* __restrict__ seq_l, uint32_t* __restrict__ err) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t l = (r + 1 < R ? seq_s[r + 1] : total) - seq_s[r];
        if (l > 0xFFFFFFFFull) atomicOr(err, 1u);
        seq_l[r] = (uint32_t)l;
    }
}

// Gate 1 and the N scan of one read: trimmed span [ts, ts + tl), internal-N flag, number of
// ACGT segments of at least MIN_K bases and their packed bytes.
struct ReadGate {
    uint32_t ts, tl, nseg, internal_n;
    uint64_t seg_bytes, seg_bases;
};


// This wave's index in the grid, as a value the compiler knows to be wave-uniform (so the
// per-read loads and the run walk below are scalar).
__device__ __forceinline__ uint64_t wave_id() {
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    return (uint64_t)blockIdx.x * (BLOCK / WAVE) + w;
}

// Calls on_run(start, len) for every maximal run of ACGT bases of the read [s, s + l), in
// order.  One wavefront per read: each 64-base chunk is one base per lane and the wave's
// ballot of "not ACGT"; the run boundaries come from that mask by count-trailing-zeros, so
// every value here is wave-uniform.  Bases past l count as not ACGT, which closes a run at
// the read's end.
template <typename F>
__device__ void wave_runs(const uint8_t* __restrict__ b, uint64_t s, uint32_t l, uint32_t lane, F&& on_run) {
    bool in_run = false;
    uint32_t run_start = 0;
    for (uint64_t c1 = 0; c1 < l; c1 += 4 * WAVE) {
        // four independent loads per lane before the first ballot (one memory round trip
        // for reads of up to 256 bases)
        uint8_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t p = c1 + j * WAVE + lane;
            v[j] = p < l ? b[s + p] : (uint8_t)'N';
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t c0 = c1 + j * WAVE;
            if (c0 >= l) break;
            const uint64_t M = __ballot(code2(v[j]) == 4u);  // past l: 'N'
            uint32_t pos = 0;
            while (pos < WAVE) {
                if (in_run) {
                    const uint64_t m = M >> pos;
                    if (!m) break;
                    const uint32_t e = pos + (uint32_t)__builtin_ctzll(m);
                    on_run(run_start, (uint32_t)(c0 + e) - run_start);
                    in_run = false;
                    pos = e;
                } else {
                    const uint64_t z = ~M >> pos;
                    if (!z) break;
                    const uint32_t st = pos + (uint32_t)__builtin_ctzll(z);
                    in_run = true;
                    run_start = (uint32_t)c0 + st;
                    pos = st;
                }
            }
        }
    }
    if (in_run) on_run(run_start, l - run_start);
}

// The trimmed span is the first run's start to the last run's end; an internal non-ACGT
// base exists exactly when there are two or more runs (packed_reads.rs:513-535, 630-690).
// Fast path: per 64-base chunk the ACGT mask gives the first and last ACGT base and the
// number of run starts by bit arithmetic; only a read with two or more runs walks them.
__device__ ReadGate wave_gate(const uint8_t* __restrict__ b, uint64_t s, uint32_t l, uint32_t lane) {
    ReadGate g{};
    uint32_t runs = 0, first = l, last_end = 0;
    uint64_t carry = 0;  // the previous chunk's last base is ACGT
    for (uint64_t c1 = 0; c1 < l; c1 += 4 * WAVE) {
        uint8_t v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t p = c1 + j * WAVE + lane;
            v[j] = p < l ? b[s + p] : (uint8_t)'N';
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t c0 = c1 + j * WAVE;
            if (c0 >= l) break;
            const uint64_t Z = __ballot(code2(v[j]) != 4u);  // ACGT bases of the chunk
            const uint64_t starts = Z & ~((Z << 1) | carry);
            runs += (uint32_t)__popcll(starts);
            if (Z) {
                if (first == l) first = (uint32_t)c0 + (uint32_t)__builtin_ctzll(Z);
                last_end = (uint32_t)c0 + 64u - (uint32_t)__builtin_clzll(Z);
            }
            carry = Z >> 63;
        }
    }
    if (runs >= 2) {
        wave_runs(b, s, l, lane, [&](uint32_t, uint32_t len) {
            if (len >= MIN_K) {
                g.nseg += 1;
                g.seg_bytes += (len + 3) / 4;
                g.seg_bases += len;
            }
        });
    } else if (runs == 1 && last_end - first >= MIN_K) {
        g.nseg = 1;
        g.seg_bytes = (last_end - first + 3) / 4;
        g.seg_bases = last_end - first;
    }
    g.ts = runs ? first : l;
    g.tl = runs ? last_end - first : 0;
    g.internal_n = runs >= 2;
    return g;
}

// Packs bases [s, s + len) (all ACGT) into out, 4 per byte, high bits first; lane q writes
// output bytes q, q + 64, ...
__device__ void wave_pack(const uint8_t* __restrict__ b, uint64_t s, uint32_t len, uint8_t* __restrict__ out,
                          uint32_t lane) {
    const uint32_t nb = (len + 3) / 4;
    for (uint32_t q = lane; q < nb; q += WAVE) {
        uint32_t v = 0;
        for (uint32_t r = 0; r < 4; ++r)
            if (4 * q + r < len) v |= code2(b[s + 4 * q + r]) << (6 - 2 * r);
        out[q] = (uint8_t)v;
    }
}

// Input read x: reads 0..R1-1 of the first file, then R1.. of the second.  One wave per read.
__global__ __launch_bounds__(BLOCK) void k_gates(const uint8_t* __restrict__ b1, const uint64_t* __restrict__ s1,
                                                 const uint32_t* __restrict__ l1, uint64_t R1,
                                                 const uint8_t* __restrict__ b2, const uint64_t* __restrict__ s2,
                                                 const uint32_t* __restrict__ l2, uint64_t R2,
                                                 ReadGate* __restrict__ gates) {
    const uint64_t total = R1 + R2;
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t x = wave_id(); x < total; x += nw) {
        const ReadGate g = x < R1 ? wave_gate(b1, s1[x], l1[x], lane) : wave_gate(b2, s2[x - R1], l2[x - R1], lane);
        if (lane == 0) gates[x] = g;
    }
}

// Gate 2 of pair i (packed_reads.rs:340-352): both mates free of internal non-ACGT bases
// and at least MIN_K long after trimming.
__global__ void k_clean(const ReadGate* __restrict__ gates, uint64_t R1, uint64_t P, uint32_t* __restrict__ clean) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < P; i += (uint64_t)gridDim.x * blockDim.x) {
        const ReadGate g = gates[i], h = gates[R1 + i];
        clean[i] = !g.internal_n && !h.internal_n && g.tl >= MIN_K && h.tl >= MIN_K;
    }
}

// Output unit of input read x.  Reads 0..P-1 of both files are pairs (when paired), the
// rest extras.  Clean pair c: units c (R1) and n_clean + c (R2); other pair d:
// 2 n_clean + 2d (+1 for R2); extra e: 2P + e (only one file has extras).
__device__ __forceinline__ uint64_t unit_of(uint64_t x, uint64_t R1, uint64_t P, const uint32_t* clean,
                                            const unsigned long long* clean_pre, uint64_t n_clean, bool& whole) {
    const bool second = x >= R1;
    const uint64_t i = second ? x - R1 : x;
    whole = false;
    if (i < P) {
        const uint64_t ci = clean_pre[i];
        if (clean[i]) {
            whole = true;
            return second ? n_clean + ci : ci;
        }
        return 2 * n_clean + 2 * (i - ci) + (second ? 1 : 0);
    }
    return 2 * P + (second ? (i - P) + (R1 - P) : i - P);
}

__global__ void k_units(const ReadGate* __restrict__ gates, uint64_t R1, uint64_t R2, uint64_t P,
                        const uint32_t* __restrict__ clean, const unsigned long long* __restrict__ clean_pre,
                        uint64_t n_clean, unsigned long long* __restrict__ unit_reads,
                        unsigned long long* __restrict__ unit_bytes, unsigned long long* __restrict__ unit_bases,
                        unsigned long long* __restrict__ ambiguous) {
    const uint64_t total = R1 + R2;
    uint32_t amb = 0;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
        const ReadGate g = gates[x];
        bool whole;
        const uint64_t u = unit_of(x, R1, P, clean, clean_pre, n_clean, whole);
        if (whole) {
            unit_reads[u] = 1;
            unit_bytes[u] = (g.tl + 3) / 4;
            if (unit_bases) unit_bases[u] = g.tl;
        } else {
            unit_reads[u] = g.nseg;
            unit_bytes[u] = g.seg_bytes;
            if (unit_bases) unit_bases[u] = g.seg_bases;
            amb += g.tl && !g.nseg;  // gate 3: nothing of the read reaches MIN_K
        }
    }
    if (amb) atomicAdd(ambiguous, (unsigned long long)amb);
}

// Where input read x's output goes: first output read, first packed byte, first quality.
struct Place {
    uint64_t rb, bb, qb;
};

__global__ void k_place(uint64_t total, uint64_t R1, uint64_t P, const uint32_t* __restrict__ clean,
                        const unsigned long long* __restrict__ clean_pre, uint64_t n_clean,
                        const unsigned long long* __restrict__ read_base,
                        const unsigned long long* __restrict__ byte_base,
                        const unsigned long long* __restrict__ base_base, Place* __restrict__ place) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
        bool whole;
        const uint64_t u = unit_of(x, R1, P, clean, clean_pre, n_clean, whole);
        place[x] = Place{read_base[u], byte_base[u], base_base ? base_base[u] : 0ull};
    }
}

// Quality characters [s, s + len) -> Phred values (character - 33, 0 below '!').
__device__ void wave_quals(const uint8_t* __restrict__ b, uint64_t s, uint32_t len, uint8_t* __restrict__ out,
                           uint32_t lane) {
    for (uint32_t i = lane; i < len; i += WAVE) {
        const uint8_t c = b[s + i];
        out[i] = c >= 33 ? (uint8_t)(c - 33) : (uint8_t)0;
    }
}

// Writes every input read's output reads (offsets, lengths, packed bytes).  One wave per
// input read: a whole read is one packed span; otherwise each run of at least MIN_K bases
// becomes an output read, in order.
__global__ __launch_bounds__(BLOCK) void k_emit(const uint8_t* __restrict__ b1, const uint64_t* __restrict__ s1,
                                                const uint32_t* __restrict__ l1, uint64_t R1,
                                                const uint8_t* __restrict__ b2, const uint64_t* __restrict__ s2,
                                                const uint32_t* __restrict__ l2, uint64_t R2, uint64_t P,
                                                const ReadGate* __restrict__ gates, const uint32_t* __restrict__ clean,
                                                const Place* __restrict__ place, uint8_t* __restrict__ data,
                                                uint64_t* __restrict__ offsets, uint32_t* __restrict__ lengths,
                                                const uint64_t* __restrict__ q1, const uint64_t* __restrict__ q2,
                                                uint8_t* __restrict__ quals, uint64_t* __restrict__ qoff) {
    const uint64_t total = R1 + R2;
    const uint32_t lane = threadIdx.x % WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * (BLOCK / WAVE);
    for (uint64_t x = wave_id(); x < total; x += nw) {
        const bool second = x >= R1;
        const uint8_t* b = second ? b2 : b1;
        const uint64_t s = second ? s2[x - R1] : s1[x];
        const uint32_t l = second ? l2[x - R1] : l1[x];
        const ReadGate g = gates[x];
        const Place pl = place[x];
        uint64_t rb = pl.rb, bb = pl.bb, qb = pl.qb;
        const uint64_t i = second ? x - R1 : x;
        const uint64_t qs = quals ? (second ? q2[x - R1] : q1[x]) : 0;  // the record's quality line
        if (i < P && clean[i]) {  // a clean pair's read: one output read of its trimmed span
            if (lane == 0) {
                offsets[rb] = bb;
                lengths[rb] = g.tl;
                if (quals) qoff[rb] = qb;
            }
            wave_pack(b, s + g.ts, g.tl, data + bb, lane);
            if (quals) wave_quals(b, qs + g.ts, g.tl, quals + qb, lane);
            continue;
        }
        if (!g.nseg) continue;
        wave_runs(b, s, l, lane, [&](uint32_t st, uint32_t len) {
            if (len < MIN_K) return;
            if (lane == 0) {
                offsets[rb] = bb;
                lengths[rb] = len;
                if (quals) qoff[rb] = qb;
            }
            wave_pack(b, s + st, len, data + bb, lane);
            if (quals) wave_quals(b, qs + st, len, quals + qb, lane);
            ++rb;
            bb += (len + 3) / 4;
            qb += len;
        });
    }
}

struct FileIndex {
    Buf buf, tiles, tile_base, nl, red, seq_s, seq_l, qual_s, err;
    Buf fa_hdr, fa_len, fa_rec, fa_off, fa_seq;  // FASTA: line flags and lengths, their scans, the sequences
    Buf ch_a, ch_b, ch_c, ch_h, ch_on;           // FASTQ header chain: scans, jumps, header lines, marks
    uint64_t n = 0, R = 0;
    bool fasta = false;
    const uint8_t* seqs() const { return fasta ? fa_seq.as<uint8_t>() : buf.as<uint8_t>(); }
};

}  // namespace

struct kc_ingest {
    int32_t device = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    FileIndex A, B;
    Buf gates, clean, clean_pre, ureads, ubytes, rbase, bbase, place, small, scan_tmp, data, off, len;
    Buf ubases, qbase, quals, qoff;  // quality side array (keep_quals)
    bool keep_quals = false;
};

namespace {

#define I_TRY(expr)                                                                     \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            g_ingest_err = std::string(#expr) + ": " + hipGetErrorString(_e);           \
            return _e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP;                     \
        }                                                                               \
    } while (0)

int fail(int code, const std::string& m) {
    g_ingest_err = m;
    return code;
}

// Exclusive scan of n values into out (64-bit sums), through the context's temp buffer.
template <typename In>
int scan_excl(kc_ingest* c, const In* in, unsigned long long* out, uint64_t n) {
    if (!n) return KC_OK;
    size_t tb = 0;
    I_TRY(rocprim::exclusive_scan(nullptr, tb, in, out, 0ull, (size_t)n, rocprim::plus<unsigned long long>(), c->st));
    I_TRY(c->scan_tmp.ensure(tb));
    I_TRY(rocprim::exclusive_scan(c->scan_tmp.p, tb, in, out, 0ull, (size_t)n, rocprim::plus<unsigned long long>(),
                                  c->st));
    return KC_OK;
}

int upload(kc_ingest* c, const uint8_t* host, uint64_t n, FileIndex& F) {
    F.n = n;
    I_TRY(F.buf.ensure(n + FILE_PAD));
    if (n) I_TRY(hipMemcpyAsync(F.buf.p, host, n, hipMemcpyHostToDevice, c->st));
    I_TRY(hipMemsetAsync(F.buf.as<uint8_t>() + n, 0, FILE_PAD, c->st));
    return KC_OK;
}

// FASTA records of one uploaded file whose first non-blank line (`first`) is a '>' header:
// the sequence lines copied together into F.fa_seq, the records' spans into it.
int index_fasta(kc_ingest* c, FileIndex& F, const char* name, uint64_t m, uint64_t first, uint64_t lines) {
    hipStream_t st = c->st;
    const uint64_t n = F.n, nlines = lines - first;
    const uint8_t* b = F.buf.as<uint8_t>();
    if (c->keep_quals) return fail(KC_EUNSUPPORTED, std::string(name) + ": FASTA has no qualities to keep");
    I_TRY(F.fa_hdr.ensure(nlines * 4 + 4));
    I_TRY(F.fa_len.ensure(nlines * 4 + 4));
    I_TRY(F.fa_rec.ensure(nlines * 8 + 8));
    I_TRY(F.fa_off.ensure(nlines * 8 + 8));
    I_TRY(F.err.ensure(64));
    const uint32_t zero = 0;
    I_TRY(hipMemcpyAsync(F.err.p, &zero, 4, hipMemcpyHostToDevice, st));
    k_fa_lines<<<grid_for(nlines), BLOCK, 0, st>>>(b, n, F.nl.as<unsigned long long>(), m, first, nlines,
                                                   F.fa_hdr.as<uint32_t>(), F.fa_len.as<uint32_t>(),
                                                   F.err.as<uint32_t>());
    I_TRY(hipGetLastError());
    int rc = scan_excl(c, F.fa_hdr.as<uint32_t>(), F.fa_rec.as<unsigned long long>(), nlines);
    if (!rc) rc = scan_excl(c, F.fa_len.as<uint32_t>(), F.fa_off.as<unsigned long long>(), nlines);
    if (rc) return rc;
    unsigned long long last[2];
    uint32_t lastv[2], err = 0;
    I_TRY(hipMemcpyAsync(&last[0], F.fa_rec.as<unsigned long long>() + nlines - 1, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&last[1], F.fa_off.as<unsigned long long>() + nlines - 1, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&lastv[0], F.fa_hdr.as<uint32_t>() + nlines - 1, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&lastv[1], F.fa_len.as<uint32_t>() + nlines - 1, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&err, F.err.p, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    if (err) return fail(KC_EINVAL, std::string(name) + ": a FASTA line of 2^32 bytes or more");
    uint64_t R = last[0] + lastv[0];  // headers; line `first` is one
    const uint64_t total = last[1] + lastv[1];
    I_TRY(F.fa_seq.ensure(total + FILE_PAD));
    I_TRY(hipMemsetAsync(F.fa_seq.as<uint8_t>() + total, 0, FILE_PAD, st));
    I_TRY(F.seq_s.ensure(R * 8 + 8));
    I_TRY(F.seq_l.ensure(R * 4 + 4));
    I_TRY(F.qual_s.ensure(R * 8 + 8));
    k_fa_copy<<<grid_for(nlines * 16), BLOCK, 0, st>>>(b, n, F.nl.as<unsigned long long>(), m, first, nlines,
                                                         F.fa_hdr.as<uint32_t>(), F.fa_len.as<uint32_t>(),
                                                         F.fa_rec.as<unsigned long long>(),
                                                         F.fa_off.as<unsigned long long>(), F.fa_seq.as<uint8_t>(),
                                                         F.seq_s.as<uint64_t>());
    I_TRY(hipGetLastError());
    k_fa_copy_long<<<(uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(nlines, 4096)), BLOCK, 0, st>>>(
        b, n, F.nl.as<unsigned long long>(), m, first, nlines, F.fa_hdr.as<uint32_t>(), F.fa_len.as<uint32_t>(),
        F.fa_off.as<unsigned long long>(), F.fa_seq.as<uint8_t>());
    I_TRY(hipGetLastError());
    k_fa_lens<<<grid_for(R), BLOCK, 0, st>>>(F.seq_s.as<uint64_t>(), R, total, F.seq_l.as<uint32_t>(),
                                             F.err.as<uint32_t>());
    I_TRY(hipGetLastError());
    uint32_t last_len = 0;
    I_TRY(hipMemcpyAsync(&last_len, F.seq_l.as<uint32_t>() + R - 1, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&err, F.err.p, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    if (err) return fail(KC_EINVAL, std::string(name) + ": a FASTA record of 2^32 bases or more");
    if (last_len == 0) --R;  // next_seq returns the last record only when it has bases
    F.fasta = true;
    F.R = R;
    return KC_OK;
}

// The header lines of a FASTQ file with blank lines between records (k_nb_init ..
// k_chain_list): F.ch_h[0 .. R).
int fastq_chain(kc_ingest* c, FileIndex& F, uint64_t m, uint64_t first, uint64_t lines, uint64_t& R) {
    hipStream_t st = c->st;
    const uint64_t L = lines;
    I_TRY(F.ch_a.ensure(L * 8 + 8));
    I_TRY(F.ch_b.ensure(L * 8 + 8));
    I_TRY(F.ch_c.ensure(L * 8 + 8));
    I_TRY(F.ch_h.ensure(L * 8 + 8));
    I_TRY(F.ch_on.ensure(L * 4 + 4));
    I_TRY(F.err.ensure(64));
    using ull = unsigned long long;
    k_nb_init<<<grid_for(L), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), F.n, F.nl.as<ull>(), m, L, F.ch_a.as<ull>());
    I_TRY(hipGetLastError());
    size_t tb = 0;
    I_TRY(rocprim::inclusive_scan(nullptr, tb, F.ch_a.as<ull>(), F.ch_b.as<ull>(), (size_t)L, rocprim::minimum<ull>(),
                                  st));
    I_TRY(c->scan_tmp.ensure(tb));
    I_TRY(rocprim::inclusive_scan(c->scan_tmp.p, tb, F.ch_a.as<ull>(), F.ch_b.as<ull>(), (size_t)L,
                                  rocprim::minimum<ull>(), st));
    k_chain_init<<<grid_for(L), BLOCK, 0, st>>>(F.ch_b.as<ull>(), L, first, F.ch_a.as<ull>(), F.ch_on.as<uint32_t>());
    I_TRY(hipGetLastError());
    ull* jump = F.ch_a.as<ull>();
    ull* jump2 = F.ch_c.as<ull>();
    for (;;) {  // each round doubles the chain's marked prefix; a round that adds none ends it
        const uint32_t zero = 0;
        uint32_t grew = 0;
        I_TRY(hipMemcpyAsync(F.err.p, &zero, 4, hipMemcpyHostToDevice, st));
        k_chain_mark<<<grid_for(L), BLOCK, 0, st>>>(jump, L, F.ch_on.as<uint32_t>(), F.err.as<uint32_t>());
        I_TRY(hipGetLastError());
        k_chain_jump<<<grid_for(L), BLOCK, 0, st>>>(jump, L, jump2);
        I_TRY(hipGetLastError());
        std::swap(jump, jump2);
        I_TRY(hipMemcpyAsync(&grew, F.err.p, 4, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        if (!grew) break;
    }
    int rc = scan_excl(c, F.ch_on.as<uint32_t>(), F.ch_b.as<ull>(), L);
    if (rc) return rc;
    ull last = 0;
    uint32_t lon = 0;
    I_TRY(hipMemcpyAsync(&last, F.ch_b.as<ull>() + L - 1, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipMemcpyAsync(&lon, F.ch_on.as<uint32_t>() + L - 1, 4, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    R = last + lon;
    k_chain_list<<<grid_for(L), BLOCK, 0, st>>>(F.ch_on.as<uint32_t>(), F.ch_b.as<ull>(), L, F.ch_h.as<ull>());
    I_TRY(hipGetLastError());
    return KC_OK;
}

// Finds one uploaded file's records' sequence spans.
int index_file(kc_ingest* c, FileIndex& F, const char* name) {
    hipStream_t st = c->st;
    const uint64_t n = F.n;
    const uint8_t* b = F.buf.as<uint8_t>();
    I_TRY(F.red.ensure(64));
    F.fasta = false;
    uint64_t m = 0;
    const uint64_t ntiles = (n + TILE - 1) / TILE;
    if (n) {
        I_TRY(F.tiles.ensure(ntiles * 4 + 4));
        I_TRY(F.tile_base.ensure(ntiles * 8 + 8));
        k_nl_count<<<(uint32_t)ntiles, BLOCK, 0, st>>>(b, n, F.tiles.as<uint32_t>());
        I_TRY(hipGetLastError());
        int rc = scan_excl(c, F.tiles.as<uint32_t>(), F.tile_base.as<unsigned long long>(), ntiles);
        if (rc) return rc;
        unsigned long long last_base = 0;
        uint32_t last_cnt = 0;
        I_TRY(hipMemcpyAsync(&last_base, F.tile_base.as<unsigned long long>() + ntiles - 1, 8, hipMemcpyDeviceToHost,
                             st));
        I_TRY(hipMemcpyAsync(&last_cnt, F.tiles.as<uint32_t>() + ntiles - 1, 4, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        m = last_base + last_cnt;
    }
    I_TRY(F.nl.ensure(m * 8 + 8));
    if (m) {
        k_nl_write<<<(uint32_t)ntiles, BLOCK, 0, st>>>(b, n, F.tile_base.as<unsigned long long>(),
                                                      F.nl.as<unsigned long long>());
        I_TRY(hipGetLastError());
    }
    // read_until semantics: a final line without a newline still counts
    uint8_t lastb = '\n';
    if (n) I_TRY(hipMemcpyAsync(&lastb, b + n - 1, 1, hipMemcpyDeviceToHost, st));
    unsigned long long red[2] = {~0ull, 0ull};
    I_TRY(hipMemcpyAsync(F.red.p, red, 16, hipMemcpyHostToDevice, st));
    I_TRY(hipStreamSynchronize(st));
    const uint64_t lines = m + (n && lastb != '\n' ? 1 : 0);
    if (lines) {
        k_lines<<<grid_for(lines, BLOCK, 4096), BLOCK, 0, st>>>(b, n, F.nl.as<unsigned long long>(), m, lines,
                                                               F.red.as<unsigned long long>());
        I_TRY(hipGetLastError());
    }
    I_TRY(hipMemcpyAsync(red, F.red.p, 16, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    if (red[0] == ~0ull) {  // empty file (or whitespace only): no records
        F.R = 0;
        return KC_OK;
    }
    const uint64_t first = red[0], end = red[1];
    uint8_t head = 0;
    {
        unsigned long long pos = 0;
        if (first) {
            I_TRY(hipMemcpyAsync(&pos, F.nl.as<unsigned long long>() + first - 1, 8, hipMemcpyDeviceToHost, st));
            I_TRY(hipStreamSynchronize(st));
        }
        const uint64_t at = first ? pos + 1 : 0;
        I_TRY(hipMemcpyAsync(&head, b + at, 1, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
    }
    if (head == '>') return index_fasta(c, F, name, m, first, lines);
    if (head != '@') return fail(KC_EINVAL, std::string(name) + ": expected '@' or '>' header");
    // Empty lines between the first and the last record are either legal empty sequence /
    // quality lines or blank separators, which FastxReader skips before a header.  With
    // none, records are 4-line aligned from the first header; with some, the headers are
    // found as FastxReader finds them (fastq_chain).
    unsigned long long empties = 0;
    I_TRY(hipMemcpyAsync(F.red.p, &empties, 8, hipMemcpyHostToDevice, st));
    k_count_empty<<<grid_for(end - first), BLOCK, 0, st>>>(b, n, F.nl.as<unsigned long long>(), m, first, end,
                                                           F.red.as<unsigned long long>());
    I_TRY(hipGetLastError());
    I_TRY(hipMemcpyAsync(&empties, F.red.p, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    const unsigned long long* hdr = nullptr;
    uint64_t R = 0;
    if (empties) {
        int rc = fastq_chain(c, F, m, first, lines, R);
        if (rc) return rc;
        hdr = F.ch_h.as<unsigned long long>();
    } else {
        // The lines after `end` are all empty, so a last record whose sequence and quality
        // are empty ends up to three lines later; the 4-line parse takes them when they exist.
        const uint64_t end4 = first + (end - first + 3) / 4 * 4;
        if (end4 > lines)
            return fail(KC_EINVAL, std::string(name) + ": malformed FASTQ: unexpected EOF inside a record");
        R = (end4 - first) / 4;
    }
    I_TRY(F.seq_s.ensure(R * 8 + 8));
    I_TRY(F.seq_l.ensure(R * 4 + 4));
    I_TRY(F.qual_s.ensure(R * 8 + 8));
    I_TRY(F.err.ensure(64));
    unsigned long long init[2] = {0ull, ~0ull};
    I_TRY(hipMemcpyAsync(F.err.p, init, 16, hipMemcpyHostToDevice, st));
    if (R)
        k_records<<<grid_for(R), BLOCK, 0, st>>>(b, n, F.nl.as<unsigned long long>(), m, first, R, hdr, lines,
                                                 F.seq_s.as<uint64_t>(), F.seq_l.as<uint32_t>(),
                                                 F.qual_s.as<uint64_t>(), F.err.as<uint32_t>(),
                                                 F.err.as<unsigned long long>() + 1);
    I_TRY(hipGetLastError());
    unsigned long long out[2];
    I_TRY(hipMemcpyAsync(out, F.err.p, 16, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    const uint32_t e = (uint32_t)out[0];
    if (e) {  // (an empty header line, bit 8, cannot occur: headers are non-empty lines)
        const char* what = (e & 9u)    ? "expected '@' header"
                           : (e & 16u) ? "unexpected EOF inside the record"
                           : (e & 2u)  ? "expected '+' separator line"
                                       : "sequence length != quality length";
        return fail(KC_EINVAL, std::string(name) + ": malformed FASTQ record " + std::to_string(out[1]) + ": " + what);
    }
    F.R = R;
    return KC_OK;
}

int ingest(kc_ingest* c, const uint8_t* r1, uint64_t n1, const uint8_t* r2, uint64_t n2, kc_packed* out) {
    I_TRY(hipSetDevice(c->device));
    hipStream_t st = c->st;
    const bool paired = r2 != nullptr;
    int rc = upload(c, r1, n1, c->A);
    if (!rc && paired) rc = upload(c, r2, n2, c->B);
    if (rc) return rc;
    I_TRY(hipEventRecord(c->ev[0], st));
    rc = index_file(c, c->A, "R1");
    if (!rc && paired) rc = index_file(c, c->B, "R2");
    if (rc) return rc;
    FileIndex& A = c->A;
    FileIndex& B = paired ? c->B : c->A;
    const uint64_t R1 = A.R, R2 = paired ? c->B.R : 0;
    const uint64_t P = paired ? std::min(R1, R2) : 0;
    const uint64_t tot = R1 + R2;
    const uint64_t U = tot;  // one output unit per input read
    I_TRY(c->gates.ensure(tot * sizeof(ReadGate) + sizeof(ReadGate)));
    I_TRY(c->clean.ensure(P * 4 + 4));
    I_TRY(c->clean_pre.ensure(P * 8 + 8));
    I_TRY(c->ureads.ensure(U * 8 + 8));
    I_TRY(c->ubytes.ensure(U * 8 + 8));
    I_TRY(c->rbase.ensure(U * 8 + 8));
    I_TRY(c->bbase.ensure(U * 8 + 8));
    I_TRY(c->place.ensure((tot + 1) * sizeof(Place)));
    const bool kq = c->keep_quals;
    if (kq) {
        I_TRY(c->ubases.ensure(U * 8 + 8));
        I_TRY(c->qbase.ensure(U * 8 + 8));
    }
    I_TRY(c->small.ensure(64));
    const unsigned long long zero = 0;
    I_TRY(hipMemcpyAsync(c->small.p, &zero, 8, hipMemcpyHostToDevice, st));
    const uint8_t* b1 = A.seqs();
    const uint8_t* b2 = B.seqs();
    ReadGate* gates = c->gates.as<ReadGate>();
    if (tot) {
        k_gates<<<grid_for(tot * WAVE), BLOCK, 0, st>>>(b1, A.seq_s.as<uint64_t>(), A.seq_l.as<uint32_t>(), R1, b2,
                                               B.seq_s.as<uint64_t>(), B.seq_l.as<uint32_t>(), R2, gates);
        I_TRY(hipGetLastError());
    }
    uint64_t n_clean = 0;
    if (P) {
        k_clean<<<grid_for(P), BLOCK, 0, st>>>(gates, R1, P, c->clean.as<uint32_t>());
        I_TRY(hipGetLastError());
        rc = scan_excl(c, c->clean.as<uint32_t>(), c->clean_pre.as<unsigned long long>(), P);
        if (rc) return rc;
        unsigned long long lp = 0;
        uint32_t lc = 0;
        I_TRY(hipMemcpyAsync(&lp, c->clean_pre.as<unsigned long long>() + P - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&lc, c->clean.as<uint32_t>() + P - 1, 4, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        n_clean = lp + lc;
    }
    if (tot) {
        k_units<<<grid_for(tot), BLOCK, 0, st>>>(gates, R1, R2, P, c->clean.as<uint32_t>(),
                                               c->clean_pre.as<unsigned long long>(), n_clean,
                                               c->ureads.as<unsigned long long>(), c->ubytes.as<unsigned long long>(),
                                               kq ? c->ubases.as<unsigned long long>() : nullptr,
                                               c->small.as<unsigned long long>());
        I_TRY(hipGetLastError());
    }
    rc = scan_excl(c, c->ureads.as<unsigned long long>(), c->rbase.as<unsigned long long>(), U);
    if (!rc) rc = scan_excl(c, c->ubytes.as<unsigned long long>(), c->bbase.as<unsigned long long>(), U);
    if (!rc && kq) rc = scan_excl(c, c->ubases.as<unsigned long long>(), c->qbase.as<unsigned long long>(), U);
    if (rc) return rc;
    unsigned long long tail[6] = {0, 0, 0, 0, 0, 0}, ambiguous = 0;
    if (U) {
        I_TRY(hipMemcpyAsync(&tail[0], c->rbase.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[1], c->ureads.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[2], c->bbase.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[3], c->ubytes.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        if (kq) {
            I_TRY(hipMemcpyAsync(&tail[4], c->qbase.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
            I_TRY(hipMemcpyAsync(&tail[5], c->ubases.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost,
                                 st));
        }
    }
    I_TRY(hipMemcpyAsync(&ambiguous, c->small.p, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    const uint64_t n_reads = tail[0] + tail[1], data_len = tail[2] + tail[3], n_quals = tail[4] + tail[5];
    I_TRY(c->data.ensure(data_len + 8));
    if (kq) {
        I_TRY(c->quals.ensure(n_quals + 8));
        I_TRY(c->qoff.ensure(n_reads * 8 + 8));
    }
    I_TRY(c->off.ensure(n_reads * 8 + 8));
    I_TRY(c->len.ensure(n_reads * 4 + 4));
    if (tot) {
        k_place<<<grid_for(tot), BLOCK, 0, st>>>(tot, R1, P, c->clean.as<uint32_t>(),
                                               c->clean_pre.as<unsigned long long>(), n_clean,
                                               c->rbase.as<unsigned long long>(), c->bbase.as<unsigned long long>(),
                                               kq ? c->qbase.as<unsigned long long>() : nullptr,
                                               c->place.as<Place>());
        I_TRY(hipGetLastError());
        k_emit<<<grid_for(tot * WAVE), BLOCK, 0, st>>>(b1, A.seq_s.as<uint64_t>(), A.seq_l.as<uint32_t>(), R1, b2,
                                                      B.seq_s.as<uint64_t>(), B.seq_l.as<uint32_t>(), R2, P, gates,
                                                      c->clean.as<uint32_t>(), c->place.as<Place>(),
                                                      c->data.as<uint8_t>(), c->off.as<uint64_t>(),
                                                      c->len.as<uint32_t>(), A.qual_s.as<uint64_t>(),
                                                      B.qual_s.as<uint64_t>(), kq ? c->quals.as<uint8_t>() : nullptr,
                                                      kq ? c->qoff.as<uint64_t>() : nullptr);
        I_TRY(hipGetLastError());
    }
    I_TRY(hipEventRecord(c->ev[1], st));
    out->data.resize(data_len);
    out->off.resize(n_reads);
    out->len.resize(n_reads);
    if (data_len) I_TRY(hipMemcpyAsync(out->data.data(), c->data.p, data_len, hipMemcpyDeviceToHost, st));
    if (n_reads) {
        I_TRY(hipMemcpyAsync(out->off.data(), c->off.p, n_reads * 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(out->len.data(), c->len.p, n_reads * 4, hipMemcpyDeviceToHost, st));
    }
    if (kq) {
        out->quals.resize(n_quals);
        out->qoff.resize(n_reads);
        if (n_quals) I_TRY(hipMemcpyAsync(out->quals.data(), c->quals.p, n_quals, hipMemcpyDeviceToHost, st));
        if (n_reads) I_TRY(hipMemcpyAsync(out->qoff.data(), c->qoff.p, n_reads * 8, hipMemcpyDeviceToHost, st));
    }
    I_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    I_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    out->info.n_reads = n_reads;
    out->info.data_len = data_len;
    out->info.pe_count = n_clean;
    out->info.n_ambiguous = ambiguous;
    out->info.r1_records = R1;
    out->info.r2_records = R2;
    out->info.ms_device = ms;
    out->info.n_quals = kq ? n_quals : 0;
    return KC_OK;
}

}  // namespace

extern "C" {

int kc_ingest_create(int32_t device, kc_ingest** out) {
    if (!out) return fail(KC_EINVAL, "null out");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(KC_EHIP, "no GPU");
    if (device < 0 || device >= ndev) return fail(KC_EINVAL, "bad device");
    I_TRY(hipSetDevice(device));
    auto* c = new kc_ingest();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&c->ev[1]);
    if (e != hipSuccess) {
        kc_ingest_destroy(c);
        return fail(KC_EHIP, hip_msg("kc_ingest_create", e));
    }
    *out = c;
    return KC_OK;
}

void kc_ingest_destroy(kc_ingest* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (hipEvent_t e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int kc_ingest_pack(kc_ingest* c, const uint8_t* r1, size_t r1_len, const uint8_t* r2, size_t r2_len,
                   kc_packed** out) {
    if (!c || !out) return fail(KC_EINVAL, "null argument");
    *out = nullptr;
    if (r1_len && !r1) return fail(KC_EINVAL, "null r1");
    if (r2 == nullptr && r2_len) return fail(KC_EINVAL, "null r2");
    auto* p = new kc_packed();
    const int rc = ingest(c, r1, r1_len, r2, r2_len, p);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return KC_OK;
}

int kc_fastq_pack(int32_t device, const uint8_t* r1, size_t r1_len, const uint8_t* r2, size_t r2_len,
                  kc_packed** out) {
    if (!out) return fail(KC_EINVAL, "null out");
    *out = nullptr;
    kc_ingest* c = nullptr;
    int rc = kc_ingest_create(device, &c);
    if (rc) return rc;
    rc = kc_ingest_pack(c, r1, r1_len, r2, r2_len, out);
    kc_ingest_destroy(c);
    return rc;
}

const char* kc_fastq_last_error(void) { return g_ingest_err.c_str(); }

int kc_packed_info_get(const kc_packed* p, kc_packed_info* info) {
    if (!p || !info) return KC_EINVAL;
    *info = p->info;
    return KC_OK;
}

int kc_packed_export(const kc_packed* p, uint8_t* data, size_t data_len, uint64_t* offsets, uint32_t* lengths,
                     size_t n_reads) {
    if (!p) return KC_EINVAL;
    if (data_len != p->data.size() || n_reads != p->len.size()) return KC_EINVAL;
    if ((data_len && !data) || (n_reads && (!offsets || !lengths))) return KC_EINVAL;
    std::copy(p->data.begin(), p->data.end(), data);
    std::copy(p->off.begin(), p->off.end(), offsets);
    std::copy(p->len.begin(), p->len.end(), lengths);
    return KC_OK;
}

int kc_ingest_keep_quals(kc_ingest* c, int32_t keep) {
    if (!c) return fail(KC_EINVAL, "null context");
    c->keep_quals = keep != 0;
    return KC_OK;
}

int kc_packed_export_quals(const kc_packed* p, uint8_t* quals, size_t n_quals, uint64_t* qoffsets, size_t n_reads) {
    if (!p) return KC_EINVAL;
    if (p->qoff.size() != p->len.size()) return fail(KC_ESTATE, "qualities were not kept (kc_ingest_keep_quals)");
    if (n_quals != p->quals.size() || n_reads != p->qoff.size()) return KC_EINVAL;
    if ((n_quals && !quals) || (n_reads && !qoffsets)) return KC_EINVAL;
    std::copy(p->quals.begin(), p->quals.end(), quals);
    std::copy(p->qoff.begin(), p->qoff.end(), qoffsets);
    return KC_OK;
}

void kc_packed_destroy(kc_packed* p) { delete p; }

}  // extern "C"
