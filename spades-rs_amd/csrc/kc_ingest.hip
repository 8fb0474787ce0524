// kc_ingest.hip — FASTQ ingestion on the device: paired (or single) file contents ->
// the reference's three-block PackedReads layout (packed_reads.rs:325-380, 490-501), with
// FastxReader's record checks (packed_reads.rs:141-217) and the three N gates
// (packed_reads.rs:513-535, 630-690).
//
// Pipeline (one stream): newline positions by a scan; per-line flags; per-record checks;
// per-read trimming / N scan / segment count; output units in three-block order (block 1
// = clean pairs' R1, block 2 = their R2, block 3 = the other pairs' segments, then the
// longer file's extra reads), whose read and byte bases come from exclusive scans; one
// thread per input read writes its output reads' offsets, lengths and packed bytes.
// The device handles the 4-line FASTQ layout (blank or whitespace-only lines allowed
// before the first record and after the last); blank lines between records and FASTA
// give KC_EUNSUPPORTED, for the host parser.
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

constexpr uint32_t MIN_K = 21;  // packed_reads.rs:38

__device__ __forceinline__ uint32_t code2(uint8_t b) {  // dna.rs:11-30, 4 = not ACGT
    const uint32_t c = b | 0x20u;
    return c == 'a' ? 0u : c == 'c' ? 1u : c == 'g' ? 2u : c == 't' ? 3u : 4u;
}

__device__ __forceinline__ bool is_ws(uint8_t b) {
    return b == ' ' || b == '\t' || b == '\n' || b == '\r' || b == '\f' || b == 0x0B;
}

__global__ void k_nl_flags(const uint8_t* __restrict__ b, uint64_t n, uint32_t* __restrict__ f) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        f[i] = b[i] == '\n';
}

__global__ void k_nl_scatter(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ pre,
                             unsigned long long* __restrict__ nl) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        if (b[i] == '\n') nl[pre[i]] = i;
}

__device__ __forceinline__ void line_span(const unsigned long long* nl, uint64_t m, uint64_t n, uint64_t j,
                                          uint64_t& s, uint64_t& e) {
    s = j ? nl[j - 1] + 1 : 0;
    e = j < m ? nl[j] : n;
}

// Per line: 1 = empty after stripping CR/LF, 2 = whitespace only.  red[0] = first line
// that is not whitespace only (FastxReader::open), red[1] = last non-empty line + 1.
__global__ void k_lines(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                        uint64_t m, uint64_t lines, unsigned long long* __restrict__ red) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < lines; j += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s, e;
        line_span(nl, m, n, j, s, e);
        uint64_t t = e;
        while (t > s && (b[t - 1] == '\r' || b[t - 1] == '\n')) --t;
        bool ws = true;
        for (uint64_t i = s; i < e && ws; ++i) ws = is_ws(b[i]);
        if (!ws) atomicMin(&red[0], (unsigned long long)j);
        if (t > s) atomicMax(&red[1], (unsigned long long)(j + 1));
    }
}

// Record r = lines first + 4r .. + 3.  Writes the sequence span (CR/LF stripped) and sets
// err bits: 1 = bad '@' header, 2 = bad '+' line, 4 = quality length, 8 = an empty
// header line (blank lines between records); bad[0] = first bad record.
__global__ void k_records(const uint8_t* __restrict__ b, uint64_t n, const unsigned long long* __restrict__ nl,
                          uint64_t m, uint64_t first, uint64_t R, uint64_t* __restrict__ seq_s,
                          uint32_t* __restrict__ seq_l, uint32_t* __restrict__ err,
                          unsigned long long* __restrict__ bad) {
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t s[4], e[4];
        for (int q = 0; q < 4; ++q) {
            line_span(nl, m, n, first + 4 * r + q, s[q], e[q]);
            while (e[q] > s[q] && (b[e[q] - 1] == '\r' || b[e[q] - 1] == '\n')) --e[q];
        }
        uint32_t f = 0;
        if (e[0] == s[0]) f |= 8u;
        else if (b[s[0]] != '@') f |= 1u;
        if (s[2] >= n || b[s[2]] != '+') f |= 2u;  // checked on the raw line
        if (e[3] - s[3] != e[1] - s[1]) f |= 4u;
        if (e[1] - s[1] > 0xFFFFFFFFull) f |= 4u;
        seq_s[r] = s[1];
        seq_l[r] = (uint32_t)(e[1] - s[1]);
        if (f) {
            atomicOr(err, f);
            atomicMin(bad, (unsigned long long)r);
        }
    }
}

// Gate 1 and the N scan of one read: trimmed span [ts, te), internal-N flag, number of
// ACGT segments of at least MIN_K bases and their packed bytes.
struct ReadGate {
    uint32_t ts, tl, nseg;
    uint8_t internal_n;
    uint64_t seg_bytes;
};

__device__ ReadGate gate(const uint8_t* __restrict__ b, uint64_t s, uint32_t l) {
    ReadGate g{};
    uint32_t a = 0, z = l;
    while (a < l && code2(b[s + a]) == 4u) ++a;
    while (z > a && code2(b[s + z - 1]) == 4u) --z;
    g.ts = a;
    g.tl = z - a;
    uint32_t i = a;
    while (i < z) {
        while (i < z && code2(b[s + i]) == 4u) {
            g.internal_n = 1;
            ++i;
        }
        uint32_t j = i;
        while (j < z && code2(b[s + j]) != 4u) ++j;
        if (j - i >= MIN_K) {
            g.nseg += 1;
            g.seg_bytes += (j - i + 3) / 4;
        }
        i = j;
    }
    return g;
}

// Per input read: its output unit and that unit's reads / bytes; pair clean flags.
// Reads 0..P-1 of both files are pairs (when paired), the rest extras.  Unit numbering:
// clean pair c: units c (R1) and n_clean + c (R2); non-clean pair d: 2 n_clean + 2d (+1);
// extra e: 2P + e.  pass 0 computes clean flags; pass 1 (with the clean prefix) the rest.
__global__ void k_units(const uint8_t* __restrict__ b1, const uint64_t* __restrict__ s1,
                        const uint32_t* __restrict__ l1, uint64_t R1, const uint8_t* __restrict__ b2,
                        const uint64_t* __restrict__ s2, const uint32_t* __restrict__ l2, uint64_t R2, uint64_t P,
                        int pass, uint32_t* __restrict__ clean, const unsigned long long* __restrict__ clean_pre,
                        uint64_t n_clean, unsigned long long* __restrict__ unit_reads,
                        unsigned long long* __restrict__ unit_bytes, unsigned long long* __restrict__ ambiguous) {
    const uint64_t total = R1 + R2;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
        const bool second = x >= R1;
        const uint64_t i = second ? x - R1 : x;
        const ReadGate g = second ? gate(b2, s2[i], l2[i]) : gate(b1, s1[i], l1[i]);
        if (i < P) {
            const ReadGate h = second ? gate(b1, s1[i], l1[i]) : gate(b2, s2[i], l2[i]);
            const bool c = !g.internal_n && !h.internal_n && g.tl >= MIN_K && h.tl >= MIN_K;
            if (pass == 0) {
                if (!second) clean[i] = c;
                continue;
            }
            const uint64_t ci = clean_pre[i];
            uint64_t u;
            if (c) {
                u = second ? n_clean + ci : ci;
                unit_reads[u] = 1;
                unit_bytes[u] = (g.tl + 3) / 4;
            } else {
                u = 2 * n_clean + 2 * (i - ci) + (second ? 1 : 0);
                unit_reads[u] = g.nseg;
                unit_bytes[u] = g.seg_bytes;
                if (g.tl && !g.nseg) atomicAdd(ambiguous, 1ull);
            }
        } else if (pass == 1) {
            const uint64_t e = second ? (i - P) + (R1 - P) : i - P;  // one file has no extras
            const uint64_t u = 2 * P + e;
            unit_reads[u] = g.nseg;
            unit_bytes[u] = g.seg_bytes;
            if (g.tl && !g.nseg) atomicAdd(ambiguous, 1ull);
        }
    }
}

__device__ void pack_run(const uint8_t* __restrict__ b, uint64_t s, uint32_t l, uint8_t* __restrict__ out) {
    for (uint32_t q = 0; q < l; q += 4) {
        uint32_t v = 0;
        for (uint32_t r = 0; r < 4; ++r)
            if (q + r < l) v |= code2(b[s + q + r]) << (6 - 2 * r);
        out[q / 4] = (uint8_t)v;
    }
}

// Writes every input read's output reads (offsets, lengths, packed bytes).
__global__ void k_emit(const uint8_t* __restrict__ b1, const uint64_t* __restrict__ s1,
                       const uint32_t* __restrict__ l1, uint64_t R1, const uint8_t* __restrict__ b2,
                       const uint64_t* __restrict__ s2, const uint32_t* __restrict__ l2, uint64_t R2, uint64_t P,
                       const uint32_t* __restrict__ clean, const unsigned long long* __restrict__ clean_pre,
                       uint64_t n_clean, const unsigned long long* __restrict__ read_base,
                       const unsigned long long* __restrict__ byte_base, uint8_t* __restrict__ data,
                       uint64_t* __restrict__ offsets, uint32_t* __restrict__ lengths) {
    const uint64_t total = R1 + R2;
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < total; x += (uint64_t)gridDim.x * blockDim.x) {
        const bool second = x >= R1;
        const uint64_t i = second ? x - R1 : x;
        const uint8_t* b = second ? b2 : b1;
        const uint64_t s = second ? s2[i] : s1[i];
        const uint32_t l = second ? l2[i] : l1[i];
        const ReadGate g = gate(b, s, l);
        uint64_t u;
        bool whole = false;
        if (i < P) {
            const uint64_t ci = clean_pre[i];
            if (clean[i]) {
                u = second ? n_clean + ci : ci;
                whole = true;
            } else {
                u = 2 * n_clean + 2 * (i - ci) + (second ? 1 : 0);
            }
        } else {
            u = 2 * P + (second ? (i - P) + (R1 - P) : i - P);
        }
        uint64_t rb = read_base[u], bb = byte_base[u];
        if (whole) {
            offsets[rb] = bb;
            lengths[rb] = g.tl;
            pack_run(b, s + g.ts, g.tl, data + bb);
            continue;
        }
        const uint32_t z = g.ts + g.tl;
        uint32_t p = g.ts;
        while (p < z) {
            while (p < z && code2(b[s + p]) == 4u) ++p;
            uint32_t q = p;
            while (q < z && code2(b[s + q]) != 4u) ++q;
            if (q - p >= MIN_K) {
                offsets[rb] = bb;
                lengths[rb] = q - p;
                pack_run(b, s + p, q - p, data + bb);
                ++rb;
                bb += (q - p + 3) / 4;
            }
            p = q;
        }
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

struct FileIndex {
    Buf buf, flags, pre, nl, red, seq_s, seq_l, err;
    uint64_t R = 0;
};

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

// Uploads one file and finds its records' sequence spans.
int index_file(const uint8_t* host, uint64_t n, hipStream_t st, FileIndex& F, const char* name) {
    I_TRY(F.buf.ensure(n + 64));
    if (n) I_TRY(hipMemcpyAsync(F.buf.p, host, n, hipMemcpyHostToDevice, st));
    I_TRY(F.flags.ensure(n * 4 + 4));
    I_TRY(F.pre.ensure(n * 8 + 8));
    I_TRY(F.red.ensure(64));
    uint64_t m = 0;
    if (n) {
        k_nl_flags<<<grid_for(n), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), n, F.flags.as<uint32_t>());
        I_TRY(hipGetLastError());
        size_t tb = 0;
        Buf tmp;
        I_TRY(rocprim::exclusive_scan(nullptr, tb, F.flags.as<uint32_t>(), F.pre.as<unsigned long long>(), 0ull,
                                      (size_t)n, rocprim::plus<unsigned long long>(), st));
        I_TRY(tmp.ensure(tb));
        I_TRY(rocprim::exclusive_scan(tmp.p, tb, F.flags.as<uint32_t>(), F.pre.as<unsigned long long>(), 0ull,
                                      (size_t)n, rocprim::plus<unsigned long long>(), st));
        uint64_t last_pre = 0;
        uint32_t last_flag = 0;
        I_TRY(hipMemcpyAsync(&last_pre, F.pre.as<unsigned long long>() + n - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&last_flag, F.flags.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        m = last_pre + last_flag;
    }
    I_TRY(F.nl.ensure(m * 8 + 8));
    if (m) {
        k_nl_scatter<<<grid_for(n), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), n, F.pre.as<unsigned long long>(),
                                                    F.nl.as<unsigned long long>());
        I_TRY(hipGetLastError());
    }
    // read_until semantics: a final line without a newline still counts
    uint8_t lastb = '\n';
    if (n) {
        I_TRY(hipMemcpyAsync(&lastb, F.buf.as<uint8_t>() + n - 1, 1, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
    }
    const uint64_t lines = m + (n && lastb != '\n' ? 1 : 0);
    unsigned long long red[2] = {~0ull, 0ull};
    I_TRY(hipMemcpyAsync(F.red.p, red, 16, hipMemcpyHostToDevice, st));
    if (lines) {
        k_lines<<<grid_for(lines), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), n, F.nl.as<unsigned long long>(), m, lines,
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
        I_TRY(hipMemcpyAsync(&head, F.buf.as<uint8_t>() + at, 1, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
    }
    if (head == '>') return fail(KC_EUNSUPPORTED, std::string(name) + ": FASTA input is parsed on the host");
    if (head != '@') return fail(KC_EINVAL, std::string(name) + ": expected '@' or '>' header");
    // Empty lines between the first and the last record are either legal empty sequence /
    // quality lines or blank separators, which FastxReader skips before a header.  With
    // none, any misalignment is a malformed record; with some, a misaligned file is left
    // to the host parser.
    unsigned long long empties = 0;
    {
        I_TRY(hipMemcpyAsync(F.red.p, &empties, 8, hipMemcpyHostToDevice, st));
        k_count_empty<<<grid_for(end - first), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), n, F.nl.as<unsigned long long>(),
                                                               m, first, end, F.red.as<unsigned long long>());
        I_TRY(hipGetLastError());
        I_TRY(hipMemcpyAsync(&empties, F.red.p, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        if ((end - first) % 4 != 0) {
            if (empties) {
                // empty sequence and quality lines are legal; blank separator lines are not
                // handled on the device
                return fail(KC_EUNSUPPORTED, std::string(name) +
                                                 ": blank lines between records: parse it on the host");
            }
            return fail(KC_EINVAL, std::string(name) + ": malformed FASTQ: unexpected EOF inside a record");
        }
    }
    const uint64_t R = (end - first) / 4;
    I_TRY(F.seq_s.ensure(R * 8 + 8));
    I_TRY(F.seq_l.ensure(R * 4 + 4));
    I_TRY(F.err.ensure(64));
    unsigned long long init[2] = {0ull, ~0ull};
    I_TRY(hipMemcpyAsync(F.err.p, init, 16, hipMemcpyHostToDevice, st));
    if (R) {
        k_records<<<grid_for(R), BLOCK, 0, st>>>(F.buf.as<uint8_t>(), n, F.nl.as<unsigned long long>(), m, first, R,
                                                 F.seq_s.as<uint64_t>(), F.seq_l.as<uint32_t>(),
                                                 F.err.as<uint32_t>(), F.err.as<unsigned long long>() + 1);
        I_TRY(hipGetLastError());
    }
    unsigned long long out[2];
    I_TRY(hipMemcpyAsync(out, F.err.p, 16, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    const uint32_t e = (uint32_t)out[0];
    if (e && ((e & 8u) || empties))
        return fail(KC_EUNSUPPORTED, std::string(name) + ": blank lines between records: parse it on the host");
    if (e) {
        const char* what = (e & 1u) ? "expected '@' header" : (e & 2u) ? "expected '+' separator line"
                                                                        : "sequence length != quality length";
        return fail(KC_EINVAL, std::string(name) + ": malformed FASTQ record " + std::to_string(out[1]) + ": " + what);
    }
    F.R = R;
    return KC_OK;
}

int scan_u64(const unsigned long long* in, unsigned long long* out, uint64_t n, hipStream_t st) {
    if (!n) return KC_OK;
    size_t tb = 0;
    Buf tmp;
    I_TRY(rocprim::exclusive_scan(nullptr, tb, in, out, 0ull, (size_t)n, rocprim::plus<unsigned long long>(), st));
    I_TRY(tmp.ensure(tb));
    I_TRY(rocprim::exclusive_scan(tmp.p, tb, in, out, 0ull, (size_t)n, rocprim::plus<unsigned long long>(), st));
    I_TRY(hipStreamSynchronize(st));
    return KC_OK;
}

int ingest(int32_t device, const uint8_t* r1, uint64_t n1, const uint8_t* r2, uint64_t n2, kc_packed* out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(KC_EHIP, "no GPU");
    if (device < 0 || device >= ndev) return fail(KC_EINVAL, "bad device");
    I_TRY(hipSetDevice(device));
    hipStream_t st = nullptr;
    I_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{st};
    FileIndex A, B;
    int rc = index_file(r1, n1, st, A, "R1");
    if (rc) return rc;
    const bool paired = r2 != nullptr;
    if (paired) {
        rc = index_file(r2, n2, st, B, "R2");
        if (rc) return rc;
    }
    const uint64_t R1 = A.R, R2 = paired ? B.R : 0;
    const uint64_t P = paired ? std::min(R1, R2) : 0;
    const uint64_t E = R1 + R2 - 2 * P;
    const uint64_t U = 2 * P + E;
    Buf clean, clean_pre, ureads, ubytes, rbase, bbase, amb;
    I_TRY(clean.ensure(P * 4 + 4));
    I_TRY(clean_pre.ensure(P * 8 + 8));
    I_TRY(ureads.ensure(U * 8 + 8));
    I_TRY(ubytes.ensure(U * 8 + 8));
    I_TRY(rbase.ensure(U * 8 + 8));
    I_TRY(bbase.ensure(U * 8 + 8));
    I_TRY(amb.ensure(8));
    const unsigned long long zero = 0;
    I_TRY(hipMemcpyAsync(amb.p, &zero, 8, hipMemcpyHostToDevice, st));
    const uint8_t* b1 = A.buf.as<uint8_t>();
    const uint8_t* b2 = paired ? B.buf.as<uint8_t>() : b1;
    const uint64_t* s2 = paired ? B.seq_s.as<uint64_t>() : A.seq_s.as<uint64_t>();
    const uint32_t* l2 = paired ? B.seq_l.as<uint32_t>() : A.seq_l.as<uint32_t>();
    const uint64_t tot = R1 + R2;
    uint64_t n_clean = 0;
    if (P) {
        k_units<<<grid_for(tot), BLOCK, 0, st>>>(b1, A.seq_s.as<uint64_t>(), A.seq_l.as<uint32_t>(), R1, b2, s2, l2, R2,
                                               P, 0, clean.as<uint32_t>(), nullptr, 0, nullptr, nullptr, nullptr);
        I_TRY(hipGetLastError());
        size_t tb = 0;
        Buf tmp;
        I_TRY(rocprim::exclusive_scan(nullptr, tb, clean.as<uint32_t>(), clean_pre.as<unsigned long long>(), 0ull,
                                      (size_t)P, rocprim::plus<unsigned long long>(), st));
        I_TRY(tmp.ensure(tb));
        I_TRY(rocprim::exclusive_scan(tmp.p, tb, clean.as<uint32_t>(), clean_pre.as<unsigned long long>(), 0ull,
                                      (size_t)P, rocprim::plus<unsigned long long>(), st));
        unsigned long long lp = 0;
        uint32_t lc = 0;
        I_TRY(hipMemcpyAsync(&lp, clean_pre.as<unsigned long long>() + P - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&lc, clean.as<uint32_t>() + P - 1, 4, hipMemcpyDeviceToHost, st));
        I_TRY(hipStreamSynchronize(st));
        n_clean = lp + lc;
    }
    if (tot) {
        k_units<<<grid_for(tot), BLOCK, 0, st>>>(b1, A.seq_s.as<uint64_t>(), A.seq_l.as<uint32_t>(), R1, b2, s2, l2, R2,
                                               P, 1, clean.as<uint32_t>(), clean_pre.as<unsigned long long>(), n_clean,
                                               ureads.as<unsigned long long>(), ubytes.as<unsigned long long>(),
                                               amb.as<unsigned long long>());
        I_TRY(hipGetLastError());
    }
    rc = scan_u64(ureads.as<unsigned long long>(), rbase.as<unsigned long long>(), U, st);
    if (rc) return rc;
    rc = scan_u64(ubytes.as<unsigned long long>(), bbase.as<unsigned long long>(), U, st);
    if (rc) return rc;
    unsigned long long tail[4] = {0, 0, 0, 0};
    if (U) {
        I_TRY(hipMemcpyAsync(&tail[0], rbase.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[1], ureads.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[2], bbase.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(&tail[3], ubytes.as<unsigned long long>() + U - 1, 8, hipMemcpyDeviceToHost, st));
    }
    unsigned long long ambiguous = 0;
    I_TRY(hipMemcpyAsync(&ambiguous, amb.p, 8, hipMemcpyDeviceToHost, st));
    I_TRY(hipStreamSynchronize(st));
    const uint64_t n_reads = tail[0] + tail[1], data_len = tail[2] + tail[3];
    Buf d_data, d_off, d_len;
    I_TRY(d_data.ensure(data_len + 8));
    I_TRY(d_off.ensure(n_reads * 8 + 8));
    I_TRY(d_len.ensure(n_reads * 4 + 4));
    if (tot) {
        k_emit<<<grid_for(tot), BLOCK, 0, st>>>(b1, A.seq_s.as<uint64_t>(), A.seq_l.as<uint32_t>(), R1, b2, s2, l2, R2,
                                              P, clean.as<uint32_t>(), clean_pre.as<unsigned long long>(), n_clean,
                                              rbase.as<unsigned long long>(), bbase.as<unsigned long long>(),
                                              d_data.as<uint8_t>(), d_off.as<uint64_t>(), d_len.as<uint32_t>());
        I_TRY(hipGetLastError());
    }
    out->data.resize(data_len);
    out->off.resize(n_reads);
    out->len.resize(n_reads);
    if (data_len) I_TRY(hipMemcpyAsync(out->data.data(), d_data.p, data_len, hipMemcpyDeviceToHost, st));
    if (n_reads) {
        I_TRY(hipMemcpyAsync(out->off.data(), d_off.p, n_reads * 8, hipMemcpyDeviceToHost, st));
        I_TRY(hipMemcpyAsync(out->len.data(), d_len.p, n_reads * 4, hipMemcpyDeviceToHost, st));
    }
    I_TRY(hipStreamSynchronize(st));
    out->info.n_reads = n_reads;
    out->info.data_len = data_len;
    out->info.pe_count = n_clean;
    out->info.n_ambiguous = ambiguous;
    out->info.r1_records = R1;
    out->info.r2_records = R2;
    return KC_OK;
}

}  // namespace

extern "C" {

int kc_fastq_pack(int32_t device, const uint8_t* r1, size_t r1_len, const uint8_t* r2, size_t r2_len,
                  kc_packed** out) {
    if (!out) return KC_EINVAL;
    *out = nullptr;
    if (r1_len && !r1) return fail(KC_EINVAL, "null r1");
    if (r2 == nullptr && r2_len) return fail(KC_EINVAL, "null r2");
    auto* p = new kc_packed();
    const int rc = ingest(device, r1, r1_len, r2, r2_len, p);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return KC_OK;
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

void kc_packed_destroy(kc_packed* p) { delete p; }

}  // extern "C"
