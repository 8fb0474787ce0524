// kc_graph.hip — compacted unitigs of a library's solid set on the GPU
// (CompactedGraph::build, graph.rs:65-234).
//
// The reference walks maximal non-branching paths one at a time from a hash set.  Its
// result is fixed by a link relation between oriented k-mers: u -> v when u has exactly
// one solid successor v (graph.rs:79-91), v has exactly one solid predecessor (then u,
// graph.rs:93-105) and v is another canonical k-mer (a walk never re-enters a visited
// k-mer, graph.rs:112-126, 138-150).  Every canonical k-mer lies on exactly one chain of
// links, a path or a cycle, and every walk emits one whole chain.  Here every step is a
// data-parallel pass over the solid k-mers:
//
//   1. index    open-addressed hash of the sorted solid values -> rank
//   2. adjacent 4 successor + 4 predecessor lookups per k-mer (both strands follow)
//   3. links    the link condition from the neighbours' degrees
//   4. rank     pointer jumping over the predecessor links of the 2n oriented k-mers:
//               the head of each chain and the distance to it, and the smallest rank on
//               each cycle; a cycle is opened before its smallest k-mer in forward
//               orientation (where the reference's pass 2 starts when it iterates in
//               ascending order, graph.rs:216-230) and ranked again
//   5. emit     one strand per chain (the one whose head has the smaller oriented id),
//               the unitig's bases and the per-position counts for its mean coverage
//
// Chains come out ordered by their head k-mer; a linear chain's strand may differ from the
// reference's walk (which depends on hash order), its reverse complement being the same
// unitig.  Odd k only: the reference's k values are odd, and at even k a palindromic k-mer
// is its own twin, which the oriented-node chains here do not model.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "../../include/kc.h"
#include "kc_device.h"
#include "kc_host.h"

using namespace kc;

namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;

template <int KW>
__device__ __forceinline__ Kv load_val(const uint64_t* __restrict__ vals, uint64_t i) {
    Kv v{0, 0, 0, 0};
    v.lo = vals[i * KW];
    if (KW >= 2) v.hi = vals[i * KW + 1];
    if (KW == 4) {
        v.w2 = vals[i * KW + 2];
        v.w3 = vals[i * KW + 3];
    }
    return v;
}

template <int KW>
__device__ __forceinline__ bool kv_eq(const Kv& a, const Kv& b) {
    return a.lo == b.lo && (KW < 2 || a.hi == b.hi) && (KW < 4 || (a.w2 == b.w2 && a.w3 == b.w3));
}

__device__ __forceinline__ uint64_t fmix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

template <int KW>
__device__ __forceinline__ uint64_t vhash(const Kv& v) {
    uint64_t h = fmix(v.lo + 0x9e3779b97f4a7c15ull);
    if (KW >= 2) h = fmix(h ^ v.hi);
    if (KW == 4) {
        h = fmix(h ^ v.w2);
        h = fmix(h ^ v.w3);
    }
    return h;
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

// Slot = (hash tag << 32) | rank; an absent key ends at an empty slot.
template <int KW>
__global__ void k_index_insert(const uint64_t* __restrict__ vals, uint64_t n, unsigned long long* __restrict__ slots,
                               uint64_t mask) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t h = vhash<KW>(load_val<KW>(vals, i));
        const unsigned long long item = ((h >> 32) << 32) | i;
        uint64_t p = h & mask;
        while (atomicCAS(&slots[p], ~0ull, item) != ~0ull) p = (p + 1) & mask;
    }
}

template <int KW>
__device__ __forceinline__ uint32_t index_find(const unsigned long long* __restrict__ slots, uint64_t mask,
                                               const uint64_t* __restrict__ vals, const Kv& c) {
    const uint64_t h = vhash<KW>(c);
    const uint32_t tag = (uint32_t)(h >> 32);
    for (uint64_t p = h & mask;; p = (p + 1) & mask) {
        const unsigned long long s = slots[p];
        if (s == ~0ull) return NONE;
        if ((uint32_t)(s >> 32) == tag && kv_eq<KW>(load_val<KW>(vals, (uint32_t)s), c)) return (uint32_t)s;
    }
}

// graph.rs:79-105 on the forward strand of every solid k-mer: how many of its four
// successors / predecessors are solid, and the last one found as an oriented id
// (2 * rank + is_rc; the only one when the degree is 1).  deg = out | in << 4.
template <int KW>
__global__ void k_adjacent(const uint64_t* __restrict__ vals, uint64_t n, uint32_t k,
                           const unsigned long long* __restrict__ slots, uint64_t mask, uint8_t* __restrict__ deg,
                           uint32_t* __restrict__ succ, uint32_t* __restrict__ pred) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const Kv v = load_val<KW>(vals, x);
        uint32_t od = 0, id = 0, s = NONE, p = NONE;
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const Kv e = extend_right<KW>(v, b, k);
            const Kv r = revcomp<KW>(e, k);
            const bool rc = !kv_le(e, r);
            const uint32_t j = index_find<KW>(slots, mask, vals, rc ? r : e);
            if (j != NONE) {
                ++od;
                s = 2 * j + rc;
            }
            const Kv f = prepend_left<KW>(v, b, k);
            const Kv q = revcomp<KW>(f, k);
            const bool qrc = !kv_le(f, q);
            const uint32_t i = index_find<KW>(slots, mask, vals, qrc ? q : f);
            if (i != NONE) {
                ++id;
                p = 2 * i + qrc;
            }
        }
        deg[x] = (uint8_t)(od | (id << 4));
        succ[x] = s;
        pred[x] = p;
    }
}

// Degrees of an oriented k-mer: the reverse strand's successors are the twins of the
// forward strand's predecessors and vice versa.
__device__ __forceinline__ uint32_t out_deg(const uint8_t* deg, uint32_t u) {
    const uint8_t d = deg[u >> 1];
    return (u & 1) ? (d >> 4) : (d & 15);
}
__device__ __forceinline__ uint32_t in_deg(const uint8_t* deg, uint32_t u) {
    const uint8_t d = deg[u >> 1];
    return (u & 1) ? (d & 15) : (d >> 4);
}

// The forward strand's outgoing and incoming links (graph.rs:112-126 forward, 138-150
// backward): nxt[x] / prv[x] are oriented ids or NONE.  The reverse strand's links are the
// twins: nxt(2x+1) = twin(prv[x]), prv(2x+1) = twin(nxt[x]).
__global__ void k_links(uint64_t n, const uint8_t* __restrict__ deg, const uint32_t* __restrict__ succ,
                        const uint32_t* __restrict__ pred, uint32_t* __restrict__ nxt, uint32_t* __restrict__ prv) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const uint8_t d = deg[x];
        uint32_t a = NONE, b = NONE;
        if ((d & 15) == 1) {
            const uint32_t s = succ[x];
            if ((s >> 1) != x && in_deg(deg, s) == 1) a = s;
        }
        if ((d >> 4) == 1) {
            const uint32_t p = pred[x];
            if ((p >> 1) != x && out_deg(deg, p) == 1) b = p;
        }
        nxt[x] = a;
        prv[x] = b;
    }
}

__device__ __forceinline__ uint32_t prv_of(const uint32_t* nxt, const uint32_t* prv, uint32_t u) {
    const uint32_t x = u >> 1;
    if (!(u & 1)) return prv[x];
    const uint32_t t = nxt[x];
    return t == NONE ? NONE : t ^ 1;
}

// Pointer-jumping state of an oriented k-mer: the farthest predecessor reached (itself
// at a chain's head), the number of links to it, and the smallest rank on the way.
struct Jump {
    uint32_t j, d, m, pad;
};

__global__ void k_jump_init(uint64_t n2, const uint32_t* __restrict__ nxt, const uint32_t* __restrict__ prv,
                            Jump* __restrict__ st) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n2; u += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = prv_of(nxt, prv, (uint32_t)u);
        st[u] = Jump{p == NONE ? (uint32_t)u : p, p == NONE ? 0u : 1u, (uint32_t)(u >> 1), 0u};
    }
}

__global__ void k_jump(uint64_t n2, const Jump* __restrict__ in, Jump* __restrict__ out) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n2; u += (uint64_t)gridDim.x * blockDim.x) {
        const Jump a = in[u];
        const Jump b = in[a.j];
        out[u] = Jump{b.j, a.d + b.d, a.m < b.m ? a.m : b.m, 0u};
    }
}

// After enough rounds every k-mer on a path points at its head (a fixed point); one
// still pointing at a non-head lies on a cycle, and m is the cycle's smallest rank.  The
// cycle is opened before that k-mer's forward strand: the link p -> (m, fwd) goes, and
// with it its twin (m, rc) -> twin(p).  counter[0] counts the cycles.
__global__ void k_open_cycles(uint64_t n, const Jump* __restrict__ st, uint32_t* __restrict__ nxt,
                              uint32_t* __restrict__ prv, unsigned long long* __restrict__ counter) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const Jump a = st[2 * x];
        if (st[a.j].j == a.j || a.m != (uint32_t)x) continue;
        const uint32_t p = prv[x];
        prv[x] = NONE;
        if (p & 1) prv[p >> 1] = NONE;  // p = (j, rc): its outgoing link is twin(prv[j])
        else nxt[p >> 1] = NONE;
        atomicAdd(counter, 1ull);
    }
}

// One strand per chain: the one whose head has the smaller oriented id.  rep[x] = head
// rank | position << 32 | strand << 63; the chain's tail writes the chain's length at its
// head's rank.
__global__ void k_represent(uint64_t n, const Jump* __restrict__ st, const uint32_t* __restrict__ nxt,
                            const uint32_t* __restrict__ prv, unsigned long long* __restrict__ rep,
                            uint32_t* __restrict__ len) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const Jump a = st[2 * x], b = st[2 * x + 1];
        const uint32_t o = b.j < a.j ? 1u : 0u;
        const Jump r = o ? b : a;
        const uint32_t head = r.j >> 1;
        rep[x] = (unsigned long long)head | ((unsigned long long)r.d << 32) | ((unsigned long long)o << 63);
        const uint32_t out = o ? prv[x] : nxt[x];
        if (out == NONE) len[head] = r.d + 1;
    }
}

// Per rank: 1 at a chain's head, and the unitig's bases (length + k - 1) there.
__global__ void k_heads(uint64_t n, uint32_t k, const unsigned long long* __restrict__ rep,
                        const uint32_t* __restrict__ len, uint32_t* __restrict__ is_head,
                        unsigned long long* __restrict__ bases) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const bool h = (uint32_t)rep[x] == (uint32_t)x && ((rep[x] >> 32) & 0x7FFFFFFFull) == 0;
        is_head[x] = h ? 1u : 0u;
        bases[x] = h ? (unsigned long long)len[x] + k - 1 : 0ull;
    }
}

// graph.rs:159-178 build_unitig: the head writes its k bases, every other k-mer the last
// base of its strand at k - 1 + position; its count goes to the chain's per-position
// coverage array (k-mer offset = base offset - (k - 1) * unitig id).
template <int KW>
__global__ void k_emit(const uint64_t* __restrict__ vals, const uint32_t* __restrict__ counts, uint64_t n, uint32_t k,
                       const unsigned long long* __restrict__ rep, const unsigned long long* __restrict__ boff,
                       const uint32_t* __restrict__ uid, uint8_t* __restrict__ seq,
                       unsigned long long* __restrict__ pos_cov) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long r = rep[x];
        const uint32_t head = (uint32_t)r, pos = (uint32_t)((r >> 32) & 0x7FFFFFFFull), o = (uint32_t)(r >> 63);
        const Kv v = load_val<KW>(vals, x);
        const Kv e = o ? revcomp<KW>(v, k) : v;
        const uint64_t b0 = boff[head];
        if (pos == 0) {
            for (uint32_t i = 0; i < k; ++i) {
                const uint32_t bit = 2 * (k - 1 - i);
                seq[b0 + i] = "ACGT"[(kv_word(e, bit >> 6) >> (bit & 63)) & 3];
            }
        } else {
            seq[b0 + k - 1 + pos] = "ACGT"[e.lo & 3];
        }
        pos_cov[b0 - (uint64_t)(k - 1) * uid[head] + pos] = counts[x];
    }
}

// Unitig u at head rank x: its base offset, k-mer count and coverage sum (from the
// inclusive scan of the per-position counts).
__global__ void k_unitig_meta(uint64_t n, uint32_t k, const uint32_t* __restrict__ is_head,
                              const uint32_t* __restrict__ uid, const unsigned long long* __restrict__ boff,
                              const uint32_t* __restrict__ len, const unsigned long long* __restrict__ cov_scan,
                              unsigned long long* __restrict__ off, unsigned long long* __restrict__ kmers,
                              unsigned long long* __restrict__ cov_sum) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        if (!is_head[x]) continue;
        const uint32_t u = uid[x];
        const uint64_t k0 = boff[x] - (uint64_t)(k - 1) * u, m = len[x];
        off[u] = boff[x];
        kmers[u] = m;
        cov_sum[u] = cov_scan[k0 + m - 1] - (k0 ? cov_scan[k0 - 1] : 0ull);
    }
}

thread_local std::string g_err;

}  // namespace

struct kc_unitigs {
    int device = 0;
    uint64_t n_unitigs = 0, n_bases = 0, n_kmers = 0;
    double ms = 0.0, ms_stage[4] = {0, 0, 0, 0};
    DevBuf seq, off, kmers, cov;
    std::string err;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace {

// Scratch buffers, held by the engine context between calls.
struct Work {
    DevBuf &vals, &counts, &idx, &tmp, &slots, &deg, &succ, &pred, &nxt, &prv, &st0, &st1, &rep, &len, &is_head,
        &bases, &boff, &uid, &pos_cov, &cov_scan;
    explicit Work(DevBuf* b)
        : vals(b[0]), counts(b[1]), idx(b[2]), tmp(b[3]), slots(b[4]), deg(b[5]), succ(b[6]), pred(b[7]), nxt(b[8]),
          prv(b[9]), st0(b[10]), st1(b[11]), rep(b[12]), len(b[13]), is_head(b[14]), bases(b[15]), boff(b[16]),
          uid(b[17]), pos_cov(b[18]), cov_scan(b[19]) {}
};
static_assert(GRAPH_SCRATCH >= 20, "scratch buffers");

#define G_TRY(expr)                                                                                          \
    do {                                                                                                     \
        hipError_t _e = (expr);                                                                              \
        if (_e != hipSuccess) return ctx_fail(ctx, _e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,          \
                                              std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

// ev[0..4]: before the decode, after the index, after the links, after the ranking, at
// the end.
template <int KW>
int build(kc_ctx* ctx, uint32_t lib, const SolidSet& S, kc_unitigs* U, Work& w, hipEvent_t* ev) {
    const uint64_t n = S.n, n2 = 2 * n;
    const uint32_t k = S.k;
    hipStream_t st = S.stream;
    G_TRY(w.vals.ensure(n * KW * 8));
    G_TRY(w.counts.ensure(n * 4));
    {
        const int rc = solid_values(ctx, lib, KW, w.vals.as<uint64_t>(), w.counts.as<uint32_t>(), w.idx, w.tmp);
        if (rc != KC_OK) return rc;
    }
    uint64_t cap = 1;
    while (cap < 4 * n) cap <<= 1;  // load <= 1/4: an absent key mostly ends in the first cache line
    G_TRY(w.slots.ensure(cap * 8));
    G_TRY(w.deg.ensure(n));
    G_TRY(w.succ.ensure(n * 4));
    G_TRY(w.pred.ensure(n * 4));
    G_TRY(w.nxt.ensure(n * 4));
    G_TRY(w.prv.ensure(n * 4));
    G_TRY(w.st0.ensure(n2 * sizeof(Jump)));
    G_TRY(w.st1.ensure(n2 * sizeof(Jump)));
    G_TRY(w.rep.ensure(n * 8));
    G_TRY(w.len.ensure(n * 4));
    G_TRY(w.is_head.ensure(n * 4));
    G_TRY(w.bases.ensure(n * 8));
    G_TRY(w.boff.ensure((n + 1) * 8));
    G_TRY(w.uid.ensure((n + 1) * 4));
    G_TRY(w.pos_cov.ensure(n * 8));
    G_TRY(w.cov_scan.ensure(n * 8 + 8));
    const uint64_t* vals = w.vals.as<uint64_t>();
    auto* slots = w.slots.as<unsigned long long>();
    k_fill<unsigned long long><<<grid_for(cap), BLOCK, 0, st>>>(slots, cap, ~0ull);
    k_index_insert<KW><<<grid_for(n), BLOCK, 0, st>>>(vals, n, slots, cap - 1);
    (void)hipEventRecord(ev[1], st);
    k_adjacent<KW><<<grid_for(n), BLOCK, 0, st>>>(vals, n, k, slots, cap - 1, w.deg.as<uint8_t>(),
                                                  w.succ.as<uint32_t>(), w.pred.as<uint32_t>());
    uint32_t* nxt = w.nxt.as<uint32_t>();
    uint32_t* prv = w.prv.as<uint32_t>();
    k_links<<<grid_for(n), BLOCK, 0, st>>>(n, w.deg.as<uint8_t>(), w.succ.as<uint32_t>(), w.pred.as<uint32_t>(), nxt,
                                           prv);
    G_TRY(hipGetLastError());
    (void)hipEventRecord(ev[2], st);
    // 2^rounds >= n bounds every chain's length, so paths reach their fixed point and
    // every cycle's window covers the whole cycle.
    uint32_t rounds = 1;
    while ((1ull << rounds) < n) ++rounds;
    Jump* a = w.st0.as<Jump>();
    Jump* b = w.st1.as<Jump>();
    auto rank = [&]() -> hipError_t {
        k_jump_init<<<grid_for(n2), BLOCK, 0, st>>>(n2, nxt, prv, a);
        for (uint32_t r = 0; r < rounds; ++r) {
            k_jump<<<grid_for(n2), BLOCK, 0, st>>>(n2, a, b);
            std::swap(a, b);
        }
        return hipGetLastError();
    };
    G_TRY(rank());
    auto* counter = w.uid.as<unsigned long long>();  // scratch before the scans
    k_fill<unsigned long long><<<1, 64, 0, st>>>(counter, 1, 0ull);
    k_open_cycles<<<grid_for(n), BLOCK, 0, st>>>(n, a, nxt, prv, counter);
    uint64_t cycles = 0;
    G_TRY(hipMemcpyAsync(&cycles, counter, 8, hipMemcpyDeviceToHost, st));
    G_TRY(hipStreamSynchronize(st));
    if (cycles) G_TRY(rank());
    (void)hipEventRecord(ev[3], st);
    auto* rep = w.rep.as<unsigned long long>();
    k_represent<<<grid_for(n), BLOCK, 0, st>>>(n, a, nxt, prv, rep, w.len.as<uint32_t>());
    k_heads<<<grid_for(n), BLOCK, 0, st>>>(n, k, rep, w.len.as<uint32_t>(), w.is_head.as<uint32_t>(),
                                           w.bases.as<unsigned long long>());
    G_TRY(hipGetLastError());
    // exclusive scans: base offsets and unitig ids (one extra element gives the totals)
    size_t tb1 = 0, tb2 = 0, tb3 = 0;
    auto* boff = w.boff.as<unsigned long long>();
    uint32_t* uid = w.uid.as<uint32_t>();
    G_TRY(rocprim::exclusive_scan(nullptr, tb1, w.bases.as<unsigned long long>(), boff, 0ull, (size_t)n,
                                  rocprim::plus<unsigned long long>(), st));
    G_TRY(rocprim::exclusive_scan(nullptr, tb2, w.is_head.as<uint32_t>(), uid, 0u, (size_t)n, rocprim::plus<uint32_t>(),
                                  st));
    G_TRY(rocprim::inclusive_scan(nullptr, tb3, w.pos_cov.as<unsigned long long>(),
                                  w.cov_scan.as<unsigned long long>(), (size_t)n, rocprim::plus<unsigned long long>(),
                                  st));
    G_TRY(w.tmp.ensure(std::max(tb1, std::max(tb2, tb3))));
    tb1 = tb2 = tb3 = w.tmp.bytes;
    G_TRY(rocprim::exclusive_scan(w.tmp.p, tb1, w.bases.as<unsigned long long>(), boff, 0ull, (size_t)n,
                                  rocprim::plus<unsigned long long>(), st));
    G_TRY(rocprim::exclusive_scan(w.tmp.p, tb2, w.is_head.as<uint32_t>(), uid, 0u, (size_t)n,
                                  rocprim::plus<uint32_t>(), st));
    // totals: last prefix + last element
    uint64_t last_b = 0, last_bo = 0;
    uint32_t last_h = 0, last_u = 0;
    G_TRY(hipMemcpyAsync(&last_b, w.bases.as<unsigned long long>() + n - 1, 8, hipMemcpyDeviceToHost, st));
    G_TRY(hipMemcpyAsync(&last_bo, boff + n - 1, 8, hipMemcpyDeviceToHost, st));
    G_TRY(hipMemcpyAsync(&last_h, w.is_head.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, st));
    G_TRY(hipMemcpyAsync(&last_u, uid + n - 1, 4, hipMemcpyDeviceToHost, st));
    G_TRY(hipStreamSynchronize(st));
    const uint64_t n_bases = last_bo + last_b, n_unitigs = (uint64_t)last_u + last_h;
    if (n_bases != n + n_unitigs * (k - 1))
        return ctx_fail(ctx, KC_EHIP, "unitig chains do not cover the solid set");
    G_TRY(U->seq.ensure(n_bases));
    G_TRY(U->off.ensure((n_unitigs + 1) * 8));
    G_TRY(U->kmers.ensure(n_unitigs * 8));
    G_TRY(U->cov.ensure(n_unitigs * 8));
    k_emit<KW><<<grid_for(n), BLOCK, 0, st>>>(vals, w.counts.as<uint32_t>(), n, k, rep, boff, uid,
                                              U->seq.as<uint8_t>(), w.pos_cov.as<unsigned long long>());
    G_TRY(hipGetLastError());
    G_TRY(rocprim::inclusive_scan(w.tmp.p, tb3, w.pos_cov.as<unsigned long long>(),
                                  w.cov_scan.as<unsigned long long>(), (size_t)n, rocprim::plus<unsigned long long>(),
                                  st));
    k_unitig_meta<<<grid_for(n), BLOCK, 0, st>>>(n, k, w.is_head.as<uint32_t>(), uid, boff, w.len.as<uint32_t>(),
                                                 w.cov_scan.as<unsigned long long>(),
                                                 U->off.as<unsigned long long>(), U->kmers.as<unsigned long long>(),
                                                 U->cov.as<unsigned long long>());
    G_TRY(hipGetLastError());
    G_TRY(hipMemcpyAsync(U->off.as<uint64_t>() + n_unitigs, &n_bases, 8, hipMemcpyHostToDevice, st));
    G_TRY(hipStreamSynchronize(st));
    U->n_unitigs = n_unitigs;
    U->n_bases = n_bases;
    U->n_kmers = n;
    return KC_OK;
}

}  // namespace

extern "C" {

int kc_unitigs_build(kc_ctx* ctx, uint32_t lib, kc_unitigs** out) {
    if (!ctx || !out) return KC_EINVAL;
    *out = nullptr;
    SolidSet S{};
    int rc = solid_set(ctx, lib, &S);
    if (rc != KC_OK) return rc;
    if (S.k % 2 == 0) return ctx_fail(ctx, KC_EUNSUPPORTED, "unitigs need an odd k");
    if (S.n >= 0x7FFFFFFFull) return ctx_fail(ctx, KC_EUNSUPPORTED, "more than 2^31 - 2 solid k-mers");
    auto* U = new (std::nothrow) kc_unitigs();
    if (!U) return ctx_fail(ctx, KC_ENOMEM, "out of host memory");
    U->device = S.device;
    if (S.n == 0) {
        *out = U;
        return KC_OK;
    }
    hipError_t e = hipSetDevice(S.device);
    if (e != hipSuccess) {
        delete U;
        return ctx_fail(ctx, KC_EHIP, hip_msg("hipSetDevice", e));
    }
    hipEvent_t ev[5] = {};
    for (hipEvent_t& x : ev) (void)hipEventCreate(&x);
    (void)hipEventRecord(ev[0], S.stream);
    Work w(graph_scratch(ctx));
    const int kw = kw_of_k(S.k);
    rc = kw == 1   ? build<1>(ctx, lib, S, U, w, ev)
         : kw == 2 ? build<2>(ctx, lib, S, U, w, ev)
                   : build<4>(ctx, lib, S, U, w, ev);
    (void)hipEventRecord(ev[4], S.stream);
    (void)hipEventSynchronize(ev[4]);
    float ms = 0.f;
    if (rc == KC_OK) {
        (void)hipEventElapsedTime(&ms, ev[0], ev[4]);
        for (int i = 0; i < 4; ++i) {
            float t = 0.f;
            (void)hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            U->ms_stage[i] = t;
        }
    }
    for (hipEvent_t x : ev) (void)hipEventDestroy(x);
    if (rc != KC_OK) {
        delete U;
        return rc;
    }
    U->ms = ms;
    *out = U;
    return KC_OK;
}

int kc_unitigs_info_get(const kc_unitigs* u, kc_unitigs_info* info) {
    if (!u || !info) return KC_EINVAL;
    info->n_unitigs = u->n_unitigs;
    info->n_bases = u->n_bases;
    info->n_kmers = u->n_kmers;
    info->ms_device = u->ms;
    for (int i = 0; i < 4; ++i) info->ms_stage[i] = u->ms_stage[i];
    return KC_OK;
}

int kc_unitigs_export(kc_unitigs* u, uint8_t* seq, size_t n_bases, uint64_t* offsets, uint64_t* kmers,
                      double* mean_cov, size_t n_unitigs) {
    if (!u) return KC_EINVAL;
    if (seq && n_bases != u->n_bases) return u->fail(KC_EINVAL, "n_bases mismatch");
    if ((offsets || kmers || mean_cov) && n_unitigs != u->n_unitigs) return u->fail(KC_EINVAL, "n_unitigs mismatch");
    if (u->n_unitigs == 0) {
        if (offsets) offsets[0] = 0;
        return KC_OK;
    }
    hipError_t e = hipSetDevice(u->device);
    if (e != hipSuccess) return u->fail(KC_EHIP, hip_msg("hipSetDevice", e));
    const uint64_t m = u->n_unitigs;
    std::vector<uint64_t> km, cs;
    if (seq && (e = hipMemcpy(seq, u->seq.p, u->n_bases, hipMemcpyDeviceToHost)) != hipSuccess)
        return u->fail(KC_EHIP, hip_msg("hipMemcpy", e));
    if (offsets && (e = hipMemcpy(offsets, u->off.p, (m + 1) * 8, hipMemcpyDeviceToHost)) != hipSuccess)
        return u->fail(KC_EHIP, hip_msg("hipMemcpy", e));
    if (kmers || mean_cov) {
        km.resize(m);
        if ((e = hipMemcpy(km.data(), u->kmers.p, m * 8, hipMemcpyDeviceToHost)) != hipSuccess)
            return u->fail(KC_EHIP, hip_msg("hipMemcpy", e));
        if (kmers) std::memcpy(kmers, km.data(), m * 8);
    }
    if (mean_cov) {
        cs.resize(m);
        if ((e = hipMemcpy(cs.data(), u->cov.p, m * 8, hipMemcpyDeviceToHost)) != hipSuccess)
            return u->fail(KC_EHIP, hip_msg("hipMemcpy", e));
        // graph.rs:171-176: f64 sum of the counts over the k-mer count
        for (uint64_t i = 0; i < m; ++i) mean_cov[i] = (double)cs[i] / (double)km[i];
    }
    return KC_OK;
}

const char* kc_unitigs_last_error(const kc_unitigs* u) { return u ? u->err.c_str() : "null unitigs"; }

void kc_unitigs_destroy(kc_unitigs* u) {
    if (!u) return;
    (void)hipSetDevice(u->device);
    u->seq.release();
    u->off.release();
    u->kmers.release();
    u->cov.release();
    delete u;
}

}  // extern "C"
