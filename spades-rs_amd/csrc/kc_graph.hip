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
//   4. rank     list ranking of the 2n oriented k-mers: sampled rulers walk their
//               sublists, pointer jumping ranks the rulers; each k-mer gets its chain's
//               head and its distance from it.  A cycle is opened before its smallest
//               k-mer in forward orientation (where the reference's pass 2 starts when
//               it iterates in ascending order, graph.rs:216-230)
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
// backward) as one word per k-mer: rec[x] = nxt | prv << 32, oriented ids or NONE.  The
// reverse strand's links are the twins: nxt(2x+1) = twin(prv), prv(2x+1) = twin(nxt).
__global__ void k_links(uint64_t n, const uint8_t* __restrict__ deg, const uint32_t* __restrict__ succ,
                        const uint32_t* __restrict__ pred, unsigned long long* __restrict__ rec) {
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
        rec[x] = (unsigned long long)a | ((unsigned long long)b << 32);
    }
}

__device__ __forceinline__ uint32_t rec_nxt(unsigned long long r, uint32_t u) {
    if (!(u & 1)) return (uint32_t)r;
    const uint32_t p = (uint32_t)(r >> 32);
    return p == NONE ? NONE : p ^ 1;
}
__device__ __forceinline__ uint32_t rec_prv(unsigned long long r, uint32_t u) {
    if (!(u & 1)) return (uint32_t)(r >> 32);
    const uint32_t t = (uint32_t)r;
    return t == NONE ? NONE : t ^ 1;
}

// ---- list ranking ------------------------------------------------------------------
// Rulers: every chain head (no incoming link) and both strands of a sampled 1/2^s of the
// k-mers (Fibonacci hash of the rank).  Each ruler walks its sublist (up to the next
// ruler) once, labelling every k-mer with (ruler, offset): O(n) dependent reads in all,
// about 2^s per walk.  Pointer jumping then ranks the short list of rulers.  A cycle with
// rulers is opened by arithmetic modulo its length; a cycle without any ruler (rare at
// 1/16 sampling unless short) is walked by each of its own k-mers.
__device__ __forceinline__ bool is_ruler(unsigned long long r, uint32_t u, uint32_t shift) {
    return rec_prv(r, u) == NONE || (((uint64_t)(u >> 1) * 0x9E3779B97F4A7C15ull) >> shift) == 0;
}

__global__ void k_ruler_flags(uint64_t n2, const unsigned long long* __restrict__ rec, uint32_t shift,
                              uint32_t* __restrict__ flag) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n2; u += (uint64_t)gridDim.x * blockDim.x)
        flag[u] = is_ruler(rec[u >> 1], (uint32_t)u, shift) ? 1u : 0u;
}

__global__ void k_ruler_list(uint64_t n2, const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rid,
                             uint32_t* __restrict__ rnode) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n2; u += (uint64_t)gridDim.x * blockDim.x)
        if (flag[u]) rnode[rid[u]] = (uint32_t)u;
}

// own[u] = ruler index | offset << 32 for every k-mer a walk reaches (~0 otherwise);
// rlen = k-mers in the sublist, rmin = its smallest rank, rnext = the next ruler or NONE.
__global__ void k_walk(uint64_t nr, const uint32_t* __restrict__ rnode, const unsigned long long* __restrict__ rec,
                       const uint32_t* __restrict__ rid, uint32_t shift, unsigned long long* __restrict__ own,
                       uint32_t* __restrict__ rlen, uint32_t* __restrict__ rmin, uint32_t* __restrict__ rnext) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nr; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t cur = rnode[i], mn = cur >> 1, off = 0, rn = NONE;
        own[cur] = i;
        unsigned long long r = rec[cur >> 1];
        for (;;) {
            const uint32_t nx = rec_nxt(r, cur);
            if (nx == NONE) break;
            const unsigned long long q = rec[nx >> 1];
            if (is_ruler(q, nx, shift)) {
                rn = rid[nx];
                break;
            }
            ++off;
            own[nx] = i | ((unsigned long long)off << 32);
            mn = min(mn, nx >> 1);
            cur = nx;
            r = q;
        }
        rlen[i] = off + 1;
        rmin[i] = mn;
        rnext[i] = rn;
    }
}

__global__ void k_ruler_prev(uint64_t nr, const uint32_t* __restrict__ rnext, uint32_t* __restrict__ rprv) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nr; i += (uint64_t)gridDim.x * blockDim.x)
        if (rnext[i] != NONE) rprv[rnext[i]] = (uint32_t)i;
}

// Pointer-jumping state of a ruler: the farthest predecessor ruler reached (itself at a
// chain's head), the k-mers from its sublist's start to this one's, and the smallest rank
// over the sublists covered.
struct Jump {
    uint32_t j, d, m, pad;
};

__global__ void k_jump_init(uint64_t nr, const uint32_t* __restrict__ rprv, const uint32_t* __restrict__ rlen,
                            const uint32_t* __restrict__ rmin, Jump* __restrict__ st) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nr; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = rprv[i];
        st[i] = Jump{p == NONE ? (uint32_t)i : p, p == NONE ? 0u : rlen[p], rmin[i], 0u};
    }
}

__global__ void k_jump(uint64_t nr, const Jump* __restrict__ in, Jump* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nr; i += (uint64_t)gridDim.x * blockDim.x) {
        const Jump a = in[i];
        const Jump b = in[a.j];
        out[i] = Jump{b.j, a.d + b.d, a.m < b.m ? a.m : b.m, 0u};
    }
}

// After enough rounds a ruler on a path points at its chain's head, which has no
// predecessor ruler; on a cycle it does not, and m is the cycle's smallest rank.  The
// cycle of oriented k-mers through (m, fwd) opens before (m, fwd), where the reference's
// walk starts it (graph.rs:216-230 in ascending order); its twin cycle opens before
// twin(p) for the link p -> (m, fwd), i.e. after (m, rc).  The ruler owning that k-mer
// becomes the head of its ruler list: cut = k-mer | its offset << 32, tail = the ruler
// before it.  A thread reads rprv only at its own pointer target and writes only its own
// entry, and the one ruler of a cycle that cuts is the only writer there.
__global__ void k_cut_cycles(uint64_t nr, const Jump* __restrict__ st, uint32_t* __restrict__ rprv,
                             const unsigned long long* __restrict__ own, const unsigned long long* __restrict__ rec,
                             unsigned long long* __restrict__ cut, uint32_t* __restrict__ tail,
                             unsigned long long* __restrict__ counter) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nr; i += (uint64_t)gridDim.x * blockDim.x) {
        const Jump a = st[i];
        if (rprv[a.j] == NONE) continue;
        const uint32_t m = a.m;
        uint32_t c = 2 * m;
        if ((uint32_t)own[c] != (uint32_t)i) {
            const uint32_t p = (uint32_t)(rec[m] >> 32);  // prv of (m, fwd): on a cycle, never NONE
            c = p == NONE ? NONE : (p ^ 1);
            if (c == NONE || (uint32_t)own[c] != (uint32_t)i) continue;
        }
        tail[i] = rprv[i];
        rprv[i] = NONE;
        cut[i] = (unsigned long long)c | ((own[c] >> 32) << 32);
        atomicAdd(counter, 1ull);
    }
}

// A k-mer that no walk reached lies on a cycle without rulers: it walks the cycle for its
// length, smallest rank and that k-mer's strand, then to the cycle's opening k-mer
// (chosen as above).  small[u] = head | position << 32.  Both walks are bounded by 2n
// steps and a walk off a chain end marks the k-mer invalid (~0), which cannot happen for
// k-mers on a cycle.
__global__ void k_small_cycles(uint64_t n2, const unsigned long long* __restrict__ own,
                               const unsigned long long* __restrict__ rec, unsigned long long* __restrict__ small) {
    for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n2; u += (uint64_t)gridDim.x * blockDim.x) {
        if (own[u] != ~0ull) continue;
        uint32_t cur = (uint32_t)u, mn = cur >> 1, mo = cur & 1;
        uint64_t len = 0;
        bool ok = true;
        do {
            cur = rec_nxt(rec[cur >> 1], cur);
            if (cur == NONE || ++len > n2) {
                ok = false;
                break;
            }
            if ((cur >> 1) < mn) {
                mn = cur >> 1;
                mo = cur & 1;
            }
        } while (cur != (uint32_t)u);
        if (!ok) {
            small[u] = ~0ull;
            continue;
        }
        const uint32_t c = mo == 0 ? 2 * mn : (((uint32_t)(rec[mn] >> 32)) ^ 1);
        uint64_t s = 0;
        for (cur = (uint32_t)u; cur != c && s <= len; ++s) cur = rec_nxt(rec[cur >> 1], cur);
        small[u] = s > len ? ~0ull : ((unsigned long long)c | ((unsigned long long)((len - s) % len) << 32));
    }
}

// Head and position of an oriented k-mer from its ruler's ranking.
__device__ __forceinline__ unsigned long long head_pos(uint32_t u, const unsigned long long* __restrict__ own,
                                                       const Jump* __restrict__ st,
                                                       const uint32_t* __restrict__ rnode,
                                                       const unsigned long long* __restrict__ cut,
                                                       const uint32_t* __restrict__ tail,
                                                       const uint32_t* __restrict__ rlen,
                                                       const unsigned long long* __restrict__ small) {
    const unsigned long long o = own[u];
    if (o == ~0ull) return small[u];
    const uint32_t i = (uint32_t)o, off = (uint32_t)(o >> 32);
    const Jump a = st[i];
    const unsigned long long cu = cut[a.j];
    const uint64_t P = (uint64_t)a.d + off;
    if (cu == ~0ull) return (unsigned long long)rnode[a.j] | (P << 32);
    const uint32_t t = tail[a.j];
    const uint64_t C = (uint64_t)st[t].d + rlen[t], co = cu >> 32;
    const uint64_t pos = P >= co ? P - co : P + C - co;
    return (cu & 0xFFFFFFFFull) | (pos << 32);
}

// One strand per chain: the one whose head has the smaller oriented id.  rep[x] = head
// rank | position << 32 | strand << 63.  A chain's two strands put x at positions p and
// L - 1 - p, so the head (position 0) writes the length L = p0 + p1 + 1.  bad[0] is set
// when a k-mer has no valid position.
__global__ void k_represent(uint64_t n, const unsigned long long* __restrict__ own, const Jump* __restrict__ st,
                            const uint32_t* __restrict__ rnode, const unsigned long long* __restrict__ cut,
                            const uint32_t* __restrict__ tail, const uint32_t* __restrict__ rlen,
                            const unsigned long long* __restrict__ small, unsigned long long* __restrict__ rep,
                            uint32_t* __restrict__ len, unsigned long long* __restrict__ bad) {
    for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < n; x += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long h0 = head_pos(2 * (uint32_t)x, own, st, rnode, cut, tail, rlen, small);
        const unsigned long long h1 = head_pos(2 * (uint32_t)x + 1, own, st, rnode, cut, tail, rlen, small);
        if (h0 == ~0ull || h1 == ~0ull) {
            atomicAdd(bad, 1ull);
            rep[x] = 0;
            continue;
        }
        const uint32_t o = (uint32_t)h1 < (uint32_t)h0 ? 1u : 0u;
        const unsigned long long h = o ? h1 : h0;
        const uint32_t pos = (uint32_t)(h >> 32);
        rep[x] = ((h & 0xFFFFFFFFull) >> 1) | ((unsigned long long)pos << 32) | ((unsigned long long)o << 63);
        if (pos == 0) len[x] = (uint32_t)(h0 >> 32) + (uint32_t)(h1 >> 32) + 1;
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
    DevBuf &vals, &counts, &idx, &tmp, &slots, &deg, &succ, &pred, &rec, &flag, &rid, &own, &small, &rnode, &rlen,
        &rmin, &rnext, &rprv, &cut, &tail, &st0, &st1, &rep, &len, &is_head, &bases, &boff, &uid, &pos_cov, &cov_scan;
    explicit Work(DevBuf* b)
        : vals(b[0]), counts(b[1]), idx(b[2]), tmp(b[3]), slots(b[4]), deg(b[5]), succ(b[6]), pred(b[7]), rec(b[8]),
          flag(b[9]), rid(b[10]), own(b[11]), small(b[12]), rnode(b[13]), rlen(b[14]), rmin(b[15]), rnext(b[16]),
          rprv(b[17]), cut(b[18]), tail(b[19]), st0(b[20]), st1(b[21]), rep(b[22]), len(b[23]), is_head(b[24]),
          bases(b[25]), boff(b[26]), uid(b[27]), pos_cov(b[28]), cov_scan(b[29]) {}
};
static_assert(GRAPH_SCRATCH >= 30, "scratch buffers");

// Sampled rulers: both strands of 1 in 2^RULER_LOG2 k-mers (walks of ~16 steps).
constexpr uint32_t RULER_LOG2 = 4;

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
    G_TRY(w.rec.ensure(n * 8));
    G_TRY(w.flag.ensure(n2 * 4));
    G_TRY(w.rid.ensure(n2 * 4));
    G_TRY(w.own.ensure(n2 * 8));
    G_TRY(w.small.ensure(n2 * 8));
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
    auto* rec = w.rec.as<unsigned long long>();
    k_links<<<grid_for(n), BLOCK, 0, st>>>(n, w.deg.as<uint8_t>(), w.succ.as<uint32_t>(), w.pred.as<uint32_t>(), rec);
    G_TRY(hipGetLastError());
    (void)hipEventRecord(ev[2], st);
    // rulers, their sublists, and the ranking of the ruler list
    const uint32_t shift = 64 - RULER_LOG2;
    uint32_t* flag = w.flag.as<uint32_t>();
    uint32_t* rid = w.rid.as<uint32_t>();
    k_ruler_flags<<<grid_for(n2), BLOCK, 0, st>>>(n2, rec, shift, flag);
    size_t tb = 0;
    G_TRY(rocprim::exclusive_scan(nullptr, tb, flag, rid, 0u, (size_t)n2, rocprim::plus<uint32_t>(), st));
    G_TRY(w.tmp.ensure(tb));
    tb = w.tmp.bytes;
    G_TRY(rocprim::exclusive_scan(w.tmp.p, tb, flag, rid, 0u, (size_t)n2, rocprim::plus<uint32_t>(), st));
    uint32_t last_f = 0, last_r = 0;
    G_TRY(hipMemcpyAsync(&last_f, flag + n2 - 1, 4, hipMemcpyDeviceToHost, st));
    G_TRY(hipMemcpyAsync(&last_r, rid + n2 - 1, 4, hipMemcpyDeviceToHost, st));
    G_TRY(hipStreamSynchronize(st));
    const uint64_t nr = (uint64_t)last_f + last_r;
    G_TRY(w.rnode.ensure(nr * 4));
    G_TRY(w.rlen.ensure(nr * 4));
    G_TRY(w.rmin.ensure(nr * 4));
    G_TRY(w.rnext.ensure(nr * 4));
    G_TRY(w.rprv.ensure(nr * 4));
    G_TRY(w.cut.ensure(nr * 8));
    G_TRY(w.tail.ensure(nr * 4));
    G_TRY(w.st0.ensure(nr * sizeof(Jump)));
    G_TRY(w.st1.ensure(nr * sizeof(Jump)));
    auto* own = w.own.as<unsigned long long>();
    uint32_t* rnode = w.rnode.as<uint32_t>();
    uint32_t* rlen = w.rlen.as<uint32_t>();
    uint32_t* rprv = w.rprv.as<uint32_t>();
    auto* cut = w.cut.as<unsigned long long>();
    auto* counters = w.idx.as<unsigned long long>();  // [0] cycles cut, [1] k-mers without a position
    k_fill<unsigned long long><<<grid_for(n2), BLOCK, 0, st>>>(own, n2, ~0ull);
    k_fill<unsigned long long><<<1, 64, 0, st>>>(counters, 2, 0ull);
    Jump* a = w.st0.as<Jump>();
    Jump* b = w.st1.as<Jump>();
    if (nr) {
        k_ruler_list<<<grid_for(n2), BLOCK, 0, st>>>(n2, flag, rid, rnode);
        k_walk<<<grid_for(nr), BLOCK, 0, st>>>(nr, rnode, rec, rid, shift, own, rlen, w.rmin.as<uint32_t>(),
                                               w.rnext.as<uint32_t>());
        k_fill<uint32_t><<<grid_for(nr), BLOCK, 0, st>>>(rprv, nr, NONE);
        k_fill<unsigned long long><<<grid_for(nr), BLOCK, 0, st>>>(cut, nr, ~0ull);
        k_ruler_prev<<<grid_for(nr), BLOCK, 0, st>>>(nr, w.rnext.as<uint32_t>(), rprv);
        // 2^rounds >= nr bounds every ruler list, so paths reach their head and every
        // cycle's window covers the whole cycle
        uint32_t rounds = 0;
        while ((1ull << rounds) < nr) ++rounds;
        auto rank = [&]() -> hipError_t {
            k_jump_init<<<grid_for(nr), BLOCK, 0, st>>>(nr, rprv, rlen, w.rmin.as<uint32_t>(), a);
            for (uint32_t r = 0; r < rounds; ++r) {
                k_jump<<<grid_for(nr), BLOCK, 0, st>>>(nr, a, b);
                std::swap(a, b);
            }
            return hipGetLastError();
        };
        G_TRY(rank());
        k_cut_cycles<<<grid_for(nr), BLOCK, 0, st>>>(nr, a, rprv, own, rec, cut, w.tail.as<uint32_t>(), counters);
        uint64_t cycles = 0;
        G_TRY(hipMemcpyAsync(&cycles, counters, 8, hipMemcpyDeviceToHost, st));
        G_TRY(hipStreamSynchronize(st));
        if (cycles) G_TRY(rank());
    }
    k_small_cycles<<<grid_for(n2), BLOCK, 0, st>>>(n2, own, rec, w.small.as<unsigned long long>());
    auto* rep = w.rep.as<unsigned long long>();
    k_represent<<<grid_for(n), BLOCK, 0, st>>>(n, own, a, rnode, cut, w.tail.as<uint32_t>(), rlen,
                                               w.small.as<unsigned long long>(), rep, w.len.as<uint32_t>(),
                                               counters + 1);
    G_TRY(hipGetLastError());
    (void)hipEventRecord(ev[3], st);
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
    uint64_t unplaced = 0;
    G_TRY(hipMemcpyAsync(&unplaced, counters + 1, 8, hipMemcpyDeviceToHost, st));
    G_TRY(hipStreamSynchronize(st));
    if (unplaced) return ctx_fail(ctx, KC_EHIP, "k-mers left without a unitig position");
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
