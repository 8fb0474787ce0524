// kc_anchor.hip — unique-anchor k-mer indexes over contigs / unitigs on the device: the
// maps that Polisher::build_kmer_index (polisher.rs:33-60),
// PairedInfoIndex::build_kmer_index (paired_info.rs:34-70) and the splitter's anchor step
// (splitter.rs:57-70) build from ASCII sequences with the same k-mer extraction as the
// counting path.
//
// The reference walks the windows in order and removes a k-mer at its second sighting
// (any second sighting, or only one in another sequence); a removed k-mer is never
// re-added.  That is order-free: a k-mer is kept when it occurs once in all, or when all
// its occurrences lie in one sequence, and its value is its first occurrence.  So one
// pass over every window (one thread each) builds a device hash table of (canonical
// k-mer -> occurrences, first window, smallest and largest sequence), a scan keeps the
// anchors, a merge sort orders them by k-mer, and a last pass turns each first window back
// into (sequence, position, is_rc).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_merge_sort.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/kc.h"
#include "kc_host.h"

using namespace kc;

namespace {

thread_local std::string g_anchor_err;

struct Buf : DevBuf {
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
};

// A k-mer value (four words, lowest first), ordered as a 256-bit number.
struct VKey {
    uint64_t v[4];
};
struct VKeyLess {
    __device__ __host__ bool operator()(const VKey& a, const VKey& b) const {
        for (int w = 3; w >= 0; --w)
            if (a.v[w] != b.v[w]) return a.v[w] < b.v[w];
        return false;
    }
};

// Largest l with starts[l] <= x (starts ascending, starts[0] == 0).
__device__ __forceinline__ uint32_t region_of(const uint64_t* __restrict__ starts, uint32_t n, uint64_t x) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= x) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Kmer256::from_bytes (dna.rs:299-321) as a value: false when a base is not ACGT/acgt.
template <int KW>
__device__ __forceinline__ bool ascii_window(const uint8_t* __restrict__ s, uint32_t k, Kv& v) {
    v = Kv{0, 0, 0, 0};
    for (uint32_t i = 0; i < k; ++i) {
        const uint32_t c = s[i] | 0x20u;
        uint64_t code;
        if (c == 'a') code = 0;
        else if (c == 'c') code = 1;
        else if (c == 'g') code = 2;
        else if (c == 't') code = 3;
        else return false;
        if (KW == 4) {
            v.w3 = (v.w3 << 2) | (v.w2 >> 62);
            v.w2 = (v.w2 << 2) | (v.hi >> 62);
        }
        if (KW >= 2) v.hi = (v.hi << 2) | (v.lo >> 62);
        v.lo = (v.lo << 2) | code;
    }
    return true;
}

template <int NW>
__global__ void k_anchor_clear(Slot<NW, uint64_t>* __restrict__ tab, uint32_t* __restrict__ smin,
                               uint32_t* __restrict__ smax, uint64_t n_slots) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_slots;
         i += (uint64_t)gridDim.x * blockDim.x) {
        Slot<NW, uint64_t> s;
#pragma unroll
        for (int w = 0; w < NW; ++w) s.k[w] = EMPTY_KEY;
        s.t0 = ~0ull;
        s.cnt = 0;
        tab[i] = s;
        smin[i] = 0xFFFFFFFFu;
        smax[i] = 0;
    }
}

// Window w of sequence p = region_of(pw0, w) starts at byte pstart[p] + (w - pw0[p]).
template <int KW>
__global__ void k_anchor_insert(const uint8_t* __restrict__ seqs, const uint64_t* __restrict__ pstart,
                                const uint64_t* __restrict__ pw0, const uint32_t* __restrict__ pseq, uint32_t n_p,
                                uint64_t n_win, uint32_t k, Slot<nw_of(KW), uint64_t>* __restrict__ tab,
                                uint64_t nbuckets, uint32_t* __restrict__ smin, uint32_t* __restrict__ smax,
                                uint32_t* __restrict__ overflow) {
    constexpr int NW = nw_of(KW);
    constexpr uint32_t SPB = slots_per_bucket<NW, uint64_t>();
    for (uint64_t w = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; w < n_win;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = region_of(pw0, n_p, w);
        Kv v;
        if (!ascii_window<KW>(seqs + pstart[p] + (w - pw0[p]), k, v)) continue;
        uint64_t key[NW];
        pack_key<NW>(canonical<KW>(v, k), key);
        const uint64_t b = home_bucket<NW>(key, nbuckets);
        uint64_t w0[SPB];
        bucket_load<NW, uint64_t>(tab, b, w0);
        const uint64_t i = table_find_or_claim<NW, uint64_t>(tab, 0, nbuckets, b, w0, key, 1u << 20);
        if (i == ~0ull) {
            atomicOr(overflow, 1u);
            continue;
        }
        atomicAdd(&tab[i].cnt, 1u);
        atomicMin(reinterpret_cast<unsigned long long*>(&tab[i].t0), (unsigned long long)w);
        atomicMin(&smin[i], pseq[p]);
        atomicMax(&smax[i], pseq[p]);
    }
}

// Keeps the anchors: mode KC_ANCHOR_PAIRED_INFO, or POLISHER at k >= 11: occurs once;
// SPLITTER, or POLISHER at k < 11: one sequence only.
template <int NW>
__global__ void k_anchor_keep(const Slot<NW, uint64_t>* __restrict__ tab, const uint32_t* __restrict__ smin,
                              const uint32_t* __restrict__ smax, uint64_t n_slots, uint32_t once,
                              VKey* __restrict__ keys, unsigned long long* __restrict__ first,
                              unsigned long long* __restrict__ counter) {
    for (uint64_t i0 = blockIdx.x * (uint64_t)blockDim.x; i0 < n_slots; i0 += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = i0 + threadIdx.x;
        bool keep = false;
        Slot<NW, uint64_t> s{};
        if (i < n_slots) {
            s = tab[i];
            keep = s.k[0] != EMPTY_KEY && (once ? s.cnt == 1 : smin[i] == smax[i]);
        }
        const uint64_t o = wave_append(keep, counter);
        if (keep) {
            const Kv v = unpack_key<NW>(s.k);
            keys[o] = VKey{{v.lo, v.hi, v.w2, v.w3}};
            first[o] = s.t0;
        }
    }
}

// Sorted anchors -> Kmer256 limbs, sequence, position, is_rc of the first occurrence.
template <int KW>
__global__ void k_anchor_values(const VKey* __restrict__ keys, const unsigned long long* __restrict__ first,
                                uint64_t n, const uint8_t* __restrict__ seqs, const uint64_t* __restrict__ pstart,
                                const uint64_t* __restrict__ pw0, const uint32_t* __restrict__ pseq, uint32_t n_p,
                                uint32_t k, uint64_t* __restrict__ limbs, uint32_t* __restrict__ seq_out,
                                uint32_t* __restrict__ pos_out, uint8_t* __restrict__ rc_out) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t w = first[e];
        const uint32_t p = region_of(pw0, n_p, w);
        const uint64_t pos = w - pw0[p];
        Kv fwd;
        (void)ascii_window<KW>(seqs + pstart[p] + pos, k, fwd);  // an ACGT window: it was inserted
        const Kv rc = revcomp<KW>(fwd, k);
        const Kv c{keys[e].v[0], keys[e].v[1], keys[e].v[2], keys[e].v[3]};
        kmer_limbs<KW>(c, k, limbs + 4 * e);
        seq_out[e] = pseq[p];
        pos_out[e] = (uint32_t)pos;
        rc_out[e] = !kv_le(fwd, rc);  // Kmer256::canonical: the reverse complement only when smaller
    }
}

// Kmer256 limbs -> value: for k <= 64 the first u128, else (lo << 2(k-64)) | hi.
__device__ __forceinline__ Kv limbs_to_value(const uint64_t* w, uint32_t k) {
    if (k <= 64) return Kv{w[0], w[1], 0, 0};
    const uint32_t s = 2 * k - 128;  // bits of the second u128
    // lo128 << s, then OR the low s bits of hi128
    Kv r{0, 0, 0, 0};
    const uint64_t lo[2] = {w[0], w[1]};
    const uint32_t q = s >> 6, sh = s & 63;
    uint64_t o[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < 2; ++i) {
        if (i + q < 4) o[i + q] |= sh ? (lo[i] << sh) : lo[i];
        if (sh && i + q + 1 < 4) o[i + q + 1] |= lo[i] >> (64 - sh);
    }
    const uint64_t h0 = s >= 64 ? w[2] : (s ? (w[2] & ((1ull << s) - 1)) : 0ull);
    const uint64_t h1 = s > 64 ? (w[3] & ((1ull << (s - 64)) - 1)) : 0ull;
    r.lo = o[0] | h0;
    r.hi = o[1] | h1;
    r.w2 = o[2];
    r.w3 = o[3];
    return r;
}

// Lookup of canonical k-mers (limbs) among the sorted anchors: index or -1.
__global__ void k_anchor_query(const VKey* __restrict__ keys, uint64_t n, const uint64_t* __restrict__ q, uint64_t nq,
                               uint32_t k, int64_t* __restrict__ idx) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < nq; t += (uint64_t)gridDim.x * blockDim.x) {
        const Kv x = limbs_to_value(q + 4 * t, k);
        const VKey y{{x.lo, x.hi, x.w2, x.w3}};
        uint64_t lo = 0, hi = n;  // first key >= y
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (VKeyLess()(keys[mid], y)) lo = mid + 1;
            else hi = mid;
        }
        const bool hit = lo < n && !VKeyLess()(y, keys[lo]);
        idx[t] = hit ? (int64_t)lo : -1;
    }
}

#define A_TRY(expr)                                                                     \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess) {                                                         \
            g_anchor_err = std::string(#expr) + ": " + hipGetErrorString(_e);           \
            return _e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP;                     \
        }                                                                               \
    } while (0)

int afail(int code, const std::string& m) {
    g_anchor_err = m;
    return code;
}

}  // namespace

struct kc_anchor {
    int32_t device = 0;
    uint32_t k = 0;
    int32_t kind = 0;
    hipStream_t st = nullptr;
    uint64_t n = 0;
    Buf seqs, pstart, pw0, pseq, keys, first, limbs, seq, pos, rc, qbuf, qidx;
    uint32_t n_p = 0;
};

namespace {

template <int KW>
int anchor_build(kc_anchor* a, uint64_t n_win) {
    constexpr int NW = nw_of(KW);
    constexpr uint64_t SPB = slots_per_bucket<NW, uint64_t>();
    hipStream_t st = a->st;
    const uint64_t slots = std::max<uint64_t>(2 * n_win, 64);
    const uint64_t nbuckets = (slots + SPB - 1) / SPB;
    const uint64_t n_slots = nbuckets * SPB;
    Buf tab, smin, smax, ctr, keys2, first2, tmp;
    A_TRY(tab.ensure(nbuckets * 64));
    A_TRY(smin.ensure(n_slots * 4));
    A_TRY(smax.ensure(n_slots * 4));
    A_TRY(ctr.ensure(64));
    A_TRY(hipMemsetAsync(ctr.p, 0, 64, st));
    k_anchor_clear<NW><<<grid_for(n_slots), BLOCK, 0, st>>>(tab.as<Slot<NW, uint64_t>>(), smin.as<uint32_t>(),
                                                            smax.as<uint32_t>(), n_slots);
    A_TRY(hipGetLastError());
    if (n_win) {
        k_anchor_insert<KW><<<grid_for(n_win), BLOCK, 0, st>>>(
            a->seqs.as<uint8_t>(), a->pstart.as<uint64_t>(), a->pw0.as<uint64_t>(), a->pseq.as<uint32_t>(), a->n_p,
            n_win, a->k, tab.as<Slot<NW, uint64_t>>(), nbuckets, smin.as<uint32_t>(), smax.as<uint32_t>(),
            ctr.as<uint32_t>());
        A_TRY(hipGetLastError());
    }
    // anchors of the table: at most one per slot
    A_TRY(keys2.ensure(n_slots * sizeof(VKey)));
    A_TRY(first2.ensure(n_slots * 8));
    const uint32_t once = a->kind == KC_ANCHOR_PAIRED_INFO || (a->kind == KC_ANCHOR_POLISHER && a->k >= 11);
    k_anchor_keep<NW><<<grid_for(n_slots), BLOCK, 0, st>>>(tab.as<Slot<NW, uint64_t>>(), smin.as<uint32_t>(),
                                                           smax.as<uint32_t>(), n_slots, once, keys2.as<VKey>(),
                                                           first2.as<unsigned long long>(),
                                                           ctr.as<unsigned long long>() + 1);
    A_TRY(hipGetLastError());
    unsigned long long h[2] = {0, 0};
    A_TRY(hipMemcpyAsync(h, ctr.p, 16, hipMemcpyDeviceToHost, st));
    A_TRY(hipStreamSynchronize(st));
    if (h[0] & 0xFFFFFFFFull) return afail(KC_ENOMEM, "anchor table overflow");
    const uint64_t n = h[1];
    a->n = n;
    A_TRY(a->keys.ensure(n * sizeof(VKey) + sizeof(VKey)));
    A_TRY(a->first.ensure(n * 8 + 8));
    A_TRY(a->limbs.ensure(n * 32 + 32));
    A_TRY(a->seq.ensure(n * 4 + 4));
    A_TRY(a->pos.ensure(n * 4 + 4));
    A_TRY(a->rc.ensure(n + 1));
    if (n) {
        size_t tb = 0;
        A_TRY(rocprim::merge_sort(nullptr, tb, keys2.as<VKey>(), a->keys.as<VKey>(), first2.as<unsigned long long>(),
                                  a->first.as<unsigned long long>(), (size_t)n, VKeyLess(), st));
        A_TRY(tmp.ensure(tb));
        A_TRY(rocprim::merge_sort(tmp.p, tb, keys2.as<VKey>(), a->keys.as<VKey>(), first2.as<unsigned long long>(),
                                  a->first.as<unsigned long long>(), (size_t)n, VKeyLess(), st));
        k_anchor_values<KW><<<grid_for(n), BLOCK, 0, st>>>(
            a->keys.as<VKey>(), a->first.as<unsigned long long>(), n, a->seqs.as<uint8_t>(), a->pstart.as<uint64_t>(),
            a->pw0.as<uint64_t>(), a->pseq.as<uint32_t>(), a->n_p, a->k, a->limbs.as<uint64_t>(), a->seq.as<uint32_t>(),
            a->pos.as<uint32_t>(), a->rc.as<uint8_t>());
        A_TRY(hipGetLastError());
    }
    A_TRY(hipStreamSynchronize(st));
    return KC_OK;
}

}  // namespace

extern "C" {

int kc_anchor_build(int32_t device, uint32_t k, int32_t kind, const uint8_t* seqs, size_t seqs_len,
                    const uint64_t* starts, const uint64_t* lengths, size_t n_seqs, kc_anchor** out) {
    if (!out) return afail(KC_EINVAL, "null out");
    *out = nullptr;
    if (k < 1 || k > 128) return afail(KC_EINVAL, "k outside 1..128");
    if (k > KC_MAX_K) return afail(KC_EUNSUPPORTED, "k = 128");
    if (kind != KC_ANCHOR_POLISHER && kind != KC_ANCHOR_PAIRED_INFO && kind != KC_ANCHOR_SPLITTER)
        return afail(KC_EINVAL, "unknown anchor index kind");
    if (n_seqs && (!starts || !lengths)) return afail(KC_EINVAL, "null starts/lengths");
    if (seqs_len && !seqs) return afail(KC_EINVAL, "null seqs");
    if (n_seqs > 0xFFFFFFFFull) return afail(KC_EUNSUPPORTED, "more than 2^32-1 sequences");
    // sequences with at least one window, and their first global window
    std::vector<uint64_t> ps, pw;
    std::vector<uint32_t> pq;
    uint64_t n_win = 0;
    for (size_t i = 0; i < n_seqs; ++i) {
        if (starts[i] > seqs_len || lengths[i] > seqs_len - starts[i])
            return afail(KC_EINVAL, "sequence " + std::to_string(i) + " extends past seqs_len");
        if (lengths[i] < k) continue;
        if (lengths[i] - k + 1 > 0xFFFFFFFFull) return afail(KC_EUNSUPPORTED, "a sequence longer than 2^32 bases");
        ps.push_back(starts[i]);
        pw.push_back(n_win);
        pq.push_back((uint32_t)i);
        n_win += lengths[i] - k + 1;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return afail(KC_EHIP, "no GPU");
    if (device < 0 || device >= ndev) return afail(KC_EINVAL, "bad device");
    A_TRY(hipSetDevice(device));
    auto* a = new kc_anchor();
    a->device = device;
    a->k = k;
    a->kind = kind;
    a->n_p = (uint32_t)pq.size();
    auto bail = [&](int rc) {
        kc_anchor_destroy(a);
        return rc;
    };
    if (hipStreamCreateWithFlags(&a->st, hipStreamNonBlocking) != hipSuccess) return bail(afail(KC_EHIP, "stream"));
    int rc = KC_OK;
    hipError_t e = a->seqs.ensure(seqs_len + 64);
    if (e == hipSuccess) e = a->pstart.ensure(ps.size() * 8 + 8);
    if (e == hipSuccess) e = a->pw0.ensure(pw.size() * 8 + 8);
    if (e == hipSuccess) e = a->pseq.ensure(pq.size() * 4 + 4);
    if (e == hipSuccess && seqs_len) e = hipMemcpyAsync(a->seqs.p, seqs, seqs_len, hipMemcpyHostToDevice, a->st);
    if (e == hipSuccess && !ps.empty()) {
        e = hipMemcpyAsync(a->pstart.p, ps.data(), ps.size() * 8, hipMemcpyHostToDevice, a->st);
        if (e == hipSuccess) e = hipMemcpyAsync(a->pw0.p, pw.data(), pw.size() * 8, hipMemcpyHostToDevice, a->st);
        if (e == hipSuccess) e = hipMemcpyAsync(a->pseq.p, pq.data(), pq.size() * 4, hipMemcpyHostToDevice, a->st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(a->st);  // the host vectors go out of scope
    if (e != hipSuccess)
        return bail(afail(e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP, hip_msg("kc_anchor_build upload", e)));
    const int kw = kw_of_k(k);
    rc = kw == 1 ? anchor_build<1>(a, n_win) : kw == 2 ? anchor_build<2>(a, n_win) : anchor_build<4>(a, n_win);
    if (rc) return bail(rc);
    *out = a;
    return KC_OK;
}

int kc_anchor_count(const kc_anchor* a, uint64_t* n) {
    if (!a || !n) return afail(KC_EINVAL, "null argument");
    *n = a->n;
    return KC_OK;
}

int kc_anchor_export(kc_anchor* a, uint64_t* keys4, uint32_t* seq_ids, uint32_t* positions, uint8_t* is_rc,
                     size_t n) {
    if (!a) return afail(KC_EINVAL, "null index");
    if (n != a->n) return afail(KC_EINVAL, "n mismatch");
    if (!n) return KC_OK;
    A_TRY(hipSetDevice(a->device));
    if (keys4) A_TRY(hipMemcpyAsync(keys4, a->limbs.p, n * 32, hipMemcpyDeviceToHost, a->st));
    if (seq_ids) A_TRY(hipMemcpyAsync(seq_ids, a->seq.p, n * 4, hipMemcpyDeviceToHost, a->st));
    if (positions) A_TRY(hipMemcpyAsync(positions, a->pos.p, n * 4, hipMemcpyDeviceToHost, a->st));
    if (is_rc) A_TRY(hipMemcpyAsync(is_rc, a->rc.p, n, hipMemcpyDeviceToHost, a->st));
    A_TRY(hipStreamSynchronize(a->st));
    return KC_OK;
}

int kc_anchor_query(kc_anchor* a, const uint64_t* keys4, size_t n, int64_t* index) {
    if (!a) return afail(KC_EINVAL, "null index");
    if (n && (!keys4 || !index)) return afail(KC_EINVAL, "null keys/index");
    if (!n) return KC_OK;
    A_TRY(hipSetDevice(a->device));
    A_TRY(a->qbuf.ensure(n * 32));
    A_TRY(a->qidx.ensure(n * 8));
    A_TRY(hipMemcpyAsync(a->qbuf.p, keys4, n * 32, hipMemcpyHostToDevice, a->st));
    k_anchor_query<<<grid_for(n), BLOCK, 0, a->st>>>(a->keys.as<VKey>(), a->n, a->qbuf.as<uint64_t>(), n, a->k,
                                                     a->qidx.as<int64_t>());
    A_TRY(hipGetLastError());
    A_TRY(hipMemcpyAsync(index, a->qidx.p, n * 8, hipMemcpyDeviceToHost, a->st));
    A_TRY(hipStreamSynchronize(a->st));
    return KC_OK;
}

const char* kc_anchor_last_error(void) { return g_anchor_err.c_str(); }

void kc_anchor_destroy(kc_anchor* a) {
    if (!a) return;
    (void)hipSetDevice(a->device);
    if (a->st) {
        (void)hipStreamSynchronize(a->st);
        (void)hipStreamDestroy(a->st);
    }
    delete a;
}

}  // extern "C"
