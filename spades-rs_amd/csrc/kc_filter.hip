// kc_filter.hip — the standalone two-tier filter (bloom.rs:13-131) and Hammer over it
// (hammer.rs:20-144) behind include/kc.h.
//
// A batch of inserts keeps the reference's sequential semantics (bloom.rs:83-107): with
// fs[p] = min(4t + j) over the batch's (event t, hash j) pairs, op j of event t found its
// bit already set iff the bit was set before the batch or fs[p] < 4t + j.

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/kc.h"
#include "kc_device.h"
#include "kc_hammer.h"
#include "kc_host.h"

using namespace kc;

namespace {

// ---------------------------------------------------------------------------
// standalone filter: sequential insert semantics over a batch of events
// ---------------------------------------------------------------------------

__device__ __forceinline__ void key_positions(const uint64_t* __restrict__ keys4, uint64_t t, const Modulus& M,
                                              uint64_t p[3]) {
    const uint64_t* w = keys4 + 4 * t;
    bloom_positions(w[0], w[1], w[2], w[3], M, p);
}

__global__ void k_ev_first_setter(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, uint32_t* __restrict__ fs) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
#pragma unroll
        for (int j = 0; j < 3; ++j) atomicMin(&fs[p[j]], (uint32_t)(4 * t + j));
    }
}

// prev_j(t) = bit set before this batch, or first set by an earlier (event, hash) of it.
__global__ void k_ev_promote(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint32_t* __restrict__ fs,
                             const uint64_t* __restrict__ seen, uint8_t* __restrict__ prom) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
        bool all = true;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const bool before = (seen[p[j] >> 6] >> (p[j] & 63)) & 1;
            all = all && (before || fs[p[j]] < (uint32_t)(4 * t + j));
        }
        prom[t] = all;
    }
}

__global__ void k_ev_apply(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint8_t* __restrict__ prom,
                           uint64_t* __restrict__ seen, uint64_t* __restrict__ solid, uint32_t* __restrict__ fs) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            atomicOr(reinterpret_cast<unsigned long long*>(&seen[p[j] >> 6]), 1ull << (p[j] & 63));
            if (prom[t]) atomicOr(reinterpret_cast<unsigned long long*>(&solid[p[j] >> 6]), 1ull << (p[j] & 63));
            fs[p[j]] = ~0u;  // restore for the next batch
        }
    }
}

__global__ void k_query_kernel(const uint64_t* __restrict__ keys4, uint64_t n, Modulus M, const uint64_t* __restrict__ solid,
                        uint8_t* __restrict__ out) {
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p[3];
        key_positions(keys4, t, M, p);
        bool all = true;
#pragma unroll
        for (int j = 0; j < 3; ++j) all = all && ((solid[p[j] >> 6] >> (p[j] & 63)) & 1);
        out[t] = all;
    }
}

// ---------------------------------------------------------------------------
// Allocation-free Hammer (kc_filter_correct_reads_noalloc): its scratch is one static
// device array per GPU, shared by every call on that GPU (calls that overlap on one GPU
// race on it).  Layout: counters (corrected reads, is_solid tests, active reads, a bounds
// error flag), the solid-tier descriptor, then the active-read list.
// ---------------------------------------------------------------------------

constexpr size_t NA_COUNTERS = 0, NA_BLOOM = 256, NA_ACTIVE = 512;
constexpr size_t NA_SCRATCH_BYTES = NA_ACTIVE + (size_t)KC_NOALLOC_MAX_READS * 4;
static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

// Counters cleared and the descriptor written (a kernel, so no memset node or host copy).
__global__ void k_na_init(uint8_t* __restrict__ sc, LibBloom lb) {
    if (threadIdx.x < 4) reinterpret_cast<unsigned long long*>(sc + NA_COUNTERS)[threadIdx.x] = 0;
    if (threadIdx.x == 0) *reinterpret_cast<LibBloom*>(sc + NA_BLOOM) = lb;
}

__global__ void k_na_copy(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Reads of at least k bases, in any order (each read is corrected on its own), appended
// one wave at a time; a read past data_len sets the error flag and is left out.
__global__ void k_na_active(const uint64_t* __restrict__ offsets, const uint32_t* __restrict__ lengths, uint64_t n,
                            uint64_t data_len, uint32_t k, uint8_t* __restrict__ sc) {
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(sc + NA_COUNTERS);
    uint32_t* act = reinterpret_cast<uint32_t*>(sc + NA_ACTIVE);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n_round = (n + stride - 1) / stride * stride;  // every lane takes part in each wave_append
    for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < n_round; r += stride) {
        bool take = false;
        if (r < n) {
            const uint64_t nb = ((uint64_t)lengths[r] + 3) / 4;
            const bool inside = offsets[r] <= data_len && nb <= data_len - offsets[r];
            if (!inside) atomicOr(&ctr[3], 1ull);
            take = inside && lengths[r] >= k;
        }
        const uint64_t slot = wave_append(take, &ctr[2]);
        if (take) act[slot] = (uint32_t)r;
    }
}

void* g_scratch[64] = {};  // per device: the address of its copy of `scratch`
std::mutex g_scratch_mu;

// The device address of this device's `scratch` (the device must be current).
hipError_t scratch_address(int device, void** out) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[device] = p;
    }
    *out = g_scratch[device];
    return hipSuccess;
}

// hamming_distance_2bit (hammer.rs:13-18): bases that differ between two 2-bit packed words.
__global__ void k_hamming(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, uint64_t n,
                          uint32_t* __restrict__ out) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t d = a[i] ^ b[i];
        out[i] = (uint32_t)__popcll((d | (d >> 1)) & 0x5555555555555555ull);
    }
}

}  // namespace

namespace kc {
// is_solid_kmer256 (bloom.rs:111-126) for n device-resident keys.
hipError_t query_solid(const uint64_t* d_keys4, uint64_t n, const Modulus& M, const uint64_t* d_solid, uint8_t* d_out,
                       hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_query_kernel<<<grid_for(n), BLOCK, 0, st>>>(d_keys4, n, M, d_solid, d_out);
    return hipGetLastError();
}
}  // namespace kc

// ---------------------------------------------------------------------------
// C ABI: standalone filter
// ---------------------------------------------------------------------------

struct kc_filter {
    int device = 0;
    hipStream_t st = nullptr;
    uint64_t num_words = 0;
    Modulus M{};
    DevBuf d_seen, d_solid, d_fs, d_keys, d_prom;
    uint8_t* scratch = nullptr;  // this device's allocation-free Hammer scratch
    std::string err;
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

#define F_TRY(expr)                                                                              \
    do {                                                                                         \
        hipError_t _e = (expr);                                                                  \
        if (_e != hipSuccess) return f->fail(_e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,    \
                                             std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

extern "C" {

int kc_filter_create(kc_filter** out, int32_t device, uint64_t capacity_bits) {
    if (!out) return KC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_EHIP;
    if (device < 0 || device >= ndev) return KC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return KC_EHIP;
    kc_filter* f = new kc_filter();
    f->device = device;
    f->num_words = std::max<uint64_t>((capacity_bits + 63) / 64, 1);
    f->M = make_modulus(f->num_words * 64);
    const uint64_t nw = f->num_words;
    if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess || f->d_seen.ensure(nw * 8) ||
        f->d_solid.ensure(nw * 8) || f->d_fs.ensure(nw * 64 * 4)) {
        kc_filter_destroy(f);
        return KC_ENOMEM;
    }
    k_fill<uint64_t><<<grid_for(nw), BLOCK, 0, f->st>>>(f->d_seen.as<uint64_t>(), nw, 0ull);
    k_fill<uint64_t><<<grid_for(nw), BLOCK, 0, f->st>>>(f->d_solid.as<uint64_t>(), nw, 0ull);
    k_fill<uint32_t><<<grid_for(nw * 64), BLOCK, 0, f->st>>>(f->d_fs.as<uint32_t>(), nw * 64, ~0u);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(f->st) != hipSuccess) {
        kc_filter_destroy(f);
        return KC_EHIP;
    }
    void* sp = nullptr;
    if (scratch_address(device, &sp) != hipSuccess) {
        kc_filter_destroy(f);
        return KC_EHIP;
    }
    f->scratch = static_cast<uint8_t*>(sp);
    *out = f;
    return KC_OK;
}

void kc_filter_destroy(kc_filter* f) {
    if (!f) return;
    (void)hipSetDevice(f->device);
    if (f->st) (void)hipStreamSynchronize(f->st);
    f->d_seen.release();
    f->d_solid.release();
    f->d_fs.release();
    f->d_keys.release();
    f->d_prom.release();
    if (f->st) (void)hipStreamDestroy(f->st);
    delete f;
}

const char* kc_filter_last_error(const kc_filter* f) { return f ? f->err.c_str() : "null filter"; }
uint64_t kc_filter_memory_usage_bytes(const kc_filter* f) { return f ? f->num_words * 8 * 2 : 0; }
uint64_t kc_filter_num_words(const kc_filter* f) { return f ? f->num_words : 0; }

int kc_filter_insert_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* promoted) {
    if (!f) return KC_EINVAL;
    if (n == 0) return KC_OK;
    if (!keys4) return f->fail(KC_EINVAL, "null keys");
    F_TRY(hipSetDevice(f->device));
    const size_t batch_max = (size_t)1 << 28;  // 4*t + 3 fits the 32-bit first-setter words
    for (size_t base = 0; base < n; base += batch_max) {
        const size_t m = std::min(batch_max, n - base);
        F_TRY(f->d_keys.ensure(m * 32));
        F_TRY(f->d_prom.ensure(m));
        F_TRY(hipMemcpyAsync(f->d_keys.p, keys4 + 4 * base, m * 32, hipMemcpyHostToDevice, f->st));
        const uint32_t g = grid_for(m);
        k_ev_first_setter<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_fs.as<uint32_t>());
        k_ev_promote<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_fs.as<uint32_t>(),
                                              f->d_seen.as<uint64_t>(), f->d_prom.as<uint8_t>());
        k_ev_apply<<<g, BLOCK, 0, f->st>>>(f->d_keys.as<uint64_t>(), m, f->M, f->d_prom.as<uint8_t>(),
                                            f->d_seen.as<uint64_t>(), f->d_solid.as<uint64_t>(), f->d_fs.as<uint32_t>());
        F_TRY(hipGetLastError());
        if (promoted) F_TRY(hipMemcpyAsync(promoted + base, f->d_prom.p, m, hipMemcpyDeviceToHost, f->st));
        F_TRY(hipStreamSynchronize(f->st));
    }
    return KC_OK;
}

int kc_filter_is_solid_kmer256(kc_filter* f, const uint64_t* keys4, size_t n, uint8_t* out) {
    if (!f) return KC_EINVAL;
    if (n == 0) return KC_OK;
    if (!keys4 || !out) return f->fail(KC_EINVAL, "null keys/out");
    F_TRY(hipSetDevice(f->device));
    F_TRY(f->d_keys.ensure(n * 32));
    F_TRY(f->d_prom.ensure(n));
    F_TRY(hipMemcpyAsync(f->d_keys.p, keys4, n * 32, hipMemcpyHostToDevice, f->st));
    F_TRY(query_solid(f->d_keys.as<uint64_t>(), n, f->M, f->d_solid.as<uint64_t>(), f->d_prom.as<uint8_t>(), f->st));
    F_TRY(hipMemcpyAsync(out, f->d_prom.p, n, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    return KC_OK;
}

int kc_filter_correct_reads(kc_filter* f, uint32_t k, const uint8_t* data, size_t data_len, const uint64_t* offsets,
                            const uint32_t* lengths, size_t n_reads, uint8_t* out, uint64_t* n_corrected) {
    if (!f) return KC_EINVAL;
    if (k < 1 || k > KC_MAX_K) return f->fail(KC_EINVAL, "k out of range");
    if (n_reads && (!offsets || !lengths)) return f->fail(KC_EINVAL, "null offsets/lengths");
    if (data_len && (!data || !out)) return f->fail(KC_EINVAL, "null data/out");
    if (n_reads > 0xFFFFFFFFull) return f->fail(KC_EUNSUPPORTED, "more than 2^32-1 reads");
    std::vector<uint32_t> active;
    for (size_t i = 0; i < n_reads; ++i) {
        const uint64_t nb = ((uint64_t)lengths[i] + 3) / 4;
        if (offsets[i] > data_len || nb > data_len - offsets[i]) return f->fail(KC_EINVAL, "read extends past data_len");
        if (lengths[i] >= k) active.push_back((uint32_t)i);
    }
    F_TRY(hipSetDevice(f->device));
    DevBuf d_data, d_corr, d_off, d_len, d_act, d_ctr, d_lib;
    F_TRY(d_data.ensure(data_len + DATA_PAD));
    F_TRY(d_corr.ensure(data_len + DATA_PAD));
    F_TRY(d_off.ensure(n_reads * 8 + 8));
    F_TRY(d_len.ensure(n_reads * 4 + 4));
    F_TRY(d_act.ensure(active.size() * 4 + 4));
    F_TRY(d_ctr.ensure(64));
    F_TRY(hipMemsetAsync(d_data.p, 0, d_data.bytes, f->st));
    F_TRY(hipMemsetAsync(d_ctr.p, 0, 64, f->st));
    if (data_len) F_TRY(hipMemcpyAsync(d_data.p, data, data_len, hipMemcpyHostToDevice, f->st));
    if (n_reads) {
        F_TRY(hipMemcpyAsync(d_off.p, offsets, n_reads * 8, hipMemcpyHostToDevice, f->st));
        F_TRY(hipMemcpyAsync(d_len.p, lengths, n_reads * 4, hipMemcpyHostToDevice, f->st));
    }
    if (!active.empty()) F_TRY(hipMemcpyAsync(d_act.p, active.data(), active.size() * 4, hipMemcpyHostToDevice, f->st));
    F_TRY(hipMemcpyAsync(d_corr.p, d_data.p, d_data.bytes, hipMemcpyDeviceToDevice, f->st));
    if (!active.empty()) {
        LibBloom lb{f->d_solid.as<uint64_t>(), f->M};
        F_TRY(d_lib.ensure(sizeof(LibBloom)));
        F_TRY(hipMemcpyAsync(d_lib.p, &lb, sizeof(LibBloom), hipMemcpyHostToDevice, f->st));
        const uint32_t g = hammer_grid(active.size());
#define KC_HAMMER_LAUNCH(KW)                                                                                  \
    k_hammer<KW><<<g, 64, 0, f->st>>>(d_data.as<uint8_t>(), d_corr.as<uint8_t>(), d_act.as<uint32_t>(),        \
                                      active.size(), d_off.as<uint64_t>(), d_len.as<uint32_t>(), nullptr,       \
                                      d_lib.as<LibBloom>(), k, d_ctr.as<unsigned long long>(), 1,         \
                                      d_ctr.as<unsigned long long>() + 1, nullptr)
        if (kw_of_k(k) == 4) KC_HAMMER_LAUNCH(4);
        else if (kw_of_k(k) == 2) KC_HAMMER_LAUNCH(2);
        else KC_HAMMER_LAUNCH(1);
#undef KC_HAMMER_LAUNCH
        F_TRY(hipGetLastError());
    }
    uint64_t ctr[2] = {0, 0};
    F_TRY(hipMemcpyAsync(ctr, d_ctr.p, 16, hipMemcpyDeviceToHost, f->st));
    if (data_len) F_TRY(hipMemcpyAsync(out, d_corr.p, data_len, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    for (DevBuf* b : {&d_data, &d_corr, &d_off, &d_len, &d_act, &d_ctr, &d_lib}) b->release();
    if (n_corrected) *n_corrected = ctr[0];
    return KC_OK;
}

int kc_filter_correct_reads_noalloc(kc_filter* f, uint32_t k, const uint8_t* d_data, size_t data_len,
                                    const uint64_t* d_offsets, const uint32_t* d_lengths, size_t n_reads,
                                    uint8_t* d_out, uint64_t* n_corrected) {
    if (!f) return KC_EINVAL;
    if (k < 1 || k > KC_MAX_K) return f->fail(KC_EINVAL, "k out of range");
    if (n_reads > KC_NOALLOC_MAX_READS)
        return f->fail(KC_EINVAL, "kc_filter_correct_reads_noalloc takes at most KC_NOALLOC_MAX_READS (" +
                                      std::to_string(KC_NOALLOC_MAX_READS) + ") reads a call");
    if (n_reads && (!d_offsets || !d_lengths)) return f->fail(KC_EINVAL, "null offsets/lengths");
    if (data_len && (!d_data || !d_out)) return f->fail(KC_EINVAL, "null data/out");
    if (!f->scratch) return f->fail(KC_ESTATE, "no scratch address for this device");
    F_TRY(hipSetDevice(f->device));
    uint8_t* sc = f->scratch;
    k_na_init<<<1, 64, 0, f->st>>>(sc, LibBloom{f->d_solid.as<uint64_t>(), f->M});
    if (data_len) k_na_copy<<<grid_for(data_len), BLOCK, 0, f->st>>>(d_data, d_out, data_len);
    if (n_reads) {
        k_na_active<<<grid_for(n_reads), BLOCK, 0, f->st>>>(d_offsets, d_lengths, n_reads, data_len, k, sc);
        unsigned long long* ctr = reinterpret_cast<unsigned long long*>(sc + NA_COUNTERS);
        const uint32_t g = hammer_grid(n_reads);
#define KC_HAMMER_LAUNCH(KW)                                                                                        \
    k_hammer<KW><<<g, 64, 0, f->st>>>(d_data, d_out, reinterpret_cast<const uint32_t*>(sc + NA_ACTIVE), n_reads,    \
                                      d_offsets, d_lengths, nullptr, reinterpret_cast<const LibBloom*>(sc + NA_BLOOM), \
                                      k, ctr, 1, ctr + 1, ctr + 2)
        if (kw_of_k(k) == 4) KC_HAMMER_LAUNCH(4);
        else if (kw_of_k(k) == 2) KC_HAMMER_LAUNCH(2);
        else KC_HAMMER_LAUNCH(1);
#undef KC_HAMMER_LAUNCH
    }
    F_TRY(hipGetLastError());
    unsigned long long ctr[4] = {0, 0, 0, 0};
    F_TRY(hipMemcpyAsync(ctr, sc + NA_COUNTERS, sizeof(ctr), hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    if (ctr[3]) return f->fail(KC_EINVAL, "read extends past data_len");
    if (n_corrected) *n_corrected = ctr[0];
    return KC_OK;
}

int kc_hamming_distance_2bit(int32_t device, const uint64_t* a, const uint64_t* b, size_t n, uint32_t* out) {
    if (n == 0) return KC_OK;
    if (!a || !b || !out) return KC_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_EHIP;
    if (device < 0 || device >= ndev) return KC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return KC_EHIP;
    DevBuf da, db, dout;
    hipStream_t st = nullptr;
    int rc = KC_OK;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return KC_EHIP;
    if (da.ensure(n * 8) || db.ensure(n * 8) || dout.ensure(n * 4)) rc = KC_ENOMEM;
    if (!rc && (hipMemcpyAsync(da.p, a, n * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(db.p, b, n * 8, hipMemcpyHostToDevice, st) != hipSuccess))
        rc = KC_EHIP;
    if (!rc) {
        k_hamming<<<grid_for(n), BLOCK, 0, st>>>(da.as<uint64_t>(), db.as<uint64_t>(), n, dout.as<uint32_t>());
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, dout.p, n * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = KC_EHIP;
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}

int kc_filter_export(kc_filter* f, uint64_t* seen, uint64_t* solid, size_t num_words) {
    if (!f) return KC_EINVAL;
    if (num_words != f->num_words) return f->fail(KC_EINVAL, "num_words mismatch");
    F_TRY(hipSetDevice(f->device));
    if (seen) F_TRY(hipMemcpyAsync(seen, f->d_seen.p, num_words * 8, hipMemcpyDeviceToHost, f->st));
    if (solid) F_TRY(hipMemcpyAsync(solid, f->d_solid.p, num_words * 8, hipMemcpyDeviceToHost, f->st));
    F_TRY(hipStreamSynchronize(f->st));
    return KC_OK;
}

}  // extern "C"
