// kc_mlp.hip — the learned first pass's network: MFMA (bf16 in, fp32 accumulate)
// inference, and training with hand-written forward/backward kernels and Adam.
//
// Inference tiles 16 rows per wave through v_mfma_f32_16x16x16_bf16: layer 1 is one
// K=16 step per 16-column block, layer 2 four K=16 steps; the bf16 first hidden layer
// goes through LDS to become the next A operand.  Training computes each sample's forward
// and backward pass in fp32 from the same bf16-rounded operands, then reduces the
// per-sample outer products per block in a fixed order (no atomics, so a training run is
// reproducible bit for bit) and sums the block partials in a second fixed-order pass.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/kc.h"
#include "kc_mlp.h"

using namespace kc;

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (NaN kept quiet).
__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) wb[i] = f2bf(w[i]);
}

constexpr int IB = 256;  // inference block: 4 waves x 16 rows

__global__ void __launch_bounds__(IB) k_mlp_infer(const float* __restrict__ X, uint64_t n, const float* __restrict__ w,
                                                  const uint16_t* __restrict__ wb, float* __restrict__ p) {
    __shared__ uint16_t sW1[MLP_IN][MLP_H];
    __shared__ uint16_t sW2[MLP_H][MLP_H];
    __shared__ float sB1[MLP_H], sB2[MLP_H], sW3[MLP_H];
    __shared__ uint16_t sH1[4][16][MLP_H];
    __shared__ float sH2[4][16][MLP_H + 1];
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    for (int i = tid; i < MLP_IN * MLP_H; i += IB) sW1[i / MLP_H][i % MLP_H] = wb[OFF_W1 + i];
    for (int i = tid; i < MLP_H * MLP_H; i += IB) sW2[i / MLP_H][i % MLP_H] = wb[OFF_W2 + i];
    for (int i = tid; i < MLP_H; i += IB) {
        sB1[i] = w[OFF_B1 + i];
        sB2[i] = w[OFF_B2 + i];
        sW3[i] = bf2f(wb[OFF_W3 + i]);
    }
    const float b3 = w[OFF_B3];
    __syncthreads();
    // operand layout of v_mfma_f32_16x16x16_bf16: lane l holds A[l%16][kq..kq+3] and
    // B[kq..kq+3][l%16] with kq = 4*(l/16); D[kq+r][l%16] for r = 0..3
    const int r16 = lane % 16, kq = (lane / 16) * 4;
    for (uint64_t base = (uint64_t)blockIdx.x * 64; base < n; base += (uint64_t)gridDim.x * 64) {
        const uint64_t row0 = base + (uint64_t)wave * 16;
        s4 a;
        {
            const uint64_t r = row0 + r16;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q] = r < n ? (short)f2bf(X[r * MLP_IN + kq + q]) : (short)0;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            s4 b;
#pragma unroll
            for (int q = 0; q < 4; ++q) b[q] = (short)sW1[kq + q][c * 16 + r16];
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = acc[r] + sB1[c * 16 + r16];
                sH1[wave][kq + r][c * 16 + r16] = f2bf(z > 0.f ? z : 0.f);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                s4 a2, b2;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a2[q] = (short)sH1[wave][r16][ks * 16 + kq + q];
                    b2[q] = (short)sW2[ks * 16 + kq + q][c * 16 + r16];
                }
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b2, acc, 0, 0, 0);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float z = acc[r] + sB2[c * 16 + r16];
                sH2[wave][kq + r][c * 16 + r16] = z > 0.f ? z : 0.f;
            }
        }
        __syncthreads();
        if (lane < 16) {
            const uint64_t r = row0 + lane;
            float z = b3;
            for (int j = 0; j < MLP_H; ++j) z = fmaf(sH2[wave][lane][j], sW3[j], z);
            if (r < n) p[r] = 1.0f / (1.0f + expf(-z));
        }
        __syncthreads();  // sH1 / sH2 are rewritten by the next tile
    }
}

// ---- training -----------------------------------------------------------------------

constexpr int GS = 32;  // samples per gradient block (threads 0..31 run one sample each)
constexpr int GT = 64;  // threads per gradient block

// Per-block gradient of the weighted mean binary cross-entropy over samples
// [blockIdx.x * GS, +GS) of the batch (sample b of the batch is row idx ? idx[b] : b).
// part[blockIdx.x][MLP_NPARAM] and loss_part[blockIdx.x] receive the block's sums.
__global__ void __launch_bounds__(GT) k_mlp_grad(const float* __restrict__ X, const uint8_t* __restrict__ y,
                                                 const float* __restrict__ sw, const uint32_t* __restrict__ idx,
                                                 uint64_t batch, float inv_n, float pos_weight,
                                                 const float* __restrict__ w, const uint16_t* __restrict__ wb,
                                                 float* __restrict__ part, float* __restrict__ loss_part) {
    __shared__ float sW1[MLP_IN][MLP_H];
    __shared__ float sW2[MLP_H][MLP_H];
    __shared__ float sB1[MLP_H], sB2[MLP_H], sW3[MLP_H];
    __shared__ float sX[GS][MLP_IN];
    __shared__ float sH1[GS][MLP_H + 1];
    __shared__ float sH2[GS][MLP_H + 1];
    __shared__ float sDZ1[GS][MLP_H + 1];
    __shared__ float sDZ2[GS][MLP_H + 1];
    __shared__ float sDZ3[GS], sLoss[GS];
    const int t = threadIdx.x;
    for (int i = t; i < MLP_IN * MLP_H; i += GT) sW1[i / MLP_H][i % MLP_H] = bf2f(wb[OFF_W1 + i]);
    for (int i = t; i < MLP_H * MLP_H; i += GT) sW2[i / MLP_H][i % MLP_H] = bf2f(wb[OFF_W2 + i]);
    for (int i = t; i < MLP_H; i += GT) {
        sB1[i] = w[OFF_B1 + i];
        sB2[i] = w[OFF_B2 + i];
        sW3[i] = bf2f(wb[OFF_W3 + i]);
    }
    const float b3 = w[OFF_B3];
    __syncthreads();
    if (t < GS) {
        const uint64_t b = (uint64_t)blockIdx.x * GS + t;
        const bool valid = b < batch;
        const uint64_t row = valid ? (idx ? idx[b] : b) : 0;
        for (int i = 0; i < MLP_IN; ++i) sX[t][i] = valid ? bf2f(f2bf(X[row * MLP_IN + i])) : 0.f;
        uint64_t m1 = 0, m2 = 0;
        for (int j = 0; j < MLP_H; ++j) {
            float z = 0.f;
            for (int i = 0; i < MLP_IN; ++i) z = fmaf(sX[t][i], sW1[i][j], z);
            z += sB1[j];
            if (z > 0.f) m1 |= 1ull << j;
            sH1[t][j] = bf2f(f2bf(z > 0.f ? z : 0.f));
        }
        for (int j = 0; j < MLP_H; ++j) {
            float z = 0.f;
            for (int i = 0; i < MLP_H; ++i) z = fmaf(sH1[t][i], sW2[i][j], z);
            z += sB2[j];
            if (z > 0.f) m2 |= 1ull << j;
            sH2[t][j] = z > 0.f ? z : 0.f;
        }
        float z3 = b3;
        for (int j = 0; j < MLP_H; ++j) z3 = fmaf(sH2[t][j], sW3[j], z3);
        const float yy = valid && y[row] ? 1.f : 0.f;
        const float wt = valid ? (sw ? sw[row] : 1.f) * (yy > 0.f ? pos_weight : 1.f) : 0.f;
        const float p = 1.0f / (1.0f + expf(-z3));
        // BCE with logits: y softplus(-z) + (1 - y) softplus(z)
        const float sp_pos = fmaxf(z3, 0.f) + log1pf(expf(-fabsf(z3)));  // softplus(z3)
        const float loss = yy > 0.f ? sp_pos - z3 : sp_pos;
        sLoss[t] = wt * loss * inv_n;
        const float dz3 = wt * (p - yy) * inv_n;
        sDZ3[t] = dz3;
        for (int j = 0; j < MLP_H; ++j) sDZ2[t][j] = (m2 >> j) & 1ull ? dz3 * sW3[j] : 0.f;
        for (int i = 0; i < MLP_H; ++i) {
            float d = 0.f;
            for (int j = 0; j < MLP_H; ++j) d = fmaf(sDZ2[t][j], sW2[i][j], d);
            sDZ1[t][i] = (m1 >> i) & 1ull ? d : 0.f;
        }
    }
    __syncthreads();
    float* out = part + (size_t)blockIdx.x * MLP_NPARAM;
    // dW2[t][j], db1[t], db2[t], dw3[t]
    for (int j = 0; j < MLP_H; ++j) {
        float s = 0.f;
        for (int q = 0; q < GS; ++q) s = fmaf(sH1[q][t], sDZ2[q][j], s);
        out[OFF_W2 + t * MLP_H + j] = s;
    }
    {
        float s1 = 0.f, s2 = 0.f, s3 = 0.f;
        for (int q = 0; q < GS; ++q) {
            s1 += sDZ1[q][t];
            s2 += sDZ2[q][t];
            s3 = fmaf(sDZ3[q], sH2[q][t], s3);
        }
        out[OFF_B1 + t] = s1;
        out[OFF_B2 + t] = s2;
        out[OFF_W3 + t] = s3;
    }
    // dW1[i][j] for i = t / 4, j = (t % 4) * 16 .. +16
    {
        const int i = t / 4, j0 = (t % 4) * 16;
        for (int j = j0; j < j0 + 16; ++j) {
            float s = 0.f;
            for (int q = 0; q < GS; ++q) s = fmaf(sX[q][i], sDZ1[q][j], s);
            out[OFF_W1 + i * MLP_H + j] = s;
        }
    }
    if (t == 0) {
        float s = 0.f, l = 0.f;
        for (int q = 0; q < GS; ++q) {
            s += sDZ3[q];
            l += sLoss[q];
        }
        out[OFF_B3] = s;
        loss_part[blockIdx.x] = l;
    }
}

// g[i] = sum over blocks of part[b][i], in block order; loss likewise (thread MLP_NPARAM).
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ loss_part, uint32_t nblk,
                             float* __restrict__ g, double* __restrict__ loss) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= MLP_NPARAM; i += gridDim.x * blockDim.x) {
        if (i == MLP_NPARAM) {
            double s = 0.0;
            for (uint32_t b = 0; b < nblk; ++b) s += loss_part[b];
            *loss = s;
        } else {
            float s = 0.f;
            for (uint32_t b = 0; b < nblk; ++b) s += part[(size_t)b * MLP_NPARAM + i];
            g[i] = s;
        }
    }
}

__global__ void k_mlp_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, uint16_t* __restrict__ wb, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_NPARAM; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

// Deterministic He-style initialisation from a 64-bit seed (splitmix64 + Box-Muller).
void init_params(uint64_t seed, std::vector<float>& w) {
    w.assign(MLP_NPARAM, 0.f);
    uint64_t s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto normal = [&]() {
        const double u1 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < MLP_IN * MLP_H; ++i) w[OFF_W1 + i] = (float)(normal() * std::sqrt(2.0 / MLP_IN));
    for (int i = 0; i < MLP_H * MLP_H; ++i) w[OFF_W2 + i] = (float)(normal() * std::sqrt(2.0 / MLP_H));
    for (int i = 0; i < MLP_H; ++i) w[OFF_W3 + i] = (float)(normal() * std::sqrt(1.0 / MLP_H));
}

#define M_TRY(expr)                                                                                   \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return m->fail(_e == hipErrorOutOfMemory ? KC_ENOMEM : KC_EHIP,                           \
                           std::string(#expr) + ": " + hipGetErrorString(_e));                       \
    } while (0)

int refresh_bf16(kc_mlp* m) {
    k_to_bf16<<<(MLP_NPARAM + 255) / 256, 256, 0, m->st>>>(m->d_w.as<float>(), m->d_wb.as<uint16_t>(), MLP_NPARAM);
    M_TRY(hipGetLastError());
    return KC_OK;
}

// Full-batch or mini-batch gradient into m->d_g (and the loss into m->d_loss) for rows
// staged in d_x / d_y / d_sw; idx (device, nullable) selects `batch` rows.
int gradient(kc_mlp* m, const uint32_t* idx, uint64_t batch, float inv_n, float pos_weight, bool has_sw) {
    const uint64_t nblk = (batch + GS - 1) / GS;
    if (nblk > 0xFFFFFFFFull) return m->fail(KC_EUNSUPPORTED, "batch too large");
    M_TRY(m->d_part.ensure(std::max<uint64_t>(nblk, 1) * (MLP_NPARAM + 1) * 4));
    float* part = m->d_part.as<float>();
    float* loss_part = part + nblk * MLP_NPARAM;
    if (nblk) {
        k_mlp_grad<<<(uint32_t)nblk, GT, 0, m->st>>>(m->d_x.as<float>(), m->d_y.as<uint8_t>(),
                                                     has_sw ? m->d_sw.as<float>() : nullptr, idx, batch, inv_n,
                                                     pos_weight, m->d_w.as<float>(), m->d_wb.as<uint16_t>(), part,
                                                     loss_part);
        M_TRY(hipGetLastError());
    }
    k_mlp_reduce<<<(MLP_NPARAM + 256) / 256, 256, 0, m->st>>>(part, loss_part, (uint32_t)nblk, m->d_g.as<float>(),
                                                               m->d_loss.as<double>());
    M_TRY(hipGetLastError());
    return KC_OK;
}

int stage(kc_mlp* m, const float* X, const uint8_t* y, const float* sw, size_t n) {
    M_TRY(m->d_x.ensure(std::max<size_t>(n, 1) * MLP_IN * 4));
    if (n) M_TRY(hipMemcpyAsync(m->d_x.p, X, n * MLP_IN * 4, hipMemcpyHostToDevice, m->st));
    if (y) {
        M_TRY(m->d_y.ensure(std::max<size_t>(n, 1)));
        if (n) M_TRY(hipMemcpyAsync(m->d_y.p, y, n, hipMemcpyHostToDevice, m->st));
    }
    if (sw) {
        M_TRY(m->d_sw.ensure(std::max<size_t>(n, 1) * 4));
        if (n) M_TRY(hipMemcpyAsync(m->d_sw.p, sw, n * 4, hipMemcpyHostToDevice, m->st));
    }
    return KC_OK;
}

}  // namespace

namespace kc {

hipError_t mlp_infer_device(const kc_mlp* m, const float* X, uint64_t n, float* p, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t tiles = (n + 63) / 64;
    const uint32_t g = (uint32_t)std::min<uint64_t>(tiles, 8192);
    k_mlp_infer<<<g, IB, 0, st>>>(X, n, m->d_w.as<float>(), m->d_wb.as<uint16_t>(), p);
    return hipGetLastError();
}

}  // namespace kc

extern "C" {

int kc_mlp_create(kc_mlp** out, int32_t device, uint64_t seed) {
    if (!out) return KC_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KC_EHIP;
    if (device < 0 || device >= ndev) return KC_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return KC_EHIP;
    kc_mlp* m = new kc_mlp();
    m->device = device;
    std::vector<float> w;
    init_params(seed, w);
    const std::vector<float> zeros(MLP_NPARAM, 0.f);
    if (hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) != hipSuccess ||
        m->d_w.ensure(MLP_NPARAM * 4) != hipSuccess || m->d_wb.ensure(MLP_NPARAM * 2) != hipSuccess ||
        m->d_m.ensure(MLP_NPARAM * 4) != hipSuccess || m->d_v.ensure(MLP_NPARAM * 4) != hipSuccess ||
        m->d_g.ensure(MLP_NPARAM * 4) != hipSuccess || m->d_loss.ensure(8) != hipSuccess ||
        hipMemcpy(m->d_w.p, w.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->d_m.p, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->d_v.p, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice) != hipSuccess ||
        refresh_bf16(m) != KC_OK || hipStreamSynchronize(m->st) != hipSuccess) {
        kc_mlp_destroy(m);
        return KC_EHIP;
    }
    *out = m;
    return KC_OK;
}

void kc_mlp_destroy(kc_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->st) (void)hipStreamSynchronize(m->st);
    kc::DevBuf* bufs[] = {&m->d_w, &m->d_wb, &m->d_m, &m->d_v, &m->d_g, &m->d_part, &m->d_x,
                          &m->d_y, &m->d_sw, &m->d_perm, &m->d_p, &m->d_loss};
    for (auto* b : bufs) b->release();
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

const char* kc_mlp_last_error(const kc_mlp* m) { return m ? m->err.c_str() : "null mlp"; }

uint32_t kc_mlp_num_params(void) { return MLP_NPARAM; }

int kc_mlp_get_params(kc_mlp* m, float* w, size_t n) {
    if (!m || !w) return KC_EINVAL;
    if (n != MLP_NPARAM) return m->fail(KC_EINVAL, "n must be kc_mlp_num_params()");
    M_TRY(hipSetDevice(m->device));
    M_TRY(hipMemcpyAsync(w, m->d_w.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->st));
    M_TRY(hipStreamSynchronize(m->st));
    return KC_OK;
}

int kc_mlp_set_params(kc_mlp* m, const float* w, size_t n) {
    if (!m || !w) return KC_EINVAL;
    if (n != MLP_NPARAM) return m->fail(KC_EINVAL, "n must be kc_mlp_num_params()");
    M_TRY(hipSetDevice(m->device));
    M_TRY(hipMemcpyAsync(m->d_w.p, w, MLP_NPARAM * 4, hipMemcpyHostToDevice, m->st));
    int rc = refresh_bf16(m);
    if (rc) return rc;
    M_TRY(hipStreamSynchronize(m->st));
    return KC_OK;
}

int kc_mlp_forward(kc_mlp* m, const float* X, size_t n, float* p) {
    if (!m || (n && (!X || !p))) return KC_EINVAL;
    if (n == 0) return KC_OK;
    M_TRY(hipSetDevice(m->device));
    int rc = stage(m, X, nullptr, nullptr, n);
    if (rc) return rc;
    M_TRY(m->d_p.ensure(n * 4));
    M_TRY(mlp_infer_device(m, m->d_x.as<float>(), n, m->d_p.as<float>(), m->st));
    M_TRY(hipMemcpyAsync(p, m->d_p.p, n * 4, hipMemcpyDeviceToHost, m->st));
    M_TRY(hipStreamSynchronize(m->st));
    return KC_OK;
}

int kc_mlp_gradient(kc_mlp* m, const float* X, const uint8_t* y, const float* sample_w, size_t n, float pos_weight,
                    float* grad, double* loss) {
    if (!m || !grad || (n && (!X || !y))) return KC_EINVAL;
    M_TRY(hipSetDevice(m->device));
    int rc = stage(m, X, y, sample_w, n);
    if (rc) return rc;
    rc = gradient(m, nullptr, n, n ? 1.0f / (float)n : 0.f, pos_weight, sample_w != nullptr);
    if (rc) return rc;
    M_TRY(hipMemcpyAsync(grad, m->d_g.p, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->st));
    double l = 0;
    M_TRY(hipMemcpyAsync(&l, m->d_loss.p, 8, hipMemcpyDeviceToHost, m->st));
    M_TRY(hipStreamSynchronize(m->st));
    if (loss) *loss = l;
    return KC_OK;
}

int kc_mlp_train(kc_mlp* m, const float* X, const uint8_t* y, const float* sample_w, size_t n, uint32_t epochs,
                 uint32_t batch, float lr, float pos_weight, uint64_t seed, double* last_loss) {
    if (!m || (n && (!X || !y))) return KC_EINVAL;
    if (batch == 0 || lr <= 0.f) return m->fail(KC_EINVAL, "batch and lr must be positive");
    if (n > 0xFFFFFFFFull) return m->fail(KC_EUNSUPPORTED, "at most 2^32-1 training rows");
    M_TRY(hipSetDevice(m->device));
    int rc = stage(m, X, y, sample_w, n);
    if (rc) return rc;
    M_TRY(m->d_perm.ensure(std::max<size_t>(n, 1) * 4));
    std::vector<uint32_t> perm(n);
    uint64_t s = seed;
    auto next = [&]() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double l = 0;
    for (uint32_t e = 0; e < epochs && n; ++e) {
        for (uint32_t i = 0; i < n; ++i) perm[i] = i;
        for (uint64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
        M_TRY(hipMemcpyAsync(m->d_perm.p, perm.data(), n * 4, hipMemcpyHostToDevice, m->st));
        for (uint64_t b0 = 0; b0 < n; b0 += batch) {
            const uint64_t nb = std::min<uint64_t>(batch, n - b0);
            rc = gradient(m, m->d_perm.as<uint32_t>() + b0, nb, 1.0f / (float)nb, pos_weight, sample_w != nullptr);
            if (rc) return rc;
            ++m->steps;
            const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
            k_mlp_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, m->st>>>(m->d_w.as<float>(), m->d_g.as<float>(),
                                                                   m->d_m.as<float>(), m->d_v.as<float>(),
                                                                   m->d_wb.as<uint16_t>(), lr, b1, b2, eps, c1, c2);
            M_TRY(hipGetLastError());
        }
        M_TRY(hipStreamSynchronize(m->st));  // the host permutation is rewritten next epoch
    }
    M_TRY(hipMemcpyAsync(&l, m->d_loss.p, 8, hipMemcpyDeviceToHost, m->st));
    M_TRY(hipStreamSynchronize(m->st));
    if (last_loss) *last_loss = l;
    return KC_OK;
}

}  // extern "C"
