// kc_mlp.h — the small MLP of the learned first pass (internal to libkc.so).
//
// Network: 16 features -> 64 (ReLU) -> 64 (ReLU) -> 1 (sigmoid).  Master weights are fp32
// on the device; the forward pass uses their bf16 roundings (bf16 weights, bf16 inputs
// and first hidden layer, fp32 accumulation, fp32 biases):
//   xb = bf16(x);  z1 = xb . bf16(W1) + b1;  h1 = bf16(relu(z1))
//   z2 = h1 . bf16(W2) + b2;  h2 = relu(z2);  z3 = h2 . bf16(w3) + b3;  p = sigmoid(z3)
// Parameter vector layout (MLP_NPARAM floats): W1[16][64], b1[64], W2[64][64], b2[64],
// w3[64], b3.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "kc_host.h"

namespace kc {

constexpr int MLP_IN = 16;
constexpr int MLP_H = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_IN * MLP_H;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = OFF_B3 + 1;  // 5313

}  // namespace kc

struct kc_mlp {
    int32_t device = 0;
    hipStream_t st = nullptr;
    std::string err;
    kc::DevBuf d_w;     // fp32 master parameters (MLP_NPARAM)
    kc::DevBuf d_wb;    // bf16 copies of W1, W2, w3 (MLP_NPARAM bf16; biases unused)
    kc::DevBuf d_m, d_v, d_g, d_part;  // Adam moments, summed gradient, per-block partials
    kc::DevBuf d_x, d_y, d_sw, d_perm, d_p, d_loss;  // data staged for train / forward / grad
    uint64_t steps = 0;  // Adam steps taken
    int fail(int code, const std::string& m) {
        err = m;
        return code;
    }
};

namespace kc {

// p[i] = network(X[i]) for n rows of MLP_IN fp32 features, on `st` (device pointers).
hipError_t mlp_infer_device(const kc_mlp* m, const float* X, uint64_t n, float* p, hipStream_t st);

}  // namespace kc
