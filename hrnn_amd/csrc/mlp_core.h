// mlp_core.h -- the learned-m network's layout and forward pass, shared by the
// training kernels (mlp.hip) and the stage-1 beam (beam.hip) so that both
// compute the same numbers.
#pragma once
#include "common.h"

namespace hrnn {

// 32-64-64-16 MLP, ReLU, softmax.  Parameter vector (FP32):
//   W1[64][32] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]
constexpr int kIn = 32, kH = 64, kOut = 16;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kParams = oB3 + kOut;
static_assert(kParams == 7312, "parameter layout");

// Forward pass of one sample by a one-wave workgroup (lane j = hidden unit j):
// x[kIn] in LDS; writes z1, h1, z2, h2 (kH each), the logits lg[kOut] and the
// probabilities p[kOut] to LDS.  Each unit sums its inputs in index order with
// FMA; lane 0 takes the softmax sequentially.
__device__ __forceinline__ void mlp_forward_wave(const float *__restrict__ P, const float *x, float *z1,
                                                 float *h1, float *z2, float *h2, float *lg, float *p,
                                                 uint32_t j) {
    {
        float a = P[oB1 + j];
        for (int i = 0; i < kIn; i++) a = __builtin_fmaf(P[oW1 + j * kIn + i], x[i], a);
        z1[j] = a;
        h1[j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    {
        float a = P[oB2 + j];
        for (int i = 0; i < kH; i++) a = __builtin_fmaf(P[oW2 + j * kH + i], h1[i], a);
        z2[j] = a;
        h2[j] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    if (j < (uint32_t)kOut) {
        float a = P[oB3 + j];
        for (int i = 0; i < kH; i++) a = __builtin_fmaf(P[oW3 + j * kH + i], h2[i], a);
        lg[j] = a;
    }
    __syncthreads();
    if (j == 0) {
        float mx = lg[0];
        for (int c = 1; c < kOut; c++) mx = fmaxf(mx, lg[c]);
        float sum = 0.f;
        for (int c = 0; c < kOut; c++) {
            p[c] = expf(lg[c] - mx);
            sum += p[c];
        }
        for (int c = 0; c < kOut; c++) p[c] /= sum;
    }
    __syncthreads();
}

// The class whose cumulative probability first reaches tau (argmax when
// tau <= 0; the last class when none does).
__device__ __forceinline__ uint32_t pick_class_of(const float *p, float tau) {
    uint32_t c = kOut - 1;
    if (tau <= 0.f) {
        c = 0;
        for (uint32_t j = 1; j < (uint32_t)kOut; j++)
            if (p[j] > p[c]) c = j;
    } else {
        float acc = 0.f;
        for (uint32_t j = 0; j < (uint32_t)kOut; j++) {
            acc += p[j];
            if (acc >= tau) {
                c = j;
                break;
            }
        }
    }
    return c;
}

}  // namespace hrnn
