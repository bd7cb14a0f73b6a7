// Tile-skip network shared by the standalone model entry points (m1_mlp.hip)
// and the batched march (m1_batch_kernels.hip): 11 inputs, three hidden
// layers of 64 ReLU units, one logit.  Weights are used in bfloat16 (rounded
// to nearest even from the float32 master copy), biases and every
// accumulation in float32.
//
// Flat parameter order (float32, m1_mlp_param_count() values):
//   W1[11][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
// with W[i][j] the weight from input i to unit j.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct m1_mlp;

namespace m1 {
namespace mlp {

constexpr int NIN = 11;
constexpr int NH = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NIN * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH;
constexpr int OFF_B4 = OFF_W4 + NH;
constexpr int NPARAM = OFF_B4 + 1;

// Parameters as the kernels read them: weights as bf16 bit patterns (the
// bias slots hold the float32 biases' bf16 neighbours and are unused),
// biases as float32.
struct DeviceParams {
    const uint16_t* w;  // NPARAM bf16 values (weights used, bias slots ignored)
    const float* p;     // NPARAM float32 values (biases used)
};

__host__ __device__ inline float bf16_to_float(uint16_t h)
{
    union {
        uint32_t u;
        float f;
    } v;
    v.u = uint32_t(h) << 16;
    return v.f;
}

// Round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t float_to_bf16(float x)
{
    union {
        float f;
        uint32_t u;
    } v;
    v.f = x;
    if ((v.u & 0x7f800000u) == 0x7f800000u && (v.u & 0x7fffffu))
        return uint16_t((v.u >> 16) | 0x40u);
    const uint32_t lsb = (v.u >> 16) & 1u;
    v.u += 0x7fffu + lsb;
    return uint16_t(v.u >> 16);
}

// Samples per workgroup of the network kernels (one wave): each thread owns
// one sample and its activations live in LDS as [unit][thread], so reads of
// a unit across the wave are conflict-free and no large per-thread arrays
// are needed.
constexpr int WG = 64;

// Forward pass of the sample of thread t with the weights staged in LDS (wl,
// bf16) and the biases in bl (float).  A and B are [NH][WG] LDS scratch;
// h1..h3 (global, nullable) receive the post-ReLU activations.  Returns the
// logit.
__device__ inline float forward_one(const uint16_t* wl, const float* bl, const float x[NIN],
                                    float (*A)[WG], float (*B)[WG], int t, float* h1, float* h2,
                                    float* h3)
{
    for (int j = 0; j < NH; ++j) {
        float s = bl[OFF_B1 + j];
#pragma unroll
        for (int i = 0; i < NIN; ++i)
            s = fmaf(x[i], bf16_to_float(wl[OFF_W1 + i * NH + j]), s);
        A[j][t] = fmaxf(s, 0.f);
    }
    for (int j = 0; j < NH; ++j) {
        float s = bl[OFF_B2 + j];
#pragma unroll 8
        for (int i = 0; i < NH; ++i)
            s = fmaf(A[i][t], bf16_to_float(wl[OFF_W2 + i * NH + j]), s);
        B[j][t] = fmaxf(s, 0.f);
    }
    if (h1)
        for (int j = 0; j < NH; ++j)
            h1[j] = A[j][t];
    for (int j = 0; j < NH; ++j) {
        float s = bl[OFF_B3 + j];
#pragma unroll 8
        for (int i = 0; i < NH; ++i)
            s = fmaf(B[i][t], bf16_to_float(wl[OFF_W3 + i * NH + j]), s);
        A[j][t] = fmaxf(s, 0.f);
    }
    if (h2)
        for (int j = 0; j < NH; ++j)
            h2[j] = B[j][t];
    float z = bl[OFF_B4];
#pragma unroll 8
    for (int i = 0; i < NH; ++i)
        z = fmaf(A[i][t], bf16_to_float(wl[OFF_W4 + i]), z);
    if (h3)
        for (int j = 0; j < NH; ++j)
            h3[j] = A[j][t];
    return z;
}

__device__ inline float sigmoid(float z) { return 1.f / (1.f + expf(-z)); }

// Launchers (m1_mlp.hip).  Arrays are device pointers; x is (n, NIN).
hipError_t launch_forward(DeviceParams prm, const float* x, int n, float* logit, float* acts,
                          hipStream_t s);
hipError_t launch_backward(DeviceParams prm, const float* acts, const float* logit,
                           const float* y, int n, float* deltas, float* loss, hipStream_t s);
hipError_t launch_weight_grad(const float* x, const float* acts, const float* deltas, int n,
                              float* grad, hipStream_t s);
hipError_t launch_adam(float* p, uint16_t* w16, const float* grad, float* m, float* v,
                       float lr, float beta1, float beta2, float eps, int step, hipStream_t s);
hipError_t launch_round_bf16(const float* p, uint16_t* w16, hipStream_t s);
// Device parameters of a model created by m1_mlp_create.
DeviceParams device_params_of(const m1_mlp* model);
int device_of(const m1_mlp* model);
// Waits for the model's queued work (training, parameter uploads).
hipError_t finish(const m1_mlp* model);

}  // namespace mlp
}  // namespace m1
