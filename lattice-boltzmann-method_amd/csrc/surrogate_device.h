// Device building blocks of the chemistry surrogate MLP shared by the
// training kernels (surrogate.hip) and the fused in-step inference kernel
// (lbm_multi.hip).
//
// v_mfma_f32_16x16x16_bf16 operand layout (64 lanes, 4 BF16 per operand and
// lane, 4 FP32 accumulators per lane):
//   A[i][k]: i = lane % 16, k = 4 (lane / 16) + 0..3
//   B[k][j]: j = lane % 16, k = 4 (lane / 16) + 0..3
//   D[i][j]: j = lane % 16, i = 4 (lane / 16) + 0..3
//
// In-step inference computes every layer transposed, Z^T = W H^T: the weights
// are the A operand (i = neuron) and the activations the B operand (j = batch
// row).  The D fragment of output tile t then holds neurons 16 t + 4 (lane / 16)
// + 0..3 of row lane % 16, which is exactly the B fragment of K-step t of the
// next layer, so activations stay in registers from layer to layer.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "surrogate.h"

namespace lbm {
namespace sur {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN and Inf keep their class).
__host__ __device__ inline uint16_t f2bf(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// exact GELU, x Phi(x), and its derivative Phi(x) + x phi(x)
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// Fragment-ordered BF16 weights: for layer L with NT = N/16 output tiles and
// KS = K/16 K-steps, element ((t KS + s) 64 + lane) 4 + q holds
// W[16 t + lane % 16][16 s + 4 (lane / 16) + q], so a wavefront reads one
// A fragment as 64 consecutive 8-byte words.
constexpr int frag_offset(int L)
{
    int off = 0;
    for (int i = 0; i < L; ++i) off += DIMS[i] * DIMS[i + 1];
    return off;
}
constexpr int FRAG_TOTAL = frag_offset(NL);      // 36864 BF16 = 72 KiB
constexpr int BIAS_TOTAL = 128 + 128 + 128 + 16;  // FP32
constexpr int bias_offset(int L)
{
    int off = 0;
    for (int i = 0; i < L; ++i) off += DIMS[i + 1];
    return off;
}

__device__ inline v4s load_frag(const uint16_t* p)
{
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    v4s r;
    r[0] = (short)(w.x & 0xffffu);
    r[1] = (short)(w.x >> 16);
    r[2] = (short)(w.y & 0xffffu);
    r[3] = (short)(w.y >> 16);
    return r;
}

// BF16(GELU(acc + bias)) for the 4 neurons of one D fragment.
__device__ inline v4s act_frag(const v4f acc, const float* bias, int n0, int lane)
{
    v4s h;
#pragma unroll
    for (int q = 0; q < 4; ++q) h[q] = (short)f2bf(gelu(acc[q] + bias[n0 + 4 * (lane >> 4) + q]));
    return h;
}

// Evaluates the network for 32 rows (two batch tiles) held by one wavefront.
// `Wf` and `bias` are the fragment-ordered weights and the biases (LDS or
// global), `xin` the BF16 inputs [32][16] of this wavefront and `yout` its
// FP32 outputs [32][16] (both LDS).
__device__ inline void mlp_rows32(const uint16_t* Wf, const float* bias, const uint16_t* xin, float* yout)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    constexpr int H = 128 / 16;  // hidden tiles
    v4s act[2][H], nxt[2][H];
    // layer 0: K = 16, one K-step
    {
        const uint16_t* W0 = Wf + frag_offset(0);
        v4s x[2];
#pragma unroll
        for (int bt = 0; bt < 2; ++bt) x[bt] = load_frag(xin + (bt * 16 + r) * 16 + kq);
#pragma unroll
        for (int t = 0; t < H; ++t) {
            const v4s a = load_frag(W0 + (t * 64 + lane) * 4);
#pragma unroll
            for (int bt = 0; bt < 2; ++bt) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, x[bt], acc, 0, 0, 0);
                act[bt][t] = act_frag(acc, bias + bias_offset(0), 16 * t, lane);
            }
        }
    }
    // layers 1 and 2: 128 -> 128
#pragma unroll
    for (int L = 1; L <= 2; ++L) {
        const uint16_t* WL = Wf + frag_offset(L);
        const float* bL = bias + bias_offset(L);
#pragma unroll
        for (int t = 0; t < H; ++t) {
            v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < H; ++s) {
                const v4s a = load_frag(WL + ((t * H + s) * 64 + lane) * 4);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[0][s], acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[1][s], acc1, 0, 0, 0);
            }
            nxt[0][t] = act_frag(acc0, bL, 16 * t, lane);
            nxt[1][t] = act_frag(acc1, bL, 16 * t, lane);
        }
#pragma unroll
        for (int t = 0; t < H; ++t) {
            act[0][t] = nxt[0][t];
            act[1][t] = nxt[1][t];
        }
    }
    // layer 3: 128 -> 16, no activation
    const uint16_t* W3 = Wf + frag_offset(3);
    const float* b3 = bias + bias_offset(3);
    v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < H; ++s) {
        const v4s a = load_frag(W3 + (s * 64 + lane) * 4);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[0][s], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[1][s], acc1, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int nrn = kq + q;
        yout[r * 16 + nrn] = acc0[q] + b3[nrn];
        yout[(16 + r) * 16 + nrn] = acc1[q] + b3[nrn];
    }
}

// Stages the fragment-ordered weights and biases of `net` into LDS.
__device__ inline void stage_net(const Mlp& net, uint16_t* Wf, float* bias)
{
    const uint4* src = reinterpret_cast<const uint4*>(net.Wf);
    uint4* dst = reinterpret_cast<uint4*>(Wf);
    for (int i = threadIdx.x; i < FRAG_TOTAL / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < BIAS_TOTAL; i += blockDim.x) {
        int L = 0;
        while (L < NL - 1 && i >= bias_offset(L + 1)) ++L;
        bias[i] = net.b[L][i - bias_offset(L)];
    }
}

}  // namespace sur
}  // namespace lbm
