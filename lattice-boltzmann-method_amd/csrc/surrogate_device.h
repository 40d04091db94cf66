// Device building blocks of the chemistry surrogate MLP shared by the
// training kernels (surrogate.hip) and the fused in-step inference kernel
// (lbm_multi.hip).
//
// v_mfma_f32_16x16x16_bf16 operand layout (64 lanes, 4 BF16 per operand and
// lane, 4 FP32 accumulators per lane):
//   A[i][k]: i = lane % 16, k = 4 (lane / 16) + 0..3
//   B[k][j]: j = lane % 16, k = 4 (lane / 16) + 0..3
//   D[i][j]: j = lane % 16, i = 4 (lane / 16) + 0..3
// v_mfma_f32_16x16x32_bf16 (gfx950, twice the K per instruction at the same
// issue cost): 8 BF16 per operand and lane, k = 8 (lane / 16) + 0..7, the same
// D layout.
//
// In-step inference computes every layer transposed, Z^T = W H^T: the weights
// are the A operand (i = neuron) and the activations the B operand (j = batch
// row).  The D fragment of output tile t then holds neurons 16 t + 4 (lane / 16)
// + 0..3 of row lane % 16.  Layer 0 (K = 16) runs 16x16x16; the 128-wide
// layers run 16x16x32 with K-step s made of the D fragments of tiles 2s and
// 2s + 1 side by side: its element e of lane group g is neuron k32_neuron(s, g,
// e).  The weight fragments are stored in that K order, so activations stay
// in registers from layer to layer and no lane exchanges data.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "surrogate.h"

namespace lbm {
namespace sur {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Neuron held by element e (0..7) of lane group g in 32-deep K-step s.
__host__ __device__ constexpr int k32_neuron(int s, int g, int e)
{
    return 32 * s + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

// FP32 -> BF16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950).
__device__ __forceinline__ uint16_t f2bf(float x)
{
    const __bf16 h = (__bf16)x;
    uint16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// GELU in its tanh form, x (1 + tanh(v)) / 2 with v = sqrt(2/pi) (x + 0.044715
// x^3), written as x / (1 + exp(-2 v)); and its derivative s + x s (1 - s) 2 v'
// with s = 1 / (1 + exp(-2 v)).
constexpr float GELU_C = 1.5957691216057308f;  // 2 sqrt(2 / pi)
constexpr float GELU_A = 0.044715f;
// (v_exp_f32 and v_rcp_f32: ~1 ulp, far below the BF16 rounding that follows;
// exp(-u) = 2^(x (K1 + K2 x^2)))
constexpr float GELU_K1 = -GELU_C * 1.4426950408889634f, GELU_K2 = -GELU_C * GELU_A * 1.4426950408889634f;
__device__ __forceinline__ float gelu(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * __builtin_fmaf(GELU_K2, x * x, GELU_K1));
    return x * __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_grad(float x)
{
    const float u = GELU_C * x * (1.0f + GELU_A * x * x);
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-u));
    return s + x * s * (1.0f - s) * (GELU_C * (1.0f + 3.0f * GELU_A * x * x));
}

// Fragment-ordered BF16 weights.  Layer 0 (K = 16): element (t 64 + lane) 4 + q
// holds W[16 t + lane % 16][4 (lane / 16) + q].  Layers with K = 32 KS:
// element ((t KS + s) 64 + lane) 8 + e holds W[16 t + lane % 16][k32_neuron(s,
// lane / 16, e)].  A wavefront reads one A fragment as 64 consecutive 8- or
// 16-byte words.
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

__device__ __forceinline__ v8bf load_frag8(const uint16_t* p)
{
    const uint4 w = *reinterpret_cast<const uint4*>(p);
    v8bf r;
    __builtin_memcpy(&r, &w, 16);
    return r;
}
// The B operand of a 32-deep K-step from the activations of two output tiles.
__device__ __forceinline__ v8bf pair_frag(const v4s lo, const v4s hi)
{
    short t[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    v8bf r;
    __builtin_memcpy(&r, t, 16);
    return r;
}

__device__ __forceinline__ v4s load_frag(const uint16_t* p)
{
    const uint2 w = *reinterpret_cast<const uint2*>(p);
    v4s r;
    r[0] = (short)(w.x & 0xffffu);
    r[1] = (short)(w.x >> 16);
    r[2] = (short)(w.y & 0xffffu);
    r[3] = (short)(w.y >> 16);
    return r;
}

// The bias of the 4 neurons of one D fragment (the accumulator's start).
__device__ __forceinline__ v4f bias_frag(const float* bias, int n0, int lane)
{
    const float* b = bias + n0 + 4 * (lane >> 4);
    return v4f{b[0], b[1], b[2], b[3]};
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// BF16(GELU(z)) for the 4 neurons of one D fragment, converted in pairs.
__device__ __forceinline__ v4s act_frag(const v4f z)
{
    const bf16x2 lo = __builtin_convertvector((f32x2){gelu(z[0]), gelu(z[1])}, bf16x2);
    const bf16x2 hi = __builtin_convertvector((f32x2){gelu(z[2]), gelu(z[3])}, bf16x2);
    uint32_t w[2];
    __builtin_memcpy(&w[0], &lo, 4);
    __builtin_memcpy(&w[1], &hi, 4);
    v4s h;
    __builtin_memcpy(&h, w, 8);
    return h;
}

// Evaluates the network for 32 rows (two batch tiles) held by one wavefront.
// `Wf` and `bias` are the fragment-ordered weights and the biases (LDS), `xin`
// the BF16 inputs [32][16] of this wavefront and `yout` its FP32 outputs
// [32][16] (LDS or global; 16-byte aligned).  Each accumulator starts from
// the layer's bias.
__device__ __forceinline__ void mlp_rows32(const uint16_t* Wf, const float* bias, const uint16_t* xin, float* yout)
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
            const v4f b = bias_frag(bias + bias_offset(0), 16 * t, lane);
#pragma unroll
            for (int bt = 0; bt < 2; ++bt)
                act[bt][t] = act_frag(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, x[bt], b, 0, 0, 0));
        }
    }
    // layers 1 and 2: 128 -> 128, 32-deep K-steps
    constexpr int KS = 128 / 32;
#pragma unroll
    for (int L = 1; L <= 2; ++L) {
        const uint16_t* WL = Wf + frag_offset(L);
        const float* bL = bias + bias_offset(L);
#pragma unroll
        for (int t = 0; t < H; ++t) {
            v4f acc0 = bias_frag(bL, 16 * t, lane), acc1 = acc0;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const v8bf a = load_frag8(WL + ((t * KS + s) * 64 + lane) * 8);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pair_frag(act[0][2 * s], act[0][2 * s + 1]), acc0,
                                                                0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pair_frag(act[1][2 * s], act[1][2 * s + 1]), acc1,
                                                                0, 0, 0);
            }
            nxt[0][t] = act_frag(acc0);
            nxt[1][t] = act_frag(acc1);
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
    v4f acc0 = bias_frag(b3, 0, lane), acc1 = acc0;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const v8bf a = load_frag8(W3 + (s * 64 + lane) * 8);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pair_frag(act[0][2 * s], act[0][2 * s + 1]), acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, pair_frag(act[1][2 * s], act[1][2 * s + 1]), acc1, 0, 0, 0);
    }
    *reinterpret_cast<v4f*>(yout + r * 16 + kq) = acc0;
    *reinterpret_cast<v4f*>(yout + (16 + r) * 16 + kq) = acc1;
}

// Stages the fragment-ordered weights and biases of `net` into LDS.
__device__ __forceinline__ void stage_net(const Mlp& net, uint16_t* Wf, float* bias)
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
