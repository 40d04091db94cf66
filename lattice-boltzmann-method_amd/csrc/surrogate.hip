// Learned chemistry surrogate: a small MLP (T, p, Y_k, dt) -> (dT, dY_k)
// trained and evaluated on the GPU with BF16 matrix-core (MFMA) GEMMs and
// FP32 accumulation, Adam on FP32 master weights.
//
// Layers: 16 (12 features, zero-padded) -> 128 -> 128 -> 128 -> 16 (10
// outputs, zero-padded), GELU on the hidden layers.  Every GEMM is
// C[m][n] = sum_k A[m][k] B[n][k] (+ bias) on 16x16 output tiles, one
// wavefront per tile, v_mfma_f32_16x16x16_bf16 with 4 BF16 of A and of B per
// lane per K-step of 16:
//   A fragment: row m0 + (lane % 16), k = k0 + 4 (lane / 16) + 0..3
//   B fragment: col n0 + (lane % 16), k = k0 + 4 (lane / 16) + 0..3
//   D fragment: col n0 + (lane % 16), row m0 + 4 (lane / 16) + 0..3
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "surrogate.h"

namespace lbm {
namespace sur {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x)
{
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// One 16x16 tile of C = A B^T with A [M][lda] and B [N][ldb] in BF16 (K
// multiple of 16, M and N multiples of 16).
__device__ inline v4f tile_mma(const uint16_t* A, int lda, const uint16_t* B, int ldb, int m0, int n0, int K)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        const uint16_t* pa = A + (size_t)(m0 + r) * lda + k0 + kq;
        const uint16_t* pb = B + (size_t)(n0 + r) * ldb + k0 + kq;
        v4s a, b;
        a[0] = (short)pa[0]; a[1] = (short)pa[1]; a[2] = (short)pa[2]; a[3] = (short)pa[3];
        b[0] = (short)pb[0]; b[1] = (short)pb[1]; b[2] = (short)pb[2]; b[3] = (short)pb[3];
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Layer forward: Z = X W^T + b (FP32, kept for the backward), H = act(Z) in
// BF16 for the next layer.  Grid: x = M/16 row tiles, y = N/16 column tiles
// (one wavefront per tile, 64 threads per block).
__global__ __launch_bounds__(64) void k_layer_fwd(const uint16_t* X, int K, const uint16_t* W, const float* b, int N,
                                                  float* Z, uint16_t* H, int act, int M)
{
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    if (m0 >= M) return;
    const v4f acc = tile_mma(X, K, W, K, m0, n0, K);
    const int lane = threadIdx.x & 63;
    const int col = n0 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = m0 + (lane >> 4) * 4 + v;
        const float z = acc[v] + b[col];
        if (Z) Z[(size_t)row * N + col] = z;
        if (H) H[(size_t)row * N + col] = f2bf(act ? gelu(z) : z);
    }
}

// dZ of the output layer for the mean-squared loss over `nout` outputs:
// dZ = 2 (z - t) / (M nout); padded outputs get 0.  Also the per-row loss.
__global__ void k_loss_grad(const float* Z, const float* T, int M, int N, int nout, uint16_t* dZb, float* loss_rows)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= M) return;
    float s = 0.f;
    for (int c = 0; c < N; ++c) {
        float g = 0.f;
        if (c < nout) {
            const float d = Z[(size_t)row * N + c] - T[(size_t)row * N + c];
            s += d * d;
            g = 2.f * d / (float)(M * nout);
        }
        dZb[(size_t)row * N + c] = f2bf(g);
    }
    loss_rows[row] = s;
}

// dX = dZ W (rows of W^T given as WT [K][N]); then for a hidden layer below,
// dZ_below = dX * gelu'(Z_below).
__global__ __launch_bounds__(64) void k_layer_bwd_x(const uint16_t* dZ, int N, const uint16_t* WT, int K,
                                                    const float* Zbelow, uint16_t* dZbelow, int M)
{
    const int m0 = blockIdx.x * 16, k0 = blockIdx.y * 16;
    if (m0 >= M) return;
    const v4f acc = tile_mma(dZ, N, WT, N, m0, k0, N);
    const int lane = threadIdx.x & 63;
    const int col = k0 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = m0 + (lane >> 4) * 4 + v;
        const float g = acc[v] * gelu_grad(Zbelow[(size_t)row * K + col]);
        dZbelow[(size_t)row * K + col] = f2bf(g);
    }
}

// Transposes a BF16 [M][N] matrix into [N][M] (for the weight gradient the
// batch is the reduction index).
__global__ void k_transpose_bf(const uint16_t* in, int M, int N, uint16_t* out)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    for (int n = 0; n < N; ++n) out[(size_t)n * M + m] = in[(size_t)m * N + n];
}

// Partial weight gradient over one chunk of rows: dW[n][k] = sum_m dZ[m][n] X[m][k]
// with dZT [N][M] and XT [K][M]; chunk c covers rows [c*chunk, (c+1)*chunk).
__global__ __launch_bounds__(64) void k_wgrad(const uint16_t* dZT, const uint16_t* XT, int M, int N, int K, int chunk,
                                              float* part)
{
    const int n0 = blockIdx.x * 16, k0 = blockIdx.y * 16, c = blockIdx.z;
    const int mbeg = c * chunk;
    const int len = min(chunk, M - mbeg);
    const v4f acc = tile_mma(dZT + mbeg, M, XT + mbeg, M, n0, k0, len);
    const int lane = threadIdx.x & 63;
    const int col = k0 + (lane & 15);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int row = n0 + (lane >> 4) * 4 + v;
        part[((size_t)c * N + row) * K + col] = acc[v];
    }
}

// Sums the chunk partials in a fixed order (deterministic) and the bias
// gradient db[n] = sum_m dZ[m][n].
__global__ void k_reduce_grad(const float* part, int nchunk, int N, int K, float* dW, const uint16_t* dZT, int M,
                              float* db)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N * K) {
        float s = 0.f;
        for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * N * K + i];
        dW[i] = s;
    }
    if (i < N) {
        float s = 0.f;
        for (int m = 0; m < M; ++m) s += bf2f(dZT[(size_t)i * M + m]);
        db[i] = s;
    }
}

// Adam on FP32 master weights; refreshes the BF16 copies W and W^T.
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t* wb, uint16_t* wbT, int rows, int cols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) {
        wb[i] = f2bf(wi);
        const int r = i / cols, c = i % cols;
        wbT[(size_t)c * rows + r] = f2bf(wi);
    }
}

__global__ void k_to_bf(const float* in, uint16_t* out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

__global__ void k_to_bf_T(const float* in, int rows, int cols, uint16_t* out, uint16_t* outT)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const uint16_t h = f2bf(in[i]);
    out[i] = h;
    outT[(size_t)(i % cols) * rows + i / cols] = h;
}

// ------------------------------------------------------------------ host
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ------------------------------------------------- fused inference kernel
// One block of 256 threads (4 wavefronts) evaluates the whole network for 64
// rows with the activations kept in LDS: each wavefront owns 16 rows and all
// output tiles of a layer.  Rows are fed by `features(row, out16)`.
constexpr int IROWS = 64;

__device__ inline void lds_layer(const uint16_t* in, int K, const uint16_t* W, const float* b, int N, uint16_t* out,
                                 float* outf, bool act)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    const int m0 = wave * 16;
    for (int n0 = 0; n0 < N; n0 += 16) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < K; k0 += 16) {
            const uint16_t* pa = in + (m0 + r) * K + k0 + kq;
            const uint16_t* pb = W + (size_t)(n0 + r) * K + k0 + kq;
            v4s a, bb;
            a[0] = (short)pa[0]; a[1] = (short)pa[1]; a[2] = (short)pa[2]; a[3] = (short)pa[3];
            bb[0] = (short)pb[0]; bb[1] = (short)pb[1]; bb[2] = (short)pb[2]; bb[3] = (short)pb[3];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
        }
        const int col = n0 + r;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = m0 + (lane >> 4) * 4 + v;
            const float z = acc[v] + b[col];
            if (out) out[row * N + col] = f2bf(act ? gelu(z) : z);
            if (outf) outf[row * N + col] = z;
        }
    }
}

__global__ __launch_bounds__(256) void k_mlp_infer(Mlp net, const float* X, int M, float* Y)
{
    __shared__ uint16_t h0[IROWS * 16];
    __shared__ uint16_t h1[IROWS * 128];
    __shared__ uint16_t h2[IROWS * 128];
    __shared__ float yout[IROWS * 16];
    const int base = blockIdx.x * IROWS;
    for (int i = threadIdx.x; i < IROWS * 16; i += 256) {
        const int row = base + i / 16;
        h0[i] = f2bf(row < M ? X[(size_t)row * 16 + (i % 16)] : 0.f);
    }
    __syncthreads();
    lds_layer(h0, 16, net.Wb[0], net.b[0], 128, h1, nullptr, true);
    __syncthreads();
    lds_layer(h1, 128, net.Wb[1], net.b[1], 128, h2, nullptr, true);
    __syncthreads();
    lds_layer(h2, 128, net.Wb[2], net.b[2], 128, h1, nullptr, true);
    __syncthreads();
    lds_layer(h1, 128, net.Wb[3], net.b[3], 16, nullptr, yout, false);
    __syncthreads();
    for (int i = threadIdx.x; i < IROWS * 16; i += 256) {
        const int row = base + i / 16;
        if (row < M) Y[(size_t)row * 16 + (i % 16)] = yout[i];
    }
}

void mlp_infer(const Mlp& net, const float* X, int M, float* Y, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_infer, dim3(cdiv(M, IROWS)), dim3(256), 0, s, net, X, M, Y);
}

void mlp_forward(const Mlp& net, const uint16_t* X, int M, Work& w, hipStream_t s)
{
    const uint16_t* in = X;
    for (int L = 0; L < NL; ++L) {
        const int K = DIMS[L], N = DIMS[L + 1];
        const bool last = (L == NL - 1);
        hipLaunchKernelGGL(k_layer_fwd, dim3(M / 16, N / 16), dim3(64), 0, s, in, K, net.Wb[L], net.b[L], N, w.Z[L],
                           last ? (uint16_t*)nullptr : w.H[L], last ? 0 : 1, M);
        in = w.H[L];
    }
}

void mlp_grad(const Mlp& net, const uint16_t* X, const float* T, int M, Work& w, hipStream_t s)
{
    mlp_forward(net, X, M, w, s);
    const int L3 = NL - 1;
    hipLaunchKernelGGL(k_loss_grad, dim3(cdiv(M, 256)), dim3(256), 0, s, w.Z[L3], T, M, DIMS[NL], NOUT, w.dZ[L3],
                       w.loss_rows);
    for (int L = NL - 1; L >= 0; --L) {
        const int K = DIMS[L], N = DIMS[L + 1];
        const uint16_t* Xin = L == 0 ? X : w.H[L - 1];
        hipLaunchKernelGGL(k_transpose_bf, dim3(cdiv(M, 256)), dim3(256), 0, s, w.dZ[L], M, N, w.dZT);
        hipLaunchKernelGGL(k_transpose_bf, dim3(cdiv(M, 256)), dim3(256), 0, s, Xin, M, K, w.XT);
        const int nchunk = cdiv(M, w.chunk);
        hipLaunchKernelGGL(k_wgrad, dim3(N / 16, K / 16, nchunk), dim3(64), 0, s, w.dZT, w.XT, M, N, K, w.chunk,
                           w.part);
        hipLaunchKernelGGL(k_reduce_grad, dim3(cdiv(N * K, 256)), dim3(256), 0, s, w.part, nchunk, N, K, w.gW[L],
                           w.dZT, M, w.gb[L]);
        if (L > 0)
            hipLaunchKernelGGL(k_layer_bwd_x, dim3(M / 16, K / 16), dim3(64), 0, s, w.dZ[L], N, net.WbT[L], K,
                               w.Z[L - 1], w.dZ[L - 1], M);
    }
}

void mlp_adam(Mlp& net, Work& w, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    for (int L = 0; L < NL; ++L) {
        const int n = DIMS[L] * DIMS[L + 1];
        hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, net.W[L], w.gW[L], net.mW[L], net.vW[L], n, lr,
                           b1, b2, eps, c1, c2, net.Wb[L], net.WbT[L], DIMS[L + 1], DIMS[L]);
        hipLaunchKernelGGL(k_adam, dim3(cdiv(DIMS[L + 1], 256)), dim3(256), 0, s, net.b[L], w.gb[L], net.mb[L],
                           net.vb[L], DIMS[L + 1], lr, b1, b2, eps, c1, c2, (uint16_t*)nullptr, (uint16_t*)nullptr, 0,
                           1);
    }
}

void to_bf16(const float* in, uint16_t* out, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_to_bf, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n);
}

void refresh_bf16(Mlp& net, hipStream_t s)
{
    for (int L = 0; L < NL; ++L) {
        const int n = DIMS[L] * DIMS[L + 1];
        hipLaunchKernelGGL(k_to_bf_T, dim3(cdiv(n, 256)), dim3(256), 0, s, net.W[L], DIMS[L + 1], DIMS[L], net.Wb[L],
                           net.WbT[L]);
    }
}

}  // namespace sur
}  // namespace lbm

namespace lbm {
namespace sur {

// Glorot-uniform initial weights from a SplitMix64 stream, zero biases.
hipError_t mlp_alloc(Mlp& net, uint64_t seed, hipStream_t s)
{
    uint64_t st = seed;
    auto rnd = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    hipError_t e;
    for (int L = 0; L < NL; ++L) {
        const int rows = DIMS[L + 1], cols = DIMS[L], n = rows * cols;
        float** f32s[] = {&net.W[L], &net.mW[L], &net.vW[L]};
        for (float** p : f32s)
            if ((e = hipMalloc((void**)p, sizeof(float) * n)) != hipSuccess) return e;
        float** bs[] = {&net.b[L], &net.mb[L], &net.vb[L]};
        for (float** p : bs)
            if ((e = hipMalloc((void**)p, sizeof(float) * rows)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&net.Wb[L], 2 * n)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&net.WbT[L], 2 * n)) != hipSuccess) return e;
        // real fan-in / fan-out (padding excluded)
        const int fin = L == 0 ? NIN : cols, fout = L == NL - 1 ? NOUT : rows;
        const double lim = sqrt(6.0 / (fin + fout));
        float* h = new float[n];
        for (int r = 0; r < rows; ++r)
            for (int c = 0; c < cols; ++c) {
                const bool pad = (L == 0 && c >= NIN) || (L == NL - 1 && r >= NOUT);
                h[r * cols + c] = pad ? 0.f : (float)((2.0 * rnd() - 1.0) * lim);
            }
        (void)hipMemcpyAsync(net.W[L], h, sizeof(float) * n, hipMemcpyHostToDevice, s);
        (void)hipStreamSynchronize(s);
        delete[] h;
        (void)hipMemsetAsync(net.mW[L], 0, sizeof(float) * n, s);
        (void)hipMemsetAsync(net.vW[L], 0, sizeof(float) * n, s);
        (void)hipMemsetAsync(net.b[L], 0, sizeof(float) * rows, s);
        (void)hipMemsetAsync(net.mb[L], 0, sizeof(float) * rows, s);
        (void)hipMemsetAsync(net.vb[L], 0, sizeof(float) * rows, s);
    }
    refresh_bf16(net, s);
    return hipStreamSynchronize(s);
}

void mlp_free(Mlp& net)
{
    for (int L = 0; L < NL; ++L) {
        void* ps[] = {net.W[L], net.b[L], net.mW[L], net.vW[L], net.mb[L], net.vb[L], net.Wb[L], net.WbT[L]};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    net = Mlp{};
}

hipError_t work_alloc(Work& w, int cap)
{
    hipError_t e;
    w.cap = cap;
    int maxd = 0;
    for (int L = 0; L <= NL; ++L) maxd = DIMS[L] > maxd ? DIMS[L] : maxd;
    for (int L = 0; L < NL; ++L) {
        const int N = DIMS[L + 1];
        if ((e = hipMalloc((void**)&w.Z[L], sizeof(float) * (size_t)cap * N)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&w.H[L], 2 * (size_t)cap * N)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&w.dZ[L], 2 * (size_t)cap * N)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&w.gW[L], sizeof(float) * DIMS[L] * N)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&w.gb[L], sizeof(float) * N)) != hipSuccess) return e;
    }
    if ((e = hipMalloc((void**)&w.dZT, 2 * (size_t)cap * maxd)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&w.XT, 2 * (size_t)cap * maxd)) != hipSuccess) return e;
    const int nchunk = (cap + w.chunk - 1) / w.chunk;
    if ((e = hipMalloc((void**)&w.part, sizeof(float) * (size_t)nchunk * maxd * maxd)) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&w.loss_rows, sizeof(float) * (size_t)cap)) != hipSuccess) return e;
    return hipSuccess;
}

void work_free(Work& w)
{
    for (int L = 0; L < NL; ++L) {
        void* ps[] = {w.Z[L], w.H[L], w.dZ[L], w.gW[L], w.gb[L]};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    void* ps[] = {w.dZT, w.XT, w.part, w.loss_rows};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    w = Work{};
}

}  // namespace sur
}  // namespace lbm
