// Learned chemistry surrogate: a small MLP (T, p, Y_k, dt) -> (dT, dY_k)
// trained and evaluated on the GPU with BF16 matrix-core (MFMA) GEMMs and
// FP32 accumulation, Adam on FP32 master weights.
//
// Layers: 16 (NIN = NS + 3 features, zero-padded) -> 128 -> 128 -> 128 -> 16
// (NOUT = NS + 1 outputs, zero-padded), GELU (tanh form) on the hidden layers.  Every training GEMM is
// C[m][n] = sum_k A[m][k] B[n][k] (+ bias) on 16x16 output tiles, one
// wavefront per tile, v_mfma_f32_16x16x16_bf16 with 4 BF16 of A and of B per
// lane per K-step of 16 (operand layout in surrogate_device.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "surrogate.h"

#include <utility>
#include "surrogate_device.h"

namespace lbm {
namespace sur {

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
// batch is the reduction index): 64 x 64 tiles through LDS, so that the
// reads run along rows of `in` and the writes along rows of `out`.
// grid (ceil(M / 64), ceil(N / 64)), 256 threads.
__global__ __launch_bounds__(256) void k_transpose_bf(const uint16_t* in, int M, int N, uint16_t* out)
{
    __shared__ uint16_t tile[64][65];
    const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4) {
        const int m = m0 + r, n = n0 + tx;
        tile[r][tx] = (m < M && n < N) ? in[(size_t)m * N + n] : (uint16_t)0;
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int n = n0 + r, m = m0 + tx;
        if (n < N && m < M) out[(size_t)n * M + m] = tile[tx][r];
    }
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

// Sums the chunk partials in a fixed order (deterministic).
__global__ void k_reduce_grad(const float* part, int nchunk, int N, int K, float* dW)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * K) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; ++c) s += part[(size_t)c * N * K + i];
    dW[i] = s;
}

// Bias gradient db[n] = sum_m dZ[m][n]: one block per n, strided partial
// sums reduced in LDS in a fixed order.
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t* dZT, int M, float* db)
{
    __shared__ float red[256];
    const int nrow = blockIdx.x;
    float s = 0.f;
    for (int m = threadIdx.x; m < M; m += 256) s += bf2f(dZT[(size_t)nrow * M + m]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[nrow] = red[0];
}

// Adam on FP32 master weights; refreshes the BF16 copies W and W^T.  The
// bias corrections of iteration *it are c12[2 it] and c12[2 it + 1].
__global__ void k_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       const float* c12, const int* it, uint16_t* wb, uint16_t* wbT, int rows, int cols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float c1 = c12[2 * *it], c2 = c12[2 * *it + 1];
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

// W[L] (FP32, [N][K]) into the fragment-ordered BF16 block of layer L: the
// 16-deep K-steps of layer 0, the 32-deep ones (k32_neuron order) of the others.
__global__ void k_to_frag(const float* W, int N, int K, uint16_t* Wf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * K) return;
    if (K == 16) {
        const int q = i & 3, lane = (i >> 2) & 63, t = i >> 8;
        const int n = 16 * t + (lane & 15), k = 4 * (lane >> 4) + q;
        Wf[i] = f2bf(W[n * K + k]);
        return;
    }
    const int KS = K / 32;
    const int e = i & 7, lane = (i >> 3) & 63, ts = i >> 9;
    const int t = ts / KS, s = ts % KS;
    const int n = 16 * t + (lane & 15), k = k32_neuron(s, lane >> 4, e);
    Wf[i] = f2bf(W[n * K + k]);
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

// ------------------------------------------------------ inference kernel
// 512 threads (8 wavefronts x 32 rows), weights and biases staged in LDS
// once per block, grid-stride over groups of 256 rows.
constexpr int IWAVES = 8, IROWS = IWAVES * 32;

__global__ __launch_bounds__(512) void k_mlp_infer(Mlp net, const float* X, int M, float* Y)
{
    __shared__ __align__(16) uint16_t Wf[FRAG_TOTAL];
    __shared__ float bias[BIAS_TOTAL];
    __shared__ __align__(16) uint16_t xin[IWAVES][32 * 16];
    __shared__ __align__(16) float yout[IWAVES][32 * 16];
    stage_net(net, Wf, bias);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int base = blockIdx.x * IROWS; base < M; base += gridDim.x * IROWS) {
        __syncthreads();
        for (int i = lane; i < 32 * 16; i += 64) {
            const int row = base + wave * 32 + i / 16;
            xin[wave][i] = f2bf(row < M ? X[(size_t)row * 16 + (i % 16)] : 0.f);
        }
        __syncthreads();
        mlp_rows32(Wf, bias, xin[wave], yout[wave]);
        __syncthreads();
        for (int i = lane; i < 32 * 16; i += 64) {
            const int row = base + wave * 32 + i / 16;
            if (row < M) Y[(size_t)row * 16 + (i % 16)] = yout[wave][i];
        }
    }
}

int infer_grid(int M)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int groups = cdiv(M, IROWS);
    return groups < cus ? (groups > 0 ? groups : 1) : cus;
}

void mlp_infer(const Mlp& net, const float* X, int M, float* Y, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_infer, dim3(infer_grid(M)), dim3(512), 0, s, net, X, M, Y);
}

// In-step network pass: inputs already BF16 in HBM, 8 wavefronts x 32 rows
// per pass of a 512-thread block, grid-stride; rows is a multiple of 256.
__global__ __launch_bounds__(512) void k_mlp_rows(Mlp net, const uint16_t* X, int64_t rows, float* Y)
{
    __shared__ __align__(16) uint16_t Wf[FRAG_TOTAL];
    __shared__ float bias[BIAS_TOTAL];
    stage_net(net, Wf, bias);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    for (int64_t base = (int64_t)blockIdx.x * IROWS; base < rows; base += (int64_t)gridDim.x * IROWS) {
        // keeps the weight fragments' LDS reads inside the loop (hoisted out
        // of it they would need 576 registers)
        asm volatile("" ::: "memory");
        const int64_t r0 = base + wave * 32;
        mlp_rows32(Wf, bias, X + r0 * 16, Y + r0 * 16);
    }
}

void mlp_infer_bf16(const Mlp& net, const uint16_t* X, int64_t rows, float* Y, hipStream_t s)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const int64_t groups = rows / IROWS;
    const int grid = (int)(groups < 2 * cus ? (groups > 0 ? groups : 1) : 2 * cus);
    hipLaunchKernelGGL(k_mlp_rows, dim3(grid), dim3(512), 0, s, net, X, rows, Y);
}

// ---------------------------------------------------- training data
// Per-block partial sum, sum of squares, min and max of the 32 capture
// columns: thread t owns column t % 32 of rows (t / 32) + 8 (block + k grid).
__global__ __launch_bounds__(256) void k_cap_stats(const float* cap, int64_t rows, double* part)
{
    __shared__ double red[256][4];
    const int q = threadIdx.x & 31, g = threadIdx.x >> 5;
    double s = 0, s2 = 0, lo = INFINITY, hi = -INFINITY;
    for (int64_t r = (int64_t)blockIdx.x * 8 + g; r < rows; r += (int64_t)gridDim.x * 8) {
        const double x = cap[r * CAP_STRIDE + q];
        s += x;
        s2 += x * x;
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
    red[threadIdx.x][0] = s;
    red[threadIdx.x][1] = s2;
    red[threadIdx.x][2] = lo;
    red[threadIdx.x][3] = hi;
    __syncthreads();
    if (g == 0) {
        for (int k = 1; k < 8; ++k) {
            const double* o = red[k * 32 + q];
            s += o[0];
            s2 += o[1];
            lo = o[2] < lo ? o[2] : lo;
            hi = o[3] > hi ? o[3] : hi;
        }
        double* p = part + ((size_t)blockIdx.x * 32 + q) * 4;
        p[0] = s;
        p[1] = s2;
        p[2] = lo;
        p[3] = hi;
    }
}

// Final reduction (fixed order) into the normalisation record.  Feature
// scales are at least 1e-3 of the feature's mean, so that rounding-level
// spread (a composition that barely varies) maps to small network inputs;
// the envelope is the captured [min, max] widened by 10 % of its width plus
// 1e-3 of the mean.  A target column with no spread (an inert species) gets
// scale 0: it is not learned and decodes to its mean.
// envelope_only: keep the normalisation (the network's input and output
// scaling) and reset only the training envelope (SS_LO, SS_HI).
__global__ void k_cap_stats_final(const double* part, int nblk, int64_t rows, float* st, int envelope_only)
{
    const int q = threadIdx.x;
    if (q >= 32) return;
    double s = 0, s2 = 0, lo = INFINITY, hi = -INFINITY;
    for (int b = 0; b < nblk; ++b) {
        const double* p = part + ((size_t)b * 32 + q) * 4;
        s += p[0];
        s2 += p[1];
        lo = p[2] < lo ? p[2] : lo;
        hi = p[3] > hi ? p[3] : hi;
    }
    const double mean = s / (double)rows;
    const double var = s2 / (double)rows - mean * mean;
    double sd = var > 0 ? sqrt(var) : 0.0;
    if (q < 16) {
        const bool used = q < NIN;
        sd = fmax(sd, 1e-3 * fabs(mean));
        if (!(sd > 1e-30)) sd = 1.0;
        const double margin = 0.1 * (hi - lo) + 1e-3 * fabs(mean);
        if (!envelope_only) {
            st[SS_FMEAN + q] = used ? (float)mean : 0.f;
            st[SS_FSTD + q] = used ? (float)sd : 1.f;
        }
        st[SS_LO + q] = used ? (float)(lo - margin) : 0.f;
        st[SS_HI + q] = used ? (float)(hi + margin) : 0.f;
    } else if (!envelope_only) {
        if (!(sd > 1e-6 * fabs(mean)) || !(sd > 1e-30)) sd = 0.0;
        const int t = q - 16;
        const bool used = t < NOUT;
        st[SS_TMEAN + t] = used ? (float)mean : 0.f;
        st[SS_TSTD + t] = used ? (float)sd : 0.f;
    }
}

void capture_stats(const float* cap, int64_t rows, double* part, float* stats, hipStream_t s, bool envelope_only)
{
    hipLaunchKernelGGL(k_cap_stats, dim3(STATS_BLOCKS), dim3(256), 0, s, cap, rows, part);
    hipLaunchKernelGGL(k_cap_stats_final, dim3(1), dim3(64), 0, s, part, STATS_BLOCKS, rows, stats,
                       envelope_only ? 1 : 0);
}

__device__ inline uint64_t splitmix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void k_gather(const float* cap, int64_t rows, const float* st, uint64_t seed, const int* itp, int B,
                         uint16_t* X, float* T)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int it = *itp;
    const uint64_t h = splitmix(seed ^ splitmix(((uint64_t)it << 32) | (uint64_t)b));
    const int64_t r = (int64_t)(h % (uint64_t)rows);
    const float* row = cap + r * CAP_STRIDE;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        X[(size_t)b * 16 + q] = f2bf(q < NIN ? (row[q] - st[SS_FMEAN + q]) / st[SS_FSTD + q] : 0.f);
        T[(size_t)b * 16 + q] =
            (q < NOUT && st[SS_TSTD + q] > 0.f) ? (row[16 + q] - st[SS_TMEAN + q]) / st[SS_TSTD + q] : 0.f;
    }
}

void gather_batch(const float* cap, int64_t rows, const float* stats, uint64_t seed, const int* it, int B,
                  uint16_t* X, float* T, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather, dim3(cdiv(B, 256)), dim3(256), 0, s, cap, rows, stats, seed, it, B, X, T);
}

__global__ void k_next_iteration(int* it) { *it += 1; }

void next_iteration(int* it, hipStream_t s) { hipLaunchKernelGGL(k_next_iteration, dim3(1), dim3(1), 0, s, it); }

// out[*idx] (out[0] without idx) = sum(lr[0..B)) * scale
__global__ __launch_bounds__(256) void k_loss_sum(const float* lr, int B, float scale, float* out, const int* idx)
{
    __shared__ float red[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < B; i += 256) s += lr[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[idx ? *idx : 0] = red[0] * scale;
}

void loss_sum(const float* loss_rows, int B, float scale, float* out, hipStream_t s, const int* idx)
{
    hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, loss_rows, B, scale, out, idx);
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
        hipLaunchKernelGGL(k_transpose_bf, dim3(cdiv(M, 64), cdiv(N, 64)), dim3(256), 0, s, w.dZ[L], M, N, w.dZT);
        hipLaunchKernelGGL(k_transpose_bf, dim3(cdiv(M, 64), cdiv(K, 64)), dim3(256), 0, s, Xin, M, K, w.XT);
        const int nchunk = cdiv(M, w.chunk);
        hipLaunchKernelGGL(k_wgrad, dim3(N / 16, K / 16, nchunk), dim3(64), 0, s, w.dZT, w.XT, M, N, K, w.chunk,
                           w.part);
        hipLaunchKernelGGL(k_reduce_grad, dim3(cdiv(N * K, 256)), dim3(256), 0, s, w.part, nchunk, N, K, w.gW[L]);
        hipLaunchKernelGGL(k_bias_grad, dim3(N), dim3(256), 0, s, w.dZT, M, w.gb[L]);
        if (L > 0)
            hipLaunchKernelGGL(k_layer_bwd_x, dim3(M / 16, K / 16), dim3(64), 0, s, w.dZ[L], N, net.WbT[L], K,
                               w.Z[L - 1], w.dZ[L - 1], M);
    }
}

void adam_corrections(int first_step, int iters, float* c12)
{
    for (int k = 0; k < iters; ++k) {
        const float step = (float)(first_step + k);
        c12[2 * k] = 1.f - powf(ADAM_B1, step);
        c12[2 * k + 1] = 1.f - powf(ADAM_B2, step);
    }
}

void mlp_adam(Mlp& net, Work& w, float lr, const float* c12, const int* it, hipStream_t s)
{
    const float b1 = ADAM_B1, b2 = ADAM_B2, eps = ADAM_EPS;
    for (int L = 0; L < NL; ++L) {
        const int n = DIMS[L] * DIMS[L + 1];
        hipLaunchKernelGGL(k_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, net.W[L], w.gW[L], net.mW[L], net.vW[L], n, lr,
                           b1, b2, eps, c12, it, net.Wb[L], net.WbT[L], DIMS[L + 1], DIMS[L]);
        hipLaunchKernelGGL(k_adam, dim3(cdiv(DIMS[L + 1], 256)), dim3(256), 0, s, net.b[L], w.gb[L], net.mb[L],
                           net.vb[L], DIMS[L + 1], lr, b1, b2, eps, c12, it, (uint16_t*)nullptr, (uint16_t*)nullptr, 0,
                           1);
        hipLaunchKernelGGL(k_to_frag, dim3(cdiv(n, 256)), dim3(256), 0, s, net.W[L], DIMS[L + 1], DIMS[L],
                           net.Wf + frag_offset(L));
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
        hipLaunchKernelGGL(k_to_frag, dim3(cdiv(n, 256)), dim3(256), 0, s, net.W[L], DIMS[L + 1], DIMS[L],
                           net.Wf + frag_offset(L));
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
    if ((e = hipMalloc((void**)&net.Wf, 2 * FRAG_TOTAL)) != hipSuccess) return e;
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

hipError_t mlp_alloc_empty(Mlp& net)
{
    hipError_t e;
    if ((e = hipMalloc((void**)&net.Wf, 2 * FRAG_TOTAL)) != hipSuccess) return e;
    for (int L = 0; L < NL; ++L) {
        const int rows = DIMS[L + 1], n = rows * DIMS[L];
        float** f32s[] = {&net.W[L], &net.mW[L], &net.vW[L]};
        for (float** p : f32s)
            if ((e = hipMalloc((void**)p, sizeof(float) * n)) != hipSuccess) return e;
        float** bs[] = {&net.b[L], &net.mb[L], &net.vb[L]};
        for (float** p : bs)
            if ((e = hipMalloc((void**)p, sizeof(float) * rows)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&net.Wb[L], 2 * n)) != hipSuccess) return e;
        if ((e = hipMalloc((void**)&net.WbT[L], 2 * n)) != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t mlp_copy(Mlp& dst, const Mlp& src, hipStream_t s)
{
    hipError_t e = hipMemcpyAsync(dst.Wf, src.Wf, 2 * FRAG_TOTAL, hipMemcpyDeviceToDevice, s);
    for (int L = 0; L < NL && e == hipSuccess; ++L) {
        const size_t rows = DIMS[L + 1], n = rows * DIMS[L];
        const std::pair<void*, const void*> f[] = {{dst.W[L], src.W[L]}, {dst.mW[L], src.mW[L]}, {dst.vW[L], src.vW[L]}};
        for (const auto& pr : f)
            if (e == hipSuccess) e = hipMemcpyAsync(pr.first, pr.second, sizeof(float) * n, hipMemcpyDeviceToDevice, s);
        const std::pair<void*, const void*> b[] = {{dst.b[L], src.b[L]}, {dst.mb[L], src.mb[L]}, {dst.vb[L], src.vb[L]}};
        for (const auto& pr : b)
            if (e == hipSuccess)
                e = hipMemcpyAsync(pr.first, pr.second, sizeof(float) * rows, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dst.Wb[L], src.Wb[L], 2 * n, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dst.WbT[L], src.WbT[L], 2 * n, hipMemcpyDeviceToDevice, s);
    }
    return e;
}

void mlp_free(Mlp& net)
{
    for (int L = 0; L < NL; ++L) {
        void* ps[] = {net.W[L], net.b[L], net.mW[L], net.vW[L], net.mb[L], net.vb[L], net.Wb[L], net.WbT[L]};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    if (net.Wf) (void)hipFree(net.Wf);
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
