// Chemistry surrogate MLP: device structures and host entry points (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mech.h"

namespace lbm {
namespace sur {

constexpr int NL = 4;                              // weight layers
constexpr int DIMS[NL + 1] = {16, 128, 128, 128, 16};
constexpr int NIN = mech::NS + 3;                  // T, p, Y_k, dt
constexpr int NOUT = mech::NS + 1;                 // dT, dY_k
// The features and targets fit the 16-wide layers (mechanisms of up to 11
// species); builds for larger mechanisms compile the surrogate but refuse its
// entry points.
constexpr bool ENABLED = NIN <= 16 && NOUT <= 12;

// FP32 master weights W[L] [DIMS[L+1]][DIMS[L]], biases, Adam moments and the
// BF16 copies (row-major and transposed) the MFMA kernels read.
struct Mlp {
    float *W[NL] = {}, *b[NL] = {}, *mW[NL] = {}, *vW[NL] = {}, *mb[NL] = {}, *vb[NL] = {};
    uint16_t *Wb[NL] = {}, *WbT[NL] = {};
    uint16_t* Wf = nullptr;  // all layers in MFMA fragment order (in-step inference)
};

// Activations and gradient scratch for batches of up to `cap` rows.
struct Work {
    int cap = 0, chunk = 256;
    float* Z[NL] = {};
    uint16_t* H[NL] = {};
    uint16_t* dZ[NL] = {};
    uint16_t *dZT = nullptr, *XT = nullptr;
    float* part = nullptr;
    float *gW[NL] = {}, *gb[NL] = {};
    float* loss_rows = nullptr;
};

// Capture rows [rows][32]: raw features T [K], p [Pa], Y_k (NS), dt [s], zero
// padding to 16 | targets dT [K], dY_k (NS) over one step, zero padding.
constexpr int CAP_STRIDE = 32;
// Normalisation and envelope record [96]: feature mean, std, target mean,
// std, feature min, max (16 each).
enum { SS_FMEAN = 0, SS_FSTD = 16, SS_TMEAN = 32, SS_TSTD = 48, SS_LO = 64, SS_HI = 80, SS_SIZE = 96 };

hipError_t mlp_alloc(Mlp& net, uint64_t seed, hipStream_t s);
// The arrays of a network, contents undefined; mlp_copy copies every array
// (weights, biases, Adam moments, BF16 copies, fragments) on stream s.
hipError_t mlp_alloc_empty(Mlp& net);
hipError_t mlp_copy(Mlp& dst, const Mlp& src, hipStream_t s);
void mlp_free(Mlp& net);
hipError_t work_alloc(Work& w, int cap);
void work_free(Work& w);
void refresh_bf16(Mlp& net, hipStream_t s);
void to_bf16(const float* in, uint16_t* out, int64_t n, hipStream_t s);
// Forward pass of M rows (M multiple of 16); the outputs are w.Z[NL-1] [M][16].
void mlp_forward(const Mlp& net, const uint16_t* X, int M, Work& w, hipStream_t s);
// Gradients of the mean-squared loss over the NOUT outputs for one batch (M
// multiple of 16) into w.gW / w.gb; per-row losses land in w.loss_rows.
void mlp_grad(const Mlp& net, const uint16_t* X, const float* T, int M, Work& w, hipStream_t s);
// One Adam step with the gradients in w; the bias corrections of the step are
// c12[2 *it], c12[2 *it + 1] (device), filled by adam_corrections (host) for
// steps first_step .. first_step + iters - 1 (first_step >= 1).
constexpr float ADAM_B1 = 0.9f, ADAM_B2 = 0.999f, ADAM_EPS = 1e-8f;
void adam_corrections(int first_step, int iters, float* c12);
void mlp_adam(Mlp& net, Work& w, float lr, const float* c12, const int* it, hipStream_t s);
// *it += 1 on the stream (the training loop's iteration counter, device).
void next_iteration(int* it, hipStream_t s);
// In-step style inference: Y [M][16] FP32 from normalised X [M][16] FP32 (any
// M), activations in registers, weights staged in LDS.
void mlp_infer(const Mlp& net, const float* X, int M, float* Y, hipStream_t s);
// The same network on BF16 inputs X [rows][16] (rows a multiple of 256) into
// FP32 outputs Y [rows][16], weights staged in LDS.
void mlp_infer_bf16(const Mlp& net, const uint16_t* X, int64_t rows, float* Y, hipStream_t s);
// Column statistics of `rows` capture rows into `stats` [SS_SIZE] (device);
// `part` holds STATS_BLOCKS * 32 * 4 doubles.  envelope_only: only the
// envelope (SS_LO, SS_HI) is rewritten, the normalisation kept.
constexpr int STATS_BLOCKS = 256;
void capture_stats(const float* cap, int64_t rows, double* part, float* stats, hipStream_t s,
                   bool envelope_only = false);
// Training batch of B rows drawn (hash of seed, iteration *it, row) from the
// capture: BF16 normalised features [B][16] and FP32 normalised targets [B][16].
void gather_batch(const float* cap, int64_t rows, const float* stats, uint64_t seed, const int* it, int B,
                  uint16_t* X, float* T, hipStream_t s);
// out[*idx] (out[0] when idx is null) = sum(loss_rows[0..B)) * scale, summed in
// a fixed order.
void loss_sum(const float* loss_rows, int B, float scale, float* out, hipStream_t s, const int* idx = nullptr);
constexpr int NPARAM = 16 * 128 + 128 + 2 * (128 * 128 + 128) + 128 * 16 + 16;

}  // namespace sur
}  // namespace lbm
