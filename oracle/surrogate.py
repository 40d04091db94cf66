"""CPU oracle of the chemistry surrogate MLP.

TEST INFRASTRUCTURE ONLY (the checker; the product never imports it).

The reference has no learned component: the network stands in for the
per-cell chemistry integration of Collide_Species (lbm-collision.cpp:370-399),
whose plain restatement is MultiState.collide_species in oracle/multicomp.py.
This module restates the network the HIP kernels implement:

  * layers 16 -> 128 -> 128 -> 128 -> 16 (12 real inputs, 10 real outputs,
    zero padding), GELU (tanh form) on the hidden layers;
  * BF16 (round to nearest even) weights, inputs and hidden activations,
    FP32 pre-activations; `use_bf16=False` evaluates the same network in exact
    float64 arithmetic (used to pin the derivative formulas by finite
    differences);
  * loss = mean over rows and the 10 outputs of (z - t)^2; the backward pass
    rounds each layer's dZ to BF16 before it enters the weight gradient and the
    next dX product, as the training kernels do;
  * Adam (beta1 0.9, beta2 0.999, eps 1e-8) on FP32 master weights;
  * Glorot-uniform initial weights from a SplitMix64 stream, zero biases;
  * the normalisation record derived from capture rows: feature and target
    means and scales (feature scales at least 1e-3 of the mean), and the
    feature envelope (captured min/max widened by 10 % of the width plus 1e-3
    of the mean); and the decoding of network outputs into the
    chemistry sources (species mass added over the step, final temperature).

Parameter layout: for each layer W [N][K] row-major, then b [N].
Pinned by tests/test_oracle_surrogate.py (finite differences of the exact
loss, BF16 rounding known answers, the closed form of Adam's first step).
"""
import math

import numpy as np

DIMS = (16, 128, 128, 128, 16)
NL = 4
NIN, NOUT = 12, 10
SS_FMEAN, SS_FSTD, SS_TMEAN, SS_TSTD, SS_LO, SS_HI, SS_SIZE = 0, 16, 32, 48, 64, 80, 96


def n_params():
    return sum(DIMS[L] * DIMS[L + 1] + DIMS[L + 1] for L in range(NL))


def unpack(params):
    """Parameter vector -> [(W [N][K], b [N])] (views, float32)."""
    out, off = [], 0
    for L in range(NL):
        K, N = DIMS[L], DIMS[L + 1]
        W = params[off:off + N * K].reshape(N, K)
        off += N * K
        b = params[off:off + N]
        off += N
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)


def bf16(x):
    """Rounds float32 values to BF16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, u & 0xFFFF0000, r)
    return (r.astype(np.uint32)).view(np.float32).reshape(x.shape)


GELU_C = 2.0 * math.sqrt(2.0 / math.pi)
GELU_A = 0.044715


def gelu(x):
    """GELU, tanh form: x (1 + tanh(v)) / 2, v = sqrt(2/pi) (x + 0.044715 x^3)."""
    return 0.5 * x * (1.0 + np.tanh(0.5 * GELU_C * (x + GELU_A * x ** 3)))


def gelu_grad(x):
    t = np.tanh(0.5 * GELU_C * (x + GELU_A * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * 0.5 * GELU_C * (1.0 + 3.0 * GELU_A * x * x)


def _splitmix_stream(seed):
    st = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    while True:
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        yield (z >> 11) * (1.0 / 9007199254740992.0)


def init_params(seed):
    """Glorot-uniform weights over the real fan-in/fan-out (padding zero),
    drawn row-major layer by layer from one SplitMix64 stream; zero biases."""
    rnd = _splitmix_stream(seed)
    layers = []
    for L in range(NL):
        K, N = DIMS[L], DIMS[L + 1]
        fin = NIN if L == 0 else K
        fout = NOUT if L == NL - 1 else N
        lim = math.sqrt(6.0 / (fin + fout))
        W = np.zeros((N, K), np.float32)
        for r in range(N):
            for c in range(K):
                pad = (L == 0 and c >= NIN) or (L == NL - 1 and r >= NOUT)
                W[r, c] = 0.0 if pad else np.float32((2.0 * next(rnd) - 1.0) * lim)
        layers.append((W, np.zeros(N, np.float32)))
    return pack(layers)


def forward(params, X, use_bf16=True):
    """Pre-activations Z_L and layer inputs H_L (H_0 = X) of every layer."""
    q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    H = [q(X).astype(np.float64)]
    Z = []
    for L, (W, b) in enumerate(unpack(np.asarray(params, np.float32 if use_bf16 else np.float64))):
        z = H[-1] @ q(W).astype(np.float64).T + b.astype(np.float64)
        if use_bf16:
            z = z.astype(np.float32).astype(np.float64)
        Z.append(z)
        if L < NL - 1:
            H.append(q(gelu(z).astype(np.float32) if use_bf16 else gelu(z)).astype(np.float64))
    return Z, H


def predict(params, X, use_bf16=True):
    return forward(params, X, use_bf16)[0][-1]


def loss_and_grad(params, X, T, use_bf16=True):
    """Mean squared error over the NOUT real outputs and its gradient in the
    parameter layout."""
    q = bf16 if use_bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(params, np.float32 if use_bf16 else np.float64))
    Z, H = forward(params, X, use_bf16)
    M = X.shape[0]
    d = Z[-1] - np.asarray(T, np.float64)
    d[:, NOUT:] = 0.0
    loss = float(np.sum(d * d) / (M * NOUT))
    dZ = q((2.0 * d / (M * NOUT)).astype(np.float32) if use_bf16 else 2.0 * d / (M * NOUT)).astype(np.float64)
    grads = [None] * NL
    for L in range(NL - 1, -1, -1):
        W, _ = layers[L]
        gW = dZ.T @ H[L]
        gb = dZ.sum(axis=0)
        grads[L] = (gW, gb)
        if L > 0:
            dH = dZ @ q(W).astype(np.float64)
            g = dH * gelu_grad(Z[L - 1])
            dZ = q(g.astype(np.float32) if use_bf16 else g).astype(np.float64)
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])


class Adam:
    def __init__(self, n, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n, np.float32)
        self.v = np.zeros(n, np.float32)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = np.float32(lr), np.float32(b1), np.float32(b2), np.float32(eps)

    def step(self, w, g):
        self.t += 1
        g = np.asarray(g, np.float32)
        self.m = self.b1 * self.m + (np.float32(1) - self.b1) * g
        self.v = self.b2 * self.v + (np.float32(1) - self.b2) * g * g
        c1 = np.float32(1) - self.b1 ** self.t
        c2 = np.float32(1) - self.b2 ** self.t
        return (w - self.lr * (self.m / c1) / (np.sqrt(self.v / c2) + self.eps)).astype(np.float32)


def norm_record(rows):
    """Normalisation and envelope record [96] from capture rows [m][32]."""
    r = np.asarray(rows, np.float64)
    mean = r.mean(axis=0)
    var = (r * r).mean(axis=0) - mean * mean
    sd = np.sqrt(np.maximum(var, 0.0))
    # features: scale at least 1e-3 of the mean; targets: no spread -> 1
    sdf = np.maximum(sd[:16], 1e-3 * np.abs(mean[:16]))
    sdf = np.where(sdf > 1e-30, sdf, 1.0)
    sdt = np.where((sd[16:] > 1e-6 * np.abs(mean[16:])) & (sd[16:] > 1e-30), sd[16:], 0.0)
    lo, hi = r[:, :16].min(axis=0), r[:, :16].max(axis=0)
    margin = 0.1 * (hi - lo) + 1e-3 * np.abs(mean[:16])
    st = np.zeros(SS_SIZE, np.float32)
    used_f = np.arange(16) < NIN
    used_t = np.arange(16) < NOUT
    st[SS_FMEAN:SS_FMEAN + 16] = np.where(used_f, mean[:16], 0.0)
    st[SS_FSTD:SS_FSTD + 16] = np.where(used_f, sdf, 1.0)
    st[SS_TMEAN:SS_TMEAN + 16] = np.where(used_t, mean[16:], 0.0)
    st[SS_TSTD:SS_TSTD + 16] = np.where(used_t, sdt, 0.0)
    st[SS_LO:SS_LO + 16] = np.where(used_f, lo - margin, 0.0)
    st[SS_HI:SS_HI + 16] = np.where(used_f, hi + margin, 0.0)
    return st


def normalise_features(raw, st):
    """Raw features [m][12] (float32) -> network inputs [m][16] and envelope flags."""
    raw = np.asarray(raw, np.float32)
    m = raw.shape[0]
    X = np.zeros((m, 16), np.float32)
    X[:, :NIN] = (raw - st[SS_FMEAN:SS_FMEAN + NIN]) / st[SS_FSTD:SS_FSTD + NIN]
    outside = np.any((raw < st[SS_LO:SS_LO + NIN]) | (raw > st[SS_HI:SS_HI + NIN]), axis=1)
    return X, outside


def normalise_targets(raw_t, st):
    raw_t = np.asarray(raw_t, np.float32)
    T = np.zeros((raw_t.shape[0], 16), np.float32)
    sd = st[SS_TSTD:SS_TSTD + NOUT]
    with np.errstate(divide="ignore", invalid="ignore"):
        T[:, :NOUT] = np.where(sd > 0, (raw_t - st[SS_TMEAN:SS_TMEAN + NOUT]) / np.where(sd > 0, sd, 1), 0)
    return T


def decode(pred, st, rho, Y0, Tsi):
    """Network outputs [m][16] -> species mass sources [NS][m] (summing to zero
    by removing the net mass in proportion to the target scales; a scale-0
    species keeps its mean) and the final temperature."""
    pred = np.asarray(pred, np.float64)
    tm = st[SS_TMEAN:SS_TMEAN + NOUT].astype(np.float64)
    ts = st[SS_TSTD:SS_TSTD + NOUT].astype(np.float64)
    src = rho[None, :] * (pred[:, 1:NOUT].T * ts[1:, None] + tm[1:, None])
    src = src - (ts[1:] / ts[1:].sum())[:, None] * src.sum(axis=0)[None, :]
    T = Tsi + (pred[:, 0] * ts[0] + tm[0])
    return src, T
