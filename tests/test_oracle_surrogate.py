"""Pins the surrogate MLP oracle (oracle/surrogate.py) on the CPU.

The derivative formulas are checked against central finite differences of
the exact (float64, no BF16) loss; the BF16 rounding against hand-worked
values; Adam's first step against its closed form; the initialiser against
the Glorot bounds; the normalisation record against numpy statistics.
"""
import math

import numpy as np
import pytest

from oracle import surrogate as S


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1 + 2**-8, 1 + 3 * 2**-8, 1 + 2**-8 + 2**-20, -2.5, 0.0, np.inf, -np.inf,
                  3.0e38, 1e-40], np.float32)
    r = S.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even
    assert r[2] == 1 + 2**-6              # tie -> even (mantissa 2)
    assert r[3] == 1 + 2**-7              # above the tie
    assert r[4] == -2.5
    assert r[5] == 0.0
    assert np.isposinf(r[6]) and np.isneginf(r[7])
    assert abs(r[8] - 3.0e38) / 3.0e38 < 2**-8
    assert np.isnan(S.bf16(np.array([np.nan], np.float32)))[0]
    # every result has 16 zero low bits and is within half an ulp
    y = np.random.default_rng(1).standard_normal(1000).astype(np.float32)
    ry = S.bf16(y)
    assert np.all(ry.view(np.uint32) & 0xFFFF == 0)
    assert np.all(np.abs(ry - y) <= np.abs(y) * 2**-8)


def _problem(seed, m=24):
    rng = np.random.default_rng(seed)
    params = S.init_params(seed).astype(np.float64)
    params += 0.05 * rng.standard_normal(params.size)  # non-zero biases and padding too
    X = rng.standard_normal((m, 16))
    T = rng.standard_normal((m, 16))
    return params, X, T


def test_gradient_matches_finite_differences():
    params, X, T = _problem(3)
    loss, g = S.loss_and_grad(params, X, T, use_bf16=False)
    assert loss > 0
    rng = np.random.default_rng(4)
    # coordinates from every layer's weights and biases
    idx, off = [], 0
    for L in range(S.NL):
        K, N = S.DIMS[L], S.DIMS[L + 1]
        idx += list(off + rng.integers(0, N * K, 6)) + list(off + N * K + rng.integers(0, N, 3))
        off += N * K + N
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (S.loss_and_grad(p1, X, T, False)[0] - S.loss_and_grad(p2, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)), (i, fd, g[i])


def test_padded_outputs_carry_no_gradient():
    params, X, T = _problem(5)
    _, g = S.loss_and_grad(params, X, T, use_bf16=False)
    W3, b3 = S.unpack(g)[3]
    assert np.all(W3[S.NOUT:] == 0) and np.all(b3[S.NOUT:] == 0)


def test_bf16_gradient_close_to_exact():
    params, X, T = _problem(6, m=64)
    l0, g0 = S.loss_and_grad(params, X, T, use_bf16=False)
    l1, g1 = S.loss_and_grad(params.astype(np.float32), X.astype(np.float32), T, use_bf16=True)
    assert abs(l1 - l0) / l0 < 2e-2
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) < 5e-2


def test_adam_first_step_closed_form():
    opt = S.Adam(3, lr=1e-2)
    w = np.array([1.0, -2.0, 0.5], np.float32)
    g = np.array([0.3, -4.0, 1e-3], np.float32)
    w1 = opt.step(w, g)
    np.testing.assert_allclose(w1, w - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-6)


def test_init_is_glorot_uniform():
    p = S.init_params(20261202)
    assert p.size == S.n_params() == 37264
    for L, (W, b) in enumerate(S.unpack(p)):
        K, N = S.DIMS[L], S.DIMS[L + 1]
        fin = S.NIN if L == 0 else K
        fout = S.NOUT if L == S.NL - 1 else N
        lim = math.sqrt(6.0 / (fin + fout))
        real = W[:S.NOUT] if L == S.NL - 1 else (W[:, :S.NIN] if L == 0 else W)
        assert np.all(np.abs(real) <= lim) and np.all(b == 0)
        assert abs(np.var(real) - lim**2 / 3) < 0.15 * lim**2 / 3
        if L == 0:
            assert np.all(W[:, S.NIN:] == 0)
        if L == S.NL - 1:
            assert np.all(W[S.NOUT:] == 0)
    assert not np.array_equal(p, S.init_params(1))


def test_norm_record_statistics():
    rng = np.random.default_rng(7)
    rows = np.zeros((500, 32), np.float32)
    rows[:, 0] = rng.uniform(300, 2000, 500)
    rows[:, 11] = 1e-9                  # constant (dt)
    rows[:, 16] = rng.standard_normal(500) * 3 + 1
    st = S.norm_record(rows)
    assert st[S.SS_FMEAN] == pytest.approx(rows[:, 0].astype(np.float64).mean(), rel=1e-6)
    assert st[S.SS_FSTD] == pytest.approx(rows[:, 0].astype(np.float64).std(), rel=1e-5)
    # constant column: scale 1e-3 of its value, envelope within 1e-3 of it
    assert st[S.SS_FSTD + 11] == np.float32(1e-12) and st[S.SS_FMEAN + 11] == np.float32(1e-9)
    assert st[S.SS_LO + 11] == np.float32(1e-9 - 1e-12) and st[S.SS_HI + 11] == np.float32(1e-9 + 1e-12)
    assert st[S.SS_FSTD + 1] == 1.0     # all-zero column
    lo, hi = rows[:, 0].astype(np.float64).min(), rows[:, 0].astype(np.float64).max()
    m = 0.1 * (hi - lo) + 1e-3 * rows[:, 0].astype(np.float64).mean()
    assert st[S.SS_LO] == np.float32(lo - m) and st[S.SS_HI] == np.float32(hi + m)
    assert st[S.SS_TSTD] == pytest.approx(rows[:, 16].astype(np.float64).std(), rel=1e-5)
    assert np.all(st[S.SS_FSTD + 12:S.SS_FSTD + 16] == 1) and np.all(st[S.SS_TSTD + 1:S.SS_TSTD + 16] == 0)
    X, outside = S.normalise_features(rows[:, :12], st)
    assert not outside.any()
    _, outside = S.normalise_features(rows[:3, :12] * 2.5, st)
    assert outside.all()


def test_decode_conserves_mass():
    rng = np.random.default_rng(8)
    st = np.zeros(96, np.float32)
    st[S.SS_TSTD:S.SS_TSTD + 16] = 1e-3
    st[S.SS_TMEAN] = 5.0
    m = 40
    pred = rng.standard_normal((m, 16)).astype(np.float32)
    Y0 = rng.dirichlet(np.ones(9), m)
    rho = rng.uniform(0.5, 1.5, m)
    src, T = S.decode(pred, st, rho, Y0, np.full(m, 1000.0))
    assert np.max(np.abs(src.sum(axis=0))) < 1e-15
    # a species with scale 0 keeps its mean exactly
    st[S.SS_TSTD + 9] = 0.0
    st[S.SS_TMEAN + 9] = 2e-3
    src, _ = S.decode(pred, st, rho, Y0, np.full(m, 1000.0))
    np.testing.assert_allclose(src[8], rho * np.float32(2e-3), rtol=1e-15)
    assert np.max(np.abs(src.sum(axis=0))) < 1e-15
    np.testing.assert_allclose(T, 1000.0 + pred[:, 0].astype(np.float64) * np.float32(1e-3) + 5.0, rtol=1e-15)
