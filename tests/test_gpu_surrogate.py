"""GPU tests of the learned chemistry surrogate, through the C ABI.

The HIP training kernels (forward, backward, Adam) and the in-step inference
kernel are compared with the numpy oracle of oracle/surrogate.py, which
emulates the BF16 roundings (pinned on the CPU in test_oracle_surrogate.py).
The in-step path is compared with the multicomponent oracle fed the oracle
network's chemistry; the fallback and audit paths must reproduce the plain
integrator bit for bit; the last test trains the network on states the
plain integrator produces and measures what it gains.

Tolerances: network outputs and gradients differ from the oracle by FP32
accumulation order and by the occasional BF16 rounding flip this causes
(one BF16 ulp = 2^-8 of an activation), so outputs are held to 2e-3 of their
scale and gradients to 1e-2 in norm per layer.
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import POP_F, POP_FPC, POP_G, FIELDS, SPECIES_FIELDS, MixtureLattice
from oracle import surrogate as S
from oracle.multicomp import MultiState, Units

pytestmark = pytest.mark.gpu

SEEDED = dict(cases.X_UNBURNT_PHI1, H=1e-4, O=1e-4, OH=1e-4, H2O=1e-3, HO2=1e-4, H2O2=1e-4)
M64 = (1 << 64) - 1


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def lattice(fl, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction)
    lat.load_case(fl)
    return lat


def oracle_state(fl):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    s = MultiState(nx + 2, ny + 2, nz + 2, Units(*cases.UNITS_FLAME))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(9, -1)
    return s


def box(n=(8, 6, 6), T_K=1400.0):
    return cases.reacting_box(*n, T_K=T_K, amp_T=0.15, amp_u=0.01, X=SEEDED)


def scaled_params(seed, scale=1.0):
    """Initial network with small random biases (so every term is exercised)."""
    p = S.init_params(seed)
    rng = np.random.default_rng(seed)
    layers = S.unpack(p.copy())
    out = []
    for L, (W, b) in enumerate(layers):
        W = W * np.float32(scale)
        b = (0.1 * rng.standard_normal(b.shape)).astype(np.float32)
        if L == S.NL - 1:
            b[S.NOUT:] = 0
        out.append((W, b))
    return S.pack(out)


# ------------------------------------------------------------- network


def test_init_params_match_oracle():
    lat = lattice(box())
    lat.surrogate_init(12345)
    np.testing.assert_array_equal(lat.surrogate_params(), S.init_params(12345))


def test_forward_matches_oracle():
    lat = lattice(box())
    params = scaled_params(7, 1.5)
    lat.set_surrogate_params(params)
    np.testing.assert_array_equal(lat.surrogate_params(), params)
    rng = np.random.default_rng(11)
    X = np.zeros((1000, 16), np.float32)  # 1000: not a multiple of the 256-row block
    X[:, :S.NIN] = rng.standard_normal((1000, S.NIN))
    Y = lat.surrogate_forward(X)
    ref = S.predict(params, X)
    exact = S.predict(params.astype(np.float64), X.astype(np.float64), use_bf16=False)
    e = relerr(Y[:, :S.NOUT], ref[:, :S.NOUT])
    assert e < 2e-3, e
    # the BF16 emulation is what the kernel computes: closer to it than to exact arithmetic
    assert e < 0.5 * relerr(Y[:, :S.NOUT], exact[:, :S.NOUT])
    assert np.all(np.abs(Y[:, S.NOUT:] - ref[:, S.NOUT:]) < 1e-6)


def test_gradient_matches_oracle():
    lat = lattice(box())
    params = scaled_params(9, 1.2)
    lat.set_surrogate_params(params)
    rng = np.random.default_rng(13)
    X = np.zeros((96, 16), np.float32)
    X[:, :S.NIN] = rng.standard_normal((96, S.NIN))
    T = np.zeros((96, 16), np.float32)
    T[:, :S.NOUT] = rng.standard_normal((96, S.NOUT))
    loss, g = lat.surrogate_gradient(X, T)
    lref, gref = S.loss_and_grad(params, X, T)
    assert abs(loss - lref) / lref < 1e-4
    for L, ((gw, gb), (rw, rb)) in enumerate(zip(S.unpack(g), S.unpack(gref))):
        ew = np.linalg.norm(gw - rw) / np.linalg.norm(rw)
        eb = np.linalg.norm(gb - rb) / np.linalg.norm(rb)
        assert ew < 1e-2 and eb < 1e-2, (L, ew, eb)


def _splitmix(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _capture_step(fl, stride=1, rows=1 << 16):
    lat = lattice(fl)
    lat.Init()
    lat.calculate_moment()
    lat.surrogate_capture(rows, stride)
    lat.Collide_Species()
    return lat


def test_capture_rows_match_oracle():
    fl = box()
    lat = _capture_step(fl)
    rows = lat.surrogate_captured()
    s = oracle_state(fl)
    s.init()
    s.moments()
    s.collide_species(True)
    c = np.nonzero(s.type == 0)[0]
    assert rows.shape == (len(c), 32)
    Y0 = s._Y_state(c)
    Tsi = s.units.si_temp(s.T[c]).astype(np.float64)
    rho_si = s.units.si_rho(s.rho[c]).astype(np.float64)
    ref = np.zeros((len(c), 32), np.float32)
    ref[:, 0] = Tsi
    ref[:, 1] = s.mech.pressure(Tsi, rho_si, Y0)
    ref[:, 2:11] = Y0
    ref[:, 11] = s.units.s
    ref[:, 16] = s.T_chem[c] - s.T_exact0[c]
    ref[:, 17:26] = (s.src_chem[:, c] / s.rho[c]).T
    key = lambda r: np.lexsort(r[:, ::-1].T)  # noqa: E731 -- capture order is not defined
    a, b = rows[key(rows)], ref[key(ref)]
    np.testing.assert_allclose(a[:, :12], b[:, :12], rtol=1e-6)
    scale = np.max(np.abs(b[:, 16:26]), axis=0) + 1e-30
    assert np.max(np.abs(a[:, 16:26] - b[:, 16:26]) / scale) < 1e-5


def test_norm_record_matches_oracle():
    lat = _capture_step(box(), stride=1)
    lat.surrogate_train(0)  # derives the normalisation record only
    np.testing.assert_allclose(lat.surrogate_norm(), S.norm_record(lat.surrogate_captured()), rtol=2e-6,
                               atol=1e-30)


def test_train_step_matches_oracle_adam():
    """One training iteration: the batch the kernels draw, the oracle's BF16
    gradient on it, and Adam's first step lr g / (|g| + eps)."""
    lat = _capture_step(box(), stride=1)
    params = scaled_params(21, 1.0)
    lat.set_surrogate_params(params)
    rows = lat.surrogate_captured()
    B, seed, lr = 256, 77, 1e-3
    lat.surrogate_train(1, batch=B, lr=lr, seed=seed)
    st = lat.surrogate_norm()
    idx = [_splitmix(seed ^ _splitmix((0 << 32) | b)) % len(rows) for b in range(B)]
    X = S.bf16(S.normalise_features(rows[idx, :12], st)[0])
    T = S.normalise_targets(rows[idx, 16:26], st)
    _, g = S.loss_and_grad(params, X, T)
    new = lat.surrogate_params()
    step = new.astype(np.float64) - params
    assert np.max(np.abs(step)) <= lr * (1 + 1e-5)
    big = np.abs(g) > 1e-3 * np.max(np.abs(g))
    assert big.sum() > 0.5 * g.size
    # first Adam step: m / c1 = g and sqrt(v / c2) = |g|, so w -= lr g / (|g| + eps); the
    # comparison is limited by the FP32 rounding of w (|w| < 1)
    np.testing.assert_allclose(step[big], -lr * g[big] / (np.abs(g[big]) + 1e-8), rtol=1e-4, atol=1.2e-7)
    # padded outputs never move; the fragment copies follow the master weights
    W3 = S.unpack(new)[3][0]
    assert np.all(W3[S.NOUT:] == 0)
    Xt = np.zeros((64, 16), np.float32)
    Xt[:, :S.NIN] = np.random.default_rng(3).standard_normal((64, S.NIN))
    assert relerr(lat.surrogate_forward(Xt)[:, :S.NOUT], S.predict(new, Xt)[:, :S.NOUT]) < 2e-3


# ------------------------------------------------------------ in the step


def _oracle_chemistry(params, st, mech, units):
    def chem(c, Tsi, rho_si, Y0, rho, u_si):
        raw = np.zeros((len(c), S.NIN), np.float32)
        raw[:, 0] = Tsi
        raw[:, 1] = mech.pressure(Tsi, rho_si, Y0)
        raw[:, 2:11] = Y0
        raw[:, 11] = units.s
        X, _ = S.normalise_features(raw, st)
        pred = S.predict(params, X).astype(np.float32)
        return S.decode(pred, st, rho, Y0, Tsi)
    return chem


def test_surrogate_step_matches_oracle():
    fl = box()
    lat = _capture_step(fl, stride=1)
    params = scaled_params(5, 1.0)
    lat.set_surrogate_params(params)
    lat.surrogate_train(0)
    st = lat.surrogate_norm()
    lat.surrogate_enable(True, audit_stride=0)
    lat.Collide_Species()
    stats = lat.surrogate_stats()
    s = oracle_state(fl)
    s.init()
    s.moments()
    s.collide_species(True, chemistry=_oracle_chemistry(params, st, s.mech, s.units))
    F = s.type == 0
    assert stats["predicted"] == F.sum() and stats["fallback"] == 0 and stats["audited"] == 0
    assert relerr(lat.get_field("HRR").ravel()[F], s.HRR[F]) < 2e-3
    # the sources enter fpc linearly: their 2e-3 network tolerance, on the fpc scale
    fpc_ref = s.fpc[:, :, F]
    tol = 2e-3 * np.max(np.abs(s.src_chem[:, F])) / np.max(np.abs(fpc_ref)) + 1e-13
    e = relerr(lat.get_populations(POP_FPC).reshape(9, 27, -1)[:, :, F], fpc_ref)
    assert e < tol, (e, tol)


def _plain_reference(fl):
    lat = lattice(fl)
    lat.Init()
    lat.calculate_moment()
    lat.Collide_Species()
    return lat.get_populations(POP_FPC), lat.get_field("HRR")


def test_outside_envelope_falls_back_to_plain_bitwise():
    fl = box()
    fpc0, hrr0 = _plain_reference(fl)
    lat = _capture_step(fl, stride=1)
    lat.surrogate_train(0)
    st = lat.surrogate_norm()
    st[S.SS_HI] = st[S.SS_LO] - 1.0  # no temperature is inside
    lat.set_surrogate_norm(st)
    lat.surrogate_enable(True, audit_stride=0)
    lat.Collide_Species()
    stats = lat.surrogate_stats()
    n = int((fl["type"] == 0).sum())
    assert stats["fallback"] == n and stats["predicted"] == 0
    assert np.array_equal(lat.get_populations(POP_FPC), fpc0)
    assert np.array_equal(lat.get_field("HRR"), hrr0)
    assert lat.diagnostics()["fallback_count"] == n


def test_failed_audits_restore_plain_bitwise():
    fl = box()
    fpc0, hrr0 = _plain_reference(fl)
    lat = _capture_step(fl, stride=1)
    lat.surrogate_train(0)
    lat.surrogate_enable(True, audit_stride=1, audit_tol=0.0)
    lat.Collide_Species()
    stats = lat.surrogate_stats()
    n = int((fl["type"] == 0).sum())
    assert stats["audited"] == n and stats["audit_failures"] == n
    assert stats["max_audit_error"] > 0
    assert np.array_equal(lat.get_populations(POP_FPC), fpc0)
    assert np.array_equal(lat.get_field("HRR"), hrr0)


def test_refresh_captures_fallbacks_and_widens_the_envelope():
    """lbm_surrogate_refresh.  While capturing, each cell the network cannot
    take joins the training set with the plain integrator's result: after a
    step whose temperature envelope is cut at the middle of the box's range,
    the new rows are exactly the plain step's rows of the cells above the cut.
    The refresh rewrites the envelope over every captured row (the oracle's
    norm_record of them), keeps the normalisation the network was trained
    with, and the same states then stop falling back."""
    fl = box()
    n = int((fl["type"] == 0).sum())
    lat = _capture_step(fl, stride=1)
    plain = lat.surrogate_captured()
    assert len(plain) == n
    lat.surrogate_train(0)
    st = lat.surrogate_norm()
    cut = np.sort(plain[:, 0])[n // 2]  # a captured binary32 temperature
    st[S.SS_HI] = cut
    lat.set_surrogate_norm(st)
    lat.surrogate_enable(True, audit_stride=0)
    lat.Collide_Species()
    k = lat.surrogate_stats()["fallback"]
    hot = plain[plain[:, 0] > cut]
    assert k == len(hot) and 0 < k < n
    rows = lat.surrogate_captured()
    assert len(rows) == n + k
    order = lambda r: r[np.lexsort(r[:, ::-1].T)]  # noqa: E731 -- capture order is not defined
    np.testing.assert_array_equal(order(rows[n:]), order(hot))
    lat.surrogate_refresh(0)
    st2 = lat.surrogate_norm()
    np.testing.assert_array_equal(st2[:S.SS_LO], st[:S.SS_LO])
    np.testing.assert_allclose(st2[S.SS_LO:], S.norm_record(rows)[S.SS_LO:], rtol=2e-6, atol=1e-30)
    assert st2[S.SS_HI] > plain[:, 0].max()
    lat.Collide_Species()
    assert lat.surrogate_stats()["fallback"] == k


def test_async_refresh_matches_the_synchronous_one():
    """lbm_surrogate_refresh_async trains a copy of the network on a second
    stream while the steps go on with the serving network (unchanged until
    the poll); applied by lbm_surrogate_refresh_poll, the network and the
    envelope are bitwise those of lbm_surrogate_refresh with the same
    iterations, batch, rate and seed from the same state: the same kernels on
    the same rows.  (One lattice, restored between the two: rows appended by
    two lattices land in orders that differ, and the batch draws by index.)"""
    fl = box()
    lat = _capture_step(fl, stride=1)
    lat.surrogate_train(20, batch=256, lr=1e-3, seed=5)
    lat.surrogate_enable(True, audit_stride=4)
    lat.Collide_Species()  # audited states join the capture
    params, norm = lat.surrogate_params().copy(), lat.surrogate_norm().copy()
    lat.set_surrogate_params(params)  # Adam moments and step count reset
    lat.surrogate_refresh(30, batch=256, lr=3e-4, seed=9)
    sync_params, sync_norm = lat.surrogate_params().copy(), lat.surrogate_norm().copy()
    assert not np.array_equal(sync_params, params) and not np.array_equal(sync_norm, norm)
    lat.set_surrogate_params(params)
    lat.set_surrogate_norm(norm)
    lat.surrogate_refresh_async(30, batch=256, lr=3e-4, seed=9)
    lat.Collide_Species()  # steps while the refresh trains (appending rows past those it reads)
    assert np.array_equal(lat.surrogate_params(), params)
    assert np.array_equal(lat.surrogate_norm(), norm)
    assert lat.surrogate_refresh_poll(wait=True)
    assert not lat.surrogate_refresh_poll(wait=True)
    np.testing.assert_array_equal(lat.surrogate_params(), sync_params)
    np.testing.assert_array_equal(lat.surrogate_norm(), sync_norm)


def test_surrogate_requires_training_and_reaction():
    from lbm_amd.capi import LbmError
    lat = lattice(box())
    with pytest.raises(LbmError, match="train"):
        lat.surrogate_enable(True)
    with pytest.raises(LbmError, match="captured"):
        lat.surrogate_train(5)
    cold = lattice(box(), reaction=False)
    with pytest.raises(LbmError, match="reaction"):
        cold.surrogate_capture(100, 1)


# ------------------------------------------------------------------ gain


def _snapshot(lat):
    snap = {k: lat.get_field(k) for k in FIELDS}
    snap.update({("sp", k): lat.get_species_field(k) for k in SPECIES_FIELDS})
    snap[("pop", POP_F)] = lat.get_populations(POP_F)
    snap[("pop", POP_G)] = lat.get_populations(POP_G)
    return snap


def _restore(lat, snap):
    for k, v in snap.items():
        if isinstance(k, str):
            lat.set_field(k, v)
        elif k[0] == "sp":
            lat.set_species_field(k[1], v)
        else:
            lat.set_populations(k[1], v)


def _mass_fractions(lat, F):
    """Y_a of the fluid cells F (the ghost cells hold no species)."""
    rho_a = lat.get_species_field("rho").reshape(9, -1)[:, F]
    return rho_a / rho_a.sum(axis=0, keepdims=True)


def test_trained_surrogate_gain_over_plain_integrator():
    """Trains on states the plain integrator produces in the run, then steps
    the same state with and without the network: the chemistry stage must be
    faster with it, at flame-temperature error <= 1 % and species L2 error
    <= 1e-3; the untrained network's source error is reported beside the
    trained one's."""
    fl = cases.mixing_layer(128, 96, 64, T_fuel=300.0, T_ox=1200.0, U_s=0.02)
    F = (fl["type"] == 0).ravel()
    lat = lattice(fl)
    lat.Init()
    lat.calculate_moment()
    lat.surrogate_capture(1 << 21, 8)
    lat.step(16)
    lat.surrogate_capture_stop()
    losses = np.concatenate([lat.surrogate_train(1500, batch=4096, lr=2e-3, seed=1),
                             lat.surrogate_train(1000, batch=4096, lr=3e-4, seed=2)])
    assert losses[-100:].mean() < 0.05 * losses[:10].mean(), (losses[:10].mean(), losses[-100:].mean())
    snap = _snapshot(lat)

    lat.profile(True)
    lat.step(1)
    hrr_plain = lat.get_field("HRR").ravel()[F]
    lat.step(9)
    t_plain, n_plain = lat.profile_read("mc_chemistry_plain")
    T_plain = lat.get_field("T").ravel()[F]
    Y_plain = _mass_fractions(lat, F)

    _restore(lat, snap)
    lat.surrogate_enable(True, audit_stride=64, audit_tol=0.5)
    lat.step(1)
    hrr_sur = lat.get_field("HRR").ravel()[F]
    lat.step(9)
    t_sur, n_sur = lat.profile_read("mc_chemistry_surrogate")
    t_fb, _ = lat.profile_read("mc_chemistry_fallback")
    T_sur = lat.get_field("T").ravel()[F]
    Y_sur = _mass_fractions(lat, F)
    stats = lat.surrogate_stats()

    untrained = lattice(fl)
    untrained.set_surrogate_norm(lat.surrogate_norm())
    _restore(untrained, snap)
    untrained.surrogate_enable(True, audit_stride=0)
    untrained.step(1)
    hrr_untrained = untrained.get_field("HRR").ravel()[F]

    speedup = (t_plain / n_plain) / ((t_sur + t_fb) / n_sur)
    err_T = float(np.max(np.abs(T_sur - T_plain) / T_plain))
    err_Y = float(np.sqrt(np.mean((Y_sur - Y_plain) ** 2)))
    err_hrr = float(np.linalg.norm(hrr_sur - hrr_plain) / np.linalg.norm(hrr_plain))
    err_hrr0 = float(np.linalg.norm(hrr_untrained - hrr_plain) / np.linalg.norm(hrr_plain))
    print(f"\nchemistry stage: plain {t_plain / n_plain:.3f} ms, surrogate {(t_sur + t_fb) / n_sur:.3f} ms "
          f"(x{speedup:.2f}); T err {err_T:.2e}, Y L2 err {err_Y:.2e}; HRR rel L2 err trained {err_hrr:.3e} "
          f"untrained {err_hrr0:.3e}; loss {losses[:10].mean():.3e} -> {losses[-100:].mean():.3e}; {stats}")
    assert speedup > 1.0
    assert err_T <= 1e-2 and err_Y <= 1e-3
    assert err_hrr < 0.1 * err_hrr0
    assert stats["fallback"] < 0.01 * stats["predicted"]
