"""GPU parity of the multicomponent path built for the two-step methane
(CH4_2S.yaml) and propane (propane_mech.yaml) mechanisms, through the C ABI
of their own libraries (liblbm_amd_ch4_2s.so, liblbm_amd_c3h8.so).

Each library is compared call by call with the numpy oracle running the same
mechanism (oracle/chem.py, pinned in test_oracle_mechanisms.py) on seeded
boxes hot enough (2000 K) for both steps to react within a few lattice steps.
Bar: max relative error 1e-10 in rho, u, T, p, rho_a and X after each step
(the FP64 parity bar), the post-collision populations on the same scale.
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_STRICT, POP_F, POP_FPC, POP_G, POP_GPC, \
    MixtureLattice
from oracle.chem import Mechanism
from oracle.multicomp import MultiStateBC, Units

pytestmark = pytest.mark.gpu

TOL = 1e-10
FUEL_AIR = {"ch4_2s": cases.X_CH4_AIR_PHI1, "c3h8": cases.X_C3H8_AIR_PHI1}
MECHS = ["ch4_2s", "c3h8"]


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def box(mech, nx=6, ny=4, nz=4, T_K=2000.0, amp_T=0.1, amp_u=0.01):
    X = dict(FUEL_AIR[mech], CO=1e-3, CO2=1e-3, H2O=1e-3)
    return cases.reacting_box(nx, ny, nz, T_K=T_K, amp_T=amp_T, amp_u=amp_u, X=X, species=cases.MECH_SPECIES[mech])


def pair(mech, fl, reaction=True, precision=LBM_FP64):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    print(f"mechanism {mech} grid {nx}x{ny}x{nz} reaction {reaction} precision {precision}", flush=True)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, mode=LBM_MODE_STRICT,
                         precision=precision, mechanism=mech)
    assert lat.SPECIES == cases.MECH_SPECIES[mech]
    lat.load_case(fl)
    s = MultiStateBC(nx + 2, ny + 2, nz + 2, Units(*cases.UNITS_FLAME), mech=Mechanism(mech))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(s.ns, -1)
    return lat, s


def check_state(lat, s, tol, where):
    F = s.type == 0
    ns = s.ns
    for k in ("rho", "u", "v", "w", "T", "p", "rhoe"):
        e = relerr(lat.get_field(k).ravel()[F], getattr(s, k)[F])
        assert e < tol, f"{where}: {k} {e:.3e}"
    e = relerr(lat.get_species_field("rho").reshape(ns, -1)[:, F], s.rho_a[:, F])
    assert e < tol, f"{where}: rho_a {e:.3e}"
    e = relerr(lat.get_species_field("X").reshape(ns, -1)[:, F], s.X[:, F])
    assert e < tol, f"{where}: X {e:.3e}"
    for d, nm in enumerate("uvw"):
        e = relerr(lat.get_species_field(nm).reshape(ns, -1)[:, F], s.u_a[d][:, F])
        assert e < max(tol, 1e-9), f"{where}: u_a[{nm}] {e:.3e}"


@pytest.mark.parametrize("mech", MECHS)
def test_init_matches_oracle(mech):
    lat, s = pair(mech, box(mech))
    lat.Init()
    s.init()
    F = s.type == 0
    assert relerr(lat.get_populations(POP_F).reshape(s.ns, 27, -1)[:, :, F], s.f[:, :, F]) < 1e-14
    assert relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F]) < 1e-14


@pytest.mark.parametrize("mech", MECHS)
def test_calls_match_oracle(mech):
    lat, s = pair(mech, box(mech))
    lat.Init()
    s.init()
    F = s.type == 0
    lat.calculate_moment()
    s.moments()
    check_state(lat, s, TOL, "moments after init")
    for step in range(1, 5):
        lat.Collide_Species()
        s.collide_species(True)
        e = relerr(lat.get_populations(POP_FPC).reshape(s.ns, 27, -1)[:, :, F], s.fpc[:, :, F])
        assert e < TOL, f"step {step} species fpc {e:.3e}"
        assert relerr(lat.get_field("HRR").ravel()[F], s.HRR[F]) < 1e-9
        lat.Collide()
        s.collide()
        e = relerr(lat.get_populations(POP_GPC).reshape(27, -1)[:, F], s.gpc[:, F])
        assert e < TOL, f"step {step} gpc {e:.3e}"
        lat.Streaming()
        s.stream()
        lat.TMS_BC()
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


@pytest.mark.parametrize("mech", MECHS)
def test_fuel_burns_like_the_oracle(mech):
    """A uniform 2000 K reactor: the fuel is consumed, CO, CO2 and H2O form,
    N2 is inert, the box stays uniform, and the species densities track the
    oracle's over 10 steps."""
    lat, s = pair(mech, box(mech, 4, 4, 4, amp_T=0.0, amp_u=0.0))
    lat.Init()
    s.init()
    lat.calculate_moment()
    s.moments()
    F = s.type == 0
    ra0 = lat.get_species_field("rho").reshape(s.ns, -1)[:, F].sum(axis=1)
    for _ in range(10):
        lat.step(1)
        s.step()
    check_state(lat, s, TOL, "after 10 steps")
    ra = lat.get_species_field("rho").reshape(s.ns, -1)[:, F]
    assert np.all(np.ptp(ra, axis=1) <= 1e-15 * np.max(ra, axis=1))
    d = dict(zip(s.mech.names, ra.sum(axis=1) - ra0))
    fuel = "CH4" if mech == "ch4_2s" else "C3H8"
    assert d[fuel] < 0 and d["O2"] < 0 and d["CO"] > 0 and d["CO2"] > 0 and d["H2O"] > 0
    assert abs(d["N2"]) < 1e-12 * ra0[s.mech.names.index("N2")]


@pytest.mark.parametrize("mech", MECHS)
def test_fp32_shifted_tracks_fp64(mech):
    """The FP32 shifted-population build of the same library stays within
    1e-5 of FP64 in rho, T and p and within 1e-5 (relative) in rho_a over 5
    reacting steps."""
    fl = box(mech)
    ns = len(cases.MECH_SPECIES[mech])
    out = {}
    for prec in (LBM_FP64, LBM_FP32_SHIFTED):
        nx, ny, nz = (n - 2 for n in fl["type"].shape)
        print(f"mechanism {mech} precision {prec}", flush=True)
        lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT, precision=prec,
                             mechanism=mech)
        lat.load_case(fl)
        lat.Init()
        lat.calculate_moment()
        lat.step(5)
        out[prec] = {k: lat.get_field(k).copy() for k in ("rho", "T", "p")}
        out[prec]["rho_a"] = lat.get_species_field("rho").reshape(ns, -1).copy()
        lat.close()
    F = fl["type"].ravel() == 0
    for k in ("rho", "T", "p"):
        assert relerr(out[LBM_FP32_SHIFTED][k].ravel()[F], out[LBM_FP64][k].ravel()[F]) < 1e-5, k
    assert relerr(out[LBM_FP32_SHIFTED]["rho_a"][:, F], out[LBM_FP64]["rho_a"][:, F]) < 1e-5


@pytest.mark.parametrize("mech", MECHS)
def test_surrogate_runs_on_six_species(mech):
    """The chemistry surrogate on a 6-species build: capture rows carry T, p,
    the 6 mass fractions and dt (NS + 3 features, zero padding to 16) and dT,
    dY (NS + 1 targets); training lowers the loss; the in-step network
    replaces the plain chemistry with every cell accounted for (predicted or
    fallback) and finite results."""
    fl = box(mech, 8, 6, 6)
    nx, ny, nz = (n - 2 for n in fl["type"].shape)
    print(f"mechanism {mech} surrogate capture", flush=True)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT, precision=LBM_FP64,
                         mechanism=mech)
    lat.load_case(fl)
    lat.Init()
    lat.calculate_moment()
    lat.surrogate_capture(1 << 16, 1)
    lat.Collide_Species()
    rows = lat.surrogate_captured()
    ns = len(lat.SPECIES)
    n = int((fl["type"] == 0).sum())
    assert rows.shape == (n, 32)
    np.testing.assert_allclose(rows[:, 2:2 + ns].sum(axis=1), 1.0, rtol=1e-5)
    assert np.all(rows[:, ns + 2] == rows[0, ns + 2]) and rows[0, ns + 2] > 0  # dt [s]
    assert np.all(rows[:, ns + 3:16] == 0) and np.all(rows[:, 16 + ns + 1:] == 0)
    fuel = lat.SPECIES.index("CH4" if mech == "ch4_2s" else "C3H8")
    assert np.all(rows[:, 17 + fuel] <= 0) and rows[:, 17 + fuel].min() < 0  # the fuel is consumed
    losses = lat.surrogate_train(300, batch=256, lr=2e-3, seed=3)
    assert np.all(np.isfinite(losses)) and losses[-20:].mean() < 0.5 * losses[:20].mean()
    lat.surrogate_capture_stop()
    lat.surrogate_enable(True, audit_stride=1, audit_tol=0.5)
    lat.Collide_Species()
    st = lat.surrogate_stats()
    assert st["predicted"] + st["fallback"] == n and st["audited"] > 0
    F = fl["type"].ravel() == 0
    assert np.all(np.isfinite(lat.get_field("HRR").ravel()[F]))
    assert np.all(np.isfinite(lat.get_populations(POP_FPC).reshape(ns, 27, -1)[:, :, F]))
