"""GPU parity of the multicomponent path built for the whole ammonia/hydrogen
mechanism (nh3.yaml: 31 species, 203 reactions with PLOG and Lindemann
rates), through the C ABI of its own library (liblbm_amd_nh3.so).

The library is compared call by call with the numpy oracle running the same
mechanism (oracle/chem.py, pinned in test_oracle_nh3.py) on seeded boxes of
an NH3/H2-air mixture hot enough (1800-2000 K) for the nitrogen chemistry to
run within a few lattice steps, and on the boundary kinds of TMS_BC: an
NH3/H2 mixing layer with NSCBC outflow, a channel with fixed-state inflow,
outflow and isothermal walls, and a column between a LODI inflow and an
NSCBC outflow (per call, from the oracle's state).  Bar: max relative error 1e-10 in rho, u, T,
p, rho_a and X after each step (the FP64 parity bar), the post-collision
populations on the same scale.  The fused and in-place modes are bitwise
equal to the strict calls.  FP32 populations and the chemistry surrogate
hold at most 16 and 11 species; this build refuses them.
"""
import json

import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT, POP_FPC, \
    POP_F, POP_G, POP_GPC, LbmError, MixtureLattice
from oracle.chem import Mechanism, mechanism_path
from oracle.multicomp import MultiStateBC, Units
from mode_compare import seed_gpu_from_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10
MECH = "nh3"
# every species above SPECIES_MIN: one set up at the floor is skipped or not in
# the first collision by the last bit of its population sum (lbm-collision.cpp:
# 512-513), where two correct implementations round apart
SEEDED = dict({n: 1e-6 for n in cases.MECH_SPECIES[MECH]}, **cases.X_NH3_H2_AIR_PHI1, H=1e-4, O=1e-4, OH=1e-4,
              H2O=1e-3, HO2=1e-5, NO=1e-4, NH2=1e-5, N2O=1e-5)


def relerr(a, b, floor=1e-300):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor))


# lattice velocity scale below which velocities are rounding noise (the
# uniform reactor's are ~1e-18 on both sides)
VEL_FLOOR = 1e-8


def box(nx=6, ny=4, nz=4, T_K=1900.0, amp_T=0.05, amp_u=0.01, X=SEEDED):
    return cases.reacting_box(nx, ny, nz, T_K=T_K, amp_T=amp_T, amp_u=amp_u, X=X, species=cases.MECH_SPECIES[MECH])


def lattice(fl, mode=LBM_MODE_STRICT, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    print(f"nh3 grid {nx}x{ny}x{nz} mode {mode} reaction {reaction}", flush=True)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, mode=mode, precision=LBM_FP64,
                         mechanism=MECH)
    assert lat.SPECIES == cases.MECH_SPECIES[MECH]
    lat.load_case(fl)
    return lat


def pair(fl, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = lattice(fl, reaction=reaction)
    s = MultiStateBC(nx + 2, ny + 2, nz + 2, Units(*cases.UNITS_FLAME), mech=Mechanism(MECH))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(s.ns, -1)
    return lat, s


def check_state(lat, s, tol, where):
    F = s.type == 0
    ns = s.ns
    for k in ("rho", "u", "v", "w", "T", "p", "rhoe"):
        e = relerr(lat.get_field(k).ravel()[F], getattr(s, k)[F], VEL_FLOOR if k in "uvw" else 1e-300)
        assert e < tol, f"{where}: {k} {e:.3e}"
    e = relerr(lat.get_species_field("rho").reshape(ns, -1)[:, F], s.rho_a[:, F])
    assert e < tol, f"{where}: rho_a {e:.3e}"
    e = relerr(lat.get_species_field("X").reshape(ns, -1)[:, F], s.X[:, F])
    assert e < tol, f"{where}: X {e:.3e}"
    for d, nm in enumerate("uvw"):
        e = relerr(lat.get_species_field(nm).reshape(ns, -1)[:, F], s.u_a[d][:, F], VEL_FLOOR)
        assert e < max(tol, 1e-9), f"{where}: u_a[{nm}] {e:.3e}"


def test_fp32_and_surrogate_are_refused():
    fl = box()
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    with pytest.raises(LbmError, match="16 species"):
        MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, precision=LBM_FP32_SHIFTED, mechanism=MECH)
    lat = lattice(fl)
    try:
        with pytest.raises(LbmError, match="11 species"):
            lat.surrogate_init(1)
    finally:
        lat.close()


def test_init_matches_oracle():
    lat, s = pair(box())
    try:
        lat.Init()
        s.init()
        F = s.type == 0
        assert relerr(lat.get_populations(POP_F).reshape(s.ns, 27, -1)[:, :, F], s.f[:, :, F]) < 1e-14
        assert relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F]) < 1e-14
    finally:
        lat.close()


def test_calls_match_oracle():
    lat, s = pair(box())
    try:
        lat.Init()
        s.init()
        F = s.type == 0
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, "moments after init")
        for step in range(1, 4):
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
    finally:
        lat.close()


def test_ammonia_burns_like_the_oracle():
    """A uniform 2000 K reactor: NH3, H2 and O2 are consumed, H2O and NH2
    form, the box stays uniform and the species densities track the
    oracle's over 10 steps.  Oxygen stays balanced to rounding; nitrogen and
    hydrogen gain what the SPECIES_MIN floor of calculate_moment adds
    (lbm-moment.cpp clamps rho_a at 1e-8; NNH is consumed down to it from
    step 2 on: +5e-7 of the nitrogen after 10 steps), the same gain as the
    oracle's."""
    lat, s = pair(box(4, 4, 4, T_K=2000.0, amp_T=0.0, amp_u=0.0))
    try:
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
        assert np.all(np.ptp(ra, axis=1) <= 1e-14 * np.max(ra, axis=1))
        d = dict(zip(s.mech.names, ra.sum(axis=1) - ra0))
        assert d["NH3"] < 0 and d["H2"] < 0 and d["O2"] < 0 and d["H2O"] > 0 and d["NH2"] > 0
        sp = json.load(open(mechanism_path(MECH)))["species"]
        moles, moles0 = ra.sum(axis=1) / s.mech.W, ra0 / s.mech.W
        oracle = s.rho_a[:, F].sum(axis=1) / s.mech.W
        nO = np.array([x["composition"].get("O", 0) for x in sp], dtype=np.float64)
        assert abs(nO @ moles - nO @ moles0) < 1e-12 * (nO @ moles0)
        assert np.any(s.rho_a[:, F] == 1e-8)  # a species at the floor
        for e in ("N", "H"):
            ne = np.array([x["composition"].get(e, 0) for x in sp], dtype=np.float64)
            gain, gain_oracle = ne @ moles - ne @ moles0, ne @ oracle - ne @ moles0
            assert gain > 0 and abs(gain / gain_oracle - 1) < 1e-6, (e, gain, gain_oracle)
    finally:
        lat.close()


@pytest.mark.parametrize("mode", [LBM_MODE_FUSED, LBM_MODE_INPLACE])
def test_fused_and_inplace_modes_are_bitwise_strict(mode):
    fl = box(6, 5, 4)
    out = {}
    for m in (LBM_MODE_STRICT, mode):
        lat = lattice(fl, mode=m)
        try:
            lat.Init()
            lat.calculate_moment()
            lat.step(3)
            out[m] = {k: lat.get_field(k).copy() for k in ("rho", "u", "v", "w", "T", "p", "rhoe")}
            out[m]["rho_a"] = lat.get_species_field("rho").copy()
            out[m]["f"] = lat.get_populations(POP_F).copy()
            out[m]["g"] = lat.get_populations(POP_G).copy()
        finally:
            lat.close()
    F = fl["type"] == 0
    for k in ("rho", "u", "v", "w", "T", "p", "rhoe"):
        assert np.array_equal(out[mode][k][F], out[LBM_MODE_STRICT][k][F]), k
    ns = len(cases.MECH_SPECIES[MECH])
    Ff = F.ravel()
    assert np.array_equal(out[mode]["rho_a"].reshape(ns, -1)[:, Ff], out[LBM_MODE_STRICT]["rho_a"].reshape(ns, -1)[:, Ff])
    assert np.array_equal(out[mode]["f"].reshape(ns, 27, -1)[:, :, Ff],
                          out[LBM_MODE_STRICT]["f"].reshape(ns, 27, -1)[:, :, Ff])
    assert np.array_equal(out[mode]["g"].reshape(27, -1)[:, Ff], out[LBM_MODE_STRICT]["g"].reshape(27, -1)[:, Ff])


@pytest.mark.parametrize("reaction", [False, True])
def test_mixing_layer_outflow_calls_match_oracle(reaction):
    """An NH3/H2 stream against hot air with NSCBC outflow on both y faces:
    every call of three steps on the GPU and in the oracle from the same
    state (the device re-seeded from the oracle before each call: fd_fuw's
    upwind choice branches on ~0 diffusion velocities next to the outflow
    faces), the outflow sweep's four launches included."""
    fl = cases.mixing_layer(4, 8, 4, T_fuel=600.0, T_ox=1600.0, U_s=0.02, X_fuel={"NH3": 0.3, "H2": 0.1, "N2": 0.6},
                            species=cases.MECH_SPECIES[MECH])
    lat, s = pair(fl, reaction=reaction)
    ns = s.ns
    try:
        lat.Init()
        s.init()
        F = s.type == 0
        for step in range(1, 4):
            seed_gpu_from_oracle(lat, s)
            lat.Collide_Species()
            s.collide_species(reaction)
            e = relerr(lat.get_populations(POP_FPC).reshape(ns, 27, -1)[:, :, F], s.fpc[:, :, F])
            assert e < TOL, f"step {step} species fpc {e:.3e}"
            seed_gpu_from_oracle(lat, s)
            lat.Collide()
            s.collide()
            e = relerr(lat.get_populations(POP_GPC).reshape(27, -1)[:, F], s.gpc[:, F])
            assert e < TOL, f"step {step} gpc {e:.3e}"
            seed_gpu_from_oracle(lat, s)
            lat.Streaming()
            s.stream()
            seed_gpu_from_oracle(lat, s)
            lat.TMS_BC()
            s.tms_bc()
            e = relerr(lat.get_populations(POP_F).reshape(ns, 27, -1)[:, :, F], s.f[:, :, F])
            assert e < TOL, f"step {step} species f after TMS {e:.3e}"
            e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
            assert e < TOL, f"step {step} g after TMS {e:.3e}"
            seed_gpu_from_oracle(lat, s)
            lat.calculate_moment()
            s.moments()
            check_state(lat, s, TOL, f"step {step}")
    finally:
        lat.close()


def per_call_steps(lat, s, reaction, steps=3, moments_first=True):
    """Collide_Species, Collide, Streaming, TMS_BC and calculate_moment on
    both sides from the oracle's state before every call; the populations
    after TMS_BC and the fields after each call that writes them checked."""
    ns = s.ns
    F = s.type == 0
    if moments_first:
        lat.calculate_moment()
        s.moments()
    for step in range(1, steps + 1):
        seed_gpu_from_oracle(lat, s)
        lat.Collide_Species()
        s.collide_species(reaction)
        e = relerr(lat.get_populations(POP_FPC).reshape(ns, 27, -1)[:, :, F], s.fpc[:, :, F])
        assert e < TOL, f"step {step} species fpc {e:.3e}"
        seed_gpu_from_oracle(lat, s)
        lat.Collide()
        s.collide()
        seed_gpu_from_oracle(lat, s)
        lat.Streaming()
        s.stream()
        seed_gpu_from_oracle(lat, s)
        lat.TMS_BC()
        s.tms_bc()
        e = relerr(lat.get_populations(POP_F).reshape(ns, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after TMS {e:.3e}"
        e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
        assert e < TOL, f"step {step} g after TMS {e:.3e}"
        check_state(lat, s, TOL, f"step {step} after TMS")
        seed_gpu_from_oracle(lat, s)
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


@pytest.mark.parametrize("reaction", [False, True])
def test_channel_fixed_state_boundaries_match_oracle(reaction):
    """TYPE_I inflow of air, TYPE_O outflow (calculate_moment_point of the
    downstream cell), isothermal TYPE_S walls, one moving, around the NH3/H2
    mixture at 1500 K: per call, from the oracle's state."""
    fl = cases.reacting_channel(8, 5, 3, T_K=1500.0, T_in_K=1400.0, T_wall_K=1200.0, X=SEEDED,
                                X_in=dict({n: 1e-6 for n in cases.MECH_SPECIES[MECH]}, O2=0.21, N2=0.79),
                                p_out_factor=0.995, species=cases.MECH_SPECIES[MECH])
    lat, s = pair(fl, reaction=reaction)
    try:
        lat.Init()
        s.init()
        per_call_steps(lat, s, reaction)
    finally:
        lat.close()


def inflow_outflow_column(nx, T_K=1500.0, u_in=0.01, seed=5):
    """Periodic y/z column between a TYPE_I_C inflow ghost (inline LODI
    target towards its fresh NH3/H2-air state at 1300 K) and a TYPE_O_C
    outflow ghost (NSCBC) at 1 atm, with a seeded +-2 % temperature and
    velocity perturbation inside."""
    rng = np.random.default_rng(seed)
    fl = cases.default_fields(nx, 3, 3)
    t = fl["type"]
    for ax in (1, 2):
        idx = [slice(None)] * 3
        for end in (0, -1):
            idx[ax] = end
            t[tuple(idx)] = cases.TYPE_P
    t[0] = cases.TYPE_I_C
    t[-1] = cases.TYPE_O_C
    ns = len(cases.MECH_SPECIES[MECH])
    X = cases.mole_vector(SEEDED, species=cases.MECH_SPECIES[MECH])
    fl["X"] = np.broadcast_to(X[:, None, None, None], (ns,) + t.shape).copy()
    fl["T"][...] = T_K / cases.LATTICE_K
    fl["T"][0] = 1300.0 / cases.LATTICE_K
    fl["p"][...] = cases.lattice_p(cases.ONE_ATM)
    fl["u"][...] = u_in
    fl["v"][...] = 0.0
    fl["w"][...] = 0.0
    m = t == cases.TYPE_F
    fl["T"][m] *= 1.0 + 0.02 * rng.uniform(-1.0, 1.0, int(m.sum()))
    fl["u"][m] *= 1.0 + 0.02 * rng.uniform(-1.0, 1.0, int(m.sum()))
    return fl


@pytest.mark.parametrize("reaction", [False, True])
def test_inflow_outflow_column_calls_match_oracle(reaction):
    """TYPE_I_C inflow (inline LODI, setState_TP density) and TYPE_O_C
    outflow (NSCBC) of the 31-species mixture: per call, from the oracle's
    state."""
    fl = inflow_outflow_column(10)
    lat, s = pair(fl, reaction=reaction)
    try:
        lat.Init()
        s.init()
        per_call_steps(lat, s, reaction)
    finally:
        lat.close()
