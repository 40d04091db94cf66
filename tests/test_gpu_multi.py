"""GPU parity of the multicomponent reacting path, through the C ABI.

The HIP kernels are compared call by call with the numpy oracle
(oracle/multicomp.py, pinned in test_oracle_multicomp.py) on the same seeded
boxes.  Bar: max relative error 1e-10 in rho, u, T and the species densities
after N steps (the FP64 parity bar); populations are compared on the
same scale.  Both sides round the lattice/SI crossings to binary32 exactly
as the reference does, so they agree far below that bar unless a value lands
on a float rounding boundary.
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_MODE_STRICT, POP_F, POP_FPC, POP_G, POP_GPC, MixtureLattice
from mode_compare import seed_gpu_from_oracle
from oracle.multicomp import MultiStateBC, Units

pytestmark = pytest.mark.gpu

SEEDED = dict(cases.X_UNBURNT_PHI1, H=1e-4, O=1e-4, OH=1e-4, H2O=1e-3, HO2=1e-4, H2O2=1e-4)
TOL = 1e-10


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(fl, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, mode=LBM_MODE_STRICT)
    lat.load_case(fl)
    s = MultiStateBC(nx + 2, ny + 2, nz + 2, Units(*cases.UNITS_FLAME))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(9, -1)
    return lat, s


def fluid(s):
    return s.type == 0


def check_state(lat, s, tol, where):
    F = fluid(s)
    for k in ("rho", "u", "v", "w", "T", "p", "rhoe"):
        e = relerr(lat.get_field(k).ravel()[F], getattr(s, k)[F])
        assert e < tol, f"{where}: {k} {e:.3e}"
    e = relerr(lat.get_species_field("rho").reshape(9, -1)[:, F], s.rho_a[:, F])
    assert e < tol, f"{where}: rho_a {e:.3e}"
    e = relerr(lat.get_species_field("X").reshape(9, -1)[:, F], s.X[:, F])
    assert e < tol, f"{where}: X {e:.3e}"
    for d, nm in enumerate("uvw"):
        e = relerr(lat.get_species_field(nm).reshape(9, -1)[:, F], s.u_a[d][:, F])
        assert e < max(tol, 1e-9), f"{where}: u_a[{nm}] {e:.3e}"


def test_mc_init_matches_oracle():
    fl = cases.reacting_box(6, 4, 4, T_K=1300.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    lat, s = pair(fl)
    lat.Init()
    s.init()
    F = fluid(s)
    assert relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F]) < 1e-14
    assert relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F]) < 1e-14
    assert relerr(lat.get_field("rho").ravel()[F], s.rho[F]) < 1e-15


@pytest.mark.parametrize("reaction", [True, False])
def test_mc_calls_match_oracle(reaction):
    fl = cases.reacting_box(6, 4, 4, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    lat, s = pair(fl, reaction=reaction)
    lat.Init()
    s.init()
    F = fluid(s)
    lat.calculate_moment()
    s.moments()
    check_state(lat, s, TOL, "moments after init")
    for step in range(1, 6):
        lat.Collide_Species()
        s.collide_species(reaction)
        e = relerr(lat.get_populations(POP_FPC).reshape(9, 27, -1)[:, :, F], s.fpc[:, :, F])
        assert e < TOL, f"step {step} species fpc {e:.3e}"
        if reaction:
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


def test_mc_uniform_reactor_heats_up_like_the_oracle():
    fl = cases.reacting_box(4, 4, 4, T_K=1400.0, X=SEEDED)
    lat, s = pair(fl)
    lat.Init()
    s.init()
    lat.calculate_moment()
    s.moments()
    for _ in range(20):
        lat.step(1)
        s.step()
    F = fluid(s)
    T = lat.get_field("T").ravel()[F]
    assert np.ptp(T) == 0.0  # the box stays uniform
    assert relerr(T, s.T[F]) < TOL
    assert T[0] * cases.LATTICE_K > 1400.0


def test_mc_species_mass_conserved_without_reaction():
    fl = cases.reacting_box(6, 4, 4, T_K=1200.0, amp_T=0.1, amp_u=0.02, X=SEEDED)
    lat, _ = pair(fl, reaction=False)
    lat.Init()
    lat.calculate_moment()
    F = fl["type"].ravel() == 0
    m0 = lat.get_species_field("rho").reshape(9, -1)[:, F].sum(axis=1)
    lat.step(10)
    m1 = lat.get_species_field("rho").reshape(9, -1)[:, F].sum(axis=1)
    np.testing.assert_allclose(m1, m0, rtol=1e-12)


def test_mc_deterministic():
    fl = cases.reacting_box(6, 4, 4, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    out = []
    for _ in range(2):
        lat, _ = pair(fl)
        lat.Init()
        lat.calculate_moment()
        lat.step(4)
        out.append((lat.get_populations(POP_F), lat.get_field("T")))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("reaction", [False, True])
def test_mc_mixing_layer_outflow_calls_match_oracle(reaction):
    """Periodic x/z, NSCBC outflow on both y faces.  Each call runs on the
    GPU and in the oracle from the same state (the device is re-seeded from
    the oracle before every call): the reference's upwind choice in fd_fuw
    branches on the sign of diffusion velocities that are ~0 next to the
    outflow faces, so states that differ by rounding drift apart over steps;
    per-call parity is what pins the kernels."""
    fl = cases.mixing_layer(4, 8, 4, T_fuel=300.0, T_ox=1200.0, U_s=0.02)
    lat, s = pair(fl, reaction=reaction)
    lat.Init()
    s.init()
    F = fluid(s)
    for step in range(1, 4):
        seed_gpu_from_oracle(lat, s)
        lat.Collide_Species()
        s.collide_species(reaction)
        e = relerr(lat.get_populations(POP_FPC).reshape(9, 27, -1)[:, :, F], s.fpc[:, :, F])
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
        e = relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after TMS {e:.3e}"
        e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
        assert e < TOL, f"step {step} g after TMS {e:.3e}"
        seed_gpu_from_oracle(lat, s)
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


def test_mc_free_slip_walls_match_oracle():
    """Free-slip walls (TYPE_FS) on both y faces of a reacting box: species
    and energy populations mirrored by dirSlip at the walls, the walls
    one-sided in the finite differences; 6 steps against the oracle."""
    fl = cases.reacting_box(6, 5, 4, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    fl["type"][:, 0, :] = cases.TYPE_FS
    fl["type"][:, -1, :] = cases.TYPE_FS
    lat, s = pair(fl)
    lat.Init()
    s.init()
    F = fluid(s)
    lat.calculate_moment()
    s.moments()
    for step in range(1, 7):
        lat.Collide_Species()
        s.collide_species(True)
        lat.Collide()
        s.collide()
        lat.Streaming()
        s.stream()
        e = relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after streaming {e:.3e}"
        e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
        assert e < TOL, f"step {step} g after streaming {e:.3e}"
        lat.TMS_BC()
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


@pytest.mark.parametrize("reaction", [False, True])
def test_mc_channel_fixed_state_boundaries_match_oracle(reaction):
    """TYPE_I inflow (ghost state, the downstream pressure after
    calculate_moment_point), TYPE_O outflow (downstream state, ghost
    pressure) and TYPE_S isothermal walls, one moving: per-call parity with
    the device re-seeded from the oracle before each call."""
    fl = cases.reacting_channel(8, 5, 3, p_out_factor=0.995)
    lat, s = pair(fl, reaction=reaction)
    lat.Init()
    s.init()
    F = fluid(s)
    lat.calculate_moment()
    s.moments()
    for step in range(1, 4):
        seed_gpu_from_oracle(lat, s)
        lat.Collide_Species()
        s.collide_species(reaction)
        seed_gpu_from_oracle(lat, s)
        lat.Collide()
        s.collide()
        seed_gpu_from_oracle(lat, s)
        lat.Streaming()
        s.stream()
        seed_gpu_from_oracle(lat, s)
        lat.TMS_BC()
        s.tms_bc()
        e = relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after TMS {e:.3e}"
        e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
        assert e < TOL, f"step {step} g after TMS {e:.3e}"
        # calculate_moment_point rewrote the fields of the cells downstream
        # of the inflow and outflow interfaces
        check_state(lat, s, TOL, f"step {step} after TMS")
        seed_gpu_from_oracle(lat, s)
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


def _run_precision(fl, precision, steps, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, precision=precision)
    lat.load_case(fl)
    lat.Init()
    lat.calculate_moment()
    lat.step(steps)
    return lat


@pytest.mark.parametrize("case", ["box", "mixing_layer"])
def test_mc_fp32_shifted_tracks_fp64(case):
    """FP32 deviation storage against the FP64 path over 10 steps.  In the
    periodic box the difference is storage rounding (~6e-8 of each deviation,
    accumulated); next to the outflow faces the reference's upwind FD branch
    (fd_fuw) flips on ~0 diffusion velocities, which amplifies any rounding
    difference, so there the bounds are the physical ones (T and rho to
    1e-4, species mass fractions to 1e-5 in L2)."""
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64
    if case == "box":
        fl = cases.reacting_box(8, 6, 6, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    else:
        # delta = 4 cells: sharper layers at this size blow up in the reference model itself
        fl = cases.mixing_layer(6, 16, 4, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=4.0)
    a = _run_precision(fl, LBM_FP64, 10)
    b = _run_precision(fl, LBM_FP32_SHIFTED, 10)
    F = (fl["type"] == 0).ravel()
    bound = 1e-5 if case == "box" else 1e-4
    for k, tol in (("rho", bound), ("T", bound), ("p", bound), ("u", 100 * bound)):
        e = relerr(b.get_field(k).ravel()[F], a.get_field(k).ravel()[F])
        assert e < tol, (k, e)
    ya = a.get_species_field("rho").reshape(9, -1)[:, F]
    yb = b.get_species_field("rho").reshape(9, -1)[:, F]
    Ya, Yb = ya / ya.sum(axis=0), yb / yb.sum(axis=0)
    assert float(np.sqrt(np.mean((Ya - Yb) ** 2))) < (1e-6 if case == "box" else 1e-5)
    # populations round-trip through the shifted storage
    fa = a.get_populations(POP_F)
    assert relerr(b.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], fa.reshape(9, 27, -1)[:, :, F]) < 10 * bound


@pytest.mark.parametrize("reaction", [False, True])
def test_mc_flame_inflow_outflow_calls_match_oracle(reaction):
    """Quasi-1-D premixed flame: TYPE_I_C inflow at x = 0 (inline LODI target)
    and TYPE_O_C outflow at x = Nx-1 (NSCBC target); per-call parity with
    the device re-seeded from the oracle before each call."""
    fl = cases.premixed_flame_1d(12, 3, 3)
    lat, s = pair(fl, reaction=reaction)
    lat.Init()
    s.init()
    F = fluid(s)
    lat.calculate_moment()
    s.moments()
    for step in range(1, 4):
        seed_gpu_from_oracle(lat, s)
        lat.Collide_Species()
        s.collide_species(reaction)
        seed_gpu_from_oracle(lat, s)
        lat.Collide()
        s.collide()
        seed_gpu_from_oracle(lat, s)
        lat.Streaming()
        s.stream()
        seed_gpu_from_oracle(lat, s)
        lat.TMS_BC()
        s.tms_bc()
        e = relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after TMS {e:.3e}"
        e = relerr(lat.get_populations(POP_G).reshape(27, -1)[:, F], s.g[:, F])
        assert e < TOL, f"step {step} g after TMS {e:.3e}"
        seed_gpu_from_oracle(lat, s)
        lat.calculate_moment()
        s.moments()
        check_state(lat, s, TOL, f"step {step}")


def test_mc_flame_runs_stably_with_the_inflow_held():
    """200 steps of the reacting flame: finite everywhere, the first cell's
    temperature stays near the 300 K inflow and the burnt side stays hot."""
    fl = cases.premixed_flame_1d(64, 3, 3)
    lat, _ = pair(fl)
    lat.Init()
    lat.calculate_moment()
    lat.step(200)
    T = lat.get_field("T")[1:-1, 1, 1] * cases.LATTICE_K
    assert np.all(np.isfinite(T))
    assert abs(T[0] - 300.0) < 30.0 and T[-1] > 1800.0
    assert lat.diagnostics()["nan_count"] == 0


def test_mc_flame_full_size_calls_match_oracle():
    """The 2048x4x4 premixed flame at its configured size (LODI inflow,
    NSCBC outflow, 9 species): two steps call by call against the oracle,
    re-seeded before each call."""
    fl = cases.premixed_flame_1d(2048, 4, 4)
    lat, s = pair(fl, reaction=True)
    lat.Init()
    s.init()
    F = fluid(s)

    def check(where):
        # velocities on the scale of the velocity vector: the transverse
        # components of this quasi-1-D flame are rounding-level
        us = max(np.max(np.abs(s.u[F])), np.max(np.abs(s.v[F])), np.max(np.abs(s.w[F])))
        for k in ("rho", "u", "v", "w", "T", "p", "rhoe"):
            ref = getattr(s, k)[F]
            scale = us if k in ("u", "v", "w") else np.max(np.abs(ref))
            e = float(np.max(np.abs(lat.get_field(k).ravel()[F] - ref)) / scale)
            assert e < TOL, f"{where}: {k} {e:.3e}"
        e = relerr(lat.get_species_field("rho").reshape(9, -1)[:, F], s.rho_a[:, F])
        assert e < TOL, f"{where}: rho_a {e:.3e}"

    lat.calculate_moment()
    s.moments()
    check("moments after init")
    for step in range(1, 3):
        print(f"flame 2048x4x4 step {step}", flush=True)
        seed_gpu_from_oracle(lat, s)
        lat.Collide_Species()
        s.collide_species(True)
        e = relerr(lat.get_populations(POP_FPC).reshape(9, 27, -1)[:, :, F], s.fpc[:, :, F])
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
        e = relerr(lat.get_populations(POP_F).reshape(9, 27, -1)[:, :, F], s.f[:, :, F])
        assert e < TOL, f"step {step} species f after TMS {e:.3e}"
        seed_gpu_from_oracle(lat, s)
        lat.calculate_moment()
        s.moments()
        check(f"step {step}")


def test_mc_fp32_refuses_unresolvable_trace_species():
    """Trace species 1e-10 of their mean (the premixed flame's default) are
    below what binary32 deviations from the mean state resolve: FP32 lbm_init
    refuses the set-up by name instead of running into NaN."""
    from lbm_amd.capi import LBM_FP32_SHIFTED, LbmError
    fl = cases.premixed_flame_1d(64, 3, 3)
    lat = MixtureLattice(64, 3, 3, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP32_SHIFTED)
    lat.load_case(fl)
    with pytest.raises(LbmError, match="cannot resolve"):
        lat.Init()


def test_mc_fp32_refuses_traces_held_only_by_the_inflow_ghost():
    """A 1e-6 trace floor in the domain but 1e-10 traces in the TYPE_I_C
    inflow ghost, whose composition TMS_BC imposes on the first plane every
    step: the FP32 init refuses it as well."""
    from lbm_amd.capi import LBM_FP32_SHIFTED, LbmError
    fl = cases.premixed_flame_1d(64, 3, 3, minim=1e-6)
    fl["X"][:, 0] = cases.mole_vector(cases.X_UNBURNT_PHI1, 1e-10)[:, None, None]
    lat = MixtureLattice(64, 3, 3, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP32_SHIFTED)
    lat.load_case(fl)
    with pytest.raises(LbmError, match="cannot resolve"):
        lat.Init()
    lat.close()
    fl["X"][:, 0] = cases.mole_vector(cases.X_UNBURNT_PHI1, 1e-6)[:, None, None]
    lat = MixtureLattice(64, 3, 3, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP32_SHIFTED)
    lat.load_case(fl)
    lat.Init()
    lat.close()


def test_mc_fp32_full_size_flame_tracks_fp64():
    """The 2048x4x4 premixed flame with a 1e-6 trace floor: FP32 shifted
    storage against FP64 over 300 steps (both finite, flame temperature and
    species close)."""
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64
    fl = cases.premixed_flame_1d(2048, 4, 4, minim=1e-6)
    out = {}
    for prec in (LBM_FP64, LBM_FP32_SHIFTED):
        print(f"flame 2048x4x4 precision {prec}", flush=True)
        lat = MixtureLattice(2048, 4, 4, *cases.UNITS_FLAME, reaction=True, precision=prec)
        lat.load_case(fl)
        lat.Init()
        lat.calculate_moment()
        lat.step(300)
        rho_a = lat.get_species_field("rho")[:, 1:-1, 2, 2]
        out[prec] = (lat.get_field("T")[1:-1, 2, 2].copy(), rho_a / rho_a.sum(axis=0))
        assert lat.diagnostics()["nan_count"] == 0
        lat.close()
    (Ta, Ya), (Tb, Yb) = out[LBM_FP64], out[LBM_FP32_SHIFTED]
    eT = float(np.max(np.abs(Tb - Ta) / Ta))
    eY = float(np.sqrt(np.mean((Yb - Ya) ** 2)))
    print(f"FP32 vs FP64 after 300 steps: T max rel {eT:.3e}, Y L2 {eY:.3e}", flush=True)
    assert eT < 5e-4 and eY < 5e-6
