"""Pins the multicomponent LBM oracle (oracle/multicomp.py).

The reference's MULTICOMP build needs Cantera and Eigen (absent), so parity
against the reference binary is unpinned; each routine is pinned instead by
what the scheme itself fixes:
  * a uniform premixed box is a set of identical 0-D constant-volume
    reactors: its temperature and composition must follow the chemistry
    oracle's own two-sub-step explicit-Euler + UV integration exactly;
  * streaming conserves every species' mass in a periodic box;
  * the species collision conserves total momentum (the Maxwell-Stefan
    friction is antisymmetric) and adds exactly the chemistry source mass;
  * a uniform non-reacting mixture at rest stays at rest;
  * the Maxwell-Stefan solve satisfies its linear system.
"""
import numpy as np
import pytest

from lbm_amd import cases
from oracle.chem import Mechanism
from oracle.multicomp import C3, CX, CY, CZ, MultiState, MultiStateBC, Units


def make_state(fl, units=None):
    units = units or Units(*cases.UNITS_FLAME)
    nx, ny, nz = fl["type"].shape
    s = MultiState(nx, ny, nz, units)
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(len(cases.H2AIR_SPECIES), -1)
    return s


SEEDED = dict(cases.X_UNBURNT_PHI1, H=1e-4, O=1e-4, OH=1e-4, H2O=1e-3, HO2=1e-4, H2O2=1e-4)


def test_uniform_box_is_a_0d_reactor():
    # minor species seeded above SPECIES_MIN so that the clamp (a separate
    # reference quirk) stays out of the comparison
    fl = cases.reacting_box(4, 4, 4, T_K=1400.0, X=SEEDED)
    s = make_state(fl)
    s.init()
    s.moments()  # the reference's set-ups call calculate_moment before loop when restarting; here: consistent X
    m, U = s.mech, s.units
    c = np.nonzero(s.type == 0)[0][0]
    # 0-D reactor with the same arithmetic: U fixed, rho fixed, 2 Euler sub-steps + UV per step
    rho = s.rho[c]
    Y = m.normalize_Y(s.rho_a[:, c] / rho)
    rho_a = s.rho_a[:, c].copy()
    T_lat = s.T[c]
    Uint = s.rhoe[c] / rho
    for step in range(6):
        Tsi = np.float64(U.si_temp(T_lat))
        rho_si = np.float64(U.si_rho(rho))
        src = np.zeros(m.ns)
        Tcur, Ycur = Tsi, Y
        for _ in range(2):
            w = m.net_production(np.array([Tcur]), np.array([rho_si]), Ycur[None])[0]
            src += 0.5 * U.rho_dot(w * m.W).astype(np.float64)
            Ycur = m.normalize_Y((rho_a + src) / rho)
            Tcur = m.T_from_UV(np.float64(U.si_energy_mass(Uint)), Ycur[None], np.array([Tcur]))[0]
        rho_a = rho_a + src
        rho = rho_a.sum()
        Y = m.normalize_Y(rho_a / rho)
        Uint = s.rhoe[c] / s.rho[c] if step == 0 else Uint
        s.step()
        T_lat = np.float64(U.temp(m.T_from_UV(np.float64(U.si_energy_mass(Uint)), Y[None],
                                               np.array([np.float64(U.si_temp(T_lat))]))[0]))
        got = s.rho_a[:, s.type == 0]
        assert np.allclose(got, got[:, :1], rtol=0, atol=1e-15), "box lost uniformity"
        np.testing.assert_allclose(s.rho_a[:, c], rho_a, rtol=1e-9, atol=1e-13)
        np.testing.assert_allclose(s.T[c], T_lat, rtol=1e-6)
    assert s.T[c] * cases.LATTICE_K > 1400.0  # the mixture has started to heat up


def test_streaming_conserves_species_mass():
    fl = cases.reacting_box(6, 4, 4, T_K=1200.0, amp_T=0.1, amp_u=0.02)
    s = make_state(fl)
    s.init()
    s.moments()
    s.collide_species()
    s.collide()
    F = s.type == 0
    before = s.fpc[:, :, F].sum(axis=(1, 2))
    gb = s.gpc[:, F].sum()
    s.stream()
    np.testing.assert_allclose(s.f[:, :, F].sum(axis=(1, 2)), before, rtol=1e-14)
    np.testing.assert_allclose(s.g[:, F].sum(), gb, rtol=1e-14)


def test_species_collision_momentum_and_mass_budget():
    fl = cases.reacting_box(6, 4, 4, T_K=1500.0, amp_T=0.1, amp_u=0.02, X=SEEDED)
    s = make_state(fl)
    s.init()
    s.moments()
    s.step()
    c = np.nonzero(s.type == 0)[0]
    f0 = s.f[:, :, c].copy()
    s.collide_species(reaction=False)
    d = s.fpc[:, :, c] - f0
    # no species is clamped here, so each species' mass is conserved exactly
    np.testing.assert_allclose(d.sum(axis=1), 0.0, atol=1e-14)
    # total momentum change = sum_a rho_a F_a = 0 by antisymmetry of the friction
    dm = np.einsum("dl,alc->dc", C3.astype(float), d)
    scale = np.abs(np.einsum("dl,alc->dc", C3.astype(float), f0)).max()
    assert np.abs(dm).max() < 1e-12 * max(scale, 1e-12)
    # with chemistry: each species gains exactly its source mass
    s.collide_species(reaction=True)
    gained = (s.fpc[:, :, c] - f0).sum(axis=1)
    assert np.abs(gained.sum(axis=0)).max() < 1e-6 * np.abs(gained).max() + 1e-18  # binary32 rho_dot


def test_rest_state_is_invariant_without_reaction():
    air = {n: 1e-6 for n in cases.H2AIR_SPECIES}
    air.update(cases.X_AIR)
    fl = cases.reacting_box(4, 4, 4, T_K=900.0, X=air)
    s = make_state(fl)
    s.init()
    s.moments()
    r0, T0 = s.rho.copy(), s.T.copy()
    for _ in range(3):
        s.collide_species(reaction=False)
        s.collide()
        s.stream()
        s.moments()
    F = s.type == 0
    np.testing.assert_allclose(s.rho[F], r0[F], rtol=1e-13)
    np.testing.assert_allclose(s.T[F], T0[F], rtol=1e-6)
    assert np.abs(s.u[F]).max() < 1e-14


def make_bc_state(fl):
    from oracle.multicomp import MultiStateBC
    nx, ny, nz = fl["type"].shape
    s = MultiStateBC(nx, ny, nz, Units(*cases.UNITS_FLAME))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(len(cases.H2AIR_SPECIES), -1)
    return s


def test_outflow_interface_cells_are_the_y_faces():
    fl = cases.mixing_layer(4, 6, 4)
    s = make_bc_state(fl)
    cells = s.interface_cells()
    assert len(cells) == 2 * 4 * 4
    ls = {l for _, l in cells}
    assert ls == {3, 4}  # +y (boundary below) and -y (boundary above)


def test_outflow_keeps_a_uniform_rest_state():
    """A uniform mixture at rest whose ghosts hold its own pressure: all LODI
    waves vanish, the NSCBC target is the current state and the TMS
    reconstruction must leave the state unchanged."""
    fl = cases.mixing_layer(4, 6, 4, T_fuel=900.0, T_ox=900.0, U_s=0.0, noise=0.0, X_fuel=cases.X_AIR)
    s = make_bc_state(fl)
    s.init()
    s.moments()
    F = s.type == 0
    r0, T0, p0 = s.rho.copy(), s.T.copy(), s.p.copy()
    for _ in range(4):
        s.step(reaction=False)
    # the NSCBC closure crosses into SI through binary32 (setState_DP of
    # si_rho/si_p), so the target state carries float rounding: ~1e-9
    np.testing.assert_allclose(s.rho[F], r0[F], rtol=1e-8)
    np.testing.assert_allclose(s.T[F], T0[F], rtol=1e-6)
    assert np.abs(s.v[F]).max() < 1e-9


def test_outflow_relaxes_toward_target_pressure():
    """With the ghosts' target pressure 0.5 % below the interior, the
    partially non-reflecting outflow must pull the boundary pressure down."""
    fl = cases.mixing_layer(4, 8, 4, T_fuel=900.0, T_ox=900.0, U_s=0.0, noise=0.0, X_fuel=cases.X_AIR)
    g = fl["type"] == cases.TYPE_O_C
    fl["p"][g] *= 0.995
    s = make_bc_state(fl)
    s.init()
    s.moments()
    nx, ny, nz = s.dims
    p_edge0 = s.p.reshape(nx, ny, nz)[1:-1, 1, 1:-1].mean()
    for _ in range(5):
        s.step(reaction=False)
    p_edge = s.p.reshape(nx, ny, nz)[1:-1, 1, 1:-1].mean()
    assert p_edge < p_edge0
    assert np.all(np.isfinite(s.p))


def _flame_state(nx=6, u_ghost=None, T_ghost=None):
    fl = cases.premixed_flame_1d(nx, 3, 3, noise=0.0, midpoint=1e3)  # unburnt everywhere
    if u_ghost is not None:
        fl["u"][0] = u_ghost
    if T_ghost is not None:
        fl["T"][0] = T_ghost / cases.LATTICE_K
    s = MultiStateBC(nx + 2, 5, 5, Units(*cases.UNITS_FLAME))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(9, -1)
    s.init()
    s.moments()
    return s, fl


def test_inflow_interface_cells_are_the_first_plane():
    s, _ = _flame_state()
    cells = s.interface_cells((5, 7))  # TYPE_I, TYPE_I_C
    assert len(cells) == 9 and all(l == 1 for _, l in cells)  # direction +x
    assert all(s._ijk(c)[0] == 1 for c, _ in cells)


def test_inflow_lodi_relaxes_velocity_towards_the_ghost():
    """With a uniform state, dp = du = 0 and the only wave is the relaxation
    L5 = K (u - u_b), K = sigma a^2 rho (1 - M^2) / N: the target velocity is
    u - sigma a (1 - M^2) (u - u_b) / (2 N) (lbm-bc.cpp:556-590, N ghost-inclusive)."""
    u0 = 4.4e-4
    s, fl = _flame_state(u_ghost=u0 * 0.5)
    c, l = s.interface_cells((5, 7))[0]
    a, _, _, _ = s._sound(c)
    u = s.u[c]
    N = fl["type"].shape[0]
    expect = u - 5.0 * a * (1 - (u / a) ** 2) * (u - 0.5 * u0) / (2 * N)
    _, _, vel_in, _ = s._inflow_lodi(c, l)
    assert abs(vel_in[0] - expect) < 1e-9 * abs(u)
    assert abs(vel_in[1]) < 1e-15 and abs(vel_in[2]) < 1e-15


def test_inflow_tms_keeps_a_matching_rest_state():
    """Ghost state equal to the interior equilibrium: the inflow target, the
    local and the target equilibria coincide, so TMS_BC changes nothing
    beyond the binary32 rounding of the unit crossings (T and rho pass
    through float several times, ~6e-8 relative each, and the corner
    directions scale with theta^3)."""
    s, fl = _flame_state(u_ghost=None)
    s.u[:] = 0.0  # interior and ghost at rest
    s.v[:] = 0.0
    s.w[:] = 0.0
    s.init()
    s.moments()
    f0, g0 = s.f.copy(), s.g.copy()
    s._tms_inflow()
    F = s.type == 0
    assert np.max(np.abs(s.f[:, :, F] - f0[:, :, F]) / np.abs(f0[:, :, F])) < 1e-5
    assert np.max(np.abs(s.g[:, F] - g0[:, F])) < 1e-5 * np.max(np.abs(g0[:, F]))


# ----------------------------------------- fixed-state TMS kinds (I, O, S)
def _channel(**kw):
    fl = cases.reacting_channel(6, 4, 3, **kw)
    s = make_bc_state(fl)
    s.init()
    s.moments()
    return s, fl


def test_moment_point_matches_calculate_moment_at_a_settled_state():
    """calculate_moment_point and calculate_moment differ only in the state
    their Maxwell-Stefan matrix and diffusivities are taken at and in the
    clamp; on a state that calculate_moment has just produced (every species
    above SPECIES_MIN), the point version must return the same moments."""
    s, _ = _channel()
    s.moments()
    c = int(np.nonzero(s.type == 0)[0][5])
    ref = {k: getattr(s, k)[..., c].copy() for k in ("rho", "u", "v", "w", "T", "p", "rhoe", "rho_a", "X", "u_a")}
    s._moment_point(c)
    for k, v in ref.items():
        np.testing.assert_allclose(getattr(s, k)[..., c], v, rtol=1e-9, atol=1e-18, err_msg=k)


def test_fixed_state_sweeps_keep_a_matching_rest_state():
    """Interior at rest in equilibrium, inflow ghost at the interior's
    velocity, temperature and composition, walls at rest at the interior
    temperature, outflow ghost at the interior pressure: every target equals
    the local state, so the three sweeps leave the populations unchanged up
    to the binary32 unit crossings."""
    s, fl = _channel(U=0.0, T_in_K=900.0, T_wall_K=900.0, X_in=cases.X_UNBURNT_PHI1)
    s.u[:] = 0.0
    s.v[:] = 0.0
    s.w[:] = 0.0
    s.init()
    s.moments()
    f0, g0 = s.f.copy(), s.g.copy()
    s.tms_bc()
    F = s.type == 0
    assert np.max(np.abs(s.f[:, :, F] - f0[:, :, F]) / np.abs(f0[:, :, F])) < 1e-5
    assert np.max(np.abs(s.g[:, F] - g0[:, F])) < 1e-5 * np.max(np.abs(g0[:, F]))


def test_wall_sweep_sets_species_density_and_velocity_targets():
    """For each species, the populations the wall sweep writes have zeroth
    moment rho_a,bb (the cell's own) and first moment rho_a,bb u_a,tgt,
    u_a,tgt the q = 1/2 interpolation between the species velocity
    downstream and the wall's velocity, averaged over the interface links:
    unknown 2 f_t - f_loc plus known f_t + f - f_loc sum to the target's
    moments because f_loc matches the local moments."""
    s, _ = _channel()
    # distinct species velocities in the fields, so that a mixture-velocity
    # target would fail
    rng = np.random.default_rng(5)
    F = np.nonzero(s.type == 0)[0]
    s.u_a[:, :, F] += 1e-3 * rng.standard_normal((3, s.ns, len(F)))
    c = s._idx(3, 1, 1)  # first layer above the lower wall
    i, j, k = s._ijk(c)
    links = []
    for l in range(27):
        up = s._idx(i - CX[l], j - CY[l], k - CZ[l])
        dn = s._idx(i + CX[l], j + CY[l], k + CZ[l])
        if s.type[up] == 1 and s.type[dn] == 0:
            links.append((up, dn))
    assert len(links) == 6  # the cz = -1 links end in the periodic ghost layer (k = 0)
    want_ra = s.f[:, :, c].sum(axis=1)
    want_va = np.zeros((s.ns, 3))
    for up, dn in links:
        uw = np.array([s.u[up], s.v[up], s.w[up]])
        want_va += (0.5 * s.u_a[:, :, dn].T + uw) / 1.5
    want_va /= len(links)
    s._tms_wall()
    got_ra = s.f[:, :, c].sum(axis=1)
    got_m = s.f[:, :, c] @ C3.T.astype(np.float64)
    np.testing.assert_allclose(got_ra, want_ra, rtol=1e-12)
    np.testing.assert_allclose(got_m / got_ra[:, None], want_va, rtol=1e-9, atol=1e-14)


def test_plain_outflow_takes_the_downstream_densities():
    """TYPE_O target: the species densities of the cell downstream,
    recomputed by calculate_moment_point from its populations (no clamp):
    after the sweep the interface cell's zeroth moments equal them."""
    s, _ = _channel(p_out_factor=0.99)
    s.collide_species(False)
    s.collide()
    s.stream()
    cells = [(c, l) for c, l in s.interface_cells() if s.type[c - s._strides()[0] * CX[l]] == 6]
    assert cells and all(l == 2 for _, l in cells)  # direction -x, ghost at i = nx+1
    c, l = cells[0]
    d = c + s._strides()[0] * CX[l]
    want = s.f[:, :, d].sum(axis=1)  # populations of d are not touched by the sweep
    s.tms_bc()
    np.testing.assert_allclose(s.rho_a[:, d], want, rtol=1e-14)  # moment_point wrote d's fields
    np.testing.assert_allclose(s.f[:, :, c].sum(axis=1), want, rtol=1e-12)


def test_plain_inflow_imposes_the_ghost_velocity():
    """TYPE_I: the target f_tgt is the equilibrium of the local density and
    composition at the ghost's velocity, so the interface cell's momentum
    after the sweep is rho_loc u_ghost."""
    s, _ = _channel()
    s.collide_species(False)
    s.collide()
    s.stream()
    cells = s.interface_cells((5, 7))
    assert cells and all(l == 1 for _, l in cells)
    c, l = [(c, l) for c, l in cells if s._ijk(c)[1] == 2][0]  # away from the walls (their sweep runs last)
    g_ = c - s._strides()[0]
    ug = np.array([s.u[g_], s.v[g_], s.w[g_]])
    s.tms_bc()
    fa = s.f[:, :, c]
    m = (fa @ C3.T.astype(np.float64)).sum(axis=0)
    np.testing.assert_allclose(m / fa.sum(), ug, rtol=1e-9, atol=1e-15)


def test_free_slip_streaming_mirrors_and_conserves():
    """Free-slip walls on the y faces (dirSlip, lbm-bc.cpp:1352-1399): every
    population of a cell next to the lower wall that comes from the wall takes
    the mirrored direction (cy -> -cy) of the neighbour at the same height,
    and streaming conserves every species' mass and the total energy."""
    fl = cases.reacting_box(4, 5, 3, T_K=1400.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    fl["type"][:, 0, :] = cases.TYPE_FS
    fl["type"][:, -1, :] = cases.TYPE_FS
    s = make_state(fl)
    s.init()
    rng = np.random.default_rng(4)
    s.fpc[:] = s.f * (1.0 + 0.01 * rng.standard_normal(s.f.shape))
    s.gpc[:] = s.g * (1.0 + 0.01 * rng.standard_normal(s.g.shape))
    F = s.type == 0
    mass0 = s.fpc[:, :, F].sum(axis=(1, 2))
    energy0 = s.gpc[:, F].sum()
    s.stream()
    np.testing.assert_allclose(s.f[:, :, F].sum(axis=(1, 2)), mass0, rtol=1e-13)
    np.testing.assert_allclose(s.g[:, F].sum(), energy0, rtol=1e-13)
    nx, ny, nz = s.dims
    idx = lambda i, j, k: (i * ny + j) * nz + k
    checked = 0
    for l in range(27):
        if CY[l] != 1:
            continue
        lp = int(np.nonzero((CX == CX[l]) & (CY == -1) & (CZ == CZ[l]))[0][0])
        for i in range(1, nx - 1):
            for k in range(1, nz - 1):
                ii = (i - CX[l] - 1) % (nx - 2) + 1
                kk = (k - CZ[l] - 1) % (nz - 2) + 1
                np.testing.assert_array_equal(s.f[:, l, idx(i, 1, k)], s.fpc[:, lp, idx(ii, 1, kk)])
                assert s.g[l, idx(i, 1, k)] == s.gpc[lp, idx(ii, 1, kk)]
                checked += 1
    assert checked == 9 * (nx - 2) * (nz - 2)
