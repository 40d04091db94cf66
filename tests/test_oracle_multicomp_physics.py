"""Closed-form pins of the multicomponent LBM oracle (oracle/multicomp.py).

The reference's MULTICOMP build needs Cantera and Eigen, both absent, so the
oracle cannot be compared with the reference binary.  These tests pin its
collision, moment and boundary arithmetic by what the mathematics fixes,
with each expected value computed from the mechanism data by the textbook
formula (tests/physics_waves.py), not by the oracle's own transport code.
The shapes are the reference's own physics cases (setup.cpp) reduced to
1 x N x 1 periodic columns; rates converge at second order in 1/N, so each
is measured at N and 2N and extrapolated (Richardson) before comparing.

Each docstring names the one-line oracle mutations that the test was seen to
fail (applied by hand to a copy of the oracle, not kept).
"""
import numpy as np
import pytest

from lbm_amd import cases
from oracle.chem import Mechanism
from oracle.multicomp import MultiState, MultiStateBC, Units

import physics_waves as pw

MECH = Mechanism()
SP = cases.H2AIR_SPECIES
P_ATM = cases.ONE_ATM


def column_state(fl, cls=MultiState, units=cases.UNITS_TRANSPORT):
    nx, ny, nz = fl["type"].shape
    s = cls(nx, ny, nz, Units(*units))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(len(SP), -1)
    s.init()
    s.moments()
    return s


def col(s, name):
    """Interior column (1, 1..N, 1) of a field."""
    return getattr(s, name).reshape(s.dims)[1, 1:-1, 1]


def run_series(fl, steps, probe):
    s = column_state(fl)
    out = []
    for _ in range(steps):
        s.step(reaction=False)
        out.append(probe(s))
    return np.array(out), s


def measured_ratio(kind, N, steps, amp, X, expected):
    fl = cases.wave_column(kind, N, amp, X=X)
    if kind == "shear":
        series, _ = run_series(fl, steps, lambda s: pw.y_mode(col(s, "u"), N))
    elif kind == "thermal":
        series, _ = run_series(fl, steps, lambda s: pw.y_mode(col(s, "T"), N))
    else:
        a = SP.index(list(X)[0])
        series, _ = run_series(fl, steps, lambda s: pw.y_mode(s.X.reshape((len(SP),) + s.dims)[a, 1, 1:-1, 1], N))
    return pw.decay_rate(series) / expected


def test_shear_wave_decays_at_the_wilke_mixture_viscosity():
    """VISCOSITY_TEST_MULTICOMP (setup.cpp:1672-1742) in air: a transverse
    velocity mode decays at nu k^2 with nu = mu_Wilke / rho.  Each species
    relaxes with omega_a = 2 X_a p / (X_a p + 2 visc_a) (lbm-collision.cpp:
    412-429); its partial pressure X_a p turns 1/omega_a - 1/2 into the
    species stress visc_a, and the visc_a sum to the Wilke viscosity.
    Fails with 2.2 visc in omega_a (ratio 0.907 -> 0.974 at N = 16)."""
    X = cases.mole_vector(cases.X_AIR)
    r = [measured_ratio("shear", N, steps, 1e-3, cases.X_AIR, pw.expected_shear_rate(MECH, N, 300.0, P_ATM, X))
         for N, steps in ((16, 200), (32, 400))]
    assert abs(r[1] - 1.0) < 0.04  # second-order lattice error at N = 32
    assert abs(pw.richardson(*r) - 1.0) < 0.005


def test_thermal_wave_decays_at_kappa_over_rho_cp():
    """CONDUCTIVITY_TEST (setup.cpp:396-461) in air: an isobaric temperature
    mode decays at kappa k^2 / (rho c_p) with the mixture-averaged
    conductivity; the energy populations relax with omega1 =
    2 p c_p / (p c_p + 2 kappa) (lbm-collision.cpp:236-254).  Fails with
    2.2 kappa in omega1 (ratio 0.835 -> 0.881 at N = 16)."""
    X = cases.mole_vector(cases.X_AIR)
    r = [measured_ratio("thermal", N, steps, 1e-3, cases.X_AIR,
                        pw.expected_thermal_rate(MECH, N, 300.0, P_ATM, X)) for N, steps in ((16, 200), (32, 400))]
    assert abs(r[1] - 1.0) < 0.06
    assert abs(pw.richardson(*r) - 1.0) < 0.006


def test_binary_diffusion_decays_at_the_binary_diffusivity():
    """TERNARY_DIFFUSION (setup.cpp:768-864) reduced to H2/N2: at uniform p and
    T a composition mode decays at D_ab k^2.  That needs the friction impulse
    -p X_a X_b / D_ab (u_a - u_b) / rho_a of the collision
    (lbm-collision.cpp:463-470) and the half-step implicit Maxwell-Stefan
    solve of the moments, A_ab = -p X_a X_b / (2 D_ab) (lbm-moment.cpp:
    313-341), to agree.  Fails with the friction scaled by 1.1 (ratio 1.014
    -> 0.846) or the solve's 1/2 dropped (-> 1.825)."""
    mix = {"H2": 0.5, "N2": 0.5}
    exp = [pw.expected_diffusion_rate(MECH, N, 300.0, P_ATM, SP.index("H2"), SP.index("N2")) for N in (16, 32)]
    r = [measured_ratio("diffusion", N, steps, 0.02, mix, e) for (N, steps), e in zip(((16, 200), (32, 300)), exp)]
    assert abs(r[1] - 1.0) < 0.02
    assert abs(pw.richardson(*r) - 1.0) < 0.005


def test_sound_wave_travels_at_the_frozen_sound_speed():
    """SOUNDSPEED_TEST (setup.cpp:463-529) in air: a standing pressure mode
    oscillates with period N / sqrt(gamma R T) (lattice units), gamma =
    c_p / c_v of the NASA7 mixture.  The pressure follows from the species
    equilibria at theta_a = R_a T and the temperature from the energy
    moments (lbm-moment.cpp:351-368), so the period pins the energy
    equilibrium's internal energy and c_v."""
    X = cases.mole_vector(cases.X_AIR)
    r = []
    for N, steps in ((16, 300), (32, 500)):
        fl = cases.wave_column("sound", N, 1e-4)
        series, _ = run_series(fl, steps, lambda s: pw.y_mode(col(s, "p"), N, "cos"))
        r.append(pw.expected_sound_period(MECH, N, 300.0, P_ATM, X) / pw.zero_crossing_period(series))
    assert abs(r[1] - 1.0) < 0.003
    assert abs(pw.richardson(*r) - 1.0) < 0.001


def test_diffusion_carries_the_species_enthalpy():
    """H2O/N2 inter-diffusion at uniform p and T keeps T uniform only when the
    energy populations carry the enthalpy flux sum h_a rho_a (u_a - u)
    (q_diff and the upwinded q_corr of lbm-collision.cpp:278-288).  The drift
    is compared with the frozen-energy estimate: the T change if the
    internal energy stayed where it was while the composition moved,
    -sum_a u_a(T) dY_a / c_v.  Away from the two rows next to the periodic
    ghosts (quirk Q2: fd_fuw returns 0 there) the drift stays below a fifth
    of that estimate.  Fails with q_diff dropped (interior drift 6 %) or
    q_corr dropped (5 %)."""
    N, steps = 32, 200
    mix = {"H2O": 0.3, "N2": 0.7}
    fl = cases.wave_column("diffusion", N, 0.02, X=mix)
    s = column_state(fl)
    T0 = col(s, "T").copy()
    Y0 = (s.rho_a / s.rho).reshape((len(SP),) + s.dims)[:, 1, 1:-1, 1].copy()
    for _ in range(steps):
        s.step(reaction=False)
    T1 = col(s, "T")
    Y1 = (s.rho_a / s.rho).reshape((len(SP),) + s.dims)[:, 1, 1:-1, 1]
    Tsi = 300.0
    u_a = (MECH.h_RT(np.array(Tsi)) - 1.0) * MECH.R * Tsi / MECH.W  # internal energies [J/kg]
    cv = pw.cp_mass(MECH, Tsi, Y0[:, 0]) - MECH.R / (1.0 / np.sum(Y0[:, 0] / MECH.W))
    frozen = np.abs(-(u_a[:, None] * (Y1 - Y0)).sum(axis=0) / cv) / Tsi
    inner = slice(2, N - 2)
    drift = np.abs(T1 / T0.mean() - 1.0)
    assert frozen[inner].max() > 0.05  # the composition moved enough for the check to bite
    assert drift[inner].max() < 0.2 * frozen[inner].max()


def test_maxwell_stefan_velocities_solve_their_system():
    """calculate_moment (lbm-moment.cpp:313-341): the species velocities solve
    A u = (rho u)_a with A_ab = -p X_a X_b / (2 D_ab) for b != a and
    A_aa = rho_a - sum_{b != a} A_ab, all at the previous state's p, X and
    (T, p) of the diffusivities.  The residual is checked against a matrix
    built here from the mechanism's D_ab, on a diffusing H2/O2/N2 column
    whose species velocities differ."""
    fl = cases.wave_column("diffusion", 16, 0.05, X={"H2": 0.3, "O2": 0.2, "N2": 0.5})
    s = column_state(fl)
    for _ in range(3):
        s.step(reaction=False)
    s.collide_species(reaction=False)
    s.collide()
    s.stream()
    F = np.nonzero(s.type == 0)[0]
    X_old, p_old, T_old, rho_old = s.X[:, F].copy(), s.p[F].copy(), s.T[F].copy(), s.rho[F].copy()
    Y_old = MECH.normalize_Y((s.rho_a[:, F] / rho_old).T)
    s.moments()
    U = s.units
    mom = np.einsum("dl,alc->dac", np.stack([cases_c(d) for d in range(3)]).astype(float), s.f[:, :, F])
    worst = 0.0
    for q, c in enumerate(F):
        Tsi = np.float64(U.si_temp(T_old[q]))
        p_si = MECH.pressure(np.array(Tsi), np.float64(U.si_rho(rho_old[q])), Y_old[q])
        D = U.nu(MECH.binary_diff(np.array(Tsi), np.array(p_si))).astype(np.float64)
        A = -0.5 * p_old[q] * np.outer(X_old[:, q], X_old[:, q]) / np.where(np.eye(len(SP)) > 0, 1.0, D)
        np.fill_diagonal(A, 0.0)
        A[np.diag_indices(len(SP))] = s.rho_a[:, c] - A.sum(axis=1)
        res = A @ s.u_a[:, :, c].T - mom[:, :, q].T
        worst = max(worst, np.abs(res).max() / np.abs(mom[:, :, q]).max())
    assert worst < 1e-12
    spread = np.ptp(s.u_a[1][:3, F], axis=0).max()
    assert spread > 1e-6  # the species really move apart


def cases_c(d):
    from oracle.multicomp import C3
    return C3[d]


@pytest.mark.parametrize("isign", [1, -1])
def test_nscbc_outflow_target_is_the_lodi_step(isign):
    """impose_NSCBC for a TYPE_O_C face (impose_nscbc.cpp:107-170, 391-426,
    479-545) on a quasi-1-D column: no transverse terms (tangential extents
    <= 3), so the target is one explicit LODI step with the outgoing wave
    from second-order one-sided derivatives and the incoming one relaxed
    towards the ghost pressure with K = 0.25 (1 - M^2) a / L, L the
    ghost-inclusive length.  TMS_BC then rebuilds the interface cell so that
    its density, momentum and species densities are the target's exactly
    (unknown directions 2 f_tgt - f_loc, known f_tgt + f - f_loc).  The
    interior carries a pressure and a velocity gradient, so both waves are
    non-zero.  Fails with the sign of the outgoing wave's rho a dv term
    flipped, with the relaxed wave put on the other side, or with sigma
    changed."""
    N = 10
    fl = cases.outflow_column(N, dp_in=2e-4, dv_in=3e-4, p_ghost_rel=0.995)
    s = column_state(fl, MultiStateBC)
    j = 1 if isign > 0 else N  # ghost-inclusive index of the interface cell
    idx = [j, j + 1, j + 2] if isign > 0 else [j, j - 1, j - 2]
    fields = {k: getattr(s, k).reshape(s.dims)[1, idx, 1].copy() for k in ("rho", "v", "p", "T")}
    Xc = s.X.reshape((len(SP),) + s.dims)[:, 1, j, 1].copy()
    _, _, Rm, gam = pw.gas_state(MECH, float(fields["T"][0] * cases.LATTICE_K), 1.0, Xc)
    a = np.sqrt(gam * Rm * fields["T"][0] * cases.LATTICE_K) * pw.lattice_scale()[1]
    p_inf = s.p.reshape(s.dims)[1, 0, 1]
    tgt = pw.lodi_outflow_target(fields, a, p_inf, isign, L=N + 2)
    assert abs(tgt["L1"]) > 1e-9 and abs(tgt["L5"]) > 1e-9
    s.tms_bc()
    s.moments()
    rho = s.rho.reshape(s.dims)[1, j, 1]
    v = s.v.reshape(s.dims)[1, j, 1]
    p = s.p.reshape(s.dims)[1, j, 1]
    # the sound speed and the DP temperature cross into SI through binary32
    # (~6e-8 per crossing); the LODI step moves p by ~2e-4 and v from 0
    assert abs(p / fields["p"][0] - 1.0) > 1e-5
    assert abs(rho / tgt["rho"] - 1.0) < 1e-9
    assert abs(v - tgt["v"]) < 1e-6 * abs(tgt["v"])
    assert abs(p / tgt["p"] - 1.0) < 5e-7
