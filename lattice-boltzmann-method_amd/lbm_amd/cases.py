"""Case set-ups producing the initial cell types and primitive fields.

Each set-up returns ghost-inclusive (nx+2, ny+2, nz+2) arrays in the
reference's index order (i, j, k), k fastest, as LBM::LBM allocates them
(lbm-main.cpp:16-35).  Fields that a set-up does not touch keep the MIXTURE
defaults of lbm.hpp:56-88 (rho = 1, u = v = w = 0, p = T = 1/3, TYPE_F).
"""
import math

import numpy as np

TYPE_F, TYPE_S, TYPE_P, TYPE_A, TYPE_FS, TYPE_I, TYPE_O, TYPE_I_C, TYPE_O_C, TYPE_Q = range(10)


def default_fields(nx, ny, nz):
    shape = (nx + 2, ny + 2, nz + 2)
    return {
        "type": np.full(shape, TYPE_F, np.uint8),
        "rho": np.ones(shape),
        "u": np.zeros(shape),
        "v": np.zeros(shape),
        "w": np.zeros(shape),
        "p": np.full(shape, 1.0 / 3.0),
        "T": np.full(shape, 1.0 / 3.0),
    }


def periodic_ghosts(fields):
    t = fields["type"]
    t[0, :, :] = TYPE_P
    t[-1, :, :] = TYPE_P
    t[:, 0, :] = TYPE_P
    t[:, -1, :] = TYPE_P
    t[:, :, 0] = TYPE_P
    t[:, :, -1] = TYPE_P


def tgv_params(N, Re=1600.0, Ma=0.1, R=1.0 / 3.0, T0=1.0, p0=1.0 / 3.0, gamma=1.4, Pr=0.71):
    """Single-component 3-D Taylor-Green parameters: U0 = Ma sqrt(gamma R T0),
    nu = U0 L / Re with L = N / (2 pi)."""
    U0 = Ma * math.sqrt(gamma * R * T0)
    nu = U0 * (N / (2.0 * math.pi)) / Re
    return dict(N=N, U0=U0, nu=nu, R=R, gamma=gamma, Pr=Pr, T0=T0, p0=p0)


def taylor_green_3d(N, U0, p0, T0, **_unused):
    """3-D Taylor-Green vortex of setup.cpp:94-148 on an N^3 periodic box.

    The analytic fields are evaluated on the ghost-inclusive index i (the
    reference's set-up does the same), ghosts are TYPE_P.
    """
    fl = default_fields(N, N, N)
    periodic_ghosts(fl)
    n = N + 2
    idx = np.arange(n, dtype=np.float64)
    a = float(N)
    x = 2.0 * math.pi / a * idx
    I, J, K = np.meshgrid(x, x, x, indexing="ij")
    I2, J2, K2 = np.meshgrid(2 * x, 2 * x, 2 * x, indexing="ij")
    m = fl["type"] == TYPE_F
    u = U0 * np.sin(I) * np.cos(J) * np.cos(K)
    v = -U0 * np.cos(I) * np.sin(J) * np.cos(K)
    p = p0 + 1.0 * U0 * U0 / 16.0 * (np.cos(I2) + np.cos(J2)) * (np.cos(K2) + 2.0)
    fl["u"][m] = u[m]
    fl["v"][m] = v[m]
    fl["w"][m] = 0.0
    fl["p"][m] = p[m]
    fl["T"][m] = T0
    return fl


def channel(nx, ny, nz, U, T0=1.0, p0=1.0 / 3.0, T_in=1.1, T_wall=0.9, p_out=None, R=1.0 / 3.0):
    """Single-component channel with every TMS_BC boundary kind.

    Fixed-state inflow ghost (TYPE_I: u = U, v = 0.05 U, T = T_in) at i = 0,
    outflow ghost (TYPE_O: p = p_out) at i = nx+1, isothermal walls (TYPE_S,
    T = T_wall) at j = 0 (at rest) and j = ny+1 (moving with u = -0.3 U),
    periodic in z; the x faces are typed last.  Interior: u = U with
    transverse waves v = 0.1 U sin(2 pi i/nx) sin(pi j/(ny+1)) and
    w = 0.05 U cos(2 pi (k-1)/nz) sin(pi j/(ny+1)), T = T0, p = p0.
    """
    fl = default_fields(nx, ny, nz)
    t = fl["type"]
    t[:, :, 0] = TYPE_P
    t[:, :, -1] = TYPE_P
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    t[0, :, :] = TYPE_I
    t[-1, :, :] = TYPE_O
    shape = t.shape
    i, j, k = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    for name in ("u", "v", "w"):
        fl[name][:] = 0.0
    fl["T"][:] = T0
    fl["p"][:] = p0
    F = t == TYPE_F
    # libm's sin/cos (math module) element by element, as a C set-up evaluates them
    msin, mcos = np.vectorize(math.sin), np.vectorize(math.cos)
    sj = msin(math.pi * j / (ny + 1))
    fl["u"][F] = U
    fl["v"][F] = (0.1 * U * msin(2.0 * math.pi * i / nx) * sj)[F]
    fl["w"][F] = (0.05 * U * mcos(2.0 * math.pi * (k - 1) / nz) * sj)[F]
    I = t == TYPE_I
    fl["u"][I] = U
    fl["v"][I] = 0.05 * U
    fl["T"][I] = T_in
    fl["p"][t == TYPE_O] = p0 if p_out is None else p_out
    S = t == TYPE_S
    fl["T"][S] = T_wall
    fl["u"][S & (j == shape[1] - 1)] = -0.3 * U
    fl["rho"][:] = fl["p"] / (R * fl["T"])
    return fl


def conjugate_box(nx, ny, nz, ns, U, T0=1.0, p0=1.0 / 3.0, T_hot=1.2, T_cold=0.9, R=1.0 / 3.0):
    """Conjugate heat-transfer box (run with the conjugate solid enabled).

    Periodic ghosts on every face; solid slabs (TYPE_S) of `ns` cells at the
    bottom (j <= ns, T_hot) and the top (j >= ny+1-ns, T_cold) of the y range,
    at rest; gas between them at T0, p0 with u = U and transverse waves
    v = 0.1 U sin(2 pi i/nx) sin(pi y), w = 0.05 U cos(2 pi (k-1)/nz) sin(pi y),
    y = (j - ns - 1/2) / (ny - 2 ns).  The slabs meet through the periodic y
    wrap, so every solid cell has only solid, periodic or gas neighbours.
    """
    fl = default_fields(nx, ny, nz)
    t = fl["type"]
    shape = t.shape
    i, j, k = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    for name in ("u", "v", "w"):
        fl[name][:] = 0.0
    fl["T"][:] = T0
    fl["p"][:] = p0
    interior = (i > 0) & (i < shape[0] - 1) & (j > 0) & (j < shape[1] - 1) & (k > 0) & (k < shape[2] - 1)
    lo = interior & (j <= ns)
    hi = interior & (j >= shape[1] - 1 - ns)
    t[lo | hi] = TYPE_S
    periodic_ghosts(fl)
    fl["T"][lo] = T_hot
    fl["T"][hi] = T_cold
    F = t == TYPE_F
    msin, mcos = np.vectorize(math.sin), np.vectorize(math.cos)
    sy = msin(math.pi * (j - ns - 0.5) / (ny - 2 * ns))
    fl["u"][F] = U
    fl["v"][F] = (0.1 * U * msin(2.0 * math.pi * i / nx) * sy)[F]
    fl["w"][F] = (0.05 * U * mcos(2.0 * math.pi * (k - 1) / nz) * sy)[F]
    fl["rho"][:] = fl["p"] / (R * fl["T"])
    return fl


def decaying_turbulence(N, seed=20261205, k0=4.0, ma_rms=0.3, t_fluct=0.05, R=1.0 / 3.0, T0=1.0, p0=1.0 / 3.0,
                        gamma=1.4, Pr=0.71, nu=None, **_unused):
    """Thermal compressible decaying turbulence on an N^3 periodic box.

    Seeded random solenoidal velocity with energy spectrum
    E(k) ~ k^4 exp(-2 k^2 / k0^2), scaled so that sqrt(<|u|^2>) / c0 = ma_rms
    with c0 = sqrt(gamma R T0); temperature T0 (1 + t_fluct * U(-1, 1));
    uniform pressure p0.  Returns (fields, params)."""
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(N, d=1.0 / N)
    kr = np.fft.rfftfreq(N, d=1.0 / N)
    KX, KY, KZ = np.meshgrid(k, k, kr, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    Kmag = np.sqrt(K2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(Kmag > 0, np.sqrt(Kmag ** 4 * np.exp(-2.0 * K2 / k0 ** 2) / (4.0 * np.pi * K2)), 0.0)
    comps = []
    for _ in range(3):
        comps.append((rng.standard_normal(KX.shape) + 1j * rng.standard_normal(KX.shape)) * amp)
    # project onto the plane normal to k (solenoidal)
    with np.errstate(divide="ignore", invalid="ignore"):
        dot = (KX * comps[0] + KY * comps[1] + KZ * comps[2]) / np.where(K2 > 0, K2, 1.0)
    comps = [comps[0] - KX * dot, comps[1] - KY * dot, comps[2] - KZ * dot]
    vel = [np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2)) for c in comps]
    c0 = math.sqrt(gamma * R * T0)
    urms = math.sqrt(float(np.mean(vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2)))
    scale = ma_rms * c0 / urms
    fl = default_fields(N, N, N)
    periodic_ghosts(fl)
    s = (slice(1, -1),) * 3
    fl["u"][s] = vel[0] * scale
    fl["v"][s] = vel[1] * scale
    fl["w"][s] = vel[2] * scale
    fl["T"][s] = T0 * (1.0 + t_fluct * rng.uniform(-1.0, 1.0, (N, N, N)))
    fl["p"][s] = p0
    if nu is None:
        nu = 1e-3
    params = dict(N=N, R=R, T0=T0, p0=p0, gamma=gamma, Pr=Pr, nu=nu, seed=seed, k0=k0, ma_rms=ma_rms)
    return fl, params


# ----------------------------------------------------------- multicomponent
H2AIR_SPECIES = ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "N2"]
# species order of the two-step mechanisms (the libraries' lbm_species_name)
MECH_SPECIES = {"h2air": H2AIR_SPECIES, "ch4_2s": ["O2", "H2O", "CH4", "CO", "CO2", "N2"],
                "c3h8": ["CO", "C3H8", "CO2", "H2O", "N2", "O2"],
                "nh3": ["AR", "N2", "HE", "H2", "H", "O2", "O", "H2O", "OH", "H2O2", "HO2", "NO", "N2O", "NO2", "HNO",
                        "HNO2", "HONO", "HONO2", "N2H2", "H2NN", "NH2OH", "HNOH", "NH3", "N2H4", "N", "NO3", "NH",
                        "NNH", "NH2", "H2NO", "N2H3"]}
# phi = 1 H2-air, unburnt mole fractions (setup.cpp:2108-2116)
X_UNBURNT_PHI1 = {"H2": 0.29586, "O2": 0.14793, "N2": 0.55621}
X_AIR = {"O2": 0.21, "N2": 0.79}
# stoichiometric fuel-air mixtures of the two-step mechanisms (air = O2 + 3.76 N2)
X_CH4_AIR_PHI1 = {"CH4": 1.0 / 10.52, "O2": 2.0 / 10.52, "N2": 7.52 / 10.52}
X_C3H8_AIR_PHI1 = {"C3H8": 1.0 / 24.8, "O2": 5.0 / 24.8, "N2": 18.8 / 24.8}
# phi = 1 ammonia/hydrogen (60/40 by mole) in air: 0.6 NH3 + 0.4 H2 + 0.65 (O2 + 3.76 N2)
X_NH3_H2_AIR_PHI1 = {"NH3": 0.6 / 4.094, "H2": 0.4 / 4.094, "O2": 0.65 / 4.094, "N2": 2.444 / 4.094}
UNITS_FLAME = (5e-6, 1e-9)  # units.set_m_kg_s(dx, dt) of the flame cases (setup.cpp:2019)
LATTICE_K = 12000.0          # temperature unit of set_m_kg_s(dx, dt) (units.hpp:36)
ONE_ATM = 101325.0


def mole_vector(d, minim=1e-10, species=None):
    x = np.array([d.get(n, minim) for n in (species or H2AIR_SPECIES)], dtype=np.float64)
    return x / x.sum()


def lattice_p(p_si, dx=UNITS_FLAME[0], dt=UNITS_FLAME[1]):
    """units.p of the flame units: p_si * m s^2 / kg with kg = m^3."""
    m = np.float32(dx)
    s = np.float32(dt)
    kg = np.float32(float(m) ** 3)
    return float(np.float32(p_si * (float(m) * float(s) ** 2) / float(kg)))


def reacting_box(nx, ny, nz, T_K=1400.0, X=None, amp_T=0.0, amp_u=0.0, seed=20261206, species=None):
    """Periodic box of premixed gas (the PERFECTLY_STIRRED_REACTOR_3D shape,
    setup.cpp:950-999, with optional smooth temperature and velocity
    perturbations).  Returns fields with species mole fractions 'X' (NS, ...)
    in the order of `species` (H2-air by default)."""
    species = species or H2AIR_SPECIES
    fl = default_fields(nx, ny, nz)
    periodic_ghosts(fl)
    X = mole_vector(X or X_UNBURNT_PHI1, species=species)
    shape = fl["type"].shape
    i, j, k = np.meshgrid(*(np.arange(s, dtype=np.float64) for s in shape), indexing="ij")
    m = fl["type"] == TYPE_F
    T = T_K * (1.0 + amp_T * np.sin(2 * np.pi * i / nx) * np.cos(2 * np.pi * j / ny))
    fl["T"][m] = (T / LATTICE_K)[m]
    fl["p"][m] = lattice_p(ONE_ATM)
    fl["u"][m] = (amp_u * np.sin(2 * np.pi * j / ny))[m]
    fl["v"][m] = (amp_u * np.cos(2 * np.pi * i / nx))[m]
    fl["w"][m] = (0.5 * amp_u * np.sin(2 * np.pi * k / nz))[m]
    fl["X"] = np.zeros((len(species),) + shape)
    fl["X"][:, m] = X[:, None]
    return fl


X_FUEL_ML = {"H2": 0.3, "N2": 0.7}  # N2-diluted hydrogen stream of the mixing layer


def _all_species(d, minim=1e-6, species=None):
    out = {n: minim for n in (species or H2AIR_SPECIES)}
    out.update(d)
    return out


def reacting_channel(nx, ny, nz, T_K=900.0, U=0.01, T_in_K=950.0, T_wall_K=800.0, X=None, X_in=None,
                     p_out_factor=1.0, species=None):
    """Multicomponent channel with the fixed-state boundary kinds of TMS_BC.

    Inflow ghost (TYPE_I: u = U, v = 0.05 U, T_in, composition X_in) at
    i = 0, outflow ghost (TYPE_O: p = p_out_factor atm) at i = nx+1,
    isothermal walls (TYPE_S at T_wall) at j = 0 (at rest) and j = ny+1
    (moving with u = -0.3 U), periodic in z; the x faces are typed last.
    Interior: u = U with transverse waves, T_K, 1 atm, composition X (every
    species present, so that the SPECIES_MIN clamp stays out of the way).
    Temperatures in K; 'X' holds mole fractions (NS, ...); `species`: the
    species order of a mechanism other than H2-air."""
    fl = default_fields(nx, ny, nz)
    t = fl["type"]
    t[:, :, 0] = TYPE_P
    t[:, :, -1] = TYPE_P
    t[:, 0, :] = TYPE_S
    t[:, -1, :] = TYPE_S
    t[0, :, :] = TYPE_I
    t[-1, :, :] = TYPE_O
    shape = t.shape
    i, j, k = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    Xv = mole_vector(_all_species(X or X_UNBURNT_PHI1, species=species), species=species)
    Xi = mole_vector(_all_species(X_in or X_AIR, species=species), species=species)
    fl["X"] = np.zeros((len(species or H2AIR_SPECIES),) + shape)
    fl["X"][:] = Xv[:, None, None, None]
    fl["T"][:] = T_K / LATTICE_K
    fl["p"][:] = lattice_p(ONE_ATM)
    for name in ("u", "v", "w"):
        fl[name][:] = 0.0
    F = t == TYPE_F
    sj = np.sin(np.pi * j / (ny + 1))
    fl["u"][F] = U
    fl["v"][F] = (0.1 * U * np.sin(2 * np.pi * i / nx) * sj)[F]
    fl["w"][F] = (0.05 * U * np.cos(2 * np.pi * (k - 1) / nz) * sj)[F]
    I = t == TYPE_I
    fl["u"][I] = U
    fl["v"][I] = 0.05 * U
    fl["T"][I] = T_in_K / LATTICE_K
    fl["X"][:, I] = Xi[:, None]
    fl["p"][t == TYPE_O] = lattice_p(ONE_ATM * p_out_factor)
    S = t == TYPE_S
    fl["T"][S] = T_wall_K / LATTICE_K
    fl["u"][S & (j == shape[1] - 1)] = -0.3 * U
    return fl


def mixing_layer(nx, ny, nz, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=None, noise=0.01, seed=20261206,
                 X_fuel=None, X_ox=None, minim=1e-6, species=None):
    """Temporal H2-air / hot-air mixing layer: periodic x and z, NSCBC outflow
    (TYPE_O_C) on both y faces (the y ghost layers are typed last, so edges and
    corners are TYPE_O_C).  Lower stream: N2-diluted H2 at T_fuel; upper
    stream: air at T_ox; u = U_s tanh((y - y0)/delta) with seeded broadband
    noise of `noise` U_s on u, v, w; p = 1 atm everywhere (also the outflow
    target pressure held by the ghosts).  `species`: the species order of a
    mechanism other than H2-air (the others at `minim`)."""
    rng = np.random.default_rng(seed)
    fl = default_fields(nx, ny, nz)
    t = fl["type"]
    t[0, :, :] = TYPE_P
    t[-1, :, :] = TYPE_P
    t[:, :, 0] = TYPE_P
    t[:, :, -1] = TYPE_P
    t[:, 0, :] = TYPE_O_C
    t[:, -1, :] = TYPE_O_C
    shape = t.shape
    delta = delta or max(ny / 16.0, 1.0)
    j = np.arange(shape[1], dtype=np.float64)
    s = 0.5 * (1.0 + np.tanh((j - 0.5 * (shape[1] - 1)) / delta))  # 0 fuel side .. 1 oxidiser side
    s3 = np.broadcast_to(s[None, :, None], shape)
    Xf = mole_vector(_all_species(X_fuel or X_FUEL_ML, minim, species), species=species)
    Xo = mole_vector(_all_species(X_ox or X_AIR, minim, species), species=species)
    fl["X"] = (1.0 - s3)[None] * Xf[:, None, None, None] + s3[None] * Xo[:, None, None, None]
    fl["T"][...] = (T_fuel + (T_ox - T_fuel) * s3) / LATTICE_K
    fl["p"][...] = lattice_p(ONE_ATM)
    fl["u"][...] = U_s * (2.0 * s3 - 1.0)
    fl["v"][...] = 0.0
    fl["w"][...] = 0.0
    m = t == TYPE_F
    for k in ("u", "v", "w"):
        fl[k][m] += noise * U_s * rng.uniform(-1.0, 1.0, int(m.sum()))
    return fl


def double_mixing_layer(nx, ny, nz, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=None, noise=0.01, seed=20261206,
                        X_fuel=None, X_ox=None, minim=1e-6, species=None):
    """Periodic twin of mixing_layer: the same streams and profiles, with the
    fuel stream in the middle half of y between two tanh layers at ny/4 and
    3 ny/4, so that the box is periodic in every direction (TYPE_P ghosts)
    and smooth across its faces; u = U_s (2 s - 1) as in mixing_layer."""
    rng = np.random.default_rng(seed)
    fl = default_fields(nx, ny, nz)
    periodic_ghosts(fl)
    t = fl["type"]
    shape = t.shape
    delta = delta or max(ny / 32.0, 1.0)
    j = np.arange(shape[1], dtype=np.float64) - 0.5  # interior cell j sits at j - 1/2 of [0, ny)
    y1, y2 = 0.25 * ny, 0.75 * ny
    s = 1.0 - 0.5 * (np.tanh((j - y1) / delta) - np.tanh((j - y2) / delta))  # 1 oxidiser .. 0 fuel
    s3 = np.broadcast_to(s[None, :, None], shape)
    Xf = mole_vector(_all_species(X_fuel or X_FUEL_ML, minim, species), species=species)
    Xo = mole_vector(_all_species(X_ox or X_AIR, minim, species), species=species)
    fl["X"] = (1.0 - s3)[None] * Xf[:, None, None, None] + s3[None] * Xo[:, None, None, None]
    fl["T"][...] = (T_fuel + (T_ox - T_fuel) * s3) / LATTICE_K
    fl["p"][...] = lattice_p(ONE_ATM)
    fl["u"][...] = U_s * (2.0 * s3 - 1.0)
    fl["v"][...] = 0.0
    fl["w"][...] = 0.0
    m = t == TYPE_F
    for k in ("u", "v", "w"):
        fl[k][m] += noise * U_s * rng.uniform(-1.0, 1.0, int(m.sum()))
    return fl


# phi = 1 burnt state (setup.cpp:2106-2116)
T_BURNT_PHI1 = 2387.637
X_BURNT_PHI1 = {"H2": 0.014565, "H": 0.00181, "O": 0.00060738, "O2": 0.0056094, "OH": 0.0072848, "H2O": 0.32437,
                "HO2": 1.2571e-06, "H2O2": 1.3391e-07, "N2": 0.64575}
S_L_PHI1 = 2.2  # m/s, laminar burning velocity of the phi = 1 H2-air flame used as the inflow speed


def _molar_masses():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mech", "h2air.json")
    with open(path) as f:
        d = json.load(f)
    return np.array([s["W"] for s in d["species"]])


def smooth(left, right, x, center, alpha):
    """math-util.cpp:89-91: left + (1 + tanh(alpha (x - center))) / 2 (right - left)."""
    return left + 0.5 * (1.0 + np.tanh(alpha * (x - center))) * (right - left)


def premixed_flame_1d(nx, ny=4, nz=4, u_in=None, noise=0.01, seed=20261204, midpoint=0.7, alpha=0.2, minim=1e-10):
    """Quasi-1-D premixed H2-air flame (the FLAME_SPEED set-up of
    setup.cpp:2014-2124 at phi = 1): periodic y/z; TYPE_I_C inflow ghost at
    i = 0 holding the unburnt state (300 K, 1 atm, u_in), TYPE_O_C outflow
    ghost at i = Nx-1 holding 1 atm; a tanh front between the unburnt and the
    burnt phi = 1 states at `midpoint` Nx (smooth(), alpha = 0.2); seeded
    +-`noise` relative white noise on the mass fractions, renormalised."""
    rng = np.random.default_rng(seed)
    fl = default_fields(nx, ny, nz)
    t = fl["type"]
    t[:, 0, :] = TYPE_P
    t[:, -1, :] = TYPE_P
    t[:, :, 0] = TYPE_P
    t[:, :, -1] = TYPE_P
    t[0, :, :] = TYPE_I_C
    t[-1, :, :] = TYPE_O_C
    shape = t.shape
    Nx = shape[0]
    if u_in is None:
        u_in = S_L_PHI1 * UNITS_FLAME[1] / UNITS_FLAME[0]
    i = np.arange(Nx, dtype=np.float64)
    Xu, Xb = mole_vector(X_UNBURNT_PHI1, minim), mole_vector(X_BURNT_PHI1, minim)
    X = smooth(Xu[:, None], Xb[:, None], i[None, :], midpoint * Nx, alpha)  # (NS, Nx)
    T = smooth(300.0, T_BURNT_PHI1, i, midpoint * Nx, alpha)
    X3 = np.broadcast_to(X[:, :, None, None], (len(H2AIR_SPECIES),) + shape).copy()
    # noise on the mass fractions of the fluid cells
    W = _molar_masses()
    m = t == TYPE_F
    Y = X3 * W[:, None, None, None]
    Y /= Y.sum(axis=0, keepdims=True)
    Y[:, m] *= 1.0 + noise * rng.uniform(-1.0, 1.0, (len(H2AIR_SPECIES), int(m.sum())))
    Y /= Y.sum(axis=0, keepdims=True)
    Xn = Y / W[:, None, None, None]
    fl["X"] = Xn / Xn.sum(axis=0, keepdims=True)
    fl["T"][...] = np.broadcast_to(T[:, None, None], shape) / LATTICE_K
    # the inflow ghosts hold the fresh mixture exactly
    fl["X"][:, 0] = Xu[:, None, None]
    fl["T"][0] = 300.0 / LATTICE_K
    fl["p"][...] = lattice_p(ONE_ATM)
    fl["u"][...] = u_in
    fl["v"][...] = 0.0
    fl["w"][...] = 0.0
    return fl


# units.set_m_kg_s(dx, dt) of the transport wave columns: dt/dx = 4e-4 keeps
# theta of the lightest species (H, 300 K) at 0.4, and dx puts tau - 1/2 of air
# near 0.9, so viscous, thermal and diffusive decays take a few hundred steps
UNITS_TRANSPORT = (5e-7, 2e-10)


def transport_column(N, X, T_K=300.0, p_si=ONE_ATM, units=UNITS_TRANSPORT, minim=1e-10):
    """Quasi-1-D periodic column of N cells along y (1 x N x 1 interior,
    periodic ghosts everywhere) holding a uniform mixture X (mole fractions by
    species name) at rest at T_K and p_si: the shape of the reference's
    VISCOSITY_TEST_MULTICOMP (setup.cpp:1672-1742, 1 x 200 x 1, periodic).
    Returns the fields and the y index of each interior cell from 0."""
    fl = default_fields(1, N, 1)
    periodic_ghosts(fl)
    shape = fl["type"].shape
    x = mole_vector(X, minim)
    fl["X"] = np.zeros((len(x),) + shape)
    fl["X"][:] = x[:, None, None, None]
    fl["T"][...] = T_K / LATTICE_K
    fl["p"][...] = lattice_p(p_si, *units)
    return fl, np.arange(N, dtype=np.float64)


def wave_column(kind, N, amp, X=None, T_K=300.0, units=UNITS_TRANSPORT, species=None, phase=0.0):
    """Single-mode waves on transport_column, phase 2 pi (j + phase) / N over
    the interior cells j = 0..N-1 (ghost-inclusive index j + 1):
      'shear'      u = amp sin           (VISCOSITY_TEST_MULTICOMP, setup.cpp:1672)
      'thermal'    T = T0 (1 + amp sin), p uniform   (CONDUCTIVITY_TEST, setup.cpp:396)
      'diffusion'  X_a = X0_a + amp sin for the first named species and
                   X0_b - amp sin for the second, p and T uniform
                   (TERNARY_DIFFUSION reduced to two species, setup.cpp:768)
      'sound'      p = p0 (1 + amp cos), T = T0 (1 + (1 - 1/gamma_guess) amp cos)
                   (SOUNDSPEED_TEST, setup.cpp:463)"""
    X = X or X_AIR
    fl, j = transport_column(N, X, T_K, units=units)
    ph = 2.0 * np.pi * (j + phase) / N
    sl = (1, slice(1, -1), 1)
    if kind == "shear":
        fl["u"][sl] = amp * np.sin(ph)
    elif kind == "thermal":
        fl["T"][sl] *= 1.0 + amp * np.sin(ph)
    elif kind == "diffusion":
        names = list(X)
        a, b = names[0], names[1]
        for q in range(N):
            xq = dict(X)
            xq[a] = X[a] + amp * np.sin(ph[q])
            xq[b] = X[b] - amp * np.sin(ph[q])
            fl["X"][:, 1, q + 1, 1] = mole_vector(xq, species=species)
    elif kind == "sound":
        fl["p"][sl] *= 1.0 + amp * np.cos(ph)
        fl["T"][sl] *= 1.0 + (1.0 - 1.0 / 1.4) * amp * np.cos(ph)
    else:
        raise ValueError(kind)
    return fl


def outflow_column(N, X=None, T_K=300.0, dp_in=0.0, dv_in=0.0, p_ghost_rel=1.0, units=UNITS_TRANSPORT):
    """Column of N cells along y at rest with NSCBC outflow ghosts (TYPE_O_C) on
    both y faces, periodic in x and z (the y boundaries of the mixing layer),
    for single TMS_BC calls: uniform T and composition; pressure
    p0 (1 + dp_in q) and velocity v = dv_in q at distance q = 0, 1, ... from
    the lower face, mirrored at the upper face; the ghosts hold p0 p_ghost_rel."""
    X = X or X_AIR
    fl, j = transport_column(N, X, T_K, units=units)
    t = fl["type"]
    t[:, 0, :] = TYPE_O_C
    t[:, -1, :] = TYPE_O_C
    p0 = fl["p"][1, 1, 1]
    q = np.minimum(j, N - 1 - j)
    sign = np.where(j < N / 2, 1.0, -1.0)
    fl["p"][1, 1:-1, 1] = p0 * (1.0 + dp_in * q)
    fl["v"][1, 1:-1, 1] = sign * dv_in * q
    fl["p"][:, 0, :] = p0 * p_ghost_rel
    fl["p"][:, -1, :] = p0 * p_ghost_rel
    return fl
