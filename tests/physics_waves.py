"""Closed forms of the multicomponent transport waves and the NSCBC outflow,
shared by the oracle tests and their GPU mirrors.

Everything here is computed from the mechanism data (NASA7 thermo, species
viscosity and conductivity fits, binary diffusivities) by the textbook
formulas, independently of the LBM arithmetic under test:
  * Wilke mixture viscosity mu = sum_a X_a mu_a / sum_b X_b phi_ab;
  * mixture-averaged conductivity 1/2 (sum X l + 1 / sum(X / l));
  * thermal diffusivity kappa / (rho c_p), binary diffusivity D_ab;
  * frozen sound speed sqrt(gamma R T / W), gamma = c_p / (c_p - R / W);
  * the 1-D LODI update of a partially non-reflecting outflow (Poinsot and
    Lele 1992) with the relaxation K = sigma (1 - M^2) a / L.
"""
import numpy as np

from lbm_amd import cases


def y_mode(a, N, phase="sin"):
    """Amplitude of the single mode sin (or cos) 2 pi j / N of a column profile."""
    ph = 2.0 * np.pi * np.arange(N) / N
    return 2.0 / N * np.sum(a * (np.sin(ph) if phase == "sin" else np.cos(ph)))


def decay_rate(series, skip=0.25):
    """Exponential decay rate per step of a positive series (least squares on
    the log after the first `skip` fraction)."""
    s = np.asarray(series, dtype=np.float64)
    t = np.arange(1, len(s) + 1)
    k = int(len(s) * skip)
    return -np.polyfit(t[k:], np.log(np.abs(s[k:])), 1)[0]


def zero_crossing_period(series):
    s = np.asarray(series, dtype=np.float64)
    zc = [i + s[i] / (s[i] - s[i + 1]) for i in range(len(s) - 1) if s[i] * s[i + 1] < 0]
    return 2.0 * float(np.mean(np.diff(zc)))


def richardson(r_coarse, r_fine):
    """Second-order extrapolation of a ratio measured at N and 2N."""
    return (4.0 * r_fine - r_coarse) / 3.0


def lattice_scale(units=cases.UNITS_TRANSPORT):
    dx, dt = units
    return dt / (dx * dx), dt / dx  # diffusivities, velocities


def mass_fractions(mech, X):
    y = X * mech.W
    return y / y.sum()


def wilke_viscosity(mech, T, X):
    mu = mech.species_viscosities(np.array(T))
    W = mech.W
    phi = np.empty((mech.ns, mech.ns))
    for a in range(mech.ns):
        for b in range(mech.ns):
            phi[a, b] = (1.0 + np.sqrt(mu[a] / mu[b]) * (W[b] / W[a]) ** 0.25) ** 2 / np.sqrt(8.0 * (1.0 + W[a] / W[b]))
    return float(np.sum(X * mu / (phi @ X)))


def mixture_conductivity(mech, T, X):
    lam = mech.species_conductivities(np.array(T))
    return float(0.5 * (np.sum(X * lam) + 1.0 / np.sum(X / lam)))


def cp_mass(mech, T, Y):
    a = mech.hi if T >= mech.Tmid[0] else mech.lo
    cpR = a[:, 0] + T * (a[:, 1] + T * (a[:, 2] + T * (a[:, 3] + T * a[:, 4])))
    return float(np.sum(Y * cpR / mech.W) * mech.R)


def gas_state(mech, T, p, X):
    """(rho, c_p, R_mix, gamma) in SI of a mixture at T, p."""
    Y = mass_fractions(mech, X)
    Wm = 1.0 / np.sum(Y / mech.W)
    R = mech.R / Wm
    cp = cp_mass(mech, T, Y)
    return p / (R * T), cp, R, cp / (cp - R)


def expected_shear_rate(mech, N, T, p, X, units=cases.UNITS_TRANSPORT):
    rho, _, _, _ = gas_state(mech, T, p, X)
    k = 2.0 * np.pi / N
    return wilke_viscosity(mech, T, X) / rho * lattice_scale(units)[0] * k * k


def expected_thermal_rate(mech, N, T, p, X, units=cases.UNITS_TRANSPORT):
    rho, cp, _, _ = gas_state(mech, T, p, X)
    k = 2.0 * np.pi / N
    return mixture_conductivity(mech, T, X) / (rho * cp) * lattice_scale(units)[0] * k * k


def expected_diffusion_rate(mech, N, T, p, a, b, units=cases.UNITS_TRANSPORT):
    k = 2.0 * np.pi / N
    D = mech.binary_diff(np.array(T), np.array(p))[a, b]
    return float(D) * lattice_scale(units)[0] * k * k


def expected_sound_period(mech, N, T, p, X, units=cases.UNITS_TRANSPORT):
    _, _, R, gam = gas_state(mech, T, p, X)
    c = np.sqrt(gam * R * T) * lattice_scale(units)[1]
    return N / c


def lodi_outflow_target(col, a, p_inf, isign, L, sigma=0.25):
    """Target (rho, v, p) of the interface cell of a 1-D outflow along y.

    col: dict of the 'rho', 'v', 'p' values of the interface cell and its
    next two inner neighbours (index 0, 1, 2 going into the domain); the face
    is below the cell (isign = +1) or above it (isign = -1).  One-sided
    second-order derivatives along +y, the characteristic waves
    L1 = (v - a)(dp - rho a dv) (towards -y), L5 = (v + a)(dp + rho a dv)
    (towards +y), Ls = v (a^2 drho - dp), the incoming one of L1/L5 replaced by
    K (p - p_inf) with K = sigma (1 - M^2) a / L, then one LODI step (dt = 1):
    rho -= (Ls + (L1 + L5) / 2) / a^2, v -= (L5 - L1) / (2 rho a),
    p -= (L1 + L5) / 2."""
    d = {k: isign * (-1.5 * col[k][0] + 2.0 * col[k][1] - 0.5 * col[k][2]) for k in ("rho", "v", "p")}
    rho, v, p = col["rho"][0], col["v"][0], col["p"][0]
    mach = abs(v) / a
    K = sigma * (1.0 - mach * mach) * a / L
    L1 = (v - a) * (d["p"] - rho * a * d["v"])
    L5 = (v + a) * (d["p"] + rho * a * d["v"])
    Ls = v * (a * a * d["rho"] - d["p"])
    if isign > 0:
        L5 = K * (p - p_inf)
    else:
        L1 = K * (p - p_inf)
    return {"rho": rho - (Ls + 0.5 * (L1 + L5)) / (a * a), "v": v - (L5 - L1) / (2.0 * a * rho),
            "p": p - 0.5 * (L1 + L5), "L1": L1, "L5": L5}
