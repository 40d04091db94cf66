"""Pins the chemistry oracle (oracle/chem.py) before it is trusted.

Cantera, which the reference calls for these quantities, is absent offline,
so each function is pinned by what the mathematics and the reference's own
data fix:
  * NASA7 against direct evaluation of the YAML coefficients (when the
    reference is mounted) and cp = d h / dT, s consistency;
  * the adiabatic flame state the reference holds (setup.cpp:2106-2116,
    phi = 1 H2-air, T_b = 2387.637 K from 300 K): enthalpy balance;
  * kinetics: mass and element conservation, near-zero net rates at that
    equilibrium composition (a wrong reverse rate or sign breaks it), and a
    brute-force per-reaction evaluation at tiny inputs;
  * transport: Wilke and mixture rules against textbook restatements; fits
    against the kinetic-theory values they were fitted to.
"""
import math
import os

import numpy as np
import pytest

from oracle.chem import Mechanism

REF_YAML = "/root/reference/src/reaction-mech/nh3.yaml"

# phi = 1 H2-air, 1 atm (setup.cpp:2106-2116): unburnt at 300 K, burnt at 2387.637 K
X_U = {"H2": 0.29586, "O2": 0.14793, "N2": 0.55621}
X_B = {"H2": 0.014565, "H": 0.00181, "O": 0.00060738, "O2": 0.0056094, "OH": 0.0072848, "H2O": 0.32437,
       "HO2": 1.2571e-06, "H2O2": 1.3391e-07, "N2": 0.64575}
T_B = 2387.637


@pytest.fixture(scope="module")
def mech():
    return Mechanism()


def vec(m, d):
    x = np.zeros(m.ns)
    for k, v in d.items():
        x[m.names.index(k)] = v
    return x / x.sum()


def test_nasa7_matches_yaml_and_is_consistent(mech):
    T = np.array([300.0, 999.0, 1001.0, 2400.0])
    cp = mech.cp_R(T)
    h = mech.h_RT(T) * T[:, None]
    # dh/dT = cp (central difference on each branch)
    for t in (500.0, 1500.0):
        d = (mech.h_RT(np.array([t + 1e-3]))[0] * (t + 1e-3) - mech.h_RT(np.array([t - 1e-3]))[0] * (t - 1e-3)) / 2e-3
        np.testing.assert_allclose(d, mech.cp_R(np.array([t]))[0], rtol=1e-7)
        ds = (mech.s_R(np.array([t + 1e-3]))[0] - mech.s_R(np.array([t - 1e-3]))[0]) / 2e-3
        np.testing.assert_allclose(ds * t, mech.cp_R(np.array([t]))[0], rtol=1e-6)
    if os.path.exists(REF_YAML):
        import yaml
        doc = yaml.safe_load(open(REF_YAML))
        by = {str(s["name"]): s for s in doc["species"]}
        for i, n in enumerate(mech.names):
            for j, t in enumerate(T):
                a = by[n]["thermo"]["data"][0 if t < 1000 else 1]
                want = a[0] + a[1] * t + a[2] * t ** 2 + a[3] * t ** 3 + a[4] * t ** 4
                assert abs(cp[j, i] - want) < 1e-12 * abs(want)
                wh = (a[0] + a[1] / 2 * t + a[2] / 3 * t ** 2 + a[3] / 4 * t ** 3 + a[4] / 5 * t ** 4 + a[5] / t) * t
                assert abs(h[j, i] - wh) < 1e-10 * max(1.0, abs(wh))
    # sanity against well-known values: cp(N2, 300 K) ~ 29.1 J/mol/K, h_f(H2O) ~ -241.8 kJ/mol
    assert abs(mech.cp_R(np.array([300.0]))[0, mech.names.index("N2")] * 8.314 - 29.1) < 0.2
    hf = mech.h_molar(np.array([298.15]))[0, mech.names.index("H2O")] / 1e6
    assert abs(hf + 241.8) < 0.3


def test_adiabatic_flame_enthalpy_balance(mech):
    """h(300 K, unburnt) = h(T_b, burnt) per unit mass at constant pressure:
    solving for T with this thermo must land on the reference's 2387.637 K
    (its 5-digit compositions and a different thermo database allow ~10 K)."""
    Yu = mech.Y_from_X(vec(mech, X_U))
    Yb = mech.Y_from_X(vec(mech, X_B))
    hu = np.sum(Yu * mech.h_RT(np.array(300.0)) / mech.W) * mech.R * 300.0
    T = 2000.0
    for _ in range(50):
        hb = np.sum(Yb * mech.h_RT(np.array(T)) / mech.W) * mech.R * T
        T -= (hb - hu) / mech.cp_mass(np.array(T), Yb)
    assert abs(T - T_B) < 10.0, T


def test_detailed_balance_at_reference_equilibrium(mech):
    """At the reference's phi = 1 equilibrium state every reaction's reverse
    rate of progress equals its forward one up to the thermo-database
    difference (the state comes from another mechanism's thermo: e.g. OH's
    heat of formation differs by ~1.5 kJ/mol, ~9 % in K_c at 2400 K).  A sign
    error in Delta G or a wrong (p/RT)^dnu factor is off by orders of
    magnitude."""
    Yb = mech.Y_from_X(vec(mech, X_B))
    rho = mech.density(np.array(T_B), 101325.0, Yb)
    qf, qr = mech.rates_of_progress(np.array([T_B]), np.array([rho]), Yb[None, :])
    ratio = qr[0] / qf[0]
    assert np.all((ratio > 0.75) & (ratio < 1.3)), ratio


def test_kinetics_conserve_mass_and_elements(mech):
    rng = np.random.default_rng(3)
    n = 64
    Y = mech.normalize_Y(rng.uniform(0, 1, (n, mech.ns)))
    T = rng.uniform(800, 2600, n)
    rho = rng.uniform(0.05, 1.0, n)
    w = mech.net_production(T, rho, Y)
    scale = np.max(np.abs(w * mech.W), axis=1)
    assert np.all(np.abs(np.sum(w * mech.W, axis=1)) < 1e-12 * scale)
    H = np.array([2, 1, 0, 0, 1, 2, 1, 2, 0])
    O = np.array([0, 0, 1, 2, 1, 1, 2, 2, 0])
    N = np.array([0, 0, 0, 0, 0, 0, 0, 0, 2])
    for el in (H, O, N):
        assert np.all(np.abs(w @ el) < 1e-10 * np.max(np.abs(w), axis=1) + 1e-300)


def test_kinetics_brute_force(mech):
    """Per-reaction loop straight from the mass-action / Troe formulas."""
    import json
    d = json.load(open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "lattice-boltzmann-method_amd", "mech", "h2air.json")))
    names = mech.names
    T, rho = 1500.0, 0.3
    Y = mech.normalize_Y(np.linspace(1, 2, mech.ns))
    C = {n: rho * Y[i] / mech.W[i] for i, n in enumerate(names)}
    g = {n: (mech.h_RT(np.array(T))[i] - mech.s_R(np.array(T))[i]) for i, n in enumerate(names)}
    wdot = {n: 0.0 for n in names}
    for e in d["reactions"]:
        k = e.get("k_inf") or e["k"]
        kf = k["A"] * T ** k["b"] * math.exp(-k["Ea_R"] / T)
        M = sum(C[n] * e.get("efficiencies", {}).get(n, 1.0) for n in names)
        if e["type"] == "three-body":
            kf *= M
        elif e["type"] == "falloff":
            k0 = e["k0"]["A"] * T ** e["k0"]["b"] * math.exp(-e["k0"]["Ea_R"] / T)
            Pr = k0 * M / kf
            t = e["troe"]
            Fc = (1 - t["A"]) * math.exp(-T / t["T3"]) + t["A"] * math.exp(-T / t["T1"]) + \
                (math.exp(-t["T2"] / T) if t["T2"] is not None else 0.0)
            c = -0.4 - 0.67 * math.log10(Fc)
            nn = 0.75 - 1.27 * math.log10(Fc)
            f1 = (math.log10(Pr) + c) / (nn - 0.14 * (math.log10(Pr) + c))
            kf = kf * Pr / (1 + Pr) * 10 ** (math.log10(Fc) / (1 + f1 * f1))
        dn = sum(e["products"].values()) - sum(e["reactants"].values())
        dG = sum(v * g[n] for n, v in e["products"].items()) - sum(v * g[n] for n, v in e["reactants"].items())
        Kc = math.exp(-dG) * (101325.0 / (mech.R * T)) ** dn
        q = kf * math.prod(C[n] ** v for n, v in e["reactants"].items()) - \
            kf / Kc * math.prod(C[n] ** v for n, v in e["products"].items())
        for n, v in e["reactants"].items():
            wdot[n] -= v * q
        for n, v in e["products"].items():
            wdot[n] += v * q
    got = mech.net_production(np.array([T]), np.array([rho]), Y[None, :])[0]
    want = np.array([wdot[n] for n in names])
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(want)))


def test_uv_inversion_round_trip(mech):
    rng = np.random.default_rng(5)
    Y = mech.normalize_Y(rng.uniform(0, 1, (32, mech.ns)))
    T = rng.uniform(250, 3000, 32)
    u = mech.u_mass(T, Y)
    T2 = mech.T_from_UV(u, Y, np.full(32, 1000.0))
    np.testing.assert_allclose(T2, T, rtol=1e-13)


def test_wilke_and_mixture_rules(mech):
    T = np.array(1200.0)
    X = vec(mech, X_B)
    mu = mech.species_viscosities(T)
    # textbook Wilke, loops
    tot = 0.0
    for a in range(mech.ns):
        den = 0.0
        for b in range(mech.ns):
            phi = (1 + math.sqrt(mu[a] / mu[b]) * (mech.W[b] / mech.W[a]) ** 0.25) ** 2 / \
                math.sqrt(8 * (1 + mech.W[a] / mech.W[b]))
            den += X[b] * phi
        tot += X[a] * mu[a] / den
    np.testing.assert_allclose(mech.mix_viscosity(T, X), tot, rtol=1e-13)
    # pure species limit
    Xp = np.zeros(mech.ns)
    Xp[mech.names.index("N2")] = 1.0
    np.testing.assert_allclose(mech.mix_viscosity(T, Xp), mu[mech.names.index("N2")], rtol=1e-14)
    lam = mech.species_conductivities(T)
    np.testing.assert_allclose(mech.mix_conductivity(T, Xp), lam[mech.names.index("N2")], rtol=1e-14)
    # binary diffusion symmetric and ~1/p
    D1 = mech.binary_diff(T, np.array(101325.0))
    D2 = mech.binary_diff(T, np.array(2 * 101325.0))
    np.testing.assert_allclose(D1, D1.T, rtol=1e-12)
    np.testing.assert_allclose(D1, 2 * D2, rtol=1e-14)
    # physical magnitudes at 300 K, 1 atm (literature: D_H2-N2 ~ 7.8e-5, mu_N2 ~ 1.79e-5)
    D = mech.binary_diff(np.array(300.0), np.array(101325.0))
    iH2, iN2 = mech.names.index("H2"), mech.names.index("N2")
    assert 6.5e-5 < D[iH2, iN2] < 9e-5
    assert 1.7e-5 < mech.species_viscosities(np.array(300.0))[iN2] < 1.9e-5


def test_transport_fits_track_kinetic_theory(mech):
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "lattice-boltzmann-method_amd", "mech"))
    gm = pytest.importorskip("gen_mech")
    if not os.path.exists(REF_YAML):
        pytest.skip("reference mechanism not mounted")
    sp = gm.species_table(gm.load(REF_YAML), gm.MECHS["h2air"]["species"])
    for T in (300.0, 1000.0, 2500.0):
        mu = mech.species_viscosities(np.array(T))
        lam = mech.species_conductivities(np.array(T))
        D = mech.binary_diff(np.array(T), np.array(101325.0))
        for i, s in enumerate(sp):
            assert abs(mu[i] / gm.viscosity(s, T) - 1) < 3e-3
            assert abs(lam[i] / gm.conductivity(s, T) - 1) < 1.5e-2
            for j, s2 in enumerate(sp):
                assert abs(D[i, j] * 101325.0 / gm.binary_diff_times_p(s, s2, T) - 1) < 3e-3
