"""Pins the oracle's full ammonia/hydrogen mechanism (nh3.yaml: 31 species,
203 reactions) before it is trusted.

What fixes each part without Cantera:
  * the tables: 31 species in the phase's order with NO read as a name (YAML
    1.1 would read it as false), 203 reactions of which 18 are PLOG, 8
    falloff (one Lindemann) and 7 three-body; the committed JSON is what the
    generator makes from the reference's file;
  * NASA7 thermo: formation enthalpies and standard entropies at 298.15 K of
    the nitrogen species, argon and helium (JANAF / ATcT tables);
  * PLOG (NH3 <=> NH2 + H, levels 0.1, 1, 10, 100 atm): the rate constant
    at a level's pressure is that level's Arrhenius rate written out from the
    published cm-mol-cal parameters, at the geometric mean of two levels it
    is the geometric mean of their rates (ln k linear in ln p), and outside
    the table the end level's rate;
  * Lindemann falloff (N2O (+M) <=> N2 + O (+M)): k_inf Pr / (1 + Pr) with
    Pr = k0 [M] / k_inf and [M] from the published efficiencies;
  * the H2/O2 module: reactions 1-23 give the same rates of progress as the
    separately pinned H2-air tables on states without argon and helium
    (their efficiencies are the only difference);
  * mass and H/O/N/Ar/He element conservation of the net production rates;
  * transport: viscosities and binary diffusivities with N2 near 300 K
    against measured values.
"""
import json
import math
import os

import numpy as np
import pytest

from oracle.chem import Mechanism, mechanism_path

REF_DIR = "/root/reference/src/reaction-mech"
R_CAL = 8.31446261815324 / 4.184  # cal/(mol K)
ATM = 101325.0

SPECIES = ["AR", "N2", "HE", "H2", "H", "O2", "O", "H2O", "OH", "H2O2", "HO2", "NO", "N2O", "NO2", "HNO", "HNO2",
           "HONO", "HONO2", "N2H2", "H2NN", "NH2OH", "HNOH", "NH3", "N2H4", "N", "NO3", "NH", "NNH", "NH2", "H2NO",
           "N2H3"]
# 298.15 K formation enthalpy [kJ/mol] and standard entropy [J/(mol K)]
HF = {"NH3": -45.90, "NO": 91.1, "N2O": 82.05, "NO2": 34.0, "N": 472.68, "AR": 0.0, "HE": 0.0, "N2H4": 95.35,
      "HONO2": -134.3, "N2": 0.0, "H2O": -241.83}
S0 = {"NH3": 192.77, "NO": 210.76, "N2O": 219.96, "NO2": 240.04, "N": 153.30, "AR": 154.85, "HE": 126.15,
      "HONO2": 266.4, "N2": 191.61, "H2O": 188.83}
# NH3 <=> NH2 + H (reaction 24): P [atm], A [1/s], b, Ea [cal/mol]
PLOG_NH3 = [(0.1, 7.23e29, -5.316, 1.108624e5), (1.0, 3.497e30, -5.224, 1.111633e5),
            (10.0, 1.975e31, -5.155, 1.118878e5), (100.0, 2.689e31, -4.92, 1.127787e5)]
# N2O (+M) <=> N2 + O (+M) (reaction 188): k0 [cm^3/mol/s], k_inf [1/s], efficiencies
N2O_LOW = (6.02e14, 0.0, 5.7444e4)
N2O_HIGH = (9.9e10, 0.0, 5.7901e4)
N2O_EFF = {"N2": 1.7, "O2": 1.4, "H2O": 12.0}
# viscosity near 300 K [uPa s]; binary diffusivity with N2 at 298 K, 1 atm [cm^2/s]
MU300 = {"NH3": 10.1, "NO": 19.2, "N2O": 14.9, "AR": 22.7, "HE": 19.9, "N2": 17.9}
D_N2 = {"NH3": 0.230, "HE": 0.687, "AR": 0.20, "H2": 0.78}


@pytest.fixture(scope="module")
def mech():
    return Mechanism("nh3")


def arr_cgs(A, b, Ea, T):
    return A * T ** b * math.exp(-Ea / (R_CAL * T))


def state(mech, T, p, X):
    """(T, rho, Y) arrays of one cell at pressure p with mole fractions X."""
    x = np.zeros(mech.ns)
    for n, v in X.items():
        x[mech.names.index(n)] = v
    Y = mech.Y_from_X(x / x.sum())
    rho = mech.density(np.array(T), p, Y)
    return np.array([T]), np.array([rho]), Y[None, :]


def test_species_and_reactions(mech):
    assert mech.names == SPECIES and mech.ns == 31 and mech.nr == 203
    kinds = {k: mech.kind.count(k) for k in set(mech.kind)}
    assert kinds == {"elementary": 170, "three-body": 7, "falloff": 8, "pressure-dependent-Arrhenius": 18}
    assert sum(1 for k, t in zip(mech.kind, mech.troe) if k == "falloff" and t is None) == 1
    d = json.load(open(mechanism_path("nh3")))
    assert d["species"][mech.names.index("NO")]["composition"] == {"N": 1, "O": 1}


def test_thermo_against_tables(mech):
    h = mech.h_molar(np.array([298.15]))[0] / 1e6
    s = mech.s_R(np.array([298.15]))[0] * mech.R / 1e3
    k = mech.names.index
    for n, want in HF.items():
        assert abs(h[k(n)] - want) < 1.0, (n, h[k(n)])
    for n, want in S0.items():
        assert abs(s[k(n)] - want) < 0.5, (n, s[k(n)])


def plog_k(mech, T, p):
    """Forward rate constant of NH3 <=> NH2 + H [1/s] at (T, p)."""
    r = 23
    assert mech.kind[r] == "pressure-dependent-Arrhenius"
    Tt, rho, Y = state(mech, T, p, {"N2": 0.99, "NH3": 0.01})
    qf, _ = mech.rates_of_progress(Tt, rho, Y)
    c_nh3 = rho[0] * Y[0, mech.names.index("NH3")] / mech.W[mech.names.index("NH3")]
    return qf[0, r] / c_nh3


def test_plog_closed_form(mech):
    for T in (900.0, 1500.0, 2200.0):
        ks = [arr_cgs(A, b, Ea, T) for _, A, b, Ea in PLOG_NH3]
        for (P, _, _, _), k in zip(PLOG_NH3, ks):
            assert abs(plog_k(mech, T, P * ATM) / k - 1) < 1e-11, (T, P)
        for i in range(3):
            pm = math.sqrt(PLOG_NH3[i][0] * PLOG_NH3[i + 1][0]) * ATM
            assert abs(plog_k(mech, T, pm) / math.sqrt(ks[i] * ks[i + 1]) - 1) < 1e-11, (T, i)
        # a quarter of the way in ln p from 1 to 10 atm
        assert abs(plog_k(mech, T, 10 ** 0.25 * ATM) / (ks[1] ** 0.75 * ks[2] ** 0.25) - 1) < 1e-11
        assert abs(plog_k(mech, T, 0.01 * ATM) / ks[0] - 1) < 1e-12
        assert abs(plog_k(mech, T, 1000.0 * ATM) / ks[3] - 1) < 1e-12


def test_lindemann_closed_form(mech):
    r = 187
    assert mech.kind[r] == "falloff" and mech.troe[r] is None
    for T, p in ((1200.0, 0.5 * ATM), (1800.0, ATM), (2400.0, 20 * ATM)):
        X = {"N2": 0.7, "O2": 0.1, "H2O": 0.1, "AR": 0.05, "N2O": 0.05}
        Tt, rho, Y = state(mech, T, p, X)
        qf, _ = mech.rates_of_progress(Tt, rho, Y)
        C = {n: rho[0] * Y[0, mech.names.index(n)] / mech.W[mech.names.index(n)] * 1e-3 for n in X}  # mol/cm^3
        M = sum(N2O_EFF.get(n, 1.0) * c for n, c in C.items())
        k0, kinf = arr_cgs(*N2O_LOW, T), arr_cgs(*N2O_HIGH, T)
        Pr = k0 * M / kinf
        want = kinf * Pr / (1 + Pr) * C["N2O"] * 1e3  # kmol/m^3/s
        assert abs(qf[0, r] / want - 1) < 1e-11, (T, p)


def test_h2o2_module_matches_h2air_tables(mech):
    """Reactions 1-23 are the H2-air mechanism's, species matched by name."""
    h2 = Mechanism("h2air")
    rng = np.random.default_rng(3)
    n = 16
    Y9 = h2.normalize_Y(rng.uniform(0.01, 1, (n, h2.ns)))
    Y = np.zeros((n, mech.ns))
    for i, s in enumerate(h2.names):
        Y[:, mech.names.index(s)] = Y9[:, i]
    T = rng.uniform(900, 2500, n)
    rho = rng.uniform(0.05, 1.0, n)
    qf, qr = mech.rates_of_progress(T, rho, Y)
    qf9, qr9 = h2.rates_of_progress(T, rho, Y9)
    np.testing.assert_allclose(qf[:, :23], qf9, rtol=1e-12)
    np.testing.assert_allclose(qr[:, :23], qr9, rtol=1e-12)


def test_conservation(mech):
    rng = np.random.default_rng(7)
    n = 64
    Y = mech.normalize_Y(rng.uniform(0, 1, (n, mech.ns)))
    T = rng.uniform(800, 2600, n)
    rho = rng.uniform(0.05, 1.0, n)
    w = mech.net_production(T, rho, Y)
    scale = np.max(np.abs(w * mech.W), axis=1)
    assert np.all(scale > 0)
    assert np.all(np.abs(np.sum(w * mech.W, axis=1)) < 1e-12 * scale)
    d = json.load(open(mechanism_path("nh3")))
    for e in ("H", "O", "N", "Ar", "He"):
        E = np.array([s["composition"].get(e, 0) for s in d["species"]], dtype=np.float64)
        assert np.all(np.abs(w @ E) < 1e-12 * np.max(np.abs(w), axis=1) * E.max()), e


def test_transport_against_measurements(mech):
    mu = mech.species_viscosities(np.array(300.0)) * 1e6
    for n, want in MU300.items():
        assert abs(mu[mech.names.index(n)] / want - 1) < 0.04, (n, mu[mech.names.index(n)])
    D = mech.binary_diff(np.array(298.0), np.array(ATM)) * 1e4
    j = mech.names.index("N2")
    for n, want in D_N2.items():
        assert abs(D[mech.names.index(n), j] / want - 1) < 0.1, (n, D[mech.names.index(n), j])
    np.testing.assert_allclose(D, D.T, rtol=1e-12)


def test_generated_tables_match_yaml():
    """The committed JSON is what the generator makes from the reference's file."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "lattice-boltzmann-method_amd", "mech"))
    gm = pytest.importorskip("gen_mech")
    path = os.path.join(REF_DIR, "nh3.yaml")
    if not os.path.exists(path):
        pytest.skip("reference mechanism not mounted")
    doc = gm.load(path)
    names = [str(n) for n in doc["phases"][0]["species"]]
    assert names == SPECIES
    rx = gm.reactions_table(doc, names, len(doc["reactions"]))
    d = json.load(open(mechanism_path("nh3")))
    assert rx == d["reactions"]
    assert [s["name"] for s in d["species"]] == SPECIES
