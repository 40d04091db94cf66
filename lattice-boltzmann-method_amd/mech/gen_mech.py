"""Generates the mechanism tables of the multicomponent path.

    python lattice-boltzmann-method_amd/mech/gen_mech.py [h2air|ch4_2s|c3h8|nh3 ...] [--yaml-dir DIR]

Mechanisms (the reference selects one at compile time, lbm-main.cpp:83-88;
here each gets its own table header and its own build of the library):

  h2air   the H2/O2 kinetics module (reactions 1-23) of the reference's
          src/reaction-mech/nh3.yaml with its NASA7 thermo and Lennard-Jones
          transport data for H2, H, O, O2, OH, H2O, HO2, H2O2 and N2.  AR and HE
          appear in that module only as third-body efficiencies; they are not
          species here and their efficiencies are dropped.
  ch4_2s  src/reaction-mech/CH4_2S.yaml (2-step CH4 BFER: O2, H2O, CH4, CO,
          CO2, N2; an irreversible step with fractional reaction orders and a
          reversible CO oxidation with a half-integer O2 coefficient).
  c3h8    src/reaction-mech/propane_mech.yaml (2-step propane: CO, C3H8, CO2,
          H2O, N2, O2).
  nh3     the whole of src/reaction-mech/nh3.yaml (Stagni et al. 2023
          ammonia/hydrogen: 31 species in the phase's order, 203 reactions
          with three-body, Troe and Lindemann falloff and
          pressure-dependent-Arrhenius (PLOG) rates).  The species NO is read
          as a name, not as YAML's boolean false.

Outputs per mechanism (committed):
  <name>.json             data for the CPU oracle and the tests
  ../csrc/mech_<name>.h   constexpr tables for the HIP kernels

Units: SI with kmol (as Cantera): A in (m^3/kmol)^(order-1)/s with order the
sum of the forward concentration exponents (explicit `orders` where the
mechanism gives them, else the reactant coefficients), Ea/R in K, W in
kg/kmol.  Transport properties are this generator's own kinetic-theory
evaluation (Chapman-Enskog with Neufeld collision integrals, Brokaw's
polar correction, the polar/non-polar combining rule and the
Warnatz/Kee thermal-conductivity model), fitted as ln(prop) = poly4(ln T)
over 250-3500 K, the CHEMKIN way (the mixture-averaged model for every
mechanism).
"""
import json
import math
import os
import re
import sys

import numpy as np
import yaml

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_YAML_DIR = "/root/reference/src/reaction-mech"
MECHS = {
    "h2air": {"yaml": "nh3.yaml", "species": ["H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "N2"],
              "reactions": 23,  # the \H2_O2 kinetics module
              "source": "nh3.yaml H2/O2 module, reactions 1-23"},
    "ch4_2s": {"yaml": "CH4_2S.yaml", "species": ["O2", "H2O", "CH4", "CO", "CO2", "N2"], "reactions": None,
               "source": "CH4_2S.yaml (CH4 BFER two-step)"},
    "c3h8": {"yaml": "propane_mech.yaml", "species": ["CO", "C3H8", "CO2", "H2O", "N2", "O2"], "reactions": None,
             "source": "propane_mech.yaml (propane two-step)"},
    "nh3": {"yaml": "nh3.yaml", "species": None, "reactions": None,
            "source": "nh3.yaml (ammonia/hydrogen, Stagni et al. 2023)"},
}

R_UNIV = 8314.46261815324  # J/(kmol K), Cantera's GasConstant
ONE_ATM = 101325.0
KB = 1.380649e-23
AVOGADRO = 6.02214076e26  # 1/kmol
CAL = 4.184
DEBYE = 1e-21 / 299792458.0  # C m
EPS0_4PI = 1.0 / 8.9875517923e9  # 4 pi eps0
ATOMIC_W = {"H": 1.008, "O": 15.999, "N": 14.007, "C": 12.011, "Ar": 39.95, "He": 4.002602}
PRESSURE_UNITS = {"atm": ONE_ATM, "Pa": 1.0, "bar": 1e5, "torr": ONE_ATM / 760.0}
T_FIT = np.linspace(250.0, 3500.0, 200)


class _Loader(yaml.SafeLoader):
    """SafeLoader without YAML 1.1 booleans: a species named NO (nitric
    oxide) stays the string "NO" in the phase list, the species records and
    the efficiency maps."""


_Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()}


def load(path):
    with open(path) as f:
        return yaml.load(f, Loader=_Loader)


def pressure_si(q):
    """'0.1 atm' (or a bare number in Pa) -> Pa."""
    if isinstance(q, (int, float)):
        return float(q)
    v, u = q.split()
    return float(v) * PRESSURE_UNITS[u]


def species_table(doc, names):
    by_name = {str(s["name"]): s for s in doc["species"]}
    out = []
    for n in names:
        s = by_name[n]
        comp = s["composition"]
        W = sum(ATOMIC_W[e] * c for e, c in comp.items())
        th = s["thermo"]
        assert th["model"] == "NASA7"
        Tr = th["temperature-ranges"]
        tr = s["transport"]
        out.append({
            "name": n, "W": W, "composition": comp,
            "nasa_lo": th["data"][0], "nasa_hi": th["data"][1], "T_lo": Tr[0], "T_mid": Tr[1], "T_hi": Tr[2],
            "geometry": tr["geometry"], "eps_k": tr["well-depth"], "sigma_A": tr["diameter"],
            "dipole_D": tr.get("dipole", 0.0), "polar_A3": tr.get("polarizability", 0.0),
            "zrot": tr.get("rotational-relaxation", 0.0),
        })
    return out


def parse_side(s):
    s = s.replace("(+M)", "").strip()
    terms = [t.strip() for t in s.split("+")]
    st = {}
    third = False
    for t in terms:
        if t == "M":
            third = True
            continue
        m = re.match(r"^(\d+(?:\.\d*)?)\s*(\S+)$", t)
        if m and not m.group(2)[0].isdigit():
            nu, sp = float(m.group(1)), m.group(2)
            nu = int(nu) if nu.is_integer() else nu
        else:
            nu, sp = 1, t
        st[sp] = st.get(sp, 0) + nu
    return st, third


def arrhenius_si(rc, order):
    # A in (cm^3/mol)^(order-1)/s -> (m^3/kmol)^(order-1)/s ; Ea cal/mol -> K
    return {"A": rc["A"] * (1e-3) ** (order - 1), "b": rc["b"], "Ea_R": rc["Ea"] * CAL * 1000.0 / R_UNIV}


def reactions_table(doc, names, count):
    out = []
    for r in doc["reactions"][:count]:
        eq = r["equation"].strip()
        if "<=>" in eq:
            lhs, rhs = eq.split("<=>")
            reversible = True
        elif "=>" in eq:
            lhs, rhs = eq.split("=>")
            reversible = False
        else:
            lhs, rhs = eq.split("=")
            reversible = True
        reac, tb_l = parse_side(lhs)
        prod, tb_r = parse_side(rhs)
        for sp in list(reac) + list(prod):
            assert sp in names, (eq, sp)
        kind = r.get("type", "elementary")
        orders = dict(reac)
        for k, v in r.get("orders", {}).items():
            assert k in names, (eq, k)
            orders[k] = v
        order = sum(orders.values())
        e = {"equation": eq, "type": kind, "reactants": reac, "products": prod,
             "duplicate": bool(r.get("duplicate", False))}
        if not reversible or "orders" in r:
            # fields absent for the reversible mass-action reactions of the H2/O2 module
            e["reversible"] = reversible
            e["orders"] = orders
        if kind == "elementary":
            e["k"] = arrhenius_si(r["rate-constant"], order)
        elif kind == "three-body":
            e["k"] = arrhenius_si(r["rate-constant"], order + 1)
        elif kind == "falloff":
            e["k_inf"] = arrhenius_si(r["high-P-rate-constant"], order)
            e["k0"] = arrhenius_si(r["low-P-rate-constant"], order + 1)
            if "Troe" in r:
                t = r["Troe"]
                e["troe"] = {"A": t["A"], "T3": t["T3"], "T1": t["T1"], "T2": t.get("T2", None)}
            # no Troe block: Lindemann form, broadening factor 1
        elif kind == "pressure-dependent-Arrhenius":
            levels = [dict(arrhenius_si(rc, order), P=pressure_si(rc["P"])) for rc in r["rate-constants"]]
            if any(lv["A"] <= 0 for lv in levels):
                raise ValueError(f"{eq}: PLOG levels with A <= 0 are not supported")
            ps = [lv["P"] for lv in levels]
            if ps != sorted(ps) or len(set(ps)) != len(ps):
                raise ValueError(f"{eq}: PLOG levels must have distinct increasing pressures")
            e["plog"] = levels
            e["k"] = {k: levels[0][k] for k in ("A", "b", "Ea_R")}
        else:
            raise ValueError(kind)
        if kind in ("three-body", "falloff"):
            eff = {k: v for k, v in r.get("efficiencies", {}).items() if k in names}
            e["efficiencies"] = eff
        out.append(e)
    return out


# ----------------------------------------------------------------- thermo
def nasa_cp_R(s, T):
    a = s["nasa_lo"] if T < s["T_mid"] else s["nasa_hi"]
    return a[0] + T * (a[1] + T * (a[2] + T * (a[3] + T * a[4])))


# ------------------------------------------------------------ transport
def _over_exp(a, x):
    """a / exp(x); 0 where exp(x) overflows (reduced temperatures of the
    light, weakly bound species, He and H, at flame temperatures)."""
    return a / math.exp(x) if x < 700.0 else 0.0


def omega11_lj(Ts):
    # Neufeld, Janzen & Aziz (1972), J. Chem. Phys. 57, 1100
    return 1.06036 / Ts ** 0.15610 + _over_exp(0.19300, 0.47635 * Ts) + _over_exp(1.03587, 1.52996 * Ts) + \
        _over_exp(1.76474, 3.89411 * Ts)


def omega22_lj(Ts):
    return 1.16145 / Ts ** 0.14874 + _over_exp(0.52487, 0.77320 * Ts) + _over_exp(2.16178, 2.43787 * Ts)


def reduced_dipole(s):
    """delta* = mu^2 / (2 (4 pi eps0) eps sigma^3)."""
    mu = s["dipole_D"] * DEBYE
    return mu * mu / (2.0 * EPS0_4PI * s["eps_k"] * KB * (s["sigma_A"] * 1e-10) ** 3)


def omega11(Ts, delta):
    return omega11_lj(Ts) + 0.19 * delta * delta / Ts  # Brokaw (1969) polar correction


def omega22(Ts, delta):
    return omega22_lj(Ts) + 0.2 * delta * delta / Ts


def pair_params(a, b):
    """epsilon, sigma, delta* of a pair; polar/non-polar correction xi (Kee et al.)."""
    da, db = reduced_dipole(a), reduced_dipole(b)
    eps = math.sqrt(a["eps_k"] * b["eps_k"])
    sig = 0.5 * (a["sigma_A"] + b["sigma_A"])
    delta = 0.5 * math.sqrt(da * db) if (da > 0 and db > 0) else 0.0
    polar_a, polar_b = a["dipole_D"] > 0, b["dipole_D"] > 0
    if polar_a != polar_b:
        p, n = (a, b) if polar_a else (b, a)
        alpha_n = n["polar_A3"] / n["sigma_A"] ** 3
        mu_p = reduced_dipole(p)  # delta*_p = mu*_p^2 / 2 ... mu*^2 = 2 delta*
        xi = 1.0 + 0.25 * alpha_n * (2.0 * mu_p) * math.sqrt(p["eps_k"] / n["eps_k"])
        eps *= xi * xi
        sig *= xi ** (-1.0 / 6.0)
    return eps, sig, delta


def viscosity(s, T):
    m = s["W"] / AVOGADRO
    Ts = T / s["eps_k"]
    sig = s["sigma_A"] * 1e-10
    return 5.0 / 16.0 * math.sqrt(math.pi * m * KB * T) / (math.pi * sig * sig * omega22(Ts, reduced_dipole(s)))


def binary_diff_times_p(a, b, T):
    eps, sig, delta = pair_params(a, b)
    mred = a["W"] * b["W"] / (a["W"] + b["W"]) / AVOGADRO
    Ts = T / eps
    s = sig * 1e-10
    return 3.0 / 16.0 * math.sqrt(2.0 * math.pi * (KB * T) ** 3 / mred) / (math.pi * s * s * omega11(Ts, delta))


def _frot(eps_over_T):
    x = eps_over_T
    return 1.0 + math.pi ** 1.5 / 2.0 * math.sqrt(x) + (math.pi ** 2 / 4.0 + 2.0) * x + math.pi ** 1.5 * x ** 1.5


def conductivity(s, T):
    """Warnatz / Kee et al. species conductivity: translational, rotational and
    internal contributions with Parker's Z_rot(T)."""
    mu = viscosity(s, T)
    W = s["W"]
    cv_R = nasa_cp_R(s, T) - 1.0
    if s["geometry"] == "atom":
        return 15.0 / 4.0 * R_UNIV / W * mu
    cv_rot = 1.0 if s["geometry"] == "linear" else 1.5
    cv_int = cv_R - 1.5 - cv_rot
    rhoD = W * ONE_ATM / (R_UNIV * T) * binary_diff_times_p(s, s, T) / ONE_ATM
    f_int = rhoD / mu
    zrot = s["zrot"] * _frot(s["eps_k"] / 298.0) / _frot(s["eps_k"] / T) if s["zrot"] > 0 else 1.0
    A = 2.5 - f_int
    B = zrot + 2.0 / math.pi * (5.0 / 3.0 * cv_rot + f_int)
    c1 = 2.0 / math.pi * A / B
    f_rot = f_int * (1.0 + c1)
    f_trans = 2.5 * (1.0 - c1 * cv_rot / 1.5)
    return mu / W * R_UNIV * (f_trans * 1.5 + f_rot * cv_rot + f_int * cv_int)


def fit_ln(fn):
    lnT = np.log(T_FIT)
    y = np.log([fn(T) for T in T_FIT])
    c = np.polyfit(lnT, y, 4)[::-1]  # ascending powers
    resid = float(np.max(np.abs(np.polyval(c[::-1], lnT) - y)))
    return [float(x) for x in c], resid


def transport_fits(sp):
    visc, cond, diff = [], [], []
    worst = 0.0
    for s in sp:
        c, r = fit_ln(lambda T, s=s: viscosity(s, T))
        visc.append(c)
        worst = max(worst, r)
        c, r = fit_ln(lambda T, s=s: conductivity(s, T))
        cond.append(c)
        worst = max(worst, r)
    n = len(sp)
    for a in range(n):
        row = []
        for b in range(n):
            c, r = fit_ln(lambda T, a=a, b=b: binary_diff_times_p(sp[a], sp[b], T))
            row.append(c)
            worst = max(worst, r)
        diff.append(row)
    return {"visc": visc, "cond": cond, "diff_p": diff, "max_ln_residual": worst,
            "T_fit": [float(T_FIT[0]), float(T_FIT[-1])]}


# ------------------------------------------------------------- emitters
def emit_header(data, path):
    sp, rx = data["species"], data["reactions"]
    names = [s["name"] for s in sp]
    ns, nr = len(sp), len(rx)
    L = []
    w = L.append
    w("// Generated by lattice-boltzmann-method_amd/mech/gen_mech.py from")
    w("// " + data["source"] + "; do not edit.  SI units with kmol.")
    w("#pragma once")
    w("namespace lbm { namespace mech {")
    w(f"constexpr int NS = {ns};")
    w(f"constexpr int NR = {nr};")
    w("constexpr double R_UNIV = %.17g;" % R_UNIV)
    w("constexpr double P_REF = %.17g;" % ONE_ATM)
    w(f'constexpr const char* MECH_NAME = "{data["name"]}";')
    w("constexpr const char* SPECIES_NAMES[NS] = {" + ", ".join(f'"{n}"' for n in names) + "};")
    for i, s in enumerate(sp):
        w(f"constexpr int {s['name']} = {i};")
    w("constexpr double W[NS] = {" + ", ".join("%.17g" % s["W"] for s in sp) + "};")
    w("constexpr double T_MID[NS] = {" + ", ".join("%.17g" % s["T_mid"] for s in sp) + "};")
    w("constexpr double NASA_LO[NS][7] = {" + ", ".join(
        "{" + ", ".join("%.17g" % x for x in s["nasa_lo"]) + "}" for s in sp) + "};")
    w("constexpr double NASA_HI[NS][7] = {" + ", ".join(
        "{" + ", ".join("%.17g" % x for x in s["nasa_hi"]) + "}" for s in sp) + "};")
    tr = data["transport"]
    w("// ln(mu [Pa s]), ln(lambda [W/m/K]), ln(p D_ab [Pa m^2/s]) = sum_i c_i (ln T)^i")
    w("constexpr double VISC_FIT[NS][5] = {" + ", ".join(
        "{" + ", ".join("%.17g" % x for x in c) + "}" for c in tr["visc"]) + "};")
    w("constexpr double COND_FIT[NS][5] = {" + ", ".join(
        "{" + ", ".join("%.17g" % x for x in c) + "}" for c in tr["cond"]) + "};")
    w("constexpr double DIFF_FIT[NS][NS][5] = {" + ", ".join(
        "{" + ", ".join("{" + ", ".join("%.17g" % x for x in c) + "}" for c in row) + "}"
        for row in tr["diff_p"]) + "};")
    # reactions: kind 0 elementary, 1 three-body, 2 falloff (Troe or Lindemann),
    # 3 pressure-dependent Arrhenius (PLOG)
    kinds = {"elementary": 0, "three-body": 1, "falloff": 2, "pressure-dependent-Arrhenius": 3}
    nu_f = [[0] * ns for _ in range(nr)]
    nu_r = [[0] * ns for _ in range(nr)]
    ord_f = [[0] * ns for _ in range(nr)]
    for r, e in enumerate(rx):
        for k, v in e["reactants"].items():
            nu_f[r][names.index(k)] = v
        for k, v in e["products"].items():
            nu_r[r][names.index(k)] = v
        for k, v in e.get("orders", e["reactants"]).items():
            ord_f[r][names.index(k)] = v

    def table(rows):
        return ", ".join("{" + ", ".join("%.17g" % x for x in row) + "}" for row in rows)

    w("constexpr int KIND[NR] = {" + ", ".join(str(kinds[e["type"]]) for e in rx) + "};")
    w("// stoichiometric coefficients; forward concentration exponents (the mechanism's")
    w("// explicit orders, else NU_F); reverse rates from K_c for REVERSIBLE reactions")
    w("constexpr double NU_F[NR][NS] = {" + table(nu_f) + "};")
    w("constexpr double NU_R[NR][NS] = {" + table(nu_r) + "};")
    w("constexpr double ORD_F[NR][NS] = {" + table(ord_f) + "};")
    w("constexpr int REVERSIBLE[NR] = {" + ", ".join(str(int(e.get("reversible", True))) for e in rx) + "};")

    def arr(e, key):
        k = e.get(key) or e.get("k")
        return k

    w("// forward rate k = A T^b exp(-Ea_R / T); falloff reactions: high-pressure limit here")
    w("constexpr double ARR_A[NR] = {" + ", ".join("%.17g" % (e.get("k_inf") or e["k"])["A"] for e in rx) + "};")
    w("constexpr double ARR_B[NR] = {" + ", ".join("%.17g" % (e.get("k_inf") or e["k"])["b"] for e in rx) + "};")
    w("constexpr double ARR_EA_R[NR] = {" + ", ".join(
        "%.17g" % (e.get("k_inf") or e["k"])["Ea_R"] for e in rx) + "};")
    w("// low-pressure limit and Troe parameters (falloff only; 0 otherwise)")
    w("constexpr double LOW_A[NR] = {" + ", ".join("%.17g" % (e["k0"]["A"] if "k0" in e else 0.0) for e in rx) + "};")
    w("constexpr double LOW_B[NR] = {" + ", ".join("%.17g" % (e["k0"]["b"] if "k0" in e else 0.0) for e in rx) + "};")
    w("constexpr double LOW_EA_R[NR] = {" + ", ".join(
        "%.17g" % (e["k0"]["Ea_R"] if "k0" in e else 0.0) for e in rx) + "};")
    w("constexpr double TROE[NR][4] = {" + ", ".join(
        "{" + ", ".join("%.17g" % (x if x is not None else 0.0) for x in (
            (e["troe"]["A"], e["troe"]["T3"], e["troe"]["T1"], e["troe"]["T2"]) if "troe" in e else (0, 0, 0, 0)))
        + "}" for e in rx) + "};")
    w("constexpr int TROE_HAS_T2[NR] = {" + ", ".join(
        str(1 if ("troe" in e and e["troe"]["T2"] is not None) else 0) for e in rx) + "};")
    w("// falloff broadening: 1 Troe, 0 Lindemann (F = 1)")
    w("constexpr int TROE_ON[NR] = {" + ", ".join(str(1 if "troe" in e else 0) for e in rx) + "};")
    # PLOG: per reaction up to PLOG_MAX levels of increasing pressure; below the
    # first (above the last) level the first (last) rate applies
    pmax = max([len(e.get("plog", [])) for e in rx] + [1])
    w("// PLOG levels: count, ln(P [Pa]), ln A, b and Ea/R of each level")
    w(f"constexpr int PLOG_MAX = {pmax};")
    w("constexpr int PLOG_N[NR] = {" + ", ".join(str(len(e.get("plog", []))) for e in rx) + "};")

    def plog(key, f=lambda x: x):
        rows = []
        for e in rx:
            lv = e.get("plog", [])
            rows.append([f(x[key]) for x in lv] + [0.0] * (pmax - len(lv)))
        return table(rows)

    w("constexpr double PLOG_LNP[NR][PLOG_MAX] = {" + plog("P", math.log) + "};")
    w("constexpr double PLOG_LNA[NR][PLOG_MAX] = {" + plog("A", math.log) + "};")
    w("constexpr double PLOG_B[NR][PLOG_MAX] = {" + plog("b") + "};")
    w("constexpr double PLOG_EA_R[NR][PLOG_MAX] = {" + plog("Ea_R") + "};")
    w("// third-body efficiencies (1 for species not listed; unused for elementary)")
    w("constexpr double EFF[NR][NS] = {" + ", ".join(
        "{" + ", ".join("%.17g" % e.get("efficiencies", {}).get(s, 1.0) for s in names) + "}" for e in rx) + "};")
    w("}}  // namespace lbm::mech")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def generate(name, yaml_dir):
    m = MECHS[name]
    doc = load(os.path.join(yaml_dir, m["yaml"]))
    names = m["species"] or [str(n) for n in doc["phases"][0]["species"]]
    sp = species_table(doc, names)
    count = m["reactions"] if m["reactions"] is not None else len(doc["reactions"])
    rx = reactions_table(doc, names, count)
    data = {"source": m["source"], "species": sp, "reactions": rx,
            "R_univ": R_UNIV, "P_ref": ONE_ATM, "transport": transport_fits(sp)}
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(data, f, indent=1)
    emit_header(dict(data, name=name), os.path.join(os.path.dirname(HERE), "csrc", f"mech_{name}.h"))
    print(name, "species", len(sp), "reactions", len(rx), "max fit residual (ln)",
          data["transport"]["max_ln_residual"])


def main():
    args = sys.argv[1:]
    yaml_dir = DEFAULT_YAML_DIR
    if "--yaml-dir" in args:
        i = args.index("--yaml-dir")
        yaml_dir = args[i + 1]
        del args[i:i + 2]
    for name in args or list(MECHS):
        generate(name, yaml_dir)


if __name__ == "__main__":
    main()
