"""Multicomponent diagnostics (lbm_diagnostics) against the oracle, through
the C ABI.

A periodic reacting box steps on the GPU and in the oracle
(oracle/multicomp.py); the kinetic energy, enstrophy, mass and energy of
output.cpp:212-276 are evaluated by the oracle's C restatement
(orc_diag_single) and compared with the device reduction; the max Mach
number uses each cell's frozen mixture gamma
c_p / (c_p - R / W) from the oracle's NASA7 thermo (oracle/chem.py), not the
configuration's 1.4.  After one step, against the sums of the oracle's fields:
1e-10 for mass and energy and 1e-8 for the kinetic energy, enstrophy and Mach
number, which square the velocity and its differences (the velocities agree
to 1e-9 of their maximum, a momentum of 1e-2 of the populations).  After five
steps, against the sums of the device's own fields: 1e-12 (summation order).
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_FP64, LBM_MODE_STRICT, MixtureLattice
from oracle.binding import SingleState
from oracle.chem import Mechanism
from oracle.multicomp import MultiState, Units

pytestmark = pytest.mark.gpu

SEEDED = dict(cases.X_UNBURNT_PHI1, H=1e-4, O=1e-4, OH=1e-4, H2O=1e-3, HO2=1e-4, H2O2=1e-4)


def test_multicomponent_diagnostics_match_oracle():
    fl = cases.reacting_box(8, 6, 5, T_K=1500.0, amp_T=0.2, amp_u=0.03, X=SEEDED)
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT, precision=LBM_FP64)
    units = Units(*cases.UNITS_FLAME)
    s = MultiState(nx + 2, ny + 2, nz + 2, units)
    F = None
    try:
        lat.load_case(fl)
        s.type[:] = fl["type"].ravel()
        for k in ("rho", "u", "v", "w", "p", "T"):
            getattr(s, k)[:] = fl[k].ravel()
        s.X[:] = fl["X"].reshape(9, -1)
        F = s.type == 0
        lat.Init()
        s.init()
        lat.calculate_moment()
        s.moments()
        # one step: against output.cpp's sums on the oracle's fields
        lat.step(1)
        s.step(True)
        check(lat.diagnostics(), {k: getattr(s, k) for k in FIELDS}, s.type, nx, ny, nz, units,
              [1e-8, 1e-8, 1e-10, 1e-10])
        # four more: against the same sums on the device's own fields (the
        # reacting box develops composition gradients while the diffusion
        # velocities that pick fd_fuw's upwind side are near zero, so its
        # state and the oracle's part at the rounding level)
        lat.step(4)
        d = lat.diagnostics()
        dev = {k: lat.get_field(k).ravel() for k in FIELDS if k != "rho_a"}
        dev["rho_a"] = lat.get_species_field("rho").reshape(9, -1)
        check(d, dev, s.type, nx, ny, nz, units, [1e-12] * 4)
        # the configuration's 1.4 would give a different number here
        u2 = dev["u"][F] ** 2 + dev["v"][F] ** 2 + dev["w"][F] ** 2
        g14 = np.max(np.sqrt(u2) / np.sqrt(1.4 * dev["p"][F] / dev["rho"][F]))
        assert abs(g14 / d["max_mach"] - 1.0) > 1e-3
        assert d["nan_count"] == 0
    finally:
        lat.close()


FIELDS = ("rho", "u", "v", "w", "rhoe", "p", "T", "rho_a")


def check(d, fields, typ, nx, ny, nz, units, bars):
    """Device diagnostics d against output.cpp's sums (the oracle's C
    restatement) and the mixture-gamma Mach number of `fields`."""
    c = SingleState(nx + 2, ny + 2, nz + 2, 1.0, 1.4, 0.7, 0.01)
    c.type[...] = typ.reshape(c.shape)
    for k in ("rho", "u", "v", "w", "rhoe"):
        getattr(c, k)[...] = fields[k].reshape(c.shape)
    got = [d["kinetic"], d["enstrophy"], d["mass"], d["energy"]]
    for name, g, w, bar in zip(("kinetic", "enstrophy", "mass", "energy"), got, c.diag(), bars):
        assert abs(g / w - 1.0) < bar, (name, g, w)
    mach = max_mach(fields, typ == 0, units)
    assert abs(d["max_mach"] / mach - 1.0) < bars[0]


def max_mach(fields, F, units):
    """max |u| / sqrt(gamma p / rho) over TYPE_F cells with the frozen mixture
    gamma c_p / (c_p - R / W) of each cell (NASA7 thermo of oracle/chem.py) at
    the SI temperature of the binary32 unit crossing (units.hpp:75-131)."""
    mech = Mechanism()
    Y = (fields["rho_a"][:, F] / fields["rho"][F]).T
    Y = Y / Y.sum(axis=1, keepdims=True)
    Tsi = units.si_temp(fields["T"][F]).astype(np.float64)
    cp = mech.cp_mass(Tsi, Y)
    gam = cp / (cp - mech.R / mech.W_mean(Y))
    u2 = fields["u"][F] ** 2 + fields["v"][F] ** 2 + fields["w"][F] ** 2
    return np.max(np.sqrt(u2) / np.sqrt(gam * fields["p"][F] / fields["rho"][F]))
