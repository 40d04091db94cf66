"""GPU parity of the SMOOTHING variant (lbm_set_smoothing; LBM::Smoothing of
lbm-smoothing.cpp:95-173 in the race-free form the oracle pins in
test_oracle_smoothing.py), through the C ABI.

Init_smooting (Init, Smoothing, calculate_moment) is compared with the oracle
on the channel (TYPE_I inflow and TYPE_O outflow ghosts, TYPE_S walls: every
boundary rule of fd_central_2der) including the inflow ghosts' populations,
and five steps of the loop with calculate_moment_smoothing on a periodic
reacting box.  Bar: 1e-13 relative on the populations after the filter (the
device contracts f + alpha lap into one fused multiply-add), 1e-10 on the
fields after the steps (the FP64 parity bar; 1e-9 for the velocity).
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT, POP_F, LbmError, MixtureLattice
from oracle.multicomp import MultiStateBC, Units

pytestmark = pytest.mark.gpu

SEEDED = dict(cases.X_UNBURNT_PHI1, H=1e-4, O=1e-4, OH=1e-4, H2O=1e-3, HO2=1e-4, H2O2=1e-4)


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def pair(fl, reaction=True, alpha=0.1):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, mode=LBM_MODE_STRICT)
    lat.set_smoothing(True, alpha)
    lat.load_case(fl)
    s = MultiStateBC(nx + 2, ny + 2, nz + 2, Units(*cases.UNITS_FLAME))
    s.type[:] = fl["type"].ravel()
    for k in ("rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[:] = fl[k].ravel()
    s.X[:] = fl["X"].reshape(9, -1)
    return lat, s


def test_init_smoothing_matches_oracle_with_inflow_outflow_and_walls():
    fl = cases.reacting_channel(8, 5, 3, p_out_factor=0.995)
    lat, s = pair(fl, alpha=0.1)
    try:
        lat.Init()
        s.init()
        s.smoothing(0.1)
        s.moments()
        kinds = np.isin(s.type, (0, 5))  # TYPE_F and TYPE_I: the cells the filter moves
        assert kinds.sum() > (s.type == 0).sum()
        f = lat.get_populations(POP_F).reshape(9, 27, -1)
        assert relerr(f[:, :, kinds], s.f[:, :, kinds]) < 1e-13
        F = s.type == 0
        for k in ("rho", "u", "v", "w", "T", "p"):
            assert relerr(lat.get_field(k).ravel()[F], getattr(s, k)[F]) < 1e-10, k
    finally:
        lat.close()


def test_loop_with_calculate_moment_smoothing_matches_oracle():
    fl = cases.reacting_box(6, 5, 4, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    lat, s = pair(fl, alpha=0.1)
    try:
        lat.Init()
        s.init()
        s.smoothing(0.1)
        s.moments()
        F = s.type == 0
        for step in range(1, 6):
            lat.step(1)
            s.collide_species(True)
            s.collide()
            s.stream()
            s.calculate_moment_smoothing(0.1)
            for k in ("rho", "u", "T", "p"):
                e = relerr(lat.get_field(k).ravel()[F], getattr(s, k)[F])
                # u: a momentum of 1e-2 of the populations, through a Laplacian
                # that cancels 3 more digits (1e-9, as the species velocities
                # of the parity tests)
                assert e < (1e-9 if k == "u" else 1e-10), f"step {step}: {k} {e:.3e}"
            e = relerr(lat.get_species_field("rho").reshape(9, -1)[:, F], s.rho_a[:, F])
            assert e < 1e-10, f"step {step}: rho_a {e:.3e}"
    finally:
        lat.close()


def test_smoothing_changes_the_state_and_is_refused_outside_strict_multicomponent():
    fl = cases.reacting_box(6, 5, 4, T_K=1500.0, amp_T=0.1, amp_u=0.01, X=SEEDED)
    on, _ = pair(fl)
    off = MixtureLattice(6, 5, 4, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT)
    try:
        off.load_case(fl)
        on.Init()
        off.Init()
        assert not np.array_equal(on.get_populations(POP_F), off.get_populations(POP_F))
    finally:
        on.close()
        off.close()
    for mode in (LBM_MODE_FUSED, LBM_MODE_INPLACE):
        lat = MixtureLattice(6, 5, 4, *cases.UNITS_FLAME, reaction=False, mode=mode)
        try:
            with pytest.raises(LbmError, match="LBM_MODE_STRICT"):
                lat.set_smoothing(True)
        finally:
            lat.close()
