"""The multicomponent fused mode (LBM_MODE_FUSED) against the strict mode,
through the C ABI.

In the fused mode Collide_Species writes every post-collision species
population straight into the slot the next Streaming would move it to, and
Streaming only swaps the species buffers and streams the energy populations.
It computes the same numbers in the same order as the strict calls (which the
oracle pins at 1e-10 in test_gpu_multi.py), so the bar here is bitwise
equality of every fluid cell's populations (f, g and the post-collision sets,
rebuilt on download), fields and species fields after each step.  The cases
cover the periodic and NSCBC faces of the mixing layer, the LODI inflow and
NSCBC outflow of the premixed flame, the fixed-state inflow, outflow and
isothermal walls of the channel, adiabatic walls, species held at the
SPECIES_MIN floor (the post-collision value a low species keeps,
lbm-collision.cpp:512-513), FP32 storage, and a restart in the middle.
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import (LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT, POP_F,
                          POP_FPC, POP_G, POP_GPC, LbmError, MixtureLattice)
from lbm_amd.cases import TYPE_A, TYPE_FS
from mode_compare import assert_same, compare as _compare, make, snapshot

pytestmark = pytest.mark.gpu


def compare(fl, steps, precision=LBM_FP64, reaction=True, tag=""):
    _compare(fl, steps, LBM_MODE_FUSED, precision, reaction, tag)


@pytest.mark.parametrize("reaction", [False, True])
def test_fused_mixing_layer_matches_strict(reaction):
    compare(cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0), 5, reaction=reaction,
            tag="mixing layer")


def test_fused_flame_matches_strict():
    compare(cases.premixed_flame_1d(24, 3, 3), 5, tag="flame")


def test_fused_channel_walls_matches_strict():
    compare(cases.reacting_channel(8, 5, 3, p_out_factor=0.995), 5, tag="channel")


def test_fused_adiabatic_walls_and_floor_species_match_strict():
    """TYPE_A walls on the y faces and species at the SPECIES_MIN floor: the
    collision leaves their populations unwritten (no chemistry source), so
    they keep the value streamed into them before."""
    fl = cases.reacting_box(6, 5, 4, T_K=1300.0, amp_T=0.1, amp_u=0.01,
                            X=dict(cases.X_UNBURNT_PHI1, H2O=1e-3))
    fl["type"][:, 0, :] = TYPE_A
    fl["type"][:, -1, :] = TYPE_A
    compare(fl, 5, reaction=False, tag="adiabatic walls")


def test_fused_fp32_mixing_layer_matches_strict():
    compare(cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0), 4,
            precision=LBM_FP32_SHIFTED, tag="mixing layer fp32")


def test_fused_restart_in_the_middle(tmp_path):
    """Three fused steps, a restart file (its post-collision records rebuilt
    from the streamed populations), a fresh fused context resumed from it and
    three more steps: the same state as six strict steps."""
    fl = cases.premixed_flame_1d(24, 3, 3)
    F = fl["type"].ravel() == 0
    strict = make(fl, LBM_MODE_STRICT)
    strict.step(6)
    fused = make(fl, LBM_MODE_FUSED)
    fused.step(3)
    path = fused.write_restart(3, str(tmp_path))
    fused.close()
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    resumed = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_FUSED)
    assert resumed.read_restart(path) == 3
    resumed.step(3)
    assert_same(snapshot(resumed, F), snapshot(strict, F), "resumed")


def test_fused_refuses_free_slip():
    fl = cases.reacting_box(4, 4, 4, T_K=1300.0)
    fl["type"][:, 0, :] = TYPE_FS
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=False, mode=LBM_MODE_FUSED)
    with pytest.raises(LbmError, match="TYPE_FS"):
        lat.load_case(fl)


def test_fused_reads_between_collide_species_and_streaming():
    """Between a fused Collide_Species (species already pushed to their
    streamed slots) and Streaming, LBM_POP_FPC is gathered from those slots
    and the side store, and equals the strict mode's post-collision set; F,
    G and GPC read as in strict mode; a population upload there is refused
    (LBM_E_STATE); stepping on after the reads stays bitwise equal to
    strict mode."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    F = fl["type"].ravel() == 0
    ns = 9
    strict = make(fl, LBM_MODE_STRICT)
    fused = make(fl, LBM_MODE_FUSED)
    try:
        strict.step(2)
        fused.step(2)  # the first fused step runs strictly; the second pushes
        for lat in (strict, fused):
            lat.Collide_Species()
            lat.Collide()
        for which in (POP_FPC, POP_F):
            a = fused.get_populations(which).reshape(ns, 27, -1)[:, :, F]
            b = strict.get_populations(which).reshape(ns, 27, -1)[:, :, F]
            assert np.array_equal(a, b), f"population set {which} differs before Streaming"
        for which in (POP_G, POP_GPC):
            assert np.array_equal(fused.get_populations(which).reshape(27, -1)[:, F],
                                  strict.get_populations(which).reshape(27, -1)[:, F])
        with pytest.raises(LbmError, match="lbm_collide_species and lbm_stream"):
            fused.set_populations(POP_F, fused.get_populations(POP_F))
        for lat in (strict, fused):
            lat.Streaming()
            lat.TMS_BC()
            lat.calculate_moment()
        assert_same(snapshot(fused, F), snapshot(strict, F), "after the mid-step reads")
        strict.step(2)
        fused.step(2)
        assert_same(snapshot(fused, F), snapshot(strict, F), "two steps later")
    finally:
        strict.close()
        fused.close()


@pytest.mark.parametrize("precision,mode", [(LBM_FP64, LBM_MODE_FUSED), (LBM_FP32_SHIFTED, LBM_MODE_FUSED),
                                            (LBM_FP32_SHIFTED, LBM_MODE_INPLACE)])
def test_overlapped_step_with_surrogate_matches_serial_calls(precision, mode, monkeypatch):
    """lbm_step in the fused and in-place modes runs the chemistry (network
    pass, fallback list) on a second stream beside the transport pass and the
    mixture collide.  With the same network, envelope and audits, four steps
    of it are bitwise those of the five calls issued one by one (a context
    created with LBM_AMD_OVERLAP=0 steps serially), surrogate counters
    included."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    F = fl["type"].ravel() == 0
    lats = []
    try:
        for env in ("1", "0"):
            monkeypatch.setenv("LBM_AMD_OVERLAP", env)
            lat = make(fl, mode, precision)
            lats.append(lat)
        ov, ser = lats
        print(f"overlapped vs serial surrogate steps precision {precision}", flush=True)
        ov.surrogate_capture(1 << 16, 1)
        ov.step(1)  # the strict first step, capturing the plain chemistry's states
        ov.surrogate_capture_stop()
        ov.surrogate_train(40, batch=256, lr=1e-3, seed=3)
        params, norm = ov.surrogate_params().copy(), ov.surrogate_norm().copy()
        ser.step(1)
        ser.set_surrogate_params(params)
        ser.set_surrogate_norm(norm)
        ov.set_surrogate_params(params)
        for lat in lats:
            lat.surrogate_enable(True, audit_stride=3, audit_tol=0.05)
        for step in range(1, 5):
            ov.step(1)
            ser.Collide_Species()
            ser.Collide()
            ser.Streaming()
            ser.TMS_BC()
            ser.calculate_moment()
            assert_same(snapshot(ov, F), snapshot(ser, F), f"step {step}")
            a, b = ov.surrogate_stats(), ser.surrogate_stats()
            a.pop("captured"), b.pop("captured")  # only `ov` captured training rows
            assert a == b
        st = ov.surrogate_stats()
        assert st["predicted"] > 0 and st["audited"] > 0
    finally:
        for lat in lats:
            lat.close()
