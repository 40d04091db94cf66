"""The multicomponent in-place mode (LBM_MODE_INPLACE) against the strict
mode, through the C ABI.

The in-place mode keeps one species population buffer and streams it by the
AA pattern: a step reads one layout and writes the other, each plain cell
writing exactly the slots it read.  Cells next to walls, in/outflow faces and
TMS_BC cells collide into a side store and are streamed by the strict
stream's rules; a species at the SPECIES_MIN floor keeps last step's
post-collision values (lbm-collision.cpp:512-513), which a save pass copies
out before the collision overwrites them.  The arithmetic is the strict
calls' (pinned against the oracle at 1e-10 in test_gpu_multi.py), so the bar
is bitwise equality of every fluid cell's populations (f, g and the
post-collision sets gathered on download), fields and species fields after
each step, in fresh contexts.
"""
import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import (LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT, POP_F,
                          POP_FPC, POP_G, POP_GPC, Lattice, LbmError, MixtureLattice)
from lbm_amd.cases import TYPE_A, TYPE_FS
from mode_compare import assert_same, compare as _compare, make, snapshot

pytestmark = pytest.mark.gpu


def compare(fl, steps, precision=LBM_FP64, reaction=True, tag=""):
    _compare(fl, steps, LBM_MODE_INPLACE, precision, reaction, tag)


@pytest.mark.parametrize("reaction", [False, True])
def test_inplace_mixing_layer_matches_strict(reaction):
    compare(cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0), 6, reaction=reaction,
            tag="mixing layer")


def test_inplace_flame_matches_strict():
    compare(cases.premixed_flame_1d(24, 3, 3), 6, tag="flame")


def test_inplace_channel_walls_matches_strict():
    compare(cases.reacting_channel(8, 5, 3, p_out_factor=0.995), 6, tag="channel")


def test_inplace_adiabatic_walls_and_floor_species_match_strict():
    """TYPE_A walls on the y faces and species at the SPECIES_MIN floor: the
    collision leaves their populations unwritten (no chemistry source), so
    they keep the value streamed into them before."""
    fl = cases.reacting_box(6, 5, 4, T_K=1300.0, amp_T=0.1, amp_u=0.01,
                            X=dict(cases.X_UNBURNT_PHI1, H2O=1e-3))
    fl["type"][:, 0, :] = TYPE_A
    fl["type"][:, -1, :] = TYPE_A
    compare(fl, 6, reaction=False, tag="adiabatic walls")


def test_inplace_low_species_list_grows():
    """A periodic box where six species sit at the SPECIES_MIN floor in every
    cell: the list of low (cell, species) pairs starts at 4096 entries and
    must grow within the first step (~12 k pairs)."""
    fl = cases.reacting_box(16, 16, 8, T_K=1300.0, amp_T=0.05, amp_u=0.01)
    compare(fl, 4, reaction=False, tag="floor species everywhere")


@pytest.mark.parametrize("precision", [LBM_FP64, LBM_FP32_SHIFTED])
def test_inplace_species_rounds_when_the_floor_list_does_not_fit(precision, monkeypatch):
    """When the list of every species' floor pairs would not fit in device
    memory, the in-place collide saves and collides one species at a time
    (LBM_AMD_LOW_MAX caps the list below the box's 6 x 120 floor pairs, above
    one species' 120): still bitwise the strict calls, with and without
    reaction."""
    monkeypatch.setenv("LBM_AMD_LOW_MAX", "300")
    fl = cases.reacting_box(6, 5, 4, T_K=1300.0, amp_T=0.1, amp_u=0.01, X={"H2": 0.3, "O2": 0.15, "N2": 0.55})
    compare(fl, 4, precision=precision, reaction=False, tag="floor list in species rounds")
    compare(fl, 4, precision=precision, reaction=True, tag="floor list in species rounds, reacting")


def test_inplace_fp32_mixing_layer_matches_strict():
    compare(cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0), 5,
            precision=LBM_FP32_SHIFTED, tag="mixing layer fp32")


def test_inplace_restart_in_the_middle(tmp_path):
    """Three in-place steps, a restart file (the post-collision records
    gathered from the buffer's layout), a fresh in-place context resumed from
    it (the uploaded post-collision set serves the first step's low species)
    and three more steps: the same state as six strict steps."""
    fl = cases.premixed_flame_1d(24, 3, 3)
    F = fl["type"].ravel() == 0
    strict = make(fl, LBM_MODE_STRICT)
    strict.step(6)
    inpl = make(fl, LBM_MODE_INPLACE)
    inpl.step(3)
    path = inpl.write_restart(3, str(tmp_path))
    inpl.close()
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    resumed = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_INPLACE)
    try:
        assert resumed.read_restart(path) == 3
        resumed.step(3)
        assert_same(snapshot(resumed, F), snapshot(strict, F), "resumed")
    finally:
        resumed.close()
        strict.close()


def test_inplace_reads_between_collide_species_and_streaming():
    """Between an in-place Collide_Species and Streaming the post-collision
    set (LBM_POP_FPC) gathers from the new layout and equals the strict
    mode's; the pre-collision species set is gone, so reading LBM_POP_F or
    uploading populations there is LBM_E_STATE; G and GPC read as in strict
    mode; stepping on stays bitwise equal to strict mode."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    F = fl["type"].ravel() == 0
    ns = 9
    strict = make(fl, LBM_MODE_STRICT)
    inpl = make(fl, LBM_MODE_INPLACE)
    try:
        strict.step(3)
        inpl.step(3)
        for lat in (strict, inpl):
            lat.Collide_Species()
            lat.Collide()
        a = inpl.get_populations(POP_FPC).reshape(ns, 27, -1)[:, :, F]
        b = strict.get_populations(POP_FPC).reshape(ns, 27, -1)[:, :, F]
        assert np.array_equal(a, b), "post-collision species set differs before Streaming"
        for which in (POP_G, POP_GPC):
            assert np.array_equal(inpl.get_populations(which).reshape(27, -1)[:, F],
                                  strict.get_populations(which).reshape(27, -1)[:, F])
        with pytest.raises(LbmError, match="pre-collision"):
            inpl.get_populations(POP_F)
        with pytest.raises(LbmError, match="lbm_collide_species and lbm_stream"):
            inpl.set_populations(POP_FPC, strict.get_populations(POP_FPC))
        with pytest.raises(LbmError, match="lbm_stream must follow"):
            inpl.Collide_Species()
        for lat in (strict, inpl):
            lat.Streaming()
            lat.TMS_BC()
            lat.calculate_moment()
        assert_same(snapshot(inpl, F), snapshot(strict, F), "after the mid-step reads")
        strict.step(3)
        inpl.step(3)
        assert_same(snapshot(inpl, F), snapshot(strict, F), "three steps later")
    finally:
        strict.close()
        inpl.close()


def test_inplace_population_upload_restarts_from_layout_zero():
    """An upload of the incoming species set mid-run (both contexts given the
    strict context's populations) continues bitwise like the strict mode."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    F = fl["type"].ravel() == 0
    strict = make(fl, LBM_MODE_STRICT)
    inpl = make(fl, LBM_MODE_INPLACE)
    try:
        strict.step(3)
        inpl.step(3)  # odd count: the buffer is in layout 1
        inpl.set_populations(POP_FPC, strict.get_populations(POP_FPC))
        inpl.set_populations(POP_F, strict.get_populations(POP_F))
        assert_same(snapshot(inpl, F), snapshot(strict, F), "after the upload")
        strict.step(2)
        inpl.step(2)
        assert_same(snapshot(inpl, F), snapshot(strict, F), "two steps after the upload")
    finally:
        strict.close()
        inpl.close()


def test_inplace_holds_one_species_buffer():
    """Before its first step (no low species list yet) the in-place
    context's device memory is the fused one's (same side store) minus one
    species population buffer, 27 NS populations per cell."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    n = fl["type"].size
    for precision, esize in ((LBM_FP64, 8), (LBM_FP32_SHIFTED, 4)):
        fused = make(fl, LBM_MODE_FUSED, precision)
        inpl = make(fl, LBM_MODE_INPLACE, precision)
        try:
            assert fused.device_bytes() - inpl.device_bytes() == n * 27 * 9 * esize
        finally:
            fused.close()
            inpl.close()


def test_inplace_refuses_free_slip_and_single_component():
    fl = cases.reacting_box(4, 4, 4, T_K=1300.0)
    fl["type"][:, 0, :] = TYPE_FS
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=False, mode=LBM_MODE_INPLACE)
    try:
        with pytest.raises(LbmError, match="TYPE_FS"):
            lat.load_case(fl)
    finally:
        lat.close()
    with pytest.raises(LbmError, match="LBM_MODE_INPLACE"):
        Lattice(4, 4, 4, 0.01, mode=LBM_MODE_INPLACE)
