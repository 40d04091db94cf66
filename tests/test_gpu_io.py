"""Output files written from the device state (lbm_write_vtk,
lbm_write_restart) against the reference's own files of the same run
(tests/golden/channel_io.npz), and resuming from the reference's restart
file (lbm_read_restart).

Tolerances: VTK arrays are float32 of FP64 fields that agree with the
reference to ~1e-12, so at most one float32 ulp apart; the XML header is
byte-identical.  Restart records: FP64 values within 1e-10 of the
reference's (normwise per quantity), cell types and header exact; the 6
padding bytes after the type are not compared (the reference leaves them
uninitialised)."""
import os

import numpy as np
import pytest

from io_format import read_restart, read_vtk
from lbm_amd import cases
from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_STRICT, Lattice, MixtureLattice

pytestmark = pytest.mark.gpu

KEYS = ("nx", "ny", "nz", "U", "nu", "R", "gamma", "Pr", "T0", "p0", "Tin", "Tw", "pout")


def _golden(golden_dir):
    gd = np.load(os.path.join(golden_dir, "channel_io.npz"))
    return gd, dict(zip(KEYS, gd["params"]))


def _lattice(p, precision=LBM_FP64):
    nx, ny, nz = int(p["nx"]), int(p["ny"]), int(p["nz"])
    fl = cases.channel(nx, ny, nz, p["U"], T0=p["T0"], p0=p["p0"], T_in=p["Tin"], T_wall=p["Tw"], p_out=p["pout"],
                       R=p["R"])
    lat = Lattice(nx, ny, nz, p["nu"], gas_const=p["R"], gamma=p["gamma"], prtl=p["Pr"], mode=LBM_MODE_STRICT,
                  precision=precision, ref_rho=1.0, ref_T=p["T0"])
    lat.load_case(fl)
    return lat


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_vtk_file_matches_reference(golden_dir, tmp_path):
    gd, p = _golden(golden_dir)
    lat = _lattice(p)
    lat.Init()
    lat.step(int(gd["steps"]))
    path = lat.OutputVTK(int(gd["steps"]), str(tmp_path))
    assert os.path.basename(path) == "field000002.vtr"
    got = open(path, "rb").read()
    ref = bytes(gd["vtk"])
    assert len(got) == len(ref)
    hg, ag = read_vtk(got)
    hr, ar = read_vtk(ref)
    assert hg == hr
    for name in ar:
        a, b = ag[name], ar[name]
        if name == "CellType" or name.startswith("Coordinate"):
            assert np.array_equal(a, b), name
        else:
            assert np.max(np.abs(a - b)) <= 1.2e-7 * np.max(np.abs(b)), name
    assert got.endswith(b"  </AppendedData>\n</VTKFile>\n")
    lat.close()


def test_restart_file_matches_reference(golden_dir, tmp_path):
    gd, p = _golden(golden_dir)
    lat = _lattice(p)
    lat.Init()
    lat.step(int(gd["steps"]))
    path = lat.write_restart(int(gd["steps"]), str(tmp_path))
    hg, rg = read_restart(open(path, "rb").read())
    hr, rr = read_restart(gd["restart"])
    assert hg == hr
    assert np.array_equal(rg["type"], rr["type"])
    F = rr["type"] == 0
    for name in ("rho", "u", "v", "w", "p", "T", "rhoe", "ptens", "q", "f", "fpc", "g", "gpc"):
        assert _rel(rg[name][F], rr[name][F]) < 1e-10, name
    for name in ("rho", "u", "v", "w", "p", "T"):
        assert _rel(rg[name][~F], rr[name][~F]) < 1e-12, name
    lat.close()


def test_resume_from_the_reference_restart_file(golden_dir, tmp_path):
    """Reading the reference's state after 2 steps and running 3 more gives
    the state of 5 steps run here from the start."""
    gd, p = _golden(golden_dir)
    path = tmp_path / "restart000002.dat"
    path.write_bytes(bytes(gd["restart"]))
    a = _lattice(p)
    assert a.read_restart(path) == 2
    a.step(3)
    b = _lattice(p)
    b.Init()
    b.step(5)
    for name in ("rho", "u", "v", "w", "p", "T"):
        assert _rel(a.get_field(name)[1:-1, 1:-1, 1:-1], b.get_field(name)[1:-1, 1:-1, 1:-1]) < 1e-10, name
    a.close()
    b.close()


def test_restart_round_trip_fp32(golden_dir, tmp_path):
    """FP32 shifted storage: writing and reading back restores the state to
    the storage's rounding."""
    _, p = _golden(golden_dir)
    a = _lattice(p, LBM_FP32_SHIFTED)
    a.Init()
    a.step(3)
    path = a.write_restart(3, str(tmp_path))
    b = _lattice(p, LBM_FP32_SHIFTED)
    assert b.read_restart(path) == 3
    a.step(2)
    b.step(2)
    for name in ("rho", "u", "T"):
        assert _rel(b.get_field(name), a.get_field(name)) < 1e-6, name
    a.close()
    b.close()


def test_multicomponent_restart_round_trip_and_vtk(tmp_path):
    """Species restart layout (MIXTURE 600 B + 9 SPECIES 472 B per cell):
    round trip to identical populations and fields; the VTK file carries HRR
    and one mole-fraction and one velocity array per species."""
    fl = cases.reacting_channel(6, 4, 3)
    nx, ny, nz = 6, 4, 3
    a = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT)
    a.load_case(fl)
    a.Init()
    a.calculate_moment()
    a.step(2)
    path = a.write_restart(2, str(tmp_path))
    hdr, rec = read_restart(open(path, "rb").read(), species=True)
    assert hdr["species"] == list(cases.H2AIR_SPECIES) and hdr["step"] == 2
    T = a.get_field("T").ravel()
    assert np.array_equal(rec["mix"]["T"], T)
    assert np.array_equal(rec["sp"]["rho"].T, a.get_species_field("rho").reshape(9, -1))
    b = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT)
    b.load_case(fl)
    assert b.read_restart(path) == 2
    from lbm_amd.capi import POP_F, POP_FPC, POP_G
    for which in (POP_F, POP_FPC, POP_G):
        assert np.array_equal(b.get_populations(which), a.get_populations(which))
    a.step(1)
    b.step(1)
    assert np.array_equal(b.get_field("T"), a.get_field("T"))
    vtk = a.OutputVTK(3, str(tmp_path), si=True, mass_fractions=True)
    head, arr = read_vtk(open(vtk, "rb").read())
    assert "HRR" in arr and all(f"Mole Fraction {s}" in arr and f"Species Velocity {s}" in arr
                                for s in cases.H2AIR_SPECIES)
    Y = a.get_species_field("rho") / a.get_field("rho")[None]
    assert np.max(np.abs(arr["Mole Fraction O2"] - Y[3].astype(np.float32))) <= 1e-7
    a.close()
    b.close()
