"""The reference's output files (tests/golden/channel_io.npz, written by its
own OutputVTK and write_restart) read back and checked against the oracle's
state of the same run: pins the readers in tests/io_format.py, i.e. the
byte layout the library's writers reproduce."""
import os

import numpy as np

from io_format import read_restart, read_vtk
from lbm_amd import cases
from oracle import binding as ob

KEYS = ("nx", "ny", "nz", "U", "nu", "R", "gamma", "Pr", "T0", "p0", "Tin", "Tw", "pout")


def _oracle_channel(p, steps):
    nx, ny, nz = int(p["nx"]), int(p["ny"]), int(p["nz"])
    fl = cases.channel(nx, ny, nz, p["U"], T0=p["T0"], p0=p["p0"], T_in=p["Tin"], T_wall=p["Tw"], p_out=p["pout"],
                       R=p["R"])
    s = ob.SingleState(nx + 2, ny + 2, nz + 2, p["R"], p["gamma"], p["Pr"], p["nu"])
    for k in ("type", "rho", "u", "v", "w", "p", "T"):
        getattr(s, k)[...] = fl[k]
    s.init()
    for _ in range(steps):
        s.step()
    return s


def _golden(golden_dir):
    gd = np.load(os.path.join(golden_dir, "channel_io.npz"))
    return gd, dict(zip(KEYS, gd["params"]))


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_golden_restart_layout_matches_the_oracle_state(golden_dir):
    gd, p = _golden(golden_dir)
    hdr, rec = read_restart(gd["restart"])
    assert (hdr["dx"], hdr["dy"], hdr["dz"], hdr["dt_sim"], hdr["step"]) == (1, 1, 1, 1, 2)
    assert (hdr["Nx"], hdr["Ny"], hdr["Nz"]) == (int(p["nx"]) + 2, int(p["ny"]) + 2, int(p["nz"]) + 2)
    assert hdr["nu"] == p["nu"] and hdr["gas_const"] == p["R"] and hdr["gamma"] == p["gamma"]
    assert hdr["prtl"] == p["Pr"] and hdr["Ra"] == 1.0
    s = _oracle_channel(p, int(gd["steps"]))
    assert np.array_equal(rec["type"], s.type.ravel())
    F = s.type.ravel() == ob.F
    n = F.size
    for name in ("rho", "u", "v", "w", "p", "T", "rhoe"):
        assert _rel(rec[name][F], getattr(s, name).ravel()[F]) < 1e-12, name
    assert _rel(rec["ptens"][F], s.ptens.reshape(9, n).T[F]) < 1e-12
    assert _rel(rec["q"][F], s.eflux.reshape(3, n).T[F]) < 1e-12
    for name in ("f", "fpc", "g", "gpc"):
        assert _rel(rec[name][F], getattr(s, name).reshape(27, n).T[F]) < 1e-12, name
    # ghosts keep their set-up and Init values
    for name in ("rho", "u", "v", "w", "p", "T"):
        assert _rel(rec[name][~F], getattr(s, name).ravel()[~F]) < 1e-15, name


def test_golden_vtk_arrays_match_the_oracle_state(golden_dir):
    gd, p = _golden(golden_dir)
    head, arr = read_vtk(gd["vtk"])
    s = _oracle_channel(p, int(gd["steps"]))
    shape = s.type.shape
    assert 'WholeExtent="0 %d 0 %d 0 %d"' % shape in head
    assert np.array_equal(arr["CellType"], s.type.astype(np.float32))
    for vtk_name, name in (("Density", "rho"), ("Temperature", "T"), ("Pressure", "p")):
        ref = getattr(s, name).astype(np.float32)
        # the oracle agrees with the reference to ~1e-15: one float32 ulp at most
        assert np.max(np.abs(arr[vtk_name] - ref) / np.abs(ref)) <= 1.2e-7, vtk_name
    vel = np.stack([s.u, s.v, s.w], axis=-1).astype(np.float32)
    assert np.max(np.abs(arr["Velocity"] - vel)) <= 1.2e-7 * np.max(np.abs(vel))
    for ax, L in zip("XYZ", shape):
        assert np.array_equal(arr["Coordinate" + ax], np.arange(L + 1, dtype=np.float32))
