"""The C++ host mirror of the reference's LBM class (host/lbm_host.hpp) drives
the GPU path through the C ABI in the reference's main -> run -> loop shape
(lbm-main.cpp:126-196), for single-component and multicomponent set-ups.

host/lbm_run reads a set-up written here, runs LBM::run (Init, then the five
step calls per step) and writes the fields, species fields and populations.
The same set-up is stepped through the Python ctypes binding; the two must be
bitwise equal, since they call the same library in the same order.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import (FIELDS, LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT,
                          POP_F, POP_G, Lattice, MixtureLattice)

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "lattice-boltzmann-method_amd", "host", "lbm_run")
SPECIES_FIELDS = ("X", "rho", "u", "v", "w")


def write_case(path, fl, ns, mode, precision, reaction, nstep, gas=(0.0, 1.0, 1.4, 0.5), ref=(1.0, 1.0),
               units=(0.0, 0.0)):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    with open(path, "wb") as fh:
        fh.write(b"LBMCASE1")
        fh.write(struct.pack("<8i", nx, ny, nz, ns, mode, precision, 1 if reaction else 0, nstep))
        fh.write(struct.pack("<8d", *gas, *ref, *units))
        fh.write(np.ascontiguousarray(fl["type"], dtype=np.uint8).tobytes())
        for k in ("rho", "u", "v", "w", "p", "T"):
            fh.write(np.ascontiguousarray(fl[k], dtype=np.float64).tobytes())
        if ns:
            fh.write(np.ascontiguousarray(fl["X"], dtype=np.float64).tobytes())


def read_out(path, n):
    with open(path, "rb") as fh:
        assert fh.read(8) == b"LBMOUT01"
        ns = struct.unpack("<i", fh.read(4))[0]
        a = np.frombuffer(fh.read(), dtype=np.float64)
    out, o = {}, 0
    for k in FIELDS:
        out[k] = a[o:o + n]
        o += n
    for k in SPECIES_FIELDS if ns else ():
        out["sp_" + k] = a[o:o + n * ns]
        o += n * ns
    m = n * 27 * max(ns, 1)
    out["f"] = a[o:o + m]
    out["g"] = a[o + m:o + m + 27 * n]
    return out


def python_run(lat, fl, nstep):
    lat.load_case(fl)
    lat.Init()
    lat.step(nstep)
    out = {k: lat.get_field(k).ravel() for k in FIELDS}
    if lat.n_species:
        for k in SPECIES_FIELDS:
            out["sp_" + k] = lat.get_species_field(k).ravel()
    out["f"] = lat.get_populations(POP_F).ravel()
    out["g"] = lat.get_populations(POP_G).ravel()
    lat.close()
    return out


def compare(cpp, py, F, ns):
    nf = F.size
    for k in cpp:
        a, b = cpp[k], py[k]
        if k.startswith("sp_"):
            a, b = a.reshape(ns, nf)[:, F], b.reshape(ns, nf)[:, F]
        elif k in ("f", "g"):
            a, b = a.reshape(-1, nf)[:, F], b.reshape(-1, nf)[:, F]
        else:
            a, b = a[F], b[F]
        assert np.array_equal(a, b, equal_nan=True), f"{k} differs between the C++ host and the ctypes path"


def run_driver(case, out):
    print(f"lbm_run {case}", flush=True)
    res = subprocess.run([DRIVER, case, out], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    print(res.stdout.strip(), flush=True)


@pytest.mark.parametrize("mode", [LBM_MODE_FUSED, LBM_MODE_STRICT])
def test_cpp_host_taylor_green_matches_ctypes(tmp_path, mode):
    N, steps = 12, 8
    pr = cases.tgv_params(N)
    fl = cases.taylor_green_3d(**pr)
    case, out = str(tmp_path / "tgv.case"), str(tmp_path / "tgv.out")
    write_case(case, fl, 0, mode, LBM_FP64, False, steps, gas=(pr["nu"], pr["R"], pr["gamma"], pr["Pr"]))
    run_driver(case, out)
    lat = Lattice(N, N, N, pr["nu"], gas_const=pr["R"], gamma=pr["gamma"], prtl=pr["Pr"], mode=mode)
    py = python_run(lat, fl, steps)
    compare(read_out(out, fl["type"].size), py, fl["type"].ravel() == 0, 0)


@pytest.mark.parametrize("mode,precision", [(LBM_MODE_FUSED, LBM_FP64), (LBM_MODE_INPLACE, LBM_FP64),
                                            (LBM_MODE_INPLACE, LBM_FP32_SHIFTED)])
def test_cpp_host_mixing_layer_matches_ctypes(tmp_path, mode, precision):
    """The reacting mixing layer (NSCBC outflow faces, chemistry on) through
    the host class's multicomponent constructor LBM(Nx, Ny, Nz, dx, dt,
    reaction)."""
    fl = cases.mixing_layer(6, 12, 5, T_fuel=300.0, T_ox=1200.0, U_s=0.02, delta=3.0)
    steps = 6
    case, out = str(tmp_path / "mix.case"), str(tmp_path / "mix.out")
    write_case(case, fl, 9, mode, precision, True, steps, units=cases.UNITS_FLAME)
    run_driver(case, out)
    lat = MixtureLattice(6, 12, 5, *cases.UNITS_FLAME, reaction=True, mode=mode, precision=precision)
    py = python_run(lat, fl, steps)
    compare(read_out(out, fl["type"].size), py, fl["type"].ravel() == 0, 9)
