"""Readers of the reference's output files, for the tests.

Restart files (restart_file.cpp:274-350): int32 dx dy dz Nx Ny Nz dt_sim
step; single component: float64 nu, gas_const, gamma, Ra, prtl, then one
1024-byte MIXTURE record per cell (lbm.hpp:56-88: short type, 6 padding
bytes, f[27], fpc[27], p_tensor[3][3], rho u v w p, g[27], gpc[27], rhoe,
energy_flux[3], temp); with species: uint64 nSpecies, (uint64 length +
bytes) per name, float64 permeability, then per cell a 600-byte MIXTURE
(no f/fpc, HRR last) and nSpecies 472-byte SPECIES records (X, rho, f, fpc,
u, v, w).  Cells in (i, j, k) order, k fastest.

VTK files (output.cpp:9-210): XML header, then '_' and raw appended
arrays, each a uint32 byte count followed by float32 values in (k, j, i)
order.
"""
import re

import numpy as np

MIX1 = np.dtype([("type", "<i2"), ("pad", "V6"), ("f", "<f8", 27), ("fpc", "<f8", 27), ("ptens", "<f8", 9),
                 ("rho", "<f8"), ("u", "<f8"), ("v", "<f8"), ("w", "<f8"), ("p", "<f8"), ("g", "<f8", 27),
                 ("gpc", "<f8", 27), ("rhoe", "<f8"), ("q", "<f8", 3), ("T", "<f8")])
MIXM = np.dtype([("type", "<i2"), ("pad", "V6"), ("ptens", "<f8", 9), ("rho", "<f8"), ("u", "<f8"), ("v", "<f8"),
                 ("w", "<f8"), ("p", "<f8"), ("g", "<f8", 27), ("gpc", "<f8", 27), ("rhoe", "<f8"), ("q", "<f8", 3),
                 ("T", "<f8"), ("HRR", "<f8")])
SPEC = np.dtype([("X", "<f8"), ("rho", "<f8"), ("f", "<f8", 27), ("fpc", "<f8", 27), ("u", "<f8"), ("v", "<f8"),
                 ("w", "<f8")])
assert MIX1.itemsize == 1024 and MIXM.itemsize == 600 and SPEC.itemsize == 472


def read_restart(data, species=False):
    """Parses restart bytes -> (header dict, records)."""
    b = bytes(data)
    hdr = dict(zip(("dx", "dy", "dz", "Nx", "Ny", "Nz", "dt_sim", "step"), np.frombuffer(b, "<i4", 8).tolist()))
    off = 32
    n = hdr["Nx"] * hdr["Ny"] * hdr["Nz"]
    if not species:
        hdr.update(zip(("nu", "gas_const", "gamma", "Ra", "prtl"), np.frombuffer(b, "<f8", 5, off).tolist()))
        off += 40
        rec = np.frombuffer(b, MIX1, n, off)
        assert off + n * MIX1.itemsize == len(b)
        return hdr, rec
    ns = int(np.frombuffer(b, "<u8", 1, off)[0])
    off += 8
    names = []
    for _ in range(ns):
        ln = int(np.frombuffer(b, "<u8", 1, off)[0])
        off += 8
        names.append(b[off:off + ln].decode())
        off += ln
    hdr["species"] = names
    hdr["permeability"] = float(np.frombuffer(b, "<f8", 1, off)[0])
    off += 8
    cell = np.dtype([("mix", MIXM), ("sp", SPEC, ns)])
    rec = np.frombuffer(b, cell, n, off)
    assert off + n * cell.itemsize == len(b)
    return hdr, rec


def read_vtk(data):
    """Parses VTK bytes -> (header text, {name: float32 array in (i, j, k)
    order}, coordinates)."""
    b = bytes(data)
    cut = b.index(b"<AppendedData encoding=\"raw\">_") + len(b"<AppendedData encoding=\"raw\">_")
    head = b[:cut].decode()
    ext = [int(x) for x in re.search(r'WholeExtent="0 (\d+) 0 (\d+) 0 (\d+)"', head).groups()]
    arrays = {}
    for m in re.finditer(r'Name="([^"]+)"(?: NumberOfComponents="(\d+)")? format="appended" offset="(\d+)"', head):
        name, nc, off = m.group(1), int(m.group(2) or 1), int(m.group(3))
        size = int(np.frombuffer(b, "<u4", 1, cut + off)[0])
        vals = np.frombuffer(b, "<f4", size // 4, cut + off + 4)
        if name.startswith("Coordinate"):
            arrays[name] = vals.copy()
        else:
            nx, ny, nz = ext
            a = vals.reshape(nz, ny, nx, nc).transpose(2, 1, 0, 3)
            arrays[name] = a[..., 0].copy() if nc == 1 else a.copy()
    return head, arrays
