"""Shared helpers of the multicomponent GPU tests: a mode (fused, in-place)
against the strict calls bitwise, and seeding the device context with the
oracle's state, through the C ABI."""
import numpy as np

from lbm_amd import cases
from lbm_amd.capi import LBM_FP64, LBM_MODE_STRICT, POP_F, POP_FPC, POP_G, POP_GPC, MixtureLattice

FIELDS = ("rho", "u", "v", "w", "T", "p", "rhoe", "HRR")
SPECIES_FIELDS = ("X", "rho", "u", "v", "w")


def make(fl, mode, precision=LBM_FP64, reaction=True):
    nx, ny, nz = (s - 2 for s in fl["type"].shape)
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=reaction, mode=mode, precision=precision)
    lat.load_case(fl)
    lat.Init()
    lat.calculate_moment()
    return lat


def snapshot(lat, F):
    ns = len(lat.SPECIES)
    out = {}
    for k in FIELDS:
        out[k] = lat.get_field(k).ravel()[F]
    for k in SPECIES_FIELDS:
        out["sp_" + k] = lat.get_species_field(k).reshape(ns, -1)[:, F]
    for which, nm in ((POP_F, "f"), (POP_FPC, "fpc")):
        out[nm] = lat.get_populations(which).reshape(ns, 27, -1)[:, :, F]
    for which, nm in ((POP_G, "g"), (POP_GPC, "gpc")):
        out[nm] = lat.get_populations(which).reshape(27, -1)[:, F]
    return out


def assert_same(a, b, where):
    for k in a:
        if not np.array_equal(a[k], b[k], equal_nan=True):
            d = np.abs(a[k] - b[k])
            raise AssertionError(f"{where}: {k} differs (max |diff| {np.nanmax(d):.3e}, "
                                 f"{int(np.count_nonzero(d))} values)")


def compare(fl, steps, mode, precision=LBM_FP64, reaction=True, tag=""):
    """`steps` steps of `mode` and of the strict calls from the same set-up;
    every fluid cell bitwise equal after each step."""
    print(f"mode {mode} vs strict {tag} grid {fl['type'].shape} steps {steps} precision {precision}", flush=True)
    F = fl["type"].ravel() == 0
    strict = make(fl, LBM_MODE_STRICT, precision, reaction)
    other = make(fl, mode, precision, reaction)
    try:
        for step in range(1, steps + 1):
            strict.step(1)
            other.step(1)
            assert_same(snapshot(other, F), snapshot(strict, F), f"{tag} step {step}")
    finally:
        strict.close()
        other.close()


def seed_gpu_from_oracle(lat, s):
    """Copies the oracle's full state into the device context."""
    shape = lat.shape
    for k in ("rho", "u", "v", "w", "p", "T", "rhoe", "HRR"):
        lat.set_field(k, getattr(s, k).reshape(shape))
    for (a, b), name in zip(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)), ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz")):
        lat.set_field(name, s.ptens[a, b].reshape(shape))
    for d, name in enumerate(("qx", "qy", "qz")):
        lat.set_field(name, s.eflux[d].reshape(shape))
    ns = s.ns
    lat.set_species_field("X", s.X.reshape((ns,) + shape))
    lat.set_species_field("rho", s.rho_a.reshape((ns,) + shape))
    for d, name in enumerate("uvw"):
        lat.set_species_field(name, s.u_a[d].reshape((ns,) + shape))
    lat.set_populations(POP_F, s.f.reshape((ns, 27) + shape))
    lat.set_populations(POP_FPC, s.fpc.reshape((ns, 27) + shape))
    lat.set_populations(POP_G, s.g.reshape((27,) + shape))
    lat.set_populations(POP_GPC, s.gpc.reshape((27,) + shape))
