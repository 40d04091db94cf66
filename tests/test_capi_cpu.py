"""C ABI checks that need no GPU: the library loads, exports every entry
point include/lbm_amd.h declares, and fails loudly without a device."""
import ctypes
import os
import re

import pytest

from lbm_amd import capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(REPO, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(lbm_[a-z_0-9]+)\s*\(", txt))
    return names


def test_header_declares_the_step_calls():
    names = declared_entry_points()
    for n in ("lbm_create", "lbm_destroy", "lbm_init", "lbm_collide_species", "lbm_collide", "lbm_stream",
              "lbm_tms_bc", "lbm_moments", "lbm_diagnostics", "lbm_sync"):
        assert n in names


@pytest.mark.parametrize("mechanism", capi.MECHANISMS)
def test_library_exports_every_declared_symbol(mechanism):
    lib = capi.load_library(mechanism=mechanism)
    missing = [n for n in sorted(declared_entry_points()) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= declared_entry_points()


@pytest.mark.parametrize("mechanism", capi.MECHANISMS)
def test_library_reports_its_mechanism(mechanism):
    from lbm_amd import cases
    from oracle.chem import Mechanism
    lib = capi.load_library(mechanism=mechanism)
    name, species = capi.mechanism_info(lib)
    assert name == mechanism
    assert species == list(cases.MECH_SPECIES[mechanism]) == Mechanism(mechanism).names
    assert lib.lbm_species_name(-1) is None and lib.lbm_species_name(len(species)) is None


def test_species_count_must_match_the_library():
    lib = capi.load_library(mechanism="ch4_2s")
    cfg = capi.Config(4, 4, 4, 9, 0, 1, 0, 0.0, 0.0, 1.4, 0.0, 1.0, 1.0, 5e-6, 1e-9, 1)
    h = ctypes.c_void_p()
    assert lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
    assert b"ch4_2s" in lib.lbm_create_error()


def test_fp32_is_refused_for_the_31_species_mechanism():
    """The FP32 moment pass stages per-species sums in LDS for at most 16
    species: the ammonia build refuses LBM_FP32_SHIFTED before touching a
    device."""
    lib = capi.load_library(mechanism="nh3")
    cfg = capi.Config(4, 4, 4, 31, capi.LBM_FP32_SHIFTED, 1, 0, 0.0, 0.0, 1.4, 0.0, 1.0, 1.0, 5e-6, 1e-9, 1)
    h = ctypes.c_void_p()
    assert lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == -4  # LBM_E_UNSUPPORTED
    assert b"16 species" in lib.lbm_create_error()


def test_config_struct_matches_header():
    txt = open(os.path.join(REPO, "include", "lbm_amd.h")).read()
    body = re.search(r"typedef struct \{(.*?)\} lbm_config;", txt, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"(int|double)\s+([A-Za-z_0-9, ]+);", body)
    flat = []
    for ty, names in fields:
        flat += [(ty, n.strip()) for n in names.split(",")]
    got = [(("int" if f[1] is ctypes.c_int else "double"), f[0]) for f in capi.Config._fields_]
    assert got == flat


def test_create_without_device_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(capi.LbmError, match="no HIP device|hip"):
        capi.Lattice(4, 4, 4, 0.01)


def test_bad_config_is_rejected_before_touching_the_device():
    lib = capi.load_library()
    cfg = capi.Config(0, 4, 4, 0, 0, 0, 0, 0.01, 1.0, 1.4, 0.7, 1.0, 1.0)
    h = ctypes.c_void_p()
    assert lib.lbm_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"grid" in lib.lbm_create_error()
