"""lbm_amd: host-side mirror of the reference's LBM class over the MI355X C ABI.

The compute path is liblbm_amd.so (hand-written HIP for gfx950, declared in
include/lbm_amd.h); this package only binds it and provides case set-ups.
"""
from .capi import (FIELDS, LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_STRICT, POP_F, POP_FPC, POP_G,
                   POP_GPC, Lattice, LbmError, MixtureLattice, load_library)
from . import cases

__all__ = ["Lattice", "MixtureLattice", "LbmError", "load_library", "cases", "FIELDS", "LBM_FP64", "LBM_FP32_SHIFTED",
           "LBM_MODE_FUSED", "LBM_MODE_STRICT", "POP_F", "POP_FPC", "POP_G", "POP_GPC"]
