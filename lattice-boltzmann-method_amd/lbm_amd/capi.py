"""ctypes binding of the HIP library and a host-side mirror of the reference's LBM class.

One library per chemical mechanism (the reference binds its Cantera
mechanism at compile time, lbm-main.cpp:83-88): liblbm_amd.so (H2-air, 9
species), liblbm_amd_ch4_2s.so (two-step methane), liblbm_amd_c3h8.so
(two-step propane); the single-component path is the same in each.

`Lattice` keeps the reference's member names and step-call structure
(lbm.hpp:104-274, lbm-main.cpp:126-196): Init, Collide_Species, Collide,
Streaming, TMS_BC, calculate_moment, run/loop, getters and setters.  Every
call goes through the C ABI of include/lbm_amd.h; there is no Python or CPU
fallback: without the built HIP library or a HIP device the constructor
raises.
"""
import ctypes
import os

import numpy as np

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_NAMES = {"h2air": "liblbm_amd.so", "ch4_2s": "liblbm_amd_ch4_2s.so", "c3h8": "liblbm_amd_c3h8.so",
             "nh3": "liblbm_amd_nh3.so"}
MECHANISMS = tuple(LIB_NAMES)
LIB_PATH = os.path.join(_PKG_DIR, LIB_NAMES["h2air"])

LBM_FP64, LBM_FP32_SHIFTED = 0, 1
LBM_MODE_FUSED, LBM_MODE_STRICT, LBM_MODE_INPLACE = 0, 1, 2
POP_F, POP_FPC, POP_G, POP_GPC = 0, 1, 2, 3

FIELDS = ["rho", "u", "v", "w", "p", "T", "rhoe", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "qx", "qy", "qz", "HRR"]
SPECIES_FIELDS = {"X": 0, "rho": 1, "u": 2, "v": 3, "w": 4}
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

EXPORTS = [
    "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_create_error", "lbm_set_types", "lbm_set_field",
    "lbm_get_field", "lbm_set_populations", "lbm_get_populations", "lbm_init", "lbm_collide_species",
    "lbm_collide", "lbm_stream", "lbm_tms_bc", "lbm_moments", "lbm_step", "lbm_diagnostics", "lbm_sync",
    "lbm_get_stream", "lbm_profile_enable", "lbm_profile_read", "lbm_device_bytes", "lbm_set_species_field",
    "lbm_get_species_field", "lbm_surrogate_param_count", "lbm_surrogate_init", "lbm_surrogate_capture",
    "lbm_surrogate_get_capture", "lbm_surrogate_train", "lbm_surrogate_enable", "lbm_surrogate_read_stats",
    "lbm_surrogate_get_params", "lbm_surrogate_set_params", "lbm_surrogate_get_norm", "lbm_surrogate_set_norm",
    "lbm_surrogate_forward", "lbm_surrogate_gradient", "lbm_write_vtk", "lbm_write_restart", "lbm_read_restart",
    "lbm_mechanism", "lbm_species_name", "lbm_set_conjugate", "lbm_set_smoothing", "lbm_surrogate_refresh", "lbm_surrogate_refresh_async", "lbm_surrogate_refresh_poll",
]

LBM_VTK_SI, LBM_VTK_MASS_FRACTIONS = 1, 2

# surrogate network (include/lbm_amd.h): layer widths; real inputs T, p, Y_k, dt
# and outputs dT, dY_k, for the species of the library's mechanism
SUR_DIMS = (16, 128, 128, 128, 16)
H2AIR_SPECIES = ("H2", "H", "O", "O2", "OH", "H2O", "HO2", "H2O2", "N2")


def surrogate_features(species):
    return ["T", "p"] + [f"Y_{s}" for s in species] + ["dt"]


def surrogate_targets(species):
    return ["dT"] + [f"dY_{s}" for s in species]


SUR_FEATURES = surrogate_features(H2AIR_SPECIES)
SUR_TARGETS = surrogate_targets(H2AIR_SPECIES)
SUR_NIN, SUR_NOUT = len(SUR_FEATURES), len(SUR_TARGETS)


class LbmError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("n_species", ctypes.c_int),
                ("precision", ctypes.c_int), ("mode", ctypes.c_int), ("device", ctypes.c_int),
                ("nu", ctypes.c_double), ("gas_const", ctypes.c_double), ("gamma", ctypes.c_double),
                ("prtl", ctypes.c_double), ("ref_rho", ctypes.c_double), ("ref_T", ctypes.c_double),
                ("dx_si", ctypes.c_double), ("dt_si", ctypes.c_double), ("reaction", ctypes.c_int)]


class SurrogateStats(ctypes.Structure):
    _fields_ = [("captured", ctypes.c_int64), ("predicted", ctypes.c_int64), ("fallback", ctypes.c_int64),
                ("audited", ctypes.c_int64), ("audit_failures", ctypes.c_int64),
                ("max_audit_error", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Diag(ctypes.Structure):
    _fields_ = [("mass", ctypes.c_double), ("energy", ctypes.c_double), ("kinetic", ctypes.c_double),
                ("enstrophy", ctypes.c_double), ("max_mach", ctypes.c_double), ("nan_count", ctypes.c_int64),
                ("tau_warnings", ctypes.c_int64), ("fallback_count", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_LIBS = {}


def library_path(mechanism="h2air"):
    if mechanism not in LIB_NAMES:
        raise LbmError(f"unknown mechanism {mechanism!r}: one of {', '.join(MECHANISMS)}")
    return os.path.join(_PKG_DIR, LIB_NAMES[mechanism])


def load_library(path=None, mechanism="h2air"):
    """Loads the library of `mechanism` (or the one at `path`); raises if it
    is not built."""
    path = path or library_path(mechanism)
    if path in _LIBS:
        return _LIBS[path]
    if not os.path.exists(path):
        raise LbmError(f"{path} is not built: run __graft_entry__.build() or `make -C lattice-boltzmann-method_amd"
                       f"{'' if mechanism == 'h2air' else ' MECH=' + mechanism}`")
    L = ctypes.CDLL(path)
    vp, i32, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    L.lbm_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(vp)]
    L.lbm_create.restype = i32
    L.lbm_destroy.argtypes = [vp]
    L.lbm_destroy.restype = None
    L.lbm_last_error.argtypes = [vp]
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_create_error.argtypes = []
    L.lbm_create_error.restype = ctypes.c_char_p
    L.lbm_set_types.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8)]
    for fn in ("lbm_set_field", "lbm_get_field", "lbm_set_populations", "lbm_get_populations",
               "lbm_set_species_field", "lbm_get_species_field"):
        getattr(L, fn).argtypes = [vp, i32, dp]
    for fn in ("lbm_init", "lbm_collide_species", "lbm_collide", "lbm_stream", "lbm_tms_bc", "lbm_moments",
               "lbm_sync"):
        getattr(L, fn).argtypes = [vp]
    L.lbm_step.argtypes = [vp, i32]
    L.lbm_diagnostics.argtypes = [vp, ctypes.POINTER(Diag)]
    L.lbm_get_stream.argtypes = [vp]
    L.lbm_get_stream.restype = vp
    L.lbm_profile_enable.argtypes = [vp, i32]
    L.lbm_profile_read.argtypes = [vp, ctypes.c_char_p, dp, ctypes.POINTER(ctypes.c_int64)]
    L.lbm_device_bytes.argtypes = [vp]
    L.lbm_device_bytes.restype = ctypes.c_int64
    fp, i64, u64 = ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.c_uint64
    L.lbm_surrogate_param_count.argtypes = []
    L.lbm_surrogate_param_count.restype = i64
    L.lbm_surrogate_init.argtypes = [vp, u64]
    L.lbm_surrogate_capture.argtypes = [vp, i64, i32]
    L.lbm_surrogate_get_capture.argtypes = [vp, fp, i64, ctypes.POINTER(i64)]
    L.lbm_surrogate_train.argtypes = [vp, i32, i32, ctypes.c_float, u64, fp]
    L.lbm_surrogate_refresh.argtypes = [vp, i32, i32, ctypes.c_float, u64, fp]
    L.lbm_surrogate_refresh_async.argtypes = [vp, i32, i32, ctypes.c_float, u64]
    L.lbm_surrogate_refresh_poll.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int)]
    L.lbm_surrogate_enable.argtypes = [vp, i32, i32, ctypes.c_double]
    L.lbm_surrogate_read_stats.argtypes = [vp, ctypes.POINTER(SurrogateStats)]
    for fn in ("lbm_surrogate_get_params", "lbm_surrogate_set_params", "lbm_surrogate_get_norm",
               "lbm_surrogate_set_norm"):
        getattr(L, fn).argtypes = [vp, fp]
    L.lbm_write_vtk.argtypes = [vp, ctypes.c_char_p, i32]
    L.lbm_write_restart.argtypes = [vp, ctypes.c_char_p, i32]
    L.lbm_read_restart.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(i32)]
    L.lbm_surrogate_forward.argtypes = [vp, fp, i64, fp]
    L.lbm_surrogate_gradient.argtypes = [vp, fp, fp, i64, fp, fp]
    L.lbm_set_conjugate.argtypes = [vp, i32]
    L.lbm_set_smoothing.argtypes = [vp, i32, ctypes.c_double]
    L.lbm_mechanism.argtypes = [ctypes.POINTER(i32)]
    L.lbm_mechanism.restype = ctypes.c_char_p
    L.lbm_species_name.argtypes = [i32]
    L.lbm_species_name.restype = ctypes.c_char_p
    _LIBS[path] = L
    return L


def mechanism_info(lib):
    """(name, species) of the mechanism a loaded library is built for."""
    n = ctypes.c_int()
    name = lib.lbm_mechanism(ctypes.byref(n)).decode()
    return name, [lib.lbm_species_name(k).decode() for k in range(n.value)]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Lattice:
    """Single-component D3Q27 lattice on one GPU, mirroring LBM(Nx, Ny, Nz, nu)."""

    def __init__(self, nx, ny, nz, nu, gas_const=1.0, gamma=1.4, prtl=0.5, precision=LBM_FP64,
                 mode=LBM_MODE_FUSED, device=0, ref_rho=1.0, ref_T=1.0):
        self._L = load_library()
        cfg = Config(nx, ny, nz, 0, precision, mode, device, nu, gas_const, gamma, prtl, ref_rho, ref_T, 0.0, 0.0, 0)
        self._create(cfg)

    def _create(self, cfg):
        h = ctypes.c_void_p()
        rc = self._L.lbm_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise LbmError(f"lbm_create failed ({rc}): {self._L.lbm_create_error().decode()}")
        self._h = h
        self.cfg = cfg
        self.shape = (cfg.nx + 2, cfg.ny + 2, cfg.nz + 2)
        self.ncell = int(np.prod(self.shape))
        self.n_species = cfg.n_species

    # -- lifetime
    def close(self):
        if getattr(self, "_h", None):
            self._L.lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise LbmError(f"{what} failed ({rc}): {self._L.lbm_last_error(self._h).decode()}")

    # -- getters mirroring lbm.hpp:200-217
    def get_Nx(self):
        return self.shape[0]

    def get_Ny(self):
        return self.shape[1]

    def get_Nz(self):
        return self.shape[2]

    def get_NX(self):
        return self.shape[0] - 2

    def get_NY(self):
        return self.shape[1] - 2

    def get_NZ(self):
        return self.shape[2] - 2

    def get_nu(self):
        return self.cfg.nu

    def get_gasconst(self):
        return self.cfg.gas_const

    def get_gamma(self):
        return self.cfg.gamma

    def get_prtl(self):
        return self.cfg.prtl

    def get_soundspeed(self, temp):
        return float(np.sqrt(self.cfg.gamma * self.cfg.gas_const * temp))

    # -- host access to cell data
    def set_types(self, types):
        t = np.ascontiguousarray(types, dtype=np.uint8).reshape(self.shape)
        self._check(self._L.lbm_set_types(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), "set_types")

    def set_field(self, name, values):
        a = np.ascontiguousarray(np.broadcast_to(values, self.shape), dtype=np.float64)
        self._check(self._L.lbm_set_field(self._h, FIELD_ID[name], _dp(a)), f"set_field({name})")

    def get_field(self, name):
        a = np.empty(self.shape, np.float64)
        self._check(self._L.lbm_get_field(self._h, FIELD_ID[name], _dp(a)), f"get_field({name})")
        return a

    def _pop_shape(self, which):
        if self.n_species and which in (POP_F, POP_FPC):
            return (self.n_species, 27) + self.shape
        return (27,) + self.shape

    def set_populations(self, which, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(self._pop_shape(which))
        self._check(self._L.lbm_set_populations(self._h, which, _dp(a)), "set_populations")

    def get_populations(self, which):
        a = np.empty(self._pop_shape(which), np.float64)
        self._check(self._L.lbm_get_populations(self._h, which, _dp(a)), "get_populations")
        return a

    def set_species_field(self, name, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape((self.n_species,) + self.shape)
        self._check(self._L.lbm_set_species_field(self._h, SPECIES_FIELDS[name], _dp(a)), f"set_species_field({name})")

    def get_species_field(self, name):
        a = np.empty((self.n_species,) + self.shape, np.float64)
        self._check(self._L.lbm_get_species_field(self._h, SPECIES_FIELDS[name], _dp(a)), f"get_species_field({name})")
        return a

    def load_case(self, fields):
        """Uploads a set-up from lbm_amd.cases (type + primitive fields, and
        species mole fractions 'X' for a multicomponent lattice)."""
        self.set_types(fields["type"])
        for k in ("rho", "u", "v", "w", "p", "T"):
            if k in fields:
                self.set_field(k, fields[k])
        if self.n_species and "X" in fields:
            self.set_species_field("X", fields["X"])

    def set_conjugate(self, on=True):
        """The CONJUGATE variant (defines.hpp:62-63): interior TYPE_S cells
        conduct heat and couple to the gas through TMS_BC (strict mode)."""
        self._check(self._L.lbm_set_conjugate(self._h, 1 if on else 0), "set_conjugate")

    def set_smoothing(self, on=True, alpha=0.1):
        """The SMOOTHING variant (lbm-smoothing.cpp:95-173; Init_smooting and
        calculate_moment_smoothing of lbm.hpp:154-174): species populations
        filtered by alpha times their Laplacian around Init and each moment
        pass (strict multicomponent contexts)."""
        self._check(self._L.lbm_set_smoothing(self._h, 1 if on else 0, float(alpha)), "set_smoothing")

    # -- the reference's step calls
    def Init(self):
        self._check(self._L.lbm_init(self._h), "Init")

    def Collide_Species(self):
        self._check(self._L.lbm_collide_species(self._h), "Collide_Species")

    def Collide(self):
        self._check(self._L.lbm_collide(self._h), "Collide")

    def Streaming(self):
        self._check(self._L.lbm_stream(self._h), "Streaming")

    def TMS_BC(self):
        self._check(self._L.lbm_tms_bc(self._h), "TMS_BC")

    def calculate_moment(self):
        self._check(self._L.lbm_moments(self._h), "calculate_moment")

    def loop(self, nstep, callback=None, tout=0):
        """LBM::loop body nstep times; `callback(step, self)` every tout steps."""
        for step in range(1, nstep + 1):
            self.Collide_Species()
            self.Collide()
            self.Streaming()
            self.TMS_BC()
            self.calculate_moment()
            if callback is not None and tout and step % tout == 0:
                callback(step, self)

    def step(self, n=1):
        self._check(self._L.lbm_step(self._h, n), "step")

    def run(self, nstep, callback=None, tout=0):
        self.Init()
        if callback is not None:
            callback(0, self)
        self.loop(nstep, callback, tout)

    # -- diagnostics, sync, profiling
    def diagnostics(self):
        d = Diag()
        self._check(self._L.lbm_diagnostics(self._h, ctypes.byref(d)), "diagnostics")
        return d.as_dict()

    def sync(self):
        self._check(self._L.lbm_sync(self._h), "sync")

    def stream_handle(self):
        return self._L.lbm_get_stream(self._h)

    def profile(self, on=True):
        self._check(self._L.lbm_profile_enable(self._h, 1 if on else 0), "profile_enable")

    def profile_read(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._check(self._L.lbm_profile_read(self._h, name.encode(), ctypes.byref(ms), ctypes.byref(n)),
                    "profile_read")
        return ms.value, n.value

    def device_bytes(self):
        return self._L.lbm_device_bytes(self._h)

    # -- output files (output.cpp, restart_file.cpp)
    def OutputVTK(self, nout, directory=".", si=False, mass_fractions=False):
        """Writes field%06d.vtr as OutputVTK does (output.cpp:9-210); returns its path."""
        path = os.path.join(directory, "field%06d.vtr" % nout)
        flags = (LBM_VTK_SI if si else 0) | (LBM_VTK_MASS_FRACTIONS if mass_fractions else 0)
        self._check(self._L.lbm_write_vtk(self._h, path.encode(), flags), "write_vtk")
        return path

    def write_restart(self, step, directory="."):
        """Writes restart%06d.dat in the reference's layout (restart_file.cpp:274-350); returns its path."""
        path = os.path.join(directory, "restart%06d.dat" % step)
        self._check(self._L.lbm_write_restart(self._h, path.encode(), int(step)), "write_restart")
        return path

    def read_restart(self, path):
        """Restores cell types, fields and populations from a restart file
        (restart_file.cpp:352-440); returns its step."""
        step = ctypes.c_int()
        self._check(self._L.lbm_read_restart(self._h, str(path).encode(), ctypes.byref(step)), "read_restart")
        return step.value

    # -- chemistry surrogate network (the network calls work on any context)
    def surrogate_init(self, seed):
        self._check(self._L.lbm_surrogate_init(self._h, seed), "surrogate_init")

    def surrogate_params(self):
        out = np.empty(self._L.lbm_surrogate_param_count(), np.float32)
        self._check(self._L.lbm_surrogate_get_params(self._h, _fp(out)), "surrogate_get_params")
        return out

    def set_surrogate_params(self, params):
        a = np.ascontiguousarray(params, np.float32)
        if a.size != self._L.lbm_surrogate_param_count():
            raise ValueError("parameter vector has the wrong length")
        self._check(self._L.lbm_surrogate_set_params(self._h, _fp(a)), "surrogate_set_params")

    def surrogate_norm(self):
        out = np.empty(96, np.float32)
        self._check(self._L.lbm_surrogate_get_norm(self._h, _fp(out)), "surrogate_get_norm")
        return out

    def set_surrogate_norm(self, norm):
        a = np.ascontiguousarray(norm, np.float32)
        if a.size != 96:
            raise ValueError("normalisation record has 96 entries")
        self._check(self._L.lbm_surrogate_set_norm(self._h, _fp(a)), "surrogate_set_norm")

    def surrogate_forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        if X.ndim != 2 or X.shape[1] != 16:
            raise ValueError("X must be [m][16]")
        Y = np.empty_like(X)
        self._check(self._L.lbm_surrogate_forward(self._h, _fp(X), X.shape[0], _fp(Y)), "surrogate_forward")
        return Y

    def surrogate_gradient(self, X, T):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        if X.shape != T.shape or X.ndim != 2 or X.shape[1] != 16:
            raise ValueError("X and T must both be [m][16]")
        g = np.empty(self._L.lbm_surrogate_param_count(), np.float32)
        loss = ctypes.c_float()
        self._check(self._L.lbm_surrogate_gradient(self._h, _fp(X), _fp(T), X.shape[0], _fp(g), ctypes.byref(loss)),
                    "surrogate_gradient")
        return float(loss.value), g


class MixtureLattice(Lattice):
    """Multicomponent lattice, mirroring LBM(Nx, Ny, Nz, species) of
    lbm-main.cpp:38-118 with units.set_m_kg_s(dx, dt), on the library of
    `mechanism` ("h2air": 9 species, "ch4_2s" and "c3h8": 6, "nh3": 31).  `SPECIES` lists
    the species in the order of every per-species array."""

    def __init__(self, nx, ny, nz, dx_si, dt_si, reaction=True, mode=LBM_MODE_STRICT, precision=LBM_FP64, device=0,
                 gamma=1.4, mechanism="h2air"):
        self._L = load_library(mechanism=mechanism)
        self.mechanism, self.SPECIES = mechanism_info(self._L)
        cfg = Config(nx, ny, nz, len(self.SPECIES), precision, mode, device, 0.0, 0.0, gamma, 0.0, 1.0, 1.0,
                     dx_si, dt_si, 1 if reaction else 0)
        self._create(cfg)

    # -- chemistry surrogate in the step
    def surrogate_capture(self, max_rows, stride=1):
        """Samples one cell in `stride` per plain chemistry call into a buffer of
        `max_rows` rows; stride 0 pauses (rows kept), max_rows 0 releases."""
        self._check(self._L.lbm_surrogate_capture(self._h, int(max_rows), int(stride)), "surrogate_capture")

    def surrogate_capture_stop(self):
        self._check(self._L.lbm_surrogate_capture(self._h, 1, 0), "surrogate_capture")

    def surrogate_captured(self, max_rows=None):
        n = ctypes.c_int64()
        self._check(self._L.lbm_surrogate_get_capture(self._h, None, 0, ctypes.byref(n)), "surrogate_get_capture")
        m = n.value if max_rows is None else min(n.value, max_rows)
        out = np.empty((max(m, 0), 32), np.float32)
        if m > 0:
            self._check(self._L.lbm_surrogate_get_capture(self._h, _fp(out), m, ctypes.byref(n)),
                        "surrogate_get_capture")
        return out

    def surrogate_train(self, iters, batch=4096, lr=1e-3, seed=1):
        losses = np.zeros(max(iters, 1), np.float32)
        self._check(self._L.lbm_surrogate_train(self._h, int(iters), int(batch), float(lr), int(seed), _fp(losses)),
                    "surrogate_train")
        return losses[:iters]

    def surrogate_refresh(self, iters, batch=4096, lr=3e-4, seed=3):
        """Widens the envelope to every captured row (the fallbacks and audits
        captured in the step included) and trains `iters` more steps."""
        losses = np.zeros(max(iters, 1), np.float32)
        self._check(self._L.lbm_surrogate_refresh(self._h, int(iters), int(batch), float(lr), int(seed),
                                                  _fp(losses)), "surrogate_refresh")
        return losses[:iters]

    def surrogate_refresh_async(self, iters, batch=4096, lr=3e-4, seed=3):
        """Starts a refresh on a second stream; the steps continue with the
        serving network until surrogate_refresh_poll applies the result."""
        self._check(self._L.lbm_surrogate_refresh_async(self._h, int(iters), int(batch), float(lr), int(seed)),
                    "surrogate_refresh_async")

    def surrogate_refresh_poll(self, wait=False):
        """Applies a finished refresh; True when this call applied one."""
        applied = ctypes.c_int(0)
        self._check(self._L.lbm_surrogate_refresh_poll(self._h, 1 if wait else 0, ctypes.byref(applied)),
                    "surrogate_refresh_poll")
        return bool(applied.value)

    def step_with_refresh(self, nsteps, window=25, max_rate=0.005, iters=300, batch=4096, lr=3e-4, overlap=False):
        """`nsteps` steps in windows of `window`; after a window in which more
        than `max_rate` of the chemistry updates fell back to (or were
        overruled by) the plain integrator, the network is refreshed on the
        captured rows (capture must be on): with overlap=False by
        lbm_surrogate_refresh between two windows, with overlap=True by
        lbm_surrogate_refresh_async on a second stream while the next windows
        step with the serving network, the result applied at the first window
        end after it finished (and at the end).  Returns (step seconds,
        refresh seconds, refresh count), each timed between device
        synchronisations; with overlap the training time is inside the steps'."""
        import time
        step_s = refresh_s = 0.0
        refreshes = 0
        done = 0
        in_flight = False
        prev = self.surrogate_stats()
        while done < nsteps:
            k = min(window, nsteps - done)
            self.sync()
            t0 = time.perf_counter()
            self.step(k)
            self.sync()
            step_s += time.perf_counter() - t0
            done += k
            st = self.surrogate_stats()
            bad = (st["fallback"] - prev["fallback"]) + (st["audit_failures"] - prev["audit_failures"])
            upd = (st["predicted"] - prev["predicted"]) + (st["fallback"] - prev["fallback"])
            prev = st
            t0 = time.perf_counter()
            if in_flight:
                if self.surrogate_refresh_poll(wait=done >= nsteps):
                    in_flight = False
                    refreshes += 1
            elif done < nsteps and upd > 0 and bad > max_rate * upd:
                if overlap:
                    self.surrogate_refresh_async(iters, batch=batch, lr=lr, seed=1000 + refreshes)
                    in_flight = True
                else:
                    self.surrogate_refresh(iters, batch=batch, lr=lr, seed=1000 + refreshes)
                    refreshes += 1
            refresh_s += time.perf_counter() - t0
        if in_flight:
            t0 = time.perf_counter()
            self.surrogate_refresh_poll(wait=True)
            refresh_s += time.perf_counter() - t0
            refreshes += 1
        return step_s, refresh_s, refreshes

    def surrogate_enable(self, on=True, audit_stride=64, audit_tol=0.5):
        self._check(self._L.lbm_surrogate_enable(self._h, 1 if on else 0, int(audit_stride), float(audit_tol)),
                    "surrogate_enable")

    def surrogate_stats(self):
        st = SurrogateStats()
        self._check(self._L.lbm_surrogate_read_stats(self._h, ctypes.byref(st)), "surrogate_read_stats")
        return st.as_dict()
