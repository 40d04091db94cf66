"""ctypes binding of the CPU oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker of the HIP path.  The
product never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F, S, P, A, FS, I, O, IC, OC, Q = range(10)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblbm_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        d3 = ctypes.POINTER(ctypes.c_double)
        L.orc_feq.restype = ctypes.c_double
        L.orc_feq.argtypes = [ctypes.c_int, ctypes.c_double, d3, ctypes.c_double, d3]
        L.orc_geq.restype = ctypes.c_double
        L.orc_geq.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, d3]
        L.orc_geq_moment.restype = ctypes.c_double
        L.orc_geq_moment.argtypes = [ctypes.c_int, ctypes.c_double, d3, d3, ctypes.c_double]
        L.orc_gstr.restype = ctypes.c_double
        L.orc_gstr.argtypes = [ctypes.c_int, ctypes.c_double, d3]
        for fn in ("orc_init_single", "orc_collide_single", "orc_stream_single", "orc_moments_single",
                   "orc_tms_single"):
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [ctypes.c_void_p]
        L.orc_diag_single.restype = None
        L.orc_diag_single.argtypes = [ctypes.c_void_p, d3]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def feq(l, rho, vel, theta, corr):
    return lib().orc_feq(l, rho, _dp(np.ascontiguousarray(vel, np.float64)), theta,
                         _dp(np.ascontiguousarray(corr, np.float64)))


def geq(l, rho, U, theta, vel):
    return lib().orc_geq(l, rho, U, theta, _dp(np.ascontiguousarray(vel, np.float64)))


def geq_moment(l, rhoe, q, Rt, theta):
    return lib().orc_geq_moment(l, rhoe, _dp(np.ascontiguousarray(q, np.float64)),
                                _dp(np.ascontiguousarray(np.asarray(Rt, np.float64).reshape(9))), theta)


def gstr(l, g, dq):
    return lib().orc_gstr(l, g, _dp(np.ascontiguousarray(dq, np.float64)))


class _OrcSingle(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("R", ctypes.c_double), ("gamma", ctypes.c_double), ("prtl", ctypes.c_double),
                ("nu", ctypes.c_double),
                ("type", ctypes.c_void_p),
                ("rho", ctypes.c_void_p), ("u", ctypes.c_void_p), ("v", ctypes.c_void_p),
                ("w", ctypes.c_void_p), ("p", ctypes.c_void_p), ("T", ctypes.c_void_p),
                ("rhoe", ctypes.c_void_p), ("ptens", ctypes.c_void_p), ("eflux", ctypes.c_void_p),
                ("f", ctypes.c_void_p), ("fpc", ctypes.c_void_p), ("g", ctypes.c_void_p),
                ("gpc", ctypes.c_void_p), ("tau_warnings", ctypes.c_int64), ("conjugate", ctypes.c_int)]


class SingleState:
    """Single-component state on a ghost-inclusive grid (nx, ny, nz), numpy-owned.

    Field defaults follow MIXTURE (lbm.hpp:56-88): rho=1, u=v=w=0, p=T=1/3,
    type=TYPE_F.
    """

    SCALARS = ("rho", "u", "v", "w", "p", "T", "rhoe")

    def __init__(self, nx, ny, nz, R, gamma, prtl, nu, conjugate=False):
        self.shape = (nx, ny, nz)
        n = nx * ny * nz
        self.type = np.zeros(self.shape, np.uint8)
        self.rho = np.ones(self.shape)
        self.u = np.zeros(self.shape)
        self.v = np.zeros(self.shape)
        self.w = np.zeros(self.shape)
        self.p = np.full(self.shape, 1.0 / 3.0)
        self.T = np.full(self.shape, 1.0 / 3.0)
        self.rhoe = np.zeros(self.shape)
        self.ptens = np.zeros((9,) + self.shape)
        self.eflux = np.zeros((3,) + self.shape)
        self.f = np.zeros((27,) + self.shape)
        self.fpc = np.zeros((27,) + self.shape)
        self.g = np.zeros((27,) + self.shape)
        self.gpc = np.zeros((27,) + self.shape)
        self._s = _OrcSingle(nx, ny, nz, R, gamma, prtl, nu)
        self._s.conjugate = 1 if conjugate else 0
        for name in ("type", "rho", "u", "v", "w", "p", "T", "rhoe", "ptens", "eflux", "f", "fpc", "g", "gpc"):
            setattr(self._s, name, getattr(self, name).ctypes.data)
        self._n = n

    @property
    def tau_warnings(self):
        return self._s.tau_warnings

    def _call(self, fn):
        getattr(lib(), fn)(ctypes.byref(self._s))

    def init(self):
        self._call("orc_init_single")

    def collide(self):
        self._call("orc_collide_single")

    def stream(self):
        self._call("orc_stream_single")

    def tms_bc(self):
        self._call("orc_tms_single")

    def moments(self):
        self._call("orc_moments_single")

    def step(self):
        self.collide()
        self.stream()
        self.tms_bc()
        self.moments()

    def diag(self):
        out = np.zeros(4)
        lib().orc_diag_single(ctypes.byref(self._s), _dp(out))
        return out
