"""Driver of the compiled multicomponent CPU step (oracle/mc_cpu, built from
oracle/mc_cpu.cpp by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: used by tests/ (against the numpy oracle) and by
the cpu_baseline leg of bench.py (the compiled all-cores baseline of the
reacting mixing layer).  The product never imports this module.
"""
import json
import os
import subprocess

import numpy as np

from .chem import Mechanism

_HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(_HERE, "mc_cpu")


def write_case(path, fl, dx, dt, reaction, state=None):
    """Case file of a periodic box: the ghost-inclusive type array and the
    primitive fields of lbm_amd.cases (rho, u, v, w, p, T; X [NS][...]).
    With `state` (an oracle.multicomp.MultiState), the run starts from that
    state's fields, species planes and populations instead of Init."""
    t = np.ascontiguousarray(fl["type"], np.uint8)
    nx, ny, nz = t.shape
    with open(path, "wb") as fh:
        fh.write(b"MCCPU01\0")
        fh.write(np.array([nx, ny, nz, 1 if reaction else 0, 0 if state is None else 1], np.int32).tobytes())
        fh.write(np.array([dx, dt], np.float64).tobytes())
        fh.write(t.tobytes())
        src = fl if state is None else {k: getattr(state, k) for k in ("rho", "u", "v", "w", "p", "T", "X")}
        for k in ("rho", "u", "v", "w", "p", "T", "X"):
            fh.write(np.ascontiguousarray(src[k], np.float64).tobytes())
        if state is not None:
            P = state.ptens
            planes = [state.rhoe, np.stack([P[0, 0], P[1, 1], P[2, 2], P[0, 1], P[0, 2], P[1, 2]]), state.rho_a,
                      state.u_a, state.f, state.g]
            for a in planes:
                fh.write(np.ascontiguousarray(a, np.float64).tobytes())


def read_out(path, n, ns=None):
    ns = ns or Mechanism().ns
    a = np.fromfile(path, np.float64)
    out, o = {}, 0
    for k, size in (("rho", n), ("u", n), ("v", n), ("w", n), ("p", n), ("T", n), ("rhoe", n), ("rho_a", ns * n),
                    ("u_a", 3 * ns * n), ("f", ns * 27 * n), ("g", 27 * n)):
        out[k] = a[o:o + size]
        o += size
    if o != a.size:
        raise RuntimeError(f"mc_cpu output holds {a.size} values, expected {o}")
    out["rho_a"] = out["rho_a"].reshape(ns, n)
    out["u_a"] = out["u_a"].reshape(3, ns, n)
    out["f"] = out["f"].reshape(ns, 27, n)
    out["g"] = out["g"].reshape(27, n)
    return out


def run(case, steps, threads, out=None, timeout=None):
    """Runs the compiled step; returns its JSON record (cells, steps, threads,
    seconds, MLUPS)."""
    if not os.path.exists(EXE):
        raise RuntimeError("oracle/mc_cpu not built: run `make -C oracle`")
    cmd = [EXE, case, str(steps), str(threads)] + ([out] if out else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"mc_cpu failed ({r.returncode}): {r.stderr.strip()}")
    return json.loads(r.stdout.strip().splitlines()[-1])
