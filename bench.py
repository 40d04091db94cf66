"""Throughput benchmark of the D3Q27 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-cpu-baseline] [--no-compare]

Prints ONE JSON line.  A step is one pass of the loop body (Collide_Species,
Collide, Streaming, TMS_BC, calculate_moment) over the whole synthetic box;
the timed region starts with the state already resident in HBM.

Workloads
  turb256   (default) thermal compressible decaying turbulence, 256^3
            periodic, single component: the FP32 shifted-population path is
            `value`; the FP64 plain path is timed beside it for the speedup
            and the FP32 error against FP64 after the timed steps.
  tgv64     3-D Taylor-Green vortex, 64^3, FP64 plain path.
"""
import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "lattice-boltzmann-method_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="turb256", choices=["turb256", "tgv64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 companion run")
    ap.add_argument("--cpu-threads", type=int, default=16)
    return ap.parse_args()


def make_lattice(fl, pr, precision, mode):
    from lbm_amd.capi import Lattice
    N = pr["N"]
    lat = Lattice(N, N, N, pr["nu"], gas_const=pr["R"], gamma=pr["gamma"], prtl=pr["Pr"], precision=precision,
                  mode=mode, ref_rho=pr["p0"] / (pr["R"] * pr["T0"]), ref_T=pr["T0"])
    lat.load_case(fl)
    lat.Init()
    return lat


def timed(lat, steps, warmup):
    """Warm-up, then exactly `steps` steps between two device synchronisations;
    kernel launches are bracketed by HIP events on the library's stream."""
    lat.step(warmup)
    lat.sync()
    lat.profile(True)
    t0 = time.perf_counter()
    lat.step(steps)
    lat.sync()
    t1 = time.perf_counter()
    kernels = {}
    for name in ("fused_collide_stream", "fused_moments", "collide_explicit"):
        ms, n = lat.profile_read(name)
        if n:
            kernels[name] = {"avg_ms": ms / n, "launches": n, "total_ms": ms}
    lat.profile(False)
    return (t1 - t0), kernels


def cpu_baseline(threads):
    """The reference's own single-component step loop (oracle/_ref/ref_timer,
    built from the reference's sources) on a bounded sample: a 128^3 periodic
    box, 20 steps."""
    exe = os.path.join(REPO, "oracle", "_ref", "ref_timer")
    if not os.path.exists(exe):
        return None
    N, steps = 128, 20
    U0 = 0.1 * math.sqrt(1.4 / 3.0)
    nu = U0 * (N / (2 * math.pi)) / 1600.0
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        out = subprocess.run([exe, "time", str(N), str(steps), repr(U0), repr(nu), repr(1 / 3), "1.4", "0.71", "1.0",
                              repr(1 / 3)], env=env, capture_output=True, text=True, timeout=300, check=True)
        r = json.loads(out.stdout.strip().splitlines()[-1])
    except Exception as e:  # the baseline is reported, never required
        return {"error": str(e)[:200]}
    cpu = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                cpu = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": r["mlups"], "unit": "MLUPS", "cores": r["threads"], "kind": "reference",
            "sample": f"reference single-component step loop (Collide, Streaming, TMS_BC, calculate_moment), "
                      f"{N}^3 periodic box, {steps} steps, {r['seconds']:.1f} s, {r['threads']} OpenMP threads, {cpu}"}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}); the multi-GPU x-slab "
              f"decomposition is a design, not a build", file=sys.stderr)
        sys.exit(2)
    from lbm_amd import cases
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED

    if args.workload == "turb256":
        N = 256
        fl, pr = cases.decaying_turbulence(N, nu=0.005)
        precision, dtype, bpl = LBM_FP32_SHIFTED, "f32", 54 * 4 * 2
        workload = "decaying_turbulence_256^3_fp32_shifted"
    else:
        N = 64
        pr = cases.tgv_params(N)
        fl = cases.taylor_green_3d(**pr)
        precision, dtype, bpl = LBM_FP64, "f64", 54 * 8 * 2
        workload = "taylor_green_64^3_fp64"
    cells = N ** 3

    lat = make_lattice(fl, pr, precision, LBM_MODE_FUSED)
    sec, kern = timed(lat, args.steps, args.warmup)
    mlups = cells * args.steps / sec / 1e6
    ms_step = sec / args.steps * 1e3

    k = kern.get("fused_collide_stream")
    roofline = None
    if k:
        achieved = bpl * cells / (k["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_fused_collide (pull-stream + collide)", "bytes_per_launch": bpl * cells,
                    "avg_ms": round(k["avg_ms"], 4)}
    extra = {"kernels": {n: {"avg_ms": round(v["avg_ms"], 4), "launches": v["launches"]} for n, v in kern.items()},
             "step_effective_GBs": round(bpl * cells * args.steps / sec / 1e9, 1)}

    if args.workload == "turb256" and not args.no_compare:
        ref = make_lattice(fl, pr, LBM_FP64, LBM_MODE_FUSED)
        sec64, kern64 = timed(ref, args.steps, args.warmup)
        mlups64 = cells * args.steps / sec64 / 1e6
        extra["fp64_mlups"] = round(mlups64, 1)
        extra["speedup_fp32_vs_fp64"] = round(mlups / mlups64, 3)
        errs = {}
        for name in ("rho", "u", "T"):
            a = lat.get_field(name)[1:-1, 1:-1, 1:-1]
            b = ref.get_field(name)[1:-1, 1:-1, 1:-1]
            errs[name] = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        da, db = lat.diagnostics(), ref.diagnostics()
        errs["kinetic_rel"] = abs(da["kinetic"] - db["kinetic"]) / abs(db["kinetic"])
        errs["enstrophy_rel"] = abs(da["enstrophy"] - db["enstrophy"]) / abs(db["enstrophy"])
        extra["fp32_vs_fp64_after_steps"] = {k2: float(f"{v:.3e}") for k2, v in errs.items()}
        extra["fp64_kernels"] = {n: {"avg_ms": round(v["avg_ms"], 4)} for n, v in kern64.items()}
        ref.close()
    diag = lat.diagnostics()
    extra["nan_count"] = diag["nan_count"]
    lat.close()

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_threads)
    rec = {
        "metric": "MLUPS", "value": round(mlups, 1), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": workload, "cells": cells, "grid": [N, N, N], "mode": "fused", "parallelism": "none"},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra, "host": platform.node(),
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
