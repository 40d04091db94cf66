"""Throughput benchmark of the D3Q27 hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-cpu-baseline] [--no-compare]

Prints ONE JSON line.  A step is one pass of the loop body (Collide_Species,
Collide, Streaming, TMS_BC, calculate_moment) over the whole synthetic box;
the timed region starts with the state already resident in HBM.

Workloads
  mixing256 reacting H2-air / hot-air temporal mixing layer, 256x128x128,
            9 species, periodic x/z, NSCBC outflow on both y faces.  `value`
            is the FP32 shifted-population path with the learned chemistry
            surrogate (trained on the GPU, inside this run, on states the
            plain integrator produced during the warm-up; training time is
            reported, not timed); the FP64 plain path is timed beside it for
            the speedup, and both runs' fields are compared after the same
            number of steps (flame temperature, T and species errors).
  turb256   thermal compressible decaying turbulence, 256^3 periodic, single
            component: the FP32 shifted-population path is `value`; the FP64
            plain path is timed beside it for the speedup and the FP32 error
            against FP64 after the timed steps.
  tgv64     3-D Taylor-Green vortex, 64^3, FP64 plain path.
  mixing_c5share  the same mixing layer on one GPU's share of the 768x384x384
            layer split into 8 x-slabs: 96x384x384 (14.2 M cells), periodic
            in x in place of the slab exchange; FP64 on the same share beside it.
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
# surrogate maintenance in the multicomponent runs: refresh after a window of
# steps in which more than REFRESH_RATE of the chemistry updates went to the
# plain integrator
REFRESH_WINDOW, REFRESH_RATE, REFRESH_ITERS = 10, 0.003, 100


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="mixing256",
                    choices=["mixing256", "mixing768", "mixing_c5share", "turb256", "turb256_report", "tgv64",
                             "copy1g"])
    ap.add_argument("--report-every", type=int, default=250, help="turb256_report: steps between error samples")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 companion run")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--refresh-window", type=int, default=REFRESH_WINDOW,
                    help="surrogate maintenance: steps per fallback-rate window")
    ap.add_argument("--refresh-rate", type=float, default=REFRESH_RATE,
                    help="surrogate maintenance: fallback + audit-failure share of a window that triggers a refresh")
    ap.add_argument("--refresh-iters", type=int, default=REFRESH_ITERS,
                    help="surrogate maintenance: training iterations per refresh")
    ap.add_argument("--no-train-graph", action="store_true",
                    help="launch the surrogate's training iterations one by one instead of replaying a HIP "
                         "graph (for rocprofv3 kernel tracing, which faults inside the graph calls)")
    ap.add_argument("--mode", default="fused", choices=["fused", "strict", "inplace"],
                    help="multicomponent workloads: the fused mode (Collide_Species writes the streamed slots; "
                         "bitwise equal to strict), one kernel per reference call, or the in-place mode (one "
                         "species buffer streamed by the AA pattern; bitwise equal to strict)")
    ap.add_argument("--chemistry", default="surrogate", choices=["surrogate", "plain"],
                    help="mixing256: learned chemistry (default) or the plain integrator in the FP32 run "
                         "(no capture or training; used for the PMC byte-counter passes)")
    return ap.parse_args()


def pmc_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 --pmc summaries (profiles/pmc_traffic_*.md, FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy's readings), or None when no summary
    covers it.  Each summary carries one line 'machine-readable: {json}'
    keyed by workload and kernel label."""
    import glob
    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "pmc_traffic_*.md"))):
        with open(path) as fh:
            for line in fh:
                if line.startswith("machine-readable:"):
                    try:
                        rec = json.loads(line.split(":", 1)[1])
                    except ValueError:
                        continue
                    hit = rec.get(workload, {}).get(kernel)
                    if hit:
                        return int(hit["traffic_bytes"])
    return None


def make_lattice(fl, pr, precision, mode):
    from lbm_amd.capi import Lattice
    N = pr["N"]
    lat = Lattice(N, N, N, pr["nu"], gas_const=pr["R"], gamma=pr["gamma"], prtl=pr["Pr"], precision=precision,
                  mode=mode, ref_rho=pr["p0"] / (pr["R"] * pr["T0"]), ref_T=pr["T0"])
    lat.load_case(fl)
    lat.Init()
    return lat


SINGLE_KERNELS = ("fused_collide_stream", "fused_moments", "collide_explicit")
MULTI_KERNELS = ("mc_chemistry_plain", "mc_chemistry_surrogate", "mc_chemistry_fallback", "mc_species_transport",
                 "mc_aa_save", "mc_collide_species", "mc_collide_species_push", "mc_collide_species_inplace",
                 "mc_collide_mixture", "mc_stream", "mc_stream_energy", "mc_tms_outflow", "mc_tms_inflow",
                 "mc_tms_wall", "mc_moments")


def timed(lat, steps, warmup, names=SINGLE_KERNELS, run=None):
    """Warm-up, then exactly `steps` steps between two device synchronisations;
    kernel launches are bracketed by HIP events on the library's stream.
    `run(steps)` replaces lat.step for the timed steps (the surrogate's
    maintained stepping)."""
    lat.step(warmup)
    lat.sync()
    lat.profile(True)
    t0 = time.perf_counter()
    (run or lat.step)(steps)
    lat.sync()
    t1 = time.perf_counter()
    kernels = {}
    for name in names:
        ms, n = lat.profile_read(name)
        if n:
            kernels[name] = {"avg_ms": ms / n, "launches": n, "total_ms": ms}
    # profile_read matches names as substrings: drop a name whose launches
    # are all those of a longer name (mc_stream of mc_stream_energy)
    for name in list(kernels):
        if any(o != name and name in o and kernels[o]["launches"] == kernels[name]["launches"]
               and kernels[o]["total_ms"] == kernels[name]["total_ms"] for o in kernels):
            del kernels[name]
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


# Algorithmic HBM bytes per cell of the multicomponent kernels: the
# populations each one reads and writes (9 species x 27 f, 27 g), per byte
# of storage.  Moment and property planes are excluded.
MC_POP_TRAFFIC = {"mc_collide_species": 2 * 9 * 27, "mc_collide_species_push": 2 * 9 * 27,
                  "mc_collide_species_inplace": 2 * 9 * 27,
                  "mc_collide_mixture": 2 * 27, "mc_stream": 2 * 10 * 27, "mc_stream_energy": 2 * 27,
                  "mc_moments": 10 * 27}


def cpu_info():
    """CPU model and the logical CPUs this process may run on."""
    model = ""
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, allowed


def cpu_baseline_mixing(threads):
    """The compiled C++/OpenMP restatement of the multicomponent step
    (oracle/mc_cpu.cpp: FP64, plain chemistry, the same thermo, kinetics and
    transport tables; pinned against the numpy oracle) on a bounded sample of
    the same workload: 64x128x128 cells (a quarter of the 256x128x128 layer)
    of its periodic twin cases.double_mixing_layer (the same streams and
    profiles, the fuel stream between two layers, so that the box is
    periodic: the restatement streams periodic boxes only; the GPU's NSCBC
    outflow pass is 3.6 % of its step), 2 steps.  The reference itself needs
    Cantera, which is absent."""
    import tempfile
    from lbm_amd import cases
    from oracle import mc_cpu
    nx, ny, nz, steps = 64, 128, 128, 2
    model, allowed = cpu_info()
    threads = max(1, min(threads, allowed))
    fl = cases.double_mixing_layer(nx, ny, nz, T_fuel=300.0, T_ox=1200.0, U_s=0.02)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "mixing_slab.bin")
        mc_cpu.write_case(path, fl, *cases.UNITS_FLAME, True)
        r = mc_cpu.run(path, steps, threads, timeout=600)
    return {"value": round(r["MLUPS"], 6), "unit": "MLUPS", "cores": r["threads"], "kind": "port",
            "sample": f"compiled C++/OpenMP multicomponent step (oracle/mc_cpu.cpp, FP64, plain chemistry), "
                      f"{nx}x{ny}x{nz} periodic twin of the reacting mixing layer, {steps} steps, "
                      f"{r['seconds']:.1f} s, {r['threads']} OpenMP threads of {allowed} allowed, {model}"}


def mixing_layer_bench(args, nx=256, ny=128, nz=128, tag="reacting_mixing_layer_256x128x128"):
    """The reacting mixing layer: FP64 plain beside FP32 + surrogate."""
    from lbm_amd import cases
    from lbm_amd.capi import (LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED, LBM_MODE_INPLACE, LBM_MODE_STRICT,
                              MixtureLattice)
    cells = nx * ny * nz
    mode = {"fused": LBM_MODE_FUSED, "strict": LBM_MODE_STRICT, "inplace": LBM_MODE_INPLACE}[args.mode]
    fl = cases.mixing_layer(nx, ny, nz, T_fuel=300.0, T_ox=1200.0, U_s=0.02)
    F = (fl["type"] == 0).ravel()
    cap_steps, cap_stride = 4, 4          # ~4.2 M training states
    train_iters, batch = 2000, 4096
    steps = args.steps

    def fields(lat):
        rho_a = lat.get_species_field("rho").reshape(9, -1)[:, F]
        return lat.get_field("T").ravel()[F], rho_a / rho_a.sum(axis=0, keepdims=True)

    # FP32 shifted populations + learned chemistry
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP32_SHIFTED, mode=mode)
    lat.load_case(fl)
    lat.Init()
    lat.calculate_moment()
    surrogate = args.chemistry == "surrogate"
    if surrogate:
        lat.surrogate_capture(cap_steps * cells // cap_stride + 4096, cap_stride)
        lat.step(cap_steps)
        lat.surrogate_capture_stop()
        lat.sync()
        t0 = time.perf_counter()
        losses = np.concatenate([lat.surrogate_train(train_iters * 3 // 4, batch=batch, lr=2e-3, seed=1),
                                 lat.surrogate_train(train_iters // 4, batch=batch, lr=3e-4, seed=2)])
        train_s = time.perf_counter() - t0
        lat.surrogate_enable(True, audit_stride=64, audit_tol=0.5)
        # maintained surrogate: the states the network cannot take (and the
        # audited ones) join the training set, and a window of REFRESH_WINDOW
        # steps with more than REFRESH_RATE of the updates falling back or
        # failing their audit triggers a refresh (envelope widened to every
        # captured state, REFRESH_ITERS more training iterations) on a second
        # stream beside the steps, inside the timed region
        lat.surrogate_capture(cap_steps * cells // cap_stride + 2 * cells, 1 << 30)
        maint = {"step_s": 0.0, "refresh_s": 0.0, "refreshes": 0}

        def run(k):
            st, rs, nr = lat.step_with_refresh(k, window=args.refresh_window, max_rate=args.refresh_rate,
                                               iters=args.refresh_iters, batch=batch, overlap=True)
            maint["step_s"] += st
            maint["refresh_s"] += rs
            maint["refreshes"] += nr
    else:
        lat.step(cap_steps)
        run = None
    sec, kern = timed(lat, steps, args.warmup, MULTI_KERNELS, run=run if surrogate else None)
    T32, Y32 = fields(lat)
    sstats = lat.surrogate_stats() if surrogate else None
    diag = lat.diagnostics()
    dev_bytes = lat.device_bytes()
    lat.close()
    mlups = cells * steps / sec / 1e6

    extra = {"kernels": {n: {"avg_ms": round(v["avg_ms"], 4), "launches": v["launches"]} for n, v in kern.items()},
             "device_bytes_fp32": dev_bytes, "nan_count": diag["nan_count"]}
    if surrogate:
        extra["surrogate"] = {"train_s": round(train_s, 2), "train_iters": train_iters, "batch": batch,
                              "loss_first": float(losses[:10].mean()), "loss_last": float(losses[-50:].mean()),
                              "captured_states": None, **sstats}
        extra["surrogate"]["captured_states"] = extra["surrogate"].pop("captured")
        extra["surrogate"]["maintenance"] = {"window": args.refresh_window, "max_rate": args.refresh_rate,
                                             "iters": args.refresh_iters, "overlapped": True,
                                             "refreshes": maint["refreshes"],
                                             "refresh_s": round(maint["refresh_s"], 3),
                                             "step_s": round(maint["step_s"], 3)}
        extra["fallback_rate"] = round((sstats["fallback"] + sstats["audit_failures"]) /
                                       max(sstats["predicted"] + sstats["fallback"], 1), 6)

    # dominant kernel's roofline (FP32 storage)
    roofline = None
    dom = max((n for n in kern if n in MC_POP_TRAFFIC), key=lambda n: kern[n]["avg_ms"], default=None)
    if dom:
        bpl = MC_POP_TRAFFIC[dom] * 4
        achieved = bpl * cells / (kern[dom]["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload, dom),
                    "kernel": dom, "bytes_per_launch": bpl * cells, "avg_ms": round(kern[dom]["avg_ms"], 4)}

    if not args.no_compare:
        def fp64_run(m):
            ref = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP64, mode=m)
            ref.load_case(fl)
            ref.Init()
            ref.calculate_moment()
            sec64, kern64 = timed(ref, steps, args.warmup + cap_steps, MULTI_KERNELS)
            T64, Y64 = fields(ref)
            ref.close()
            return cells * steps / sec64 / 1e6, kern64, T64, Y64

        # FP64 plain chemistry in the same mode; in the fused mode also the
        # strict calls (the two modes give bitwise equal states)
        mlups64, kern64, T64, Y64 = fp64_run(mode)
        extra["fp64_mlups"] = round(mlups64, 2)
        extra["speedup_fp32_surrogate_vs_fp64"] = round(mlups / mlups64, 3)
        extra["fp64_kernels"] = {n: {"avg_ms": round(v["avg_ms"], 4)} for n, v in kern64.items()}
        if mode != LBM_MODE_STRICT:
            mlups64s, kern64s, _, _ = fp64_run(LBM_MODE_STRICT)
            extra["fp64_strict_mlups"] = round(mlups64s, 2)
            extra["speedup_fp32_surrogate_vs_fp64_strict"] = round(mlups / mlups64s, 3)
        extra["error_vs_fp64"] = {
            "flame_T_rel": float(f"{abs(T32.max() - T64.max()) / T64.max():.3e}"),
            "T_max_rel": float(f"{np.max(np.abs(T32 - T64) / T64):.3e}"),
            "Y_L2": float(f"{np.sqrt(np.mean((Y32 - Y64) ** 2)):.3e}"),
            "Y_max_abs": float(f"{np.max(np.abs(Y32 - Y64)):.3e}"),
        }
    rec = {
        "metric": "MLUPS", "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(sec / steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": tag + "_9sp_fp32_shifted_" + args.chemistry, "cells": cells,
                   "grid": [nx, ny, nz], "mode": args.mode, "parallelism": "none"},
        "roofline": roofline, "cpu_baseline": None, "extra": extra, "host": platform.node(),
    }
    return rec


def block_mean(a, nx, ny, nz):
    """2 x 2 x 2 block means of the interior of a ghost-inclusive field."""
    a = a[..., 1:-1, 1:-1, 1:-1]
    lead = a.shape[:-3]
    return a.reshape(lead + (nx // 2, 2, ny // 2, 2, nz // 2, 2)).mean(axis=(-5, -3, -1))


def mixing_c5_bench(args):
    """The combined large variant: the reacting mixing layer on 768x384x384
    (113 M cells) in the in-place mode (one species buffer), FP32 shifted
    populations and the learned chemistry, on one GPU.  Then the FP64 plain
    run of the 384x192x192 twin (same physical set-up: dx and dt doubled, the
    tanh thickness and the speeds in lattice units unchanged) for half the
    steps, and the fine run's fields averaged over 2x2x2 blocks against it:
    the error of resolution and precision together."""
    from lbm_amd import cases
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_INPLACE, MixtureLattice
    nx, ny, nz = 768, 384, 384
    cells = nx * ny * nz
    steps = args.steps

    def log(msg):
        print(f"mixing768: {msg}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    fl = cases.mixing_layer(nx, ny, nz, T_fuel=300.0, T_ox=1200.0, U_s=0.02)
    log(f"set-up built in {time.perf_counter() - t0:.0f} s")
    lat = MixtureLattice(nx, ny, nz, *cases.UNITS_FLAME, reaction=True, precision=LBM_FP32_SHIFTED,
                         mode=LBM_MODE_INPLACE)
    lat.load_case(fl)
    del fl
    lat.Init()
    lat.calculate_moment()
    log(f"initialised, {lat.device_bytes() / 1e9:.1f} GB on the device")
    cap_steps = 4
    cap_stride = max(4, cap_steps * cells // 4_200_000)   # ~4.2 M training states
    train_iters, batch = 2000, 4096
    lat.surrogate_capture(cap_steps * cells // cap_stride + 65536, cap_stride)
    lat.step(cap_steps)
    lat.surrogate_capture_stop()
    lat.sync()
    t1 = time.perf_counter()
    lat.surrogate_train(train_iters * 3 // 4, batch=batch, lr=2e-3, seed=1)
    lat.surrogate_train(train_iters // 4, batch=batch, lr=3e-4, seed=2)
    train_s = time.perf_counter() - t1
    lat.surrogate_enable(True, audit_stride=64, audit_tol=0.5)
    log(f"surrogate trained in {train_s:.1f} s; timing {steps} steps")
    # maintained surrogate as in mixing256 (fallback and audited states join a
    # training set of up to 16 M rows; a window of bad updates above
    # REFRESH_RATE refreshes the network beside the steps), run in chunks of
    # 100 steps with a progress line after each
    lat.surrogate_capture(cap_steps * cells // cap_stride + 16_000_000, 1 << 30)
    maint = {"step_s": 0.0, "refresh_s": 0.0, "refreshes": 0}

    def run(k):
        done = 0
        while done < k:
            m = min(100, k - done)
            st, rs, nr = lat.step_with_refresh(m, window=args.refresh_window, max_rate=args.refresh_rate,
                                               iters=args.refresh_iters, batch=batch, overlap=True)
            maint["step_s"] += st
            maint["refresh_s"] += rs
            maint["refreshes"] += nr
            done += m
            if k > 100:
                log(f"step {done} of {k}, {maint['refreshes']} refreshes")

    sec, kern = timed(lat, steps, args.warmup, MULTI_KERNELS, run=run)
    mlups = cells * steps / sec / 1e6
    log(f"{mlups:.1f} MLUPS")
    sstats = lat.surrogate_stats()
    diag = lat.diagnostics()
    dev_bytes = lat.device_bytes()
    T32 = lat.get_field("T")
    rho32 = lat.get_field("rho")
    ra32 = lat.get_species_field("rho")
    lat.close()
    fine_steps = cap_steps + args.warmup + steps
    T_f = block_mean(T32, nx, ny, nz)
    Y_f = block_mean(ra32 / rho32[None], nx, ny, nz)
    del T32, rho32, ra32

    tx, ty, tz = nx // 2, ny // 2, nz // 2
    twin_units = (2 * cases.UNITS_FLAME[0], 2 * cases.UNITS_FLAME[1])
    ft = cases.mixing_layer(tx, ty, tz, T_fuel=300.0, T_ox=1200.0, U_s=0.02)  # the twin's own noise draw
    ref = MixtureLattice(tx, ty, tz, *twin_units, reaction=True, precision=LBM_FP64, mode=LBM_MODE_INPLACE)
    ref.load_case(ft)
    ref.Init()
    ref.calculate_moment()
    twin_steps = fine_steps // 2
    t2 = time.perf_counter()
    ref.step(twin_steps)
    ref.sync()
    twin_s = time.perf_counter() - t2
    T64 = ref.get_field("T")[1:-1, 1:-1, 1:-1]
    Y64 = (ref.get_species_field("rho") / ref.get_field("rho")[None])[:, 1:-1, 1:-1, 1:-1]
    ref.close()
    log(f"twin: {twin_steps} FP64 steps in {twin_s:.1f} s")
    # the FP32 + surrogate path on the twin's own grid and steps: its error
    # against the FP64 twin is that of precision and chemistry alone
    lt = MixtureLattice(tx, ty, tz, *twin_units, reaction=True, precision=LBM_FP32_SHIFTED, mode=LBM_MODE_INPLACE)
    lt.load_case(ft)
    lt.Init()
    lt.calculate_moment()
    tcells = tx * ty * tz
    lt.surrogate_capture(cap_steps * tcells // 4 + 65536, 4)
    lt.step(cap_steps)
    lt.surrogate_capture_stop()
    lt.surrogate_train(train_iters * 3 // 4, batch=batch, lr=2e-3, seed=1)
    lt.surrogate_train(train_iters // 4, batch=batch, lr=3e-4, seed=2)
    lt.surrogate_enable(True, audit_stride=64, audit_tol=0.5)
    lt.step(twin_steps - cap_steps)
    T32t = lt.get_field("T")[1:-1, 1:-1, 1:-1]
    Y32t = (lt.get_species_field("rho") / lt.get_field("rho")[None])[:, 1:-1, 1:-1, 1:-1]
    lt.close()
    del ft

    dom = "mc_collide_species_inplace"
    roofline = None
    if dom in kern:
        bpl = MC_POP_TRAFFIC[dom] * 4
        achieved = bpl * cells / (kern[dom]["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload, dom),
                    "kernel": dom, "bytes_per_launch": bpl * cells, "avg_ms": round(kern[dom]["avg_ms"], 4)}
    rel = np.abs(T_f - T64) / T64
    extra = {"kernels": {n: {"avg_ms": round(v["avg_ms"], 4), "launches": v["launches"]} for n, v in kern.items()},
             "device_bytes_fp32": dev_bytes, "nan_count": diag["nan_count"],
             "surrogate": {"train_s": round(train_s, 2), "capture_stride": cap_stride, **sstats,
                           "maintenance": {"window": args.refresh_window, "max_rate": args.refresh_rate,
                                           "iters": args.refresh_iters, "overlapped": True,
                                           "refreshes": maint["refreshes"],
                                           "refresh_s": round(maint["refresh_s"], 3)}},
             "fallback_rate_over_run": round((sstats["fallback"] + sstats["audit_failures"]) /
                                             max(sstats["predicted"] + sstats["fallback"], 1), 6),
             "twin": {"grid": [tx, ty, tz], "precision": "f64", "chemistry": "plain", "steps": twin_steps,
                      "fine_steps": fine_steps, "units": twin_units, "seconds": round(twin_s, 1)},
             "error_fine_fp32_block_mean_vs_fp64_twin": {"flame_T_rel": float(f"{abs(T_f.max() - T64.max()) / T64.max():.3e}"),
                                    "T_max_rel": float(f"{rel.max():.3e}"),
                                    "T_rms_rel": float(f"{np.sqrt(np.mean(rel * rel)):.3e}"),
                                    "Y_L2": float(f"{np.sqrt(np.mean((Y_f - Y64) ** 2)):.3e}"),
                                    "Y_max_abs": float(f"{np.max(np.abs(Y_f - Y64)):.3e}")},
             "error_fp32_surrogate_vs_fp64_on_twin_grid": {
                 "flame_T_rel": float(f"{abs(T32t.max() - T64.max()) / T64.max():.3e}"),
                 "T_max_rel": float(f"{np.max(np.abs(T32t - T64) / T64):.3e}"),
                 "Y_L2": float(f"{np.sqrt(np.mean((Y32t - Y64) ** 2)):.3e}"),
                 "Y_max_abs": float(f"{np.max(np.abs(Y32t - Y64)):.3e}")}}
    return {
        "metric": "MLUPS", "value": round(mlups, 2), "unit": "MLUPS", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": round(sec / steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "reacting_mixing_layer_768x384x384_9sp_fp32_shifted_surrogate", "cells": cells,
                   "grid": [nx, ny, nz], "mode": "inplace", "parallelism": "none"},
        "roofline": roofline, "cpu_baseline": None, "extra": extra, "host": platform.node(),
    }


def turbulence_report(args):
    """The 256^3 decaying-turbulence precision variant: FP32 shifted
    populations and FP64 plain, both fused, stepped side by side from the same
    set-up.  The configured state (Ma_rms 0.3, nu 0.005) for up to --steps
    steps: every step for the first 20, then every --report-every steps, the
    kinetic energy and enstrophy of each run (output.cpp:212-276 sums), their
    relative difference, and the FP32 temperature field's error against FP64
    (max and RMS relative over the interior).  Once a run holds a non-finite
    cell it stops, and the step is reported for each precision.  The
    reference's model does not survive this set-up (Init gives the mass
    populations zero momentum while the energy populations carry the flow's
    heat flux, lbm-init.cpp:22-60; fd_central treats periodic ghosts as
    walls), so the same record is taken, every 50 steps, for lower Mach
    numbers and a larger viscosity, to show where each turns non-finite.
    Progress lines go to stderr; one JSON line to stdout."""
    from lbm_amd import cases
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED
    N = 256

    def run(ma, nu, steps, every_of):
        fl, pr = cases.decaying_turbulence(N, nu=nu, ma_rms=ma)
        f32 = make_lattice(fl, pr, LBM_FP32_SHIFTED, LBM_MODE_FUSED)
        f64 = make_lattice(fl, pr, LBM_FP64, LBM_MODE_FUSED)
        rows, done = [], 0
        diverged = {"fp32": None, "fp64": None}
        t0 = time.perf_counter()
        while done < steps and not (diverged["fp32"] and diverged["fp64"]):
            k = min(every_of(done), steps - done)
            for name, lat in (("fp32", f32), ("fp64", f64)):
                if not diverged[name]:
                    lat.step(k)
            done += k
            a, b = f32.diagnostics(), f64.diagnostics()
            for name, d in (("fp32", a), ("fp64", b)):
                if not diverged[name] and (d["nan_count"] or not np.isfinite(d["kinetic"])):
                    diverged[name] = done
            row = {"step": done, "ke_fp32": a["kinetic"], "ke_fp64": b["kinetic"], "ens_fp32": a["enstrophy"],
                   "ens_fp64": b["enstrophy"], "max_mach_fp64": b["max_mach"]}
            if not (diverged["fp32"] or diverged["fp64"]):
                row["ke_rel"] = abs(a["kinetic"] - b["kinetic"]) / abs(b["kinetic"])
                row["ens_rel"] = abs(a["enstrophy"] - b["enstrophy"]) / abs(b["enstrophy"])
                T32 = f32.get_field("T")[1:-1, 1:-1, 1:-1]
                T64 = f64.get_field("T")[1:-1, 1:-1, 1:-1]
                rel = np.abs(T32 - T64) / np.abs(T64)
                row["T_max_rel"] = float(rel.max())
                row["T_rms_rel"] = float(np.sqrt(np.mean(rel * rel)))
            rows.append(row)
            print(f"turb256 report ma {ma} nu {nu}: step {done}: " +
                  ", ".join(f"{key} {val:.4g}" for key, val in row.items() if key != "step"), file=sys.stderr,
                  flush=True)
        wall = time.perf_counter() - t0
        f32.close()
        f64.close()
        return {"ma_rms": ma, "nu": nu, "params": pr, "steps_requested": steps, "steps_run": done,
                "nonfinite_first_seen_at": diverged, "wall_s": round(wall, 1), "rows": rows}

    stated = run(0.3, 0.005, args.steps, lambda done: 1 if done < 20 else args.report_every)
    variants = [run(ma, nu, args.steps, lambda done: 50)
                for ma, nu in ((0.1, 0.005), (0.1, 0.02), (0.05, 0.005), (0.05, 0.02), (0.02, 0.005), (0.02, 0.02))]
    return {"workload": "decaying_turbulence_256^3_fp32_vs_fp64", "grid": [N, N, N], "configured": stated,
            "variants": variants}


def main():
    args = parse()
    if args.no_train_graph:
        os.environ["LBM_AMD_TRAIN_GRAPH"] = "0"
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}); the multi-GPU x-slab "
              f"decomposition is a design, not a build", file=sys.stderr)
        sys.exit(2)
    from lbm_amd import cases
    from lbm_amd.capi import LBM_FP32_SHIFTED, LBM_FP64, LBM_MODE_FUSED

    if args.workload == "copy1g":
        # calibration for the PMC byte counters: a plain 1 GiB device copy
        import torch
        n = 1 << 28
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        print(json.dumps({"workload": "copy1g", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                          "copies": args.steps, "GBs": 8 * n * args.steps / sec / 1e9}))
        return
    if args.workload == "mixing768":
        print(json.dumps(mixing_c5_bench(args)))
        return
    if args.workload == "turb256_report":
        print(json.dumps(turbulence_report(args)))
        return
    if args.workload == "mixing_c5share":
        print(json.dumps(mixing_layer_bench(args, 96, 384, 384, "reacting_mixing_layer_c5_share_96x384x384")))
        return
    if args.workload == "mixing256":
        rec = mixing_layer_bench(args)
        if not args.no_cpu_baseline:
            try:
                rec["cpu_baseline"] = cpu_baseline_mixing(args.cpu_threads)
            except Exception as e:  # the baseline is reported, never required
                rec["cpu_baseline"] = {"error": str(e)[:200]}
        print(json.dumps(rec))
        return
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
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(args.workload, "fused_collide_stream"),
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
