"""Writes the golden fixtures of this directory from the reference's own code.

Runs oracle/_ref/ref_driver (the reference's single-component sources
compiled unmodified by oracle/build_ref.sh, driven through its public LBM
class) and stores its outputs as small .npz fixtures.  Only data is stored:
inputs and the reference's outputs.

    python tests/golden/make_golden.py

Fixtures
  fn_vectors.npz     per-call vectors: calculate_feq, calculate_geq (both
                     overloads), calculate_gstr at seeded random inputs
  tgv6.npz           6^3 periodic Taylor-Green box: interior moments and
                     populations after Init (step 0) and steps 1, 2, 10, 40,
                     100; post-collision populations of step 1
  tgv16_series.npz   16^3 box: KE, enstrophy, mass, energy for 200 steps and
                     the moment fields at steps 30 and 50 (the reference's
                     TGV turns unstable near step 40 and is NaN by step 100)
  tgv64_series.npz   64^3 box (the 64^3 Re=1600 Ma=0.1 configuration): KE,
                     enstrophy, mass, energy every step up to 140 (the
                     reference diverges to NaN in this window)
  channel.npz        10x5x3 channel with a fixed-state inflow ghost (TYPE_I),
                     an outflow ghost (TYPE_O) and isothermal walls (TYPE_S,
                     one moving), periodic in z: interior moments after
                     steps 0, 1, 2, 10, 40, 100; populations after steps 1,
                     2, 10; and the state of step 2 before TMS_BC
  channel_io.npz     6x3x3 channel after 2 steps, written by the reference's
                     OutputVTK (field000002.vtr) and write_restart
                     (restart000002.dat): the two files' bytes
  conjugate.npz      6x10x3 periodic box with two 2-cell solid slabs (hot
                     below, cold above) and gas between them, from the
                     reference built with -DCONJUGATE (ref_driver_conj):
                     interior moments and populations after steps 1, 5, 60,
                     and the post-collision populations of step 1
"""
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver")
DRIVER_CONJ = os.path.join(REPO, "oracle", "_ref", "ref_driver_conj")

# 3-D TGV parameters of the 64^3 configuration: R = 1/3, T0 = 1, p0 = 1/3,
# gamma = 1.4, Pr = 0.71, Ma = 0.1 -> U0 = 0.1 sqrt(gamma R T0), Re = 1600 on
# L = N / (2 pi).
R_GAS, T0, P0, GAMMA, PR = 1.0 / 3.0, 1.0, 1.0 / 3.0, 1.4, 0.71
U0 = 0.1 * math.sqrt(GAMMA * R_GAS * T0)


def tgv_params(N, Re=1600.0):
    nu = U0 * (N / (2.0 * math.pi)) / Re
    return dict(N=N, U0=U0, nu=nu, R=R_GAS, gamma=GAMMA, Pr=PR, T0=T0, p0=P0)


def _pargs(p):
    return [repr(p["U0"]), repr(p["nu"]), repr(p["R"]), repr(p["gamma"]), repr(p["Pr"]), repr(p["T0"]), repr(p["p0"])]


def run_tgv(p, nsteps, snaps, env):
    with tempfile.TemporaryDirectory() as d:
        cmd = [DRIVER, "tgv", str(p["N"]), str(nsteps), ",".join(map(str, snaps)), d] + _pargs(p)
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        series = np.fromfile(os.path.join(d, "series.bin")).reshape(-1, 4)
        out = {}
        n = p["N"] ** 3
        for s in snaps:
            out[s] = np.fromfile(os.path.join(d, f"snap_{s}.bin")).reshape(n, 127)
        return series, out


CHANNEL = dict(nx=10, ny=5, nz=3, U=0.05, nu=0.02, R=R_GAS, gamma=GAMMA, Pr=PR, T0=1.0, p0=1.0 / 3.0,
               Tin=1.1, Tw=0.9, pout=0.999 / 3.0)
CHANNEL_KEYS = ("nx", "ny", "nz", "U", "nu", "R", "gamma", "Pr", "T0", "p0", "Tin", "Tw", "pout")


CONJ = dict(nx=6, ny=10, nz=3, ns=2, U=0.03, nu=0.02, R=R_GAS, gamma=GAMMA, Pr=PR, T0=1.0, p0=1.0 / 3.0,
            Thot=1.2, Tcold=0.9)
CONJ_KEYS = ("nx", "ny", "nz", "ns", "U", "nu", "R", "gamma", "Pr", "T0", "p0", "Thot", "Tcold")


def run_conj(p, nsteps, snaps, env):
    with tempfile.TemporaryDirectory() as d:
        cmd = [DRIVER_CONJ, "conj", str(p["nx"]), str(p["ny"]), str(p["nz"]), str(p["ns"]), str(nsteps),
               ",".join(map(str, snaps)), d] + [repr(p[k]) for k in CONJ_KEYS[4:]]
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        n = p["nx"] * p["ny"] * p["nz"]
        return {s: np.fromfile(os.path.join(d, ("post_%d" % -s if s < 0 else "snap_%d" % s) + ".bin")).reshape(n, 127)
                for s in snaps}


def run_channel(p, nsteps, snaps, env):
    with tempfile.TemporaryDirectory() as d:
        cmd = [DRIVER, "channel", str(p["nx"]), str(p["ny"]), str(p["nz"]), str(nsteps),
               ",".join(map(str, snaps)), d] + [repr(p[k]) for k in CHANNEL_KEYS[3:]]
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        n = p["nx"] * p["ny"] * p["nz"]
        return {s: np.fromfile(os.path.join(d, ("pre_%d" % -s if s < 0 else "snap_%d" % s) + ".bin")).reshape(n, 127)
                for s in snaps}


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: oracle/build_ref.sh")
    env = dict(os.environ, OMP_NUM_THREADS="8")
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "fn.bin")
        subprocess.run([DRIVER, "fn", "20261202", "540", fn], check=True, env=env)
        rec = np.fromfile(fn).reshape(-1, 30)
    np.savez_compressed(os.path.join(HERE, "fn_vectors.npz"), rec=rec,
                        columns=np.array("l rho vx vy vz theta cx cy cz U dqx dqy dqz rhoe qx qy qz "
                                         "R00 R01 R02 R10 R11 R12 R20 R21 R22 feq geq gstr geq_moment"))

    p6 = tgv_params(6)
    snaps = [0, 1, 2, 10, 40, 100]
    series, sn = run_tgv(p6, 100, snaps, env)
    arrays = {"series": series, "params": np.array([p6[k] for k in ("N", "U0", "nu", "R", "gamma", "Pr", "T0", "p0")])}
    for s in snaps:
        # columns: rho u v w p T rhoe | ptens 9 | eflux 3 | f 27 | g 27 | fpc 27 | gpc 27
        arrays[f"mom_{s}"] = sn[s][:, :7]
        arrays[f"f_{s}"] = sn[s][:, 19:46]
        arrays[f"g_{s}"] = sn[s][:, 46:73]
    arrays["ptens_1"] = sn[1][:, 7:16]
    arrays["fpc_1"] = sn[1][:, 73:100]
    arrays["gpc_1"] = sn[1][:, 100:127]
    np.savez_compressed(os.path.join(HERE, "tgv6.npz"), **arrays)

    p16 = tgv_params(16)
    series, sn = run_tgv(p16, 200, [30, 50], env)
    np.savez_compressed(os.path.join(HERE, "tgv16_series.npz"), series=series, mom_30=sn[30][:, :7],
                        mom_50=sn[50][:, :7],
                        params=np.array([p16[k] for k in ("N", "U0", "nu", "R", "gamma", "Pr", "T0", "p0")]))

    p64 = tgv_params(64)
    series, _ = run_tgv(p64, 140, [], env)
    np.savez_compressed(os.path.join(HERE, "tgv64_series.npz"), series=series,
                        params=np.array([p64[k] for k in ("N", "U0", "nu", "R", "gamma", "Pr", "T0", "p0")]))
    snaps = [0, 1, -2, 2, 10, 40, 100]
    sn = run_channel(CHANNEL, 100, snaps, env)
    arrays = {"params": np.array([CHANNEL[k] for k in CHANNEL_KEYS], np.float64)}
    for s in snaps:
        key = "pre2" if s < 0 else str(s)
        arrays[f"mom_{key}"] = sn[s][:, :7]
        if s in (1, -2, 2, 10):
            arrays[f"f_{key}"] = sn[s][:, 19:46]
            arrays[f"g_{key}"] = sn[s][:, 46:73]
    np.savez_compressed(os.path.join(HERE, "channel.npz"), **arrays)
    io = dict(CHANNEL, nx=6, ny=3, nz=3)
    with tempfile.TemporaryDirectory() as d:
        cmd = [DRIVER, "io", str(io["nx"]), str(io["ny"]), str(io["nz"]), "2", d] + [repr(io[k]) for k in CHANNEL_KEYS[3:]]
        subprocess.run(cmd, check=True, env=env, stdout=subprocess.DEVNULL)
        vtk = np.frombuffer(open(os.path.join(d, "field000002.vtr"), "rb").read(), np.uint8)
        rst = np.frombuffer(open(os.path.join(d, "restart000002.dat"), "rb").read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "channel_io.npz"), vtk=vtk, restart=rst, steps=np.array(2),
                        params=np.array([io[k] for k in CHANNEL_KEYS], np.float64))
    sn = run_conj(CONJ, 60, [-1, 1, 5, 60], env)
    arrays = {"params": np.array([CONJ[k] for k in CONJ_KEYS], np.float64)}
    for s in (1, 5, 60):
        arrays[f"mom_{s}"] = sn[s][:, :7]
        arrays[f"f_{s}"] = sn[s][:, 19:46]
        arrays[f"g_{s}"] = sn[s][:, 46:73]
    arrays["fpc_post1"] = sn[-1][:, 73:100]
    arrays["gpc_post1"] = sn[-1][:, 100:127]
    np.savez_compressed(os.path.join(HERE, "conjugate.npz"), **arrays)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
