"""Initial-condition writer: data/snap0000.h5 for a problem directory, to run
with `spooky --input-dir DIR -r 0`.  It takes the place of the reference's h5py
scripts (problems/*/create_initial_conditions.py, tests/*/check-*.py
createICs): grid, box and fields come from the problem's spooky.cfg as the
driver resolves it (spooky --print-params).

    python spooky-git_amd/write_ic.py --input-dir DIR [--output-dir DIR] IC [options]

IC:
  zero           every field 0 (plus --b0 along x for the magnetic field)
  tg-yz          v = (0, sin(2 pi y/ly) cos(2 pi z/lz), -cos(2 pi y/ly) sin(2 pi z/lz))
                 (tests/taylor-green-vortex)
  tg3d           v = (sin x cos y cos z, -cos x sin y cos z, 0) with x scaled by 2 pi/L
  mhd-random     random-phase solenoidal u and B, E(k) ~ k^4 exp(-2 (k/k0)^2), rms 1 (--seed, --k0)
  mti-noise      u = 0, B = (b0, 0, 0), theta white noise uniform in [-A, A] (--seed, --amplitude, --b0)
  hbi-noise      u_x, u_y random amplitude A U(-1/2, 1/2) sqrt(27/8 N) and random phase at every
                 scale, B = (0, b0, 0), theta = 0 (problems/hbi2d; --seed, --amplitude, --b0)
"""
import argparse
import importlib
import json
import os
import subprocess
import sys

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(PKG))
snapshot = importlib.import_module("spooky-git_amd.snapshot")
initial = importlib.import_module("spooky-git_amd.initial")
SPOOKY = os.path.join(PKG, "bin", "spooky")


def resolve(input_dir, output_dir=None):
    cmd = [SPOOKY, "--input-dir", input_dir, "--print-params"]
    if output_dir:
        cmd += ["--output-dir", output_dir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("spooky --print-params failed:\n" + r.stdout + r.stderr)
    return json.loads(r.stdout.strip().splitlines()[-1])


def real_fields(p, ic, seed=0, amplitude=1e-4, b0=0.0, k0=4.0):
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    lx, ly, lz = p["lx"], p["ly"], p["lz"]
    names = p["fields"]
    out = {n: np.zeros((nx, ny, nz)) for n in names}
    X, Y, Z = initial.grid_coords(nx, ny, nz, lx, ly, lz)
    if ic == "zero":
        pass
    elif ic == "tg-yz":
        out["vy"] = np.sin(2 * np.pi * Y / ly) * np.cos(2 * np.pi * Z / lz)
        out["vz"] = -np.cos(2 * np.pi * Y / ly) * np.sin(2 * np.pi * Z / lz)
    elif ic == "tg3d":
        kx, ky, kz = 2 * np.pi / lx, 2 * np.pi / ly, 2 * np.pi / lz
        out["vx"] = np.sin(kx * X) * np.cos(ky * Y) * np.cos(kz * Z)
        out["vy"] = -np.cos(kx * X) * np.sin(ky * Y) * np.cos(kz * Z)
    elif ic == "mhd-random":
        c = initial.mhd_random((nx, ny, nz), (lx, ly, lz), seed=seed, k0=k0)
        for i, n in enumerate(("vx", "vy", "vz", "bx", "by", "bz")):
            if n in out:
                out[n] = np.fft.irfftn(c[i], s=(nx, ny, nz))
    elif ic == "mti-noise":
        c = initial.mti_white_noise((nx, ny, nz), (lx, ly, lz), seed=seed, amp=amplitude, b0=b0)
        out["th"] = np.fft.irfftn(c[6], s=(nx, ny, nz))
    elif ic == "hbi-noise":
        c = initial.hbi_noise((nx, ny, nz), (lx, ly, lz), seed=seed, amp=amplitude, b0=b0)
        for i, n in enumerate(("vx", "vy", "vz", "bx", "by", "bz", "th")):
            if n in out:
                out[n] = np.fft.irfftn(c[i], s=(nx, ny, nz), axes=(0, 1, 2))
    else:
        raise ValueError("unknown initial condition " + ic)
    if "bx" in out and ic in ("zero", "tg-yz", "tg3d"):
        out["bx"] = out["bx"] + b0
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("ic", choices=["zero", "tg-yz", "tg3d", "mhd-random", "mti-noise", "hbi-noise"])
    ap.add_argument("--input-dir", default="./")
    ap.add_argument("--output-dir", default=None)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--amplitude", type=float, default=1e-4)
    ap.add_argument("--b0", type=float, default=0.0)
    ap.add_argument("--k0", type=float, default=4.0)
    a = ap.parse_args(argv)
    p = resolve(a.input_dir, a.output_dir)
    data = os.path.join(p["output_dir"], "data")
    os.makedirs(data, exist_ok=True)
    f = real_fields(p, a.ic, a.seed, a.amplitude, a.b0, a.k0)
    path = snapshot.snapshot_path(data, 0)
    snapshot.write_snapshot(path, f, t_start=p["t_initial"], t_save=p["t_initial"])
    print("wrote", path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
