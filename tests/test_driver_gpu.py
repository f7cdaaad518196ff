"""The reference's end-to-end tests (tests/*/check-*.py with their own
spooky.cfg, kept under tests/golden/cfg/) run through the `spooky` driver on
the GPU exactly as the reference's ctest does: the initial snapshot is written
(here by the snapshot library instead of h5py), the driver runs with
`--input-dir <cfg> --output-dir <tmp> -r 0`, and the snapshots it writes are
read back and compared with the closed form at the reference's tolerance.
Also: a restart reproduces an uninterrupted run, and timevar.spooky carries
the configured columns."""
import importlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import reference_cases as rc
from spkmod import REPO

pytestmark = pytest.mark.gpu
snapshot = importlib.import_module("spooky-git_amd.snapshot")
SPOOKY = os.path.join(REPO, "spooky-git_amd", "bin", "spooky")
CFG = os.path.join(REPO, "tests", "golden", "cfg")


def run(cfgdir, outdir, *extra, timeout=600):
    r = subprocess.run([SPOOKY, "--input-dir", cfgdir, "--output-dir", outdir] + list(extra), capture_output=True,
                       text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    return r.stdout


def write_ic(outdir, fields):
    os.makedirs(os.path.join(outdir, "data"), exist_ok=True)
    snapshot.write_snapshot(snapshot.snapshot_path(os.path.join(outdir, "data"), 0), fields)


def read_all(outdir, names, shape):
    data = os.path.join(outdir, "data")
    out = []
    for i in range(snapshot.last_snapshot(data) + 1):
        hdr, f = snapshot.read_snapshot(snapshot.snapshot_path(data, i), names, shape)
        out.append((hdr, f))
    return out


def test_taylor_green_vortex(tmp_path):
    """tests/taylor-green-vortex: L1 < 1e-7 (check-taylor-green.py:34) at every snapshot."""
    out = str(tmp_path)
    write_ic(out, rc.taylor_green_yz(0.0))
    run(os.path.join(CFG, "taylor-green-vortex"), out, "-r", "0")
    snaps = read_all(out, ["vx", "vy", "vz"], (64, 64, 64))
    assert len(snaps) == 11  # t = 0, 1, ..., 10 (snapshot_step = 1, t_final = 10)
    for hdr, f in snaps:
        err = rc.l1(f, rc.taylor_green_yz(hdr["t_save"]), ["vx", "vy", "vz"])
        assert err < 1e-7, (hdr, err)
    assert snaps[-1][0]["t_save"] == pytest.approx(10.0, abs=1e-12)
    # timevar.spooky: the configured columns, one row per timevar_step; ev is the
    # kinetic energy 1/4 exp(-4 nu (2 pi)^2 t); the flow lies in the (y, z) plane
    # and its curl along x, so the helicity hv is 0
    lines = open(os.path.join(out, "data", "timevar.spooky")).read().splitlines()
    assert lines[1].split() == ["##", "t", "ev", "Kx", "Ky", "Kz", "w2", "hv"]
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[2:]])
    assert rows.shape[0] >= 100 and np.all(np.diff(rows[:, 0]) > 0)
    ev = 0.25 * np.exp(-4 * (1 / 200.0) * rc.TWO_PI ** 2 * rows[:, 0])
    assert np.max(np.abs(rows[:, 1] - ev) / ev) < 1e-7
    assert np.max(np.abs(rows[:, 6])) < 1e-12


def test_taylor_green_vortex_fp32_bulk(tmp_path):
    """The same case with `--fp32-bulk` (transforms and products in FP32, state in
    FP64): L1 < 1e-5 against the closed form at every snapshot."""
    out = str(tmp_path)
    write_ic(out, rc.taylor_green_yz(0.0))
    log = run(os.path.join(CFG, "taylor-green-vortex"), out, "-r", "0", "--fp32-bulk")
    assert "FP32 bulk" in log
    snaps = read_all(out, ["vx", "vy", "vz"], (64, 64, 64))
    assert len(snaps) == 11
    for hdr, f in snaps:
        err = rc.l1(f, rc.taylor_green_yz(hdr["t_save"]), ["vx", "vy", "vz"])
        assert err < 1e-5, (hdr, err)


def test_alfven_wave(tmp_path):
    """tests/alfven-wave: L1 < 1e-7 at t_final = 1 (check-alfven-wave.py:38, one period)."""
    out = str(tmp_path)
    names = ["vx", "vy", "vz", "bx", "by", "bz"]
    write_ic(out, rc.alfven_wave(0.0))
    run(os.path.join(CFG, "alfven-wave"), out, "-r", "0")
    snaps = read_all(out, names, (64, 64, 64))
    hdr, f = snaps[-1]
    assert hdr["t_save"] == pytest.approx(1.0, abs=1e-12)
    assert rc.l1(f, rc.alfven_wave(hdr["t_save"]), names) < 1e-7


def test_mti_eigenmode(tmp_path):
    """tests/mti-eigenmode (512x4x4, anisotropic conduction): L1 < 1e-7
    (check-mti-eigenmode.py:47) at every snapshot up to t_final = 2."""
    out = str(tmp_path)
    names = ["vx", "vy", "vz", "bx", "by", "bz", "th"]
    write_ic(out, rc.mti_eigenmode(0.0))
    run(os.path.join(CFG, "mti-eigenmode"), out, "-r", "0")
    snaps = read_all(out, names, (512, 4, 4))
    assert snaps[-1][0]["t_save"] == pytest.approx(2.0, abs=1e-12)
    for hdr, f in snaps:
        err = rc.l1(f, rc.mti_eigenmode(hdr["t_save"]), names)
        assert err < 1e-7, (hdr, err)


def test_restart_reproduces_uninterrupted_run(tmp_path):
    """Snapshots 0-2 of one run; a second run restarted from snapshot 1 (-r 1)
    writes the same snapshot 2 (ReadDataFile, hdf5_io.cu:110-252)."""
    cfg = tmp_path / "cfg"
    cfg.mkdir()
    text = open(os.path.join(CFG, "taylor-green-vortex", "spooky.cfg")).read()
    text = text.replace("t_final = 10.0;", "t_final = 2.0;")
    assert "t_final = 2.0;" in text
    (cfg / "spooky.cfg").write_text(text)
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    write_ic(a, rc.taylor_green_yz(0.0))
    run(str(cfg), a, "-r", "0")
    os.makedirs(os.path.join(b, "data"))
    for i in (0, 1):
        shutil.copy(snapshot.snapshot_path(os.path.join(a, "data"), i), os.path.join(b, "data"))
    run(str(cfg), b, "-r", "1")
    names, shape = ["vx", "vy", "vz"], (64, 64, 64)
    ha, fa = snapshot.read_snapshot(snapshot.snapshot_path(os.path.join(a, "data"), 2), names, shape)
    hb, fb = snapshot.read_snapshot(snapshot.snapshot_path(os.path.join(b, "data"), 2), names, shape)
    assert ha["step"] == hb["step"] and ha["t_save"] == pytest.approx(hb["t_save"], rel=1e-13)
    for k in names:
        assert np.max(np.abs(fa[k] - fb[k])) < 1e-12


def test_heat_diffusion(tmp_path):
    """tests/heat-diffusion (check-heat-diffusion.py, test_rk3_heatdiff.py): the
    double top-hat th_0 of tanh edges (a = 0.01) diffusing with nu_th = 1e-3 on
    1024x2x2, run with -r 0 to t_final = 10 with a snapshot at t = 5.  Pinned by
    the exact RK3 amplification 1 + z + z^2/2 + z^3/6 (z = -nu_th kx^2 dt) of every
    mode over the dt sequence of the heat-equation branch of compute_dt
    (dt = cfl cfl_par / (nu_th kxmax^2), compute_dt.cu:40-57, clamped to the
    snapshot and final times, :288-293).  Every y-z column uniform (:150-155).
    Tolerance: L1 (mean absolute) error 1e-12.  The reference's 1e-15
    (check-heat-diffusion.py:29, :160) compares two codes with the same operation
    order; over these 45818 steps the product of exact amplifications and Wray's
    two-register form, both in numpy FP64, already differ by 1.2e-13 in L1, so a
    bound below that would test rounding order, not the algorithm."""
    nx, chi, cfl, cfl_par, t_final, t_out = 1024, 1e-3, 1.5, 1.5, 10.0, 5.0
    out = str(tmp_path)
    x = -0.5 + np.arange(nx) / nx
    th0 = 1 + 0.5 * (np.tanh((x + 0.375) / 0.01) - np.tanh((x + 0.125) / 0.01)) \
        + 0.5 * (np.tanh((x - 0.125) / 0.01) - np.tanh((x - 0.375) / 0.01))
    write_ic(out, {"th": np.broadcast_to(th0[:, None, None], (nx, 2, 2)).copy()})
    run(os.path.join(CFG, "heat-diffusion"), out, "-r", "0")
    snaps = read_all(out, ["th"], (nx, 2, 2))
    assert [round(h["t_save"], 12) for h, _ in snaps] == [0.0, 5.0, 10.0]
    # the expected states: every mode times the RK3 amplification of each step
    kx = 2 * np.pi * np.fft.rfftfreq(nx, 1.0 / nx)
    kxmax = 2 * np.pi * (nx // 2 - 1)
    dt0 = cfl * cfl_par / (chi * kxmax ** 2)
    th_hat = np.fft.rfft(th0)
    t, t_last, step, expect = 0.0, 0.0, 0, {}
    while t < t_final:
        dt = dt0
        if t + dt > t_final:
            dt = t_final - t
        elif t + dt - t_last > t_out:
            dt = t_out - t + t_last
        z = -chi * kx ** 2 * dt
        th_hat = th_hat * (1 + z + z * z / 2 + z * z * z / 6)
        t += dt
        step += 1
        if t - t_last >= t_out:
            t_last += t_out
            expect[round(t_last, 9)] = (step, np.fft.irfft(th_hat, n=nx))
    for hdr, f in snaps[1:]:
        th = f["th"]
        assert np.all(th == th[:, :1, :1])  # uniform along y and z
        st, ref = expect[round(hdr["t_save"], 9)]
        assert hdr["step"] == st, (hdr, st)
        l1 = np.mean(np.abs(th[:, 0, 0] - ref))
        assert l1 < 1e-12, (hdr["t_save"], l1)


@pytest.mark.parametrize("nx", [32, 128, 512])
def test_heat_diffusion_rkl2(tmp_path, nx):
    """tests/heat-diffusion-sts, rkl2 column (check-heat-diffusion-sts.py:60-122):
    a Gaussian (sigma 0.1, lx = 4) diffusing with nu_th = 1, run with
    --ngrid nx 2 2 -r 0.  With this configuration the reference takes one
    dt = t_final step (its sts cap sits in the incompressible branch of
    compute_dt.cu:272-283, and the t_final clamp pre-empts the snapshot clamp,
    :288-293), i.e. one RKL2 cycle of many stages.  The last snapshot equals
    the FP64 oracle's state to 1e-12."""
    import sys
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    cfg = os.path.join(CFG, "heat-diffusion-sts-rkl2")
    out = str(tmp_path)
    lx, sig = 4.0, 0.1
    X = np.broadcast_to((-0.5 * lx + lx * np.arange(nx) / nx)[:, None, None], (nx, 2, 2))
    T0 = 1 / np.sqrt(2 * np.pi * sig ** 2) * np.exp(-0.5 * X ** 2 / sig ** 2)
    write_ic(out, {"th": T0})
    run(cfg, out, "--ngrid", str(nx), "2", "2", "-r", "0")
    hdr, f = read_all(out, ["th"], (nx, 2, 2))[-1]
    assert hdr["t_save"] == pytest.approx(0.5, abs=1e-14)
    p = json.loads(subprocess.run([SPOOKY, "--input-dir", cfg, "--print-params"], capture_output=True,
                                  text=True).stdout.strip().splitlines()[-1])
    q = {k: p[k] for k in orc.DEFAULTS if k in p}
    q.update(nx=nx, ny=2, nz=2)
    sim = orc.Sim(**q)
    sim.set_fields(np.fft.rfftn(T0)[None])
    while sim.time()["t"] < 0.5 - 1e-14:
        sim.main_loop(1)
    assert sim.time()["step"] == hdr["step"]
    ref = np.fft.irfftn(sim.fields[0], s=(nx, 2, 2), axes=(0, 1, 2))
    assert np.max(np.abs(f["th"] - ref)) < 1e-12 * np.max(np.abs(ref))


def test_advected_mhd_vortex(tmp_path):
    """tests/advected-mhd-vortex (512x512x2): L1 < 1e-4 (check-advected-mhd-vortex.py:47)
    at every snapshot."""
    out = str(tmp_path)
    names = ["vx", "vy", "vz", "bx", "by", "bz"]
    write_ic(out, rc.advected_mhd_vortex(0.0))
    run(os.path.join(CFG, "advected-mhd-vortex"), out, "-r", "0")
    snaps = read_all(out, names, (512, 512, 2))
    # snapshots every 0.2 up to t = 1.0; the run ends at t_final = 1.1 between two
    # snapshot times, so no snapshot is written there (CheckSnapshot, check_output.cu:62)
    assert len(snaps) == 6 and snaps[-1][0]["t_save"] == pytest.approx(1.0, abs=1e-12)
    for hdr, f in snaps:
        err = rc.l1(f, rc.advected_mhd_vortex(hdr["t_save"]), names)
        assert err < 1e-4, (hdr, err)


def test_shearing_wave(tmp_path):
    """tests/shearing-wave (shearing box, S = 1.5): the snapshots hold the
    unsheared fields (UnshearOutput, unshear_output.cu:21-54) and match the
    sheared, viscously decaying wave with L1 < 1e-7 (check-shearing-wave.py:38,
    :104-118) at every snapshot up to t_final = 10."""
    out = str(tmp_path)
    names = ["vx", "vy", "vz"]
    write_ic(out, rc.shearing_wave(0.0))
    run(os.path.join(CFG, "shearing-wave"), out, "-r", "0")
    snaps = read_all(out, names, (64, 64, 64))
    assert len(snaps) == 11 and snaps[-1][0]["t_save"] == pytest.approx(10.0, abs=1e-12)
    for hdr, f in snaps:
        err = rc.l1(f, rc.shearing_wave(hdr["t_save"]), names)
        assert err < 1e-7, (hdr, err)


def shell_box(a, b, n, L, nb):
    """Spectrum1d (cuda_kernels_generic.cu:441-471) in numpy on an n, L box:
    Re(a b*) / N^2 per half-space mode (k_z > 0 twice), shells of deltak = 2 pi / min L."""
    k = [rc.TWO_PI / Li * np.fft.fftfreq(ni, 1.0 / ni) for ni, Li in zip(n[:2], L[:2])]
    KX, KY, KZ = np.meshgrid(k[0], k[1], rc.TWO_PI / L[2] * np.arange(n[2] // 2 + 1), indexing="ij")
    pw = np.real(a * np.conj(b)) / float(np.prod(n)) ** 2
    pw[:, :, 1:] *= 2.0
    bins = (np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) / (rc.TWO_PI / min(L)) + 0.5).astype(int)
    keep = bins < nb
    return np.bincount(bins[keep], weights=pw[keep], minlength=nb)[:nb]



def test_mri_eigenmode_and_spectrum_file(tmp_path):
    """tests/mri-eigenmode (shearing box with B0z = 0.1): L1 < 1e-7
    (check-mri-eigenmode.py:76, :118-135) at every snapshot.  The same run,
    with spectra and user quantities added to its output section, writes
    data/spectrum.spooky and data/user-timevar.spooky as the reference does
    (WriteSpectrumOutputHeader / writeSpectrumHelper,
    spooky_output_spectrum.cu:403-480; user_output_timevar.cu:18-112), read
    here the way tests/mti-eigenmode/plot_spectrum.py reads it."""
    cfg = tmp_path / "cfg"
    cfg.mkdir()
    text = open(os.path.join(CFG, "mri-eigenmode", "spooky.cfg")).read()
    anchor = '"bxby","bxbz","bybz");'
    assert anchor in text
    spectra = ["Kx", "Mz", "bxbz", "lorentz_y", "Eth", "emfpower_y"]
    text = text.replace(anchor, anchor + '\n\tspectra_vars = (%s);\n\tuser_timevar_vars = ("kpartheta", "uservar1");'
                        % ", ".join('"%s"' % s for s in spectra))
    (cfg / "spooky.cfg").write_text(text)
    out = str(tmp_path / "run")
    names = ["vx", "vy", "vz", "bx", "by", "bz"]
    ic = rc.mri_eigenmode(0.0)
    write_ic(out, ic)
    run(str(cfg), out, "-r", "0")
    snaps = read_all(out, names, (64, 64, 64))
    assert snaps[-1][0]["t_save"] == pytest.approx(1.0, abs=1e-12)
    for hdr, f in snaps:
        err = rc.l1(f, rc.mri_eigenmode(hdr["t_save"]), names)
        assert err < 1e-7, (hdr, err)

    lines = open(os.path.join(out, "data", "spectrum.spooky")).read().splitlines(True)
    assert lines[0].startswith("## This file contains the 1d (shell-integrated) energy spectral densities")
    assert lines[1].lstrip("##").rstrip("\n").split("\t")[1:-1] == spectra
    n, L = (64, 64, 64), (1.0, 1.5, 1.0)
    kmax = np.sqrt(sum((2.0 / 3.0 * rc.TWO_PI / Li * (ni // 2 - 1)) ** 2 for ni, Li in zip(n, L)))
    nb = int(kmax / (rc.TWO_PI) + 0.5)
    assert lines[2].startswith("wavevector")
    kvec = np.array(lines[2].rstrip("\n").split("\t")[1:-1], dtype="float64")
    assert np.allclose(kvec, rc.TWO_PI * np.arange(nb), rtol=1e-8)
    rows = {}
    for ln in lines[3:]:
        sp = ln.rstrip("\n").split("\t")
        assert sp[0] == "t" and sp[-1] == ""
        rows.setdefault(sp[2], []).append((float(sp[1]), np.array(sp[3:-1], dtype="float64")))
    # Eth needs the heat modules: no line, as in the reference
    assert sorted(rows) == sorted(s for s in spectra if s != "Eth")
    times = [t for t, _ in rows["Kx"]]
    # one row per timevar_step crossing (0.1, check_output.cu:31) from t = 0
    assert len(times) >= 10 and times[0] == 0.0 and np.all(np.diff(times) > 0) and times[-1] <= 1.0 + 1e-12
    c = {k: np.fft.rfftn(ic[k]) for k in names}
    # t = 0: the single mode (kx, kz) = (2 pi, 4 pi) sits in bin round(sqrt 5) = 2
    # with <vx^2> = a^2 / 2
    kx0 = rows["Kx"][0][1]
    assert kx0.shape == (nb,) and kx0[2] == pytest.approx(0.5e-8, rel=1e-7)
    assert np.abs(np.delete(kx0, 2)).max() < 1e-20
    for name, (a, b) in (("Kx", ("vx", "vx")), ("bxbz", ("bx", "by")), ("Mz", ("bz", "bz"))):
        ref = shell_box(c[a], c[b], n, L, nb)
        got = rows[name][0][1]
        assert np.abs(got - ref).max() <= 1e-7 * np.abs(ref).max(), name  # 8 significant digits in the file
    # lorentz_y and emfpower_y carry the same energy transfer (their bin sums agree)
    for (t1, l), (t2, e) in zip(rows["lorentz_y"], rows["emfpower_y"]):
        assert t1 == t2 and l.sum() == pytest.approx(e.sum(), rel=1e-6, abs=1e-30)
    user = open(os.path.join(out, "data", "user-timevar.spooky")).read().splitlines()
    assert user[1].split() == ["##", "kpartheta", "uservar1"]
    assert len(user) == 2 + len(times)
    for ln in user[2:]:
        assert ln.split() == ["-1.00000000e+00", "0.00000000e+00"]  # kpartheta needs anisotropic conduction


@pytest.mark.parametrize("problem, ic, t_end", [
    ("problem-generic", ["tg-yz"], 0.05),
    ("problem-hbi2d", ["hbi-noise", "--amplitude", "0.01", "--b0", "1e-5", "--seed", "98765"], 0.5),
    ("problem-mti3d", ["tg-yz", "--b0", "1e-4"], 0.02),
])
def test_reference_problem_runs_unchanged(tmp_path, problem, ic, t_end):
    """problems/{generic, hbi2d, mti3d}/spooky.cfg (kept under tests/golden/cfg)
    run through the driver from the initial state of each problem's
    initial-condition script (written by write_ic.py), with only t_final and
    snapshot_step shortened to t_end; the final snapshot equals the FP64
    oracle's state after the same steps to 1e-10, and the timevar, spectrum and
    user files carry rows."""
    import re
    import sys
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc
    write_ic_mod = importlib.import_module("spooky-git_amd.write_ic")
    cfg = tmp_path / "cfg"
    cfg.mkdir()
    text = open(os.path.join(CFG, problem, "spooky.cfg")).read()
    text, n1 = re.subn(r"t_final\s*=\s*[0-9.eE+-]+;", "t_final = %r;" % t_end, text)
    text, n2 = re.subn(r"snapshot_step\s*=\s*[0-9.eE+-]+;", "snapshot_step = %r;" % t_end, text)
    assert n1 == 1 and n2 == 1
    (cfg / "spooky.cfg").write_text(text)
    out = str(tmp_path / "run")
    assert write_ic_mod.main(ic + ["--input-dir", str(cfg), "--output-dir", out]) == 0
    p = json.loads(subprocess.run([SPOOKY, "--input-dir", str(cfg), "--output-dir", out, "--print-params"],
                                  capture_output=True, text=True).stdout.strip().splitlines()[-1])
    names = p["fields"]
    shape = (p["nx"], p["ny"], p["nz"])
    _, f0 = snapshot.read_snapshot(snapshot.snapshot_path(os.path.join(out, "data"), 0), names, shape)
    run(str(cfg), out, "-r", "0")
    snaps = read_all(out, names, shape)
    hdr, f = snaps[-1]
    assert hdr["t_save"] == pytest.approx(t_end, abs=1e-12)
    q = {k: p[k] for k in orc.DEFAULTS if k in p}
    sim = orc.Sim(**q)
    sim.set_fields(np.stack([np.fft.rfftn(f0[k]) for k in names]))
    while sim.time()["t"] < t_end - 1e-12:
        sim.main_loop(1)
    assert sim.time()["step"] == hdr["step"]
    ref = {k: np.fft.irfftn(sim.fields[i], s=shape, axes=(0, 1, 2)) for i, k in enumerate(names)}
    scale = max(np.max(np.abs(ref[k])) for k in names)
    for k in names:
        assert np.max(np.abs(f[k] - ref[k])) < 1e-10 * scale, (k, np.max(np.abs(f[k] - ref[k])))
    data = os.path.join(out, "data")
    tv = open(os.path.join(data, "timevar.spooky")).read().splitlines()
    assert len(tv) >= 3 and all(len(ln.split()) == len(tv[1].split()) - 1 for ln in tv[2:])
    sp = open(os.path.join(data, "spectrum.spooky")).read().splitlines()
    assert sp[2].startswith("wavevector") and (len(sp) > 3) == (p["spectra_vars"] != [])
    if p["user_timevar_vars"]:
        assert len(open(os.path.join(data, "user-timevar.spooky")).read().splitlines()) == len(tv)
