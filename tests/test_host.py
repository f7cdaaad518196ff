"""Host side of a run (no GPU): the spooky.cfg reader and Parameters mapping
(checked through `spooky --print-params`, which resolves a configuration the
way the driver does and exits before touching the GPU), the command line, and
the HDF5 snapshot layout of the reference (hdf5_io.cu)."""
import importlib
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from spkmod import REPO

SPOOKY = os.path.join(REPO, "spooky-git_amd", "bin", "spooky")
CFG = os.path.join(REPO, "tests", "golden", "cfg")
H5LS = shutil.which("h5ls") or ("/opt/conda/bin/h5ls" if os.path.exists("/opt/conda/bin/h5ls") else None)

pytestmark = pytest.mark.skipif(not os.path.exists(SPOOKY), reason="spooky driver not built (spooky-git_amd/build.py)")


def params(cfgdir, *extra):
    r = subprocess.run([SPOOKY, "--input-dir", cfgdir, "--print-params"] + list(extra), capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def write_cfg(tmp_path, text):
    d = tmp_path / "cfg"
    d.mkdir(exist_ok=True)
    (d / "spooky.cfg").write_text(text)
    return str(d)


def test_reference_test_configs():
    """Values the reference's own test configurations set (tests/*/spooky.cfg)."""
    tg = params(os.path.join(CFG, "taylor-green-vortex"))
    assert (tg["nx"], tg["ny"], tg["nz"]) == (64, 64, 64)
    assert tg["nu"] == 1.0 / 200.0 and tg["fields"] == ["vx", "vy", "vz"]
    assert tg["t_final"] == 10.0 and tg["toutput_flow"] == 1.0 and tg["toutput_time"] == 0.1
    assert tg["timevar_vars"] == ["t", "ev", "Kx", "Ky", "Kz", "w2", "hv"]
    assert tg["configname"] == "Taylor-Green vortex test problem"
    mti = params(os.path.join(CFG, "mti-eigenmode"))
    assert (mti["nx"], mti["ny"], mti["nz"]) == (512, 4, 4)
    assert mti["reynolds_ani"] == 100.0 and mti["N2"] == 0.1 and mti["OmegaT2"] == 1.0
    assert mti["nu"] == 1e-5 and mti["nu_m"] == 1e-5 and mti["nu_th"] == 1e-6
    assert mti["fields"] == ["vx", "vy", "vz", "bx", "by", "bz", "th"] and mti["sts_mask"] == 0
    sts = params(os.path.join(CFG, "heat-diffusion-sts-rkl2"))
    assert sts["heat_equation"] == 1 and sts["fields"] == ["th"]
    assert sts["supertimestepping"] == 1 and sts["sts_algorithm"] == "rkl2" and sts["sts_mask"] == 1
    assert sts["lx"] == 4.0
    sw = params(os.path.join(CFG, "shearing-wave"))
    assert sw["shearing"] == 1 and sw["shear"] == 1.5 and sw["ly"] == 1.5
    # checkParameters: snapshots every ly / (shear lx) when shearing (parameters.cu:548-562)
    assert sw["toutput_flow"] == 1.0
    for name in os.listdir(CFG):
        if name != "heat-diffusion-sts-sts":
            params(os.path.join(CFG, name))


def test_unavailable_sts_algorithm_is_refused():
    r = subprocess.run([SPOOKY, "--input-dir", os.path.join(CFG, "heat-diffusion-sts-sts"), "--print-params"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "rkl2" in r.stderr


def test_grammar_and_lookup_semantics(tmp_path):
    d = write_cfg(tmp_path, r'''
# comment
/* block
   comment */
configname = "two " "parts";   // adjacent strings concatenate
modules: { incompressible = true; mhd = TRUE, boussinesq = false; };
physics:
{
    boxsize = (2.0, 3.0, 0.5);
    gridsize = [16, 8, 4];
    reynolds = 100;           // an int: libconfig without auto-conversion keeps the default
    reynolds_magnetic = 250.0;
    brunt_vaissala_squared = 1e-1;
    shear = -0.25;
};
code: { cfl = 1.; t_final = 2.5e0; safety_sts = 0x10; };
output: { timevar_vars = ["t", "ev"]; spectra_vars = ("Kx"); };
''')
    p = params(d)
    assert p["configname"] == "two parts"
    assert p["incompressible"] == 1 and p["mhd"] == 1 and p["boussinesq"] == 0
    assert (p["lx"], p["ly"], p["lz"]) == (2.0, 3.0, 0.5) and (p["nx"], p["ny"], p["nz"]) == (16, 8, 4)
    assert p["nu"] == 1.0 and p["nu_m"] == 1.0 / 250.0 and p["N2"] == 0.1 and p["shear"] == -0.25
    assert p["cfl"] == 1.0 and p["t_final"] == 2.5 and p["safety_sts"] == 16
    assert p["timevar_vars"] == ["t", "ev"] and p["spectra_vars"] == ["Kx"]
    assert p["fields"] == ["vx", "vy", "vz", "bx", "by", "bz"]
    # defaults of parameters.cu for everything absent
    assert p["cfl_par"] == 1.0 and p["safety_source"] == 0.2 and p["strat_direction"] == 2
    assert p["toutput_flow"] == 1.0 and p["toutput_time"] == 1.0 and p["antialiasing"] == 1


def test_syntax_error_names_the_line(tmp_path):
    d = write_cfg(tmp_path, "modules:\n{\n  mhd = true;\n  incompressible = \n};\n")
    r = subprocess.run([SPOOKY, "--input-dir", d, "--print-params"], capture_output=True, text=True)
    assert r.returncode == 1 and "line 5" in r.stderr


def test_command_line(tmp_path):
    d = write_cfg(tmp_path, "modules: { incompressible = true; };\noutput: { output_dir = \"/nowhere\"; };\n")
    p = params(d, "-n", "16", "8", "4", "--with-2d", "--output-dir", str(tmp_path / "o"))
    assert (p["nx"], p["ny"], p["nz"]) == (16, 8, 4) and p["with2d"] == 1
    assert p["output_dir"] == str(tmp_path / "o")
    p = params(d, "-t", "1", "30", "0")
    assert p["max_walltime_elapsed"] == pytest.approx(0.95 * 1.5)
    assert p["bulk"] == "f64" and params(d, "--fp32-bulk")["bulk"] == "f32"
    for bad in (["--bogus"], ["-n", "16", "8"], ["-r", "x"]):
        r = subprocess.run([SPOOKY, "--input-dir", d] + bad, capture_output=True, text=True)
        assert r.returncode == 1 and "usage" in r.stderr


def test_check_parameters(tmp_path):
    d = write_cfg(tmp_path, "modules: { incompressible = true; heat_equation = true; };\n")
    r = subprocess.run([SPOOKY, "--input-dir", d, "--print-params"], capture_output=True, text=True)
    assert r.returncode == 1 and "mutually exclusive" in r.stdout


def test_snapshot_layout_and_round_trip(tmp_path):
    snapshot = importlib.import_module("spooky-git_amd.snapshot")
    rng = np.random.default_rng(3)
    f = {n: rng.standard_normal((8, 4, 6)) for n in ("vx", "vy", "vz", "th")}
    path = str(tmp_path / "snap0000.h5")
    snapshot.write_snapshot(path, f, t_start=0.5, t_save=1.25, t_lastsnap=1.0, t_lastvar=1.2, step=17)
    hdr, g = snapshot.read_snapshot(path, list(f), (8, 4, 6))
    assert hdr == dict(t_save=1.25, t_lastsnap=1.0, t_lastvar=1.2, step=17)
    for n in f:
        assert np.array_equal(f[n], g[n])
    with pytest.raises(RuntimeError):
        snapshot.read_snapshot(path, ["bx"], (8, 4, 6))
    with pytest.raises(RuntimeError):
        snapshot.read_snapshot(path, ["vx"], (8, 4, 5))
    assert snapshot.last_snapshot(str(tmp_path)) == 0
    if H5LS:
        # the reference's datasets: float64 scalars of shape (1,), int32 step,
        # fields flattened to (NX*NY*NZ,) in C order
        out = subprocess.run([H5LS, path], capture_output=True, text=True).stdout
        ds = {ln.split()[0]: ln.split("{")[1].split("}")[0] for ln in out.splitlines()}
        assert ds == {"step": "1", "t_lastsnap": "1", "t_lastvar": "1", "t_save": "1", "t_start": "1",
                      "vx": "192", "vy": "192", "vz": "192", "th": "192"}
