"""GPU parity at the headline size: the 256^3 MHD workload of the benchmark
(random-phase solenoidal u and B, E(k) ~ k^4 exp(-2 (k/4)^2), rms 1 each, seed
2026; nu = eta = 1e-3, CFL 0.5, 2 pi box) advanced 2 main-loop steps by the HIP
path and by the FP64 oracle from the same state (reference: Supervisor main loop,
supervisor.cu:169-190; RungeKutta3, hydro_mhd_advance.cu:82-189).

Both back ends run: the fused kernels (forward x pass + epilogue + RK3 + next
inverse x pass in one kernel, the benchmark's choice at 256^3) and the unfused
pipeline.  Tolerances: full-state relative L2 <= 1e-10 (FP64 bulk) / 2e-4 (FP32
bulk); energy and cross helicity <= 1e-12 / 1e-6 relative; the six shell
spectra Kx..Mz <= 1e-10 / 1e-4 relative on bins above 1e-6 of their peak.

Also at full size: the 3D transforms at 256^3 (x and y lines of 256 points)
against the oracle's transform, and the reference's own FFT pin, a 512^3
round trip at relative L2 < 1e-12 (src/tests/tests.cu:18-136, :121).
"""
import math

import numpy as np
import pytest

import pyoracle as orc
from spkmod import spk

pytestmark = pytest.mark.gpu

ini = __import__("importlib").import_module("spooky-git_amd.initial")

N = 256
NSTEPS = 2
PARAMS = dict(nx=N, ny=N, nz=N, incompressible=1, mhd=1, lx=2 * math.pi, ly=2 * math.pi, lz=2 * math.pi,
              nu=1e-3, nu_m=1e-3, cfl=0.5, t_final=1e9, toutput_flow=1e9)
TOL = {"f64": dict(state=1e-10, scalar=1e-12, spec=1e-10), "f32": dict(state=2e-4, scalar=1e-6, spec=1e-4)}


def relerr(a, b):
    return float(np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def state0():
    return ini.mhd_random((N, N, N), seed=2026)


@pytest.fixture(scope="module")
def oracle(state0):
    print("\n# oracle: 256^3 MHD, %d steps" % NSTEPS, flush=True)
    sim = orc.Sim(**PARAMS)
    sim.set_fields(state0)
    sim.main_loop(NSTEPS)
    res = dict(fields=sim.fields.copy(), time=sim.time())
    res["energy"] = sum(sim.energy(f) for f in range(6))
    sim.complex2real()
    res["hc"] = sum(sim.correlation(i, 3 + i) for i in range(3))
    res["spec"] = [sim.spectrum(f, f) for f in range(6)]
    del sim
    return res


def spectra_err(got, ref):
    err = 0.0
    for a, b in zip(got, ref):
        keep = b > 1e-6 * b.max()
        err = max(err, float(np.max(np.abs(a[keep] - b[keep]) / b[keep])))
    return err


@pytest.mark.parametrize("prec,fused", [("f64", 2), ("f64", 0), ("f32", 2)])
def test_mhd256_steps_vs_oracle(ctx, state0, oracle, prec, fused):
    print("# gpu: 256^3 MHD, %s, fused mode %d" % (prec, fused), flush=True)
    sol = spk.Solver(ctx, prec, **PARAMS)
    sol.set_fused(fused)
    assert sol.pipeline()[0] == (fused == 2)
    sol.set_fields(state0)
    sol.main_loop(NSTEPS)
    got = sol.get_fields()
    tol = TOL[prec]
    e_state = relerr(got, oracle["fields"])
    assert e_state < tol["state"], (prec, fused, e_state)
    t = sol.time()
    assert t["step"] == NSTEPS
    assert abs(t["t"] - oracle["time"]["t"]) <= (1e-13 if prec == "f64" else 1e-6) * oracle["time"]["t"]
    e = sum(sol.energy(f) for f in range(6))
    assert abs(e - oracle["energy"]) <= tol["scalar"] * oracle["energy"], (e, oracle["energy"])
    sol.complex2real()
    hc = sum(sol.correlation(i, 3 + i) for i in range(3))
    # |hc| is small against E for independent u and B: scale by the energy
    assert abs(hc - oracle["hc"]) <= tol["scalar"] * oracle["energy"], (hc, oracle["hc"])
    es = spectra_err([sol.spectrum(f, f) for f in range(6)], oracle["spec"])
    assert es < tol["spec"], es
    sol.destroy()


def test_mhd64_100_steps_vs_oracle(ctx):
    """The 64^3 x 100-step FP64 check of the same MHD workload (seed 2026): state,
    time and energy against the oracle after every 25 steps."""
    n = 64
    p = dict(PARAMS, nx=n, ny=n, nz=n)
    u0 = ini.mhd_random((n, n, n), seed=2026)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sol = spk.Solver(ctx, "f64", **p)
    sol.set_fields(u0)
    for chunk in range(4):
        sim.main_loop(25)
        sol.main_loop(25)
        e_state = relerr(sol.get_fields(), sim.fields)
        assert e_state < 1e-10, (chunk, e_state)
        e_ref = sum(sim.energy(f) for f in range(6))
        e = sum(sol.energy(f) for f in range(6))
        assert abs(e - e_ref) <= 1e-12 * e_ref, (chunk, e, e_ref)
        assert abs(sol.time()["t"] - sim.time()["t"]) <= 1e-13 * sim.time()["t"]
    assert sol.time()["step"] == 100
    sol.destroy()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fft_256_vs_oracle(ctx, prec):
    """256^3 R2C and C2R (x and y lines of 256 points) against the oracle's transform."""
    rng = np.random.default_rng(17)
    r = rng.standard_normal((N, N, N))
    c_orc = orc.r2c(r)
    plan = spk.FftPlan(ctx, N, N, N, prec, 1)
    rdt, cdt = (np.float64, np.complex128) if prec == "f64" else (np.float32, np.complex64)
    buf = spk.padded_real(r).astype(rdt)
    d = ctx.alloc(buf.nbytes)
    d.upload(buf)
    plan.r2c(d.ptr, d.ptr)
    c_gpu = d.download(cdt, c_orc.shape).astype(np.complex128)
    tol = 1e-13 if prec == "f64" else 2e-6
    assert relerr(c_gpu, c_orc) < tol
    d.upload(np.ascontiguousarray(c_orc.astype(cdt)))
    plan.c2r(d.ptr, d.ptr)
    back = d.download(rdt, buf.shape)[..., :N].astype(np.float64)
    assert relerr(back, r) < tol
    assert relerr(back, orc.c2r(c_orc, N)) < tol
    plan.destroy()


def test_fft_512_roundtrip(ctx):
    """The reference's FFT check: forward then inverse at 512^3, FP64, relative
    L2 error < 1e-12 (src/tests/tests.cu:121)."""
    n = 512
    rng = np.random.default_rng(23)
    buf = np.zeros((n, n, 2 * (n // 2 + 1)))
    buf[..., :n] = rng.standard_normal((n, n, n))
    ref = buf[..., :n].copy()
    plan = spk.FftPlan(ctx, n, n, n, "f64", 1)
    d = ctx.alloc(buf.nbytes)
    d.upload(buf)
    del buf
    plan.r2c(d.ptr, d.ptr)
    plan.c2r(d.ptr, d.ptr)
    out = d.download(np.float64, (n, n, 2 * (n // 2 + 1)))[..., :n]
    assert relerr(out, ref) < 1e-12
    plan.destroy()


# MTI 256^3 (the bench's MTI workload: spooky-git_amd.initial.mti_params and
# mti_white_noise, seed 7): Boussinesq MHD with stratification along z and
# anisotropic conduction of theta advanced by RKL2 (14 stages per cycle).  Two
# main-loop steps cover both orders of HydroMHDAdvance (hydro_mhd_advance.cu:
# 18-75): step 0 runs RK3 and then the RKL2 cycle, step 1 the cycle and then
# RK3.  Fused mode 2 runs the stratified fused RK3 kernel (temperature and
# velocity groups in one block) and the fused RKL2 stage; mode 0 the unfused
# kernels.  Each field group (u, B, theta) against the FP64 oracle to 1e-10
# relative L2: u starts at zero and is driven by buoyancy alone, so a group's
# own norm is the scale.
MTI_STEPS = 2


@pytest.fixture(scope="module")
def mti_case():
    p = ini.mti_params(N)
    u0 = ini.mti_white_noise((N, N, N), seed=7)
    print("\n# oracle: 256^3 MTI, %d steps" % MTI_STEPS, flush=True)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sim.main_loop(MTI_STEPS)
    res = dict(params=p, u0=u0, fields=sim.fields.copy(), time=sim.time())
    del sim
    return res


@pytest.mark.parametrize("fused", [2, 0])
def test_mti256_steps_vs_oracle(ctx, mti_case, fused):
    print("# gpu: 256^3 MTI, fused mode %d" % fused, flush=True)
    sol = spk.Solver(ctx, "f64", **mti_case["params"])
    sol.set_fused(fused)
    assert sol.pipeline()[0] == (fused == 2)
    sol.set_fields(mti_case["u0"])
    sol.main_loop(MTI_STEPS)
    got, ref = sol.get_fields(), mti_case["fields"]
    v = sol.vars
    groups = {"u": slice(v["VX"], v["VX"] + 3), "B": slice(v["BX"], v["BX"] + 3), "theta": slice(v["TH"], v["TH"] + 1)}
    for name, sl in groups.items():
        err = relerr(got[sl], ref[sl])
        assert err < 1e-10, (fused, name, err)
    assert np.linalg.norm(np.ravel(ref[groups["u"]])) > 0.0
    t = sol.time()
    assert t["step"] == MTI_STEPS
    assert abs(t["t"] - mti_case["time"]["t"]) <= 1e-13 * mti_case["time"]["t"]
    sol.destroy()
