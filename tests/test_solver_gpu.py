"""GPU parity of the fused solver (libspk, through its C ABI) against the FP64
oracle on identical seeded inputs.

Tolerances (relative L2 over all fields): FP64 bulk <= 1e-11 for one RHS,
<= 1e-10 after several steps; FP32 bulk <= 3e-5 for one RHS, <= 2e-4 after
several steps (FFTs and products in FP32, state in FP64)."""
import numpy as np
import pytest

import pyoracle as orc
from physics_cases import CASES, FUSED_CASES, PLANE_CASES, ini, initial_state, relerr
from spkmod import spk

pytestmark = pytest.mark.gpu

TOL_RHS = {"f64": 1e-11, "f32": 3e-5}
TOL_STEPS = {"f64": 1e-10, "f32": 2e-4}


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


def make_pair(ctx, name, prec, seed=1):
    p = CASES[name]["params"]
    sim = orc.Sim(**p)
    sol = spk.Solver(ctx, prec, **p)
    assert sol.nf == sim.nf and sol.vars == sim.vars
    u0 = initial_state(name, seed)
    sim.set_fields(u0)
    sol.set_fields(u0)
    return sim, sol, u0


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_compute_dfield(ctx, name, prec):
    sim, sol, _ = make_pair(ctx, name, prec)
    sim.complex2real()
    sim.compute_dfield(0)
    sol.compute_dfield()
    ref = sim.dfields.copy()
    got = sol.get_dfields()
    assert relerr(got, ref) < TOL_RHS[prec], (name, prec, relerr(got, ref))


@pytest.mark.parametrize("name", sorted(CASES))
def test_compute_dt(ctx, name):
    sim, sol, _ = make_pair(ctx, name, "f64")
    sim.compute_dt()
    sol.compute_dt()
    a, b = sim.time(), sol.time()
    for k in ("dt", "dt_hyp", "dt_par", "dt_sts"):
        if np.isfinite(a[k]) and a[k] != 0.0:
            assert abs(b[k] - a[k]) <= 1e-12 * abs(a[k]), (name, k, a[k], b[k])
        else:
            assert a[k] == b[k] or (np.isinf(a[k]) and np.isinf(b[k])), (name, k, a[k], b[k])


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_main_loop_steps(ctx, name, prec):
    sim, sol, _ = make_pair(ctx, name, prec)
    nsteps = 6
    sim.main_loop(nsteps)
    sol.main_loop(nsteps)
    e = relerr(sol.get_fields(), sim.fields)
    assert e < TOL_STEPS[prec], (name, prec, e)
    ta, tb = sim.time(), sol.time()
    assert abs(tb["t"] - ta["t"]) <= 1e-12 * max(abs(ta["t"]), 1e-300) * (1 if prec == "f64" else 1e5)
    assert tb["step"] == ta["step"] == nsteps


def test_shearing_remap_crossing(ctx):
    """Start just before the remap time so the first steps cross it
    (hydro_mhd_advance.cu:168-180, shearing_remapping.cu:15-59)."""
    name = "shear_mhd"
    sim, sol, _ = make_pair(ctx, name, "f64")
    p = CASES[name]["params"]
    t0 = p["ly"] / (2 * p["shear"] * p["lx"]) - 0.004
    sim.set_time(t0)
    sol.set_time(t0)
    sim.main_loop(5)
    sol.main_loop(5)
    assert relerr(sol.get_fields(), sim.fields) < 1e-10
    assert abs(sol.time()["tremap"] - sim.time()["tremap"]) < 1e-12
    assert sim.time()["tremap"] < 0  # the remap happened


def test_diagnostics(ctx):
    sim, sol, _ = make_pair(ctx, "mhd", "f64")
    v = sim.vars
    for f in range(sim.nf):
        assert abs(sol.energy(f) - sim.energy(f)) <= 1e-12 * abs(sim.energy(f))
    for f0 in (v["VEL"], v["MAG"]):
        assert abs(sol.enstrophy(f0) - sim.enstrophy(f0)) <= 1e-12 * abs(sim.enstrophy(f0))
    sim.complex2real()
    sol.complex2real()
    assert relerr(sol.get_realbuf(), sim.realbuf) < 1e-13
    for fa, fb in ((v["VX"], v["BX"]), (v["VY"], v["VY"]), (v["VZ"], v["BY"])):
        ref = sim.correlation(fa, fb)
        assert abs(sol.correlation(fa, fb) - ref) <= 1e-12 * max(abs(ref), 1e-3)
    for fa, fb in ((v["VX"], v["VX"]), (v["BY"], v["BZ"])):
        ref = sim.spectrum(fa, fb)
        got = sol.spectrum(fa, fb)
        assert got.shape == ref.shape
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_rkl2_cycle(ctx):
    name = "mhd_bouss_rkl2"
    sim, sol, _ = make_pair(ctx, name, "f64")
    sim.compute_dt()
    sol.compute_dt()
    sim.rkl2_cycle()
    sol.rkl2_cycle()
    assert relerr(sol.get_fields(), sim.fields) < 1e-12
    s, _ = orc.rkl2_coeffs(sim.time()["dt"], sim.time()["dt_sts"])
    assert s > 2  # the cycle really is multi-stage here


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("device_plan", [False, True])
def test_fused_rkl2_stage_matches_unfused_and_oracle(ctx, prec, device_plan):
    """Anisotropic conduction of theta by RKL2 on a grid the fused x kernel takes
    (NX = 32): one cycle, and four main-loop steps (cycles before and after RK3),
    with the fused stage (forward x pass of the flux + update + next inverse x
    pass in one kernel) against the unfused stages and the oracle; host- and
    device-planned cycles."""
    name = "mhd_bouss_rkl2_x32"
    p = CASES[name]["params"]
    u0 = initial_state(name, 3)
    fus = spk.Solver(ctx, prec, **p)
    unf = spk.Solver(ctx, prec, **p)
    unf.set_fused(0)
    for s in (fus, unf):
        s.set_rkl_device(device_plan)
        s.set_fields(u0)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sim.compute_dt()
    for s in (fus, unf):
        s.compute_dt()
        s.rkl2_cycle()
    sim.rkl2_cycle()
    n, _ = orc.rkl2_coeffs(sim.time()["dt"], sim.time()["dt_sts"])
    assert n > 2
    tol_pair = 1e-13 if prec == "f64" else 2e-5
    tol_orc = 1e-12 if prec == "f64" else 2e-4
    a, b = fus.get_fields(), unf.get_fields()
    assert relerr(a, b) < tol_pair, relerr(a, b)
    assert relerr(a, sim.fields) < tol_orc, relerr(a, sim.fields)
    for s in (fus, unf):
        s.set_fields(u0)
        s.set_time(0.0)
        s.main_loop(4)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sim.main_loop(4)
    a, b = fus.get_fields(), unf.get_fields()
    assert relerr(a, b) < (1e-12 if prec == "f64" else 2e-5), relerr(a, b)
    assert relerr(a, sim.fields) < TOL_STEPS[prec], relerr(a, sim.fields)
    fus.destroy()
    unf.destroy()


def test_taylor_green_64_100_steps(ctx):
    """Config 1: HD Taylor-Green, 64^3, 2pi box, nu=0.01, CFL 0.5, 2/3 dealiasing,
    100 steps; kinetic energy and dissipation 2 nu <w^2> against the oracle."""
    ini = __import__("importlib").import_module("spooky-git_amd.initial")
    p = dict(nx=64, ny=64, nz=64, incompressible=1, lx=2 * np.pi, ly=2 * np.pi, lz=2 * np.pi, nu=0.01,
             cfl=0.5, t_final=1e9, toutput_flow=1e9)
    u0 = ini.taylor_green(64)
    sim = orc.Sim(**p)
    sol = spk.Solver(ctx, "f64", **p)
    sim.set_fields(u0)
    sol.set_fields(u0)
    for chunk in range(4):
        sim.main_loop(25)
        sol.main_loop(25)
        ev_ref = sum(sim.energy(f) for f in range(3))
        ev = sum(sol.energy(f) for f in range(3))
        eps_ref = 2 * 0.01 * sim.enstrophy(0)
        eps = 2 * 0.01 * sol.enstrophy(0)
        assert abs(ev - ev_ref) <= 1e-12 * ev_ref, (chunk, ev, ev_ref)
        assert abs(eps - eps_ref) <= 1e-12 * eps_ref, (chunk, eps, eps_ref)
    assert sol.time()["step"] == 100


@pytest.mark.parametrize("points", [4, 8])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", FUSED_CASES)
def test_fused_path_matches_unfused_and_oracle(ctx, name, prec, points):
    """The fused forward-x + epilogue + projection + next-inverse-x kernel, at 4 and
    at 8 points per thread, gives the unfused pipeline's results (to rounding) and
    the oracle's, over several steps (the W buffer handed from one step to the next
    is exercised from step 2 on)."""
    sim, sol, u0 = make_pair(ctx, name, prec)
    sol.set_fused(2)
    sol.set_fused_points(points)
    assert sol.pipeline() == (True, points)
    ref2 = spk.Solver(ctx, prec, **CASES[name]["params"])
    ref2.set_fused(0)
    ref2.set_fields(u0)
    nsteps = 5
    sim.main_loop(nsteps)
    sol.main_loop(nsteps)
    ref2.main_loop(nsteps)
    a, b = sol.get_fields(), ref2.get_fields()
    assert relerr(a, b) < (1e-12 if prec == "f64" else 2e-5), relerr(a, b)
    assert relerr(a, sim.fields) < TOL_STEPS[prec]
    assert sol.time()["t"] == pytest.approx(sim.time()["t"], rel=1e-12 if prec == "f64" else 1e-6)
    # two calls in a row (fresh W at the start of each call) equal one longer call
    sol2 = spk.Solver(ctx, prec, **CASES[name]["params"])
    sol2.set_fused(2)
    sol2.set_fused_points(points)
    sol2.set_fields(u0)
    sol2.main_loop(2)
    sol2.main_loop(3)
    assert relerr(sol2.get_fields(), a) < (1e-13 if prec == "f64" else 1e-6)
    t = sol.kernel_timings(2)
    assert t["fwdx_epilogue"] is not None and t["fwdx_epilogue"] > 0


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", ["mhd_x32", "hd_strat_x32", "mhd_bouss_rkl2_x32", "hd_strat_sts_u_x32",
                                  "q2d_hbi_y512"])
def test_main_loop_resume(ctx, name, prec):
    """spk_main_loop_resume continues the previous call: one step and then four
    resumed single steps give the state and clock of one five-step call, bit for
    bit (super-time-stepped cases cross both cycle placements); after another
    entry point (compute_dfield, which rewrites the transform buffers) a resumed
    call starts fresh, as a plain call does."""
    p = CASES[name]["params"]
    u0 = initial_state(name)
    sols = []
    for _ in range(4):
        sol = spk.Solver(ctx, prec, **p)
        sol.set_fused(2)
        sol.set_fused_points(8)  # the same tile in every solver compared (the autotune may differ)
        sol.set_fields(u0)
        sols.append(sol)
    a, b, c, d = sols
    a.main_loop(5)
    b.main_loop(1)
    for _ in range(4):
        b.main_loop(1, resume=True)
    assert np.array_equal(a.get_fields(), b.get_fields())
    assert a.time() == b.time()
    c.main_loop(1)
    c.compute_dfield()
    c.main_loop(1, resume=True)
    d.main_loop(1)
    d.main_loop(1)
    assert np.array_equal(c.get_fields(), d.get_fields())
    # resuming from a changed state (set_fields) starts fresh too
    a.set_fields(u0)
    a.set_time(0.0)
    a.main_loop(1, resume=True)
    e = spk.Solver(ctx, prec, **p)
    e.set_fused(2)
    e.set_fused_points(8)
    e.set_fields(u0)
    e.main_loop(1)
    assert np.array_equal(a.get_fields(), e.get_fields())
    for s_ in sols + [e]:
        s_.destroy()


@pytest.mark.parametrize("setup", ["unfused", "forced"])
def test_main_loop_resume_without_continuation(ctx, setup):
    """Where no step leaves a continuation (the unfused back end; forcing, which
    changes the state after the step), resumed single steps equal one long call
    bit for bit."""
    name = "mhd_x32"
    p = CASES[name]["params"]
    u0 = initial_state(name)
    nx, ny, nzh = p["nx"], p["ny"], p["nz"] // 2 + 1
    out = []
    for resume in (False, True):
        sol = spk.Solver(ctx, "f64", **p)
        if setup == "unfused":
            sol.set_fused(0)
        else:
            sol.set_fused(2)
            sol.set_fused_points(8)
            idx = np.array([(1 * ny + 2) * nzh + 1, (3 * ny + 0) * nzh + 2], dtype=np.int64)
            sol.set_forcing(idx, [[0.1, 0.2j, 0.0], [0.0, 0.05, -0.1j]])
        sol.set_fields(u0)
        if resume:
            sol.main_loop(1)
            for _ in range(3):
                sol.main_loop(1, resume=True)
        else:
            sol.main_loop(4)
        out.append((sol.get_fields(), sol.time()))
        sol.destroy()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("fused", [2, 0])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(PLANE_CASES))
def test_plane_kernels_match_three_kernel_path(ctx, name, prec, fused):
    """On quasi-2D grids the plane kernel (inverse y pass, z-row products and
    forward y pass of the products in one kernel per x plane; for the RKL2 stages
    the anisotropic flux) gives the state of the separate y-pass and z kernels to
    rounding, and the oracle's, over several steps of both back ends."""
    sim, sol, u0 = make_pair(ctx, name, prec)
    sol.set_fused(fused)
    ref = spk.Solver(ctx, prec, **CASES[name]["params"])
    ref.set_fused(fused)
    ref.set_plane(False)
    ref.set_fields(u0)
    rk3, rkl2 = PLANE_CASES[name]
    assert sol.plane() == (rk3, rkl2 and fused == 2)
    assert ref.plane() == (False, False)
    nsteps = 6
    sim.main_loop(nsteps)
    sol.main_loop(nsteps)
    ref.main_loop(nsteps)
    a, b = sol.get_fields(), ref.get_fields()
    assert relerr(a, b) < (1e-13 if prec == "f64" else 1e-6), relerr(a, b)
    assert relerr(a, sim.fields) < TOL_STEPS[prec], relerr(a, sim.fields)
    assert sol.time()["t"] == pytest.approx(sim.time()["t"], rel=1e-12 if prec == "f64" else 1e-6)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fused_matches_unfused_at_256(ctx, prec):
    """At 256^3 MHD (2064 column tiles of 16 per field group) the fused kernel
    gives the unfused pipeline's state after a two-step call and a one-step call
    (the second call starts from a fresh W)."""
    n = 256
    p = dict(nx=n, ny=n, nz=n, incompressible=1, mhd=1, lx=2 * np.pi, ly=2 * np.pi, lz=2 * np.pi, nu=1e-3,
             nu_m=1e-3, cfl=0.5, t_final=1e9, toutput_flow=1e9)
    shape = (n, n, n)
    u0 = np.concatenate([ini.random_solenoidal(shape, seed=11), ini.random_solenoidal(shape, seed=12, rms=0.5)])
    out = {}
    for mode in (0, 2):
        sol = spk.Solver(ctx, prec, **p)
        sol.set_fused(mode)
        sol.set_fields(u0)
        sol.main_loop(2)
        sol.main_loop(1)
        out[mode] = sol.get_fields()
        sol.destroy()
    tol = 1e-12 if prec == "f64" else 2e-5
    assert relerr(out[2], out[0]) < tol, relerr(out[2], out[0])
