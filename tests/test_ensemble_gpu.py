"""Batched ensemble (spk_ensemble) and device-planned RKL2 cycles on the GPU.

- A device-planned RKL2 cycle (the recurrence of rkl.cu:466-585 evaluated by a
  one-thread kernel, extra stages returning at once) gives the host-planned
  cycle's state, and its stage count is the recurrence's.
- Every box of an ensemble (uneven grids and physics, two steps per graph
  replay) ends where the same solver run on its own ends, with the graph and
  without it, and where the FP64 oracle ends.
- Batched launches (one launch per kernel for every box of a shape group, the
  box in blockIdx.z) give every box the bits that box-by-box launches give it,
  for any cut of the groups into batches.
- The layout autotune leaves every box's state and clock as it found them.
"""
import numpy as np
import pytest

import pyoracle as orc
from physics_cases import CASES, initial_state, relerr
from spkmod import spk

pytestmark = pytest.mark.gpu
ini = __import__("importlib").import_module("spooky-git_amd.initial")


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


def rkl_stages(r):
    """Stage count of rkl.cu:473-475 for r = dt / dt_sts."""
    return 1 + int(4.0 * (1.0 + 2.0 * r) / (1.0 + np.sqrt(9.0 + 16.0 * r)))


def test_device_rkl_plan_matches_host(ctx):
    name = "mhd_bouss_rkl2"
    p = CASES[name]["params"]
    u0 = initial_state(name)
    a = spk.Solver(ctx, "f64", **p)
    b = spk.Solver(ctx, "f64", **p)
    b.set_rkl_device(True)
    for s in (a, b):
        s.set_fields(u0)
    for _ in range(3):
        a.main_loop(2)
        b.main_loop(2)
        fa, fb = a.get_fields(), b.get_fields()
        assert np.max(np.abs(fa - fb)) <= 1e-14 * np.max(np.abs(fa))
    t = b.time()
    stages, overflow, smax = b.rkl_plan()
    assert overflow == 0 and smax == rkl_stages(p["safety_sts"] * (1 + 1e-12)) == 9
    assert stages == rkl_stages(t["dt"] / t["dt_sts"]) and stages > 2
    a.destroy()
    b.destroy()
    # refused where the stage count has no bound
    h = spk.Solver(ctx, "f64", **CASES["heat"]["params"])
    with pytest.raises(RuntimeError):
        h.set_rkl_device(True)
    h.destroy()


def boxes():
    """Uneven boxes: MHD 16^3 and 32x16x16 with different coefficients, and two
    64x64x2 Boussinesq-MHD boxes with anisotropic conduction and RKL2 on theta
    (quasi-2D: their RK3 and RKL2 stages run the plane kernels)."""
    T = 2 * np.pi
    big = dict(t_final=1e9, toutput_flow=1e9)
    mhd = dict(incompressible=1, mhd=1, lx=T, ly=T, lz=T, cfl=0.5, **big)
    out = [dict(mhd, nx=16, ny=16, nz=16, nu=3e-3, nu_m=1e-2),
           dict(mhd, nx=32, ny=16, nz=16, nu=1e-2, nu_m=2e-3),
           dict(mhd, nx=16, ny=16, nz=16, nu=5e-3, nu_m=5e-3),
           ini.hbi_params(64, nu=2e-3, eta=1e-3, chi=0.05),
           ini.hbi_params(64, nu=1e-3, eta=3e-3, chi=0.02)]
    return out


def box_state(p, seed):
    shape = (p["nx"], p["ny"], p["nz"])
    L = (p["lx"], p["ly"], p["lz"])
    if p.get("boussinesq"):
        st = ini.hbi_noise(shape, L, seed=seed, amp=0.5, b0=1.0)
        st[6] = 0.1 * np.fft.rfftn(np.random.default_rng(seed + 1).standard_normal(shape)) * ini.dealias_mask(
            *shape, *L)
        return st
    return np.concatenate([0.5 * ini.random_solenoidal(shape, L, seed=seed, k0=2.0),
                           0.5 * ini.random_solenoidal(shape, L, seed=seed + 7, k0=2.0)])


@pytest.mark.parametrize("graph", [True, False])
def test_ensemble_matches_standalone_and_oracle(ctx, graph):
    bx = boxes()
    states = [box_state(p, 100 + i) for i, p in enumerate(bx)]
    ens = spk.Ensemble(bx, "f64", lanes=2)
    ens.set_graph(graph)
    for i, st in enumerate(states):
        ens.set_fields(i, st)
    ens.run(6)
    got = [ens.get_fields(i) for i in range(len(bx))]
    times = [ens.time(i) for i in range(len(bx))]
    with pytest.raises(RuntimeError):
        ens.run(3)  # a replay advances every box by two steps
    ens.destroy()
    for i, (p, st) in enumerate(zip(bx, states)):
        s = spk.Solver(ctx, "f64", **p)
        if p.get("supertimestepping"):
            s.set_rkl_device(True)
        s.set_fields(st)
        for _ in range(3):
            s.main_loop(2)
        ref = s.get_fields()
        assert np.max(np.abs(got[i] - ref)) <= 1e-14 * np.max(np.abs(ref)), i
        assert times[i]["step"] == 6 and times[i]["t"] == s.time()["t"]
        s.destroy()
        sim = orc.Sim(**p)
        sim.set_fields(st)
        for _ in range(6):
            sim.main_loop(1)
        assert relerr(got[i], sim.fields) < 1e-10, (i, relerr(got[i], sim.fields))


def test_ensemble_fp32_and_recapture(ctx):
    """FP32 bulk within the solver's FP32 tolerance of the oracle; a box stepped
    outside the graph (parity change) makes the next run capture again."""
    bx = boxes()
    states = [box_state(p, 200 + i) for i, p in enumerate(bx)]
    ens = spk.Ensemble(bx, "f32", lanes=4)
    for i, st in enumerate(states):
        ens.set_fields(i, st)
    ens.run(2)
    ens.sync()
    ens.boxes[0].main_loop(1)
    ens.run(2)
    got = [ens.get_fields(i) for i in range(len(bx))]
    assert ens.time(0)["step"] == 5 and all(ens.time(i)["step"] == 4 for i in range(1, len(bx)))
    ens.destroy()
    for i, (p, st) in enumerate(zip(bx, states)):
        sim = orc.Sim(**p)
        sim.set_fields(st)
        for _ in range(5 if i == 0 else 4):
            sim.main_loop(1)
        assert relerr(got[i], sim.fields) < 2e-4, (i, relerr(got[i], sim.fields))


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batched_launches_match_box_by_box(ctx, prec):
    """Shape groups of boxes() (two 16^3 MHD boxes, one 32x16x16, two 64x64x2
    HBI boxes with RKL2): one batch per group, two batches per group, and box by
    box give bit-identical states and clocks after 6 steps (graph on, then 2
    more steps without it)."""
    bx = boxes()
    states = [box_state(p, 400 + i) for i, p in enumerate(bx)]
    results = {}
    # one ensemble throughout: every box keeps the back end its group's first box
    # was tuned to, so the three launch layouts run the same arithmetic
    ens = spk.Ensemble(bx, prec, lanes=2)
    assert ens.layout() == dict(groups=3, split=1, batches=3, lanes=3)
    for split in (1, 2, 0):
        print("# split %d, %s" % (split, prec), flush=True)
        ens.set_split(split)
        lay = ens.layout()
        assert lay["split"] == split and lay["batches"] == {1: 3, 2: 5, 0: 0}[split]
        ens.set_graph(True)
        for i, st in enumerate(states):
            ens.set_fields(i, st)
            ens.boxes[i].set_time(0.0)
        ens.run(6)
        ens.set_graph(False)
        ens.run(2)
        results[split] = ([ens.get_fields(i) for i in range(len(bx))], [ens.time(i) for i in range(len(bx))])
    ens.destroy()
    for split in (1, 2):
        for i in range(len(bx)):
            assert np.array_equal(results[split][0][i], results[0][0][i]), (split, i)
            assert results[split][1][i] == results[0][1][i], (split, i)
    # the HBI boxes took RKL2 cycles of their own stage counts
    t3, t4 = results[1][1][3], results[1][1][4]
    assert t3["dt_sts"] != t4["dt_sts"]


def test_ensemble_restarts_fresh_after_new_state(ctx):
    """Replays continue from the inverse x pass the previous replay left; new
    states set between runs make the next replay start fresh, so an ensemble
    stepped, reloaded and stepped again equals a new ensemble stepped from the
    reloaded states, bit for bit."""
    bx = boxes()

    def pinned(e):
        # the same back end and tile in both ensembles (each creation autotunes its own)
        for s_ in e.boxes:
            s_.set_fused(2)
            s_.set_fused_points(8)
        return e

    ens = pinned(spk.Ensemble(bx, "f64", lanes=2))
    for i, p in enumerate(bx):
        ens.set_fields(i, box_state(p, 700 + i))
    ens.run(4)
    new = [box_state(p, 800 + i) for i, p in enumerate(bx)]
    for i, st in enumerate(new):
        ens.set_fields(i, st)
        ens.boxes[i].set_time(0.0)
    ens.run(4)
    ref = pinned(spk.Ensemble(bx, "f64", lanes=2))
    for i, st in enumerate(new):
        ref.set_fields(i, st)
    ref.run(4)
    for i in range(len(bx)):
        assert np.array_equal(ens.get_fields(i), ref.get_fields(i)), i
        assert ens.time(i) == ref.time(i), i
    ens.destroy()
    ref.destroy()


def test_ensemble_tune_restores_state(ctx):
    bx = boxes()
    states = [box_state(p, 300 + i) for i, p in enumerate(bx)]
    ens = spk.Ensemble(bx, "f32")
    for s_ in ens.boxes:  # the same back end and tile as the reference ensemble below
        s_.set_fused(2)
        s_.set_fused_points(8)
    for i, st in enumerate(states):
        ens.set_fields(i, st)
    ens.run(2)
    before = [ens.get_fields(i) for i in range(len(bx))]
    tb = [ens.time(i) for i in range(len(bx))]
    split, ms = ens.tune(probe=2)
    assert split in (1, 2, 4) and ms > 0 and ens.layout()["split"] == split
    for i in range(len(bx)):
        assert np.array_equal(ens.get_fields(i), before[i])
        t = ens.time(i)
        assert t["t"] == tb[i]["t"] and t["step"] == tb[i]["step"]
    # and it keeps stepping from there as an untuned ensemble would after its
    # states were set again (the restored states start afresh: the probes
    # overwrote the transform buffers the next replay would otherwise continue from)
    ens.run(2)
    after = ens.get_fields(3)
    ens.destroy()
    ref = spk.Ensemble(bx, "f32", lanes=1)
    for s_ in ref.boxes:
        s_.set_fused(2)
        s_.set_fused_points(8)
    for i, st in enumerate(states):
        ref.set_fields(i, st)
    ref.run(2)
    for i in range(len(bx)):
        ref.set_fields(i, ref.get_fields(i))
    ref.run(2)
    assert np.max(np.abs(ref.get_fields(3) - after)) <= 1e-14 * np.max(np.abs(after))
    ref.destroy()


def test_ensemble_refuses_shearing_box(ctx):
    p = dict(CASES["shear_mhd"]["params"])
    with pytest.raises(RuntimeError):
        spk.Ensemble([p], "f64")


def boxes_at_size(seed=2026):
    """One box of each kind of the ensemble workload at its own size: MHD 64^3 and
    128^3 (2 pi box, CFL 0.5) and a 512x512x2 Boussinesq-MHD box with anisotropic
    conduction and RKL2 on theta (problems/hbi2d physics), with nu, eta drawn
    log-uniform in [1e-4, 1e-2] and chi_par uniform in [1e-3, 1e-1]."""
    rng = np.random.default_rng(seed)
    T = 2 * np.pi
    out = []
    for n in (64, 128):
        nu, eta = 10.0 ** rng.uniform(-4, -2, 2)
        out.append(dict(nx=n, ny=n, nz=n, incompressible=1, mhd=1, lx=T, ly=T, lz=T, nu=float(nu), nu_m=float(eta),
                        cfl=0.5, t_final=1e9, toutput_flow=1e9))
    nu, eta = 10.0 ** rng.uniform(-4, -2, 2)
    out.append(ini.hbi_params(512, nu=float(nu), eta=float(eta), chi=float(rng.uniform(1e-3, 1e-1))))
    return out


def state_at_size(p, seed):
    shape = (p["nx"], p["ny"], p["nz"])
    if p.get("boussinesq"):
        return ini.hbi_noise(shape, seed=seed)
    return ini.mhd_random(shape, seed=seed)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ensemble_boxes_at_size_vs_oracle(ctx, prec):
    """The ensemble workload's box kinds at their own sizes, advanced together
    (graph on, 4 steps: two RKL2 cycles on the HBI box) against the FP64 oracle
    box by box: state rel-L2 <= 1e-10 (FP64 bulk) / 2e-4 (FP32 bulk)."""
    bx = boxes_at_size()
    states = [state_at_size(p, 7000 + i) for i, p in enumerate(bx)]
    ens = spk.Ensemble(bx, prec, lanes=3)
    for i, st in enumerate(states):
        ens.set_fields(i, st)
    ens.run(4)
    got = [ens.get_fields(i) for i in range(len(bx))]
    tms = [ens.time(i) for i in range(len(bx))]
    ens.destroy()
    tol = 1e-10 if prec == "f64" else 2e-4
    for i, (p, st) in enumerate(zip(bx, states)):
        print("# box %d: %dx%dx%d vs oracle" % (i, p["nx"], p["ny"], p["nz"]), flush=True)
        sim = orc.Sim(**p)
        sim.set_fields(st)
        sim.main_loop(4)
        e = relerr(got[i], sim.fields)
        assert e < tol, (i, prec, e)
        assert tms[i]["step"] == 4
        assert abs(tms[i]["t"] - sim.time()["t"]) <= (1e-13 if prec == "f64" else 1e-5) * sim.time()["t"]
        if p.get("supertimestepping"):
            t = sim.time()
            assert rkl_stages(t["dt"] / t["dt_sts"]) > 2  # the RKL2 cycles are multi-stage
        del sim
