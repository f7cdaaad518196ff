"""Benchmark: 3D incompressible MHD RK3 steps/s at 256^3 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--n 256]

Workload (the 256^3 MHD configuration): 6 fields (u, B), random-phase solenoidal
initial state with E(k) ~ k^4 exp(-2 (k/4)^2), rms 1 each (seeded), 2 pi box,
nu = eta = 1e-3, CFL 0.5, 2/3 dealiasing.  A step is one iteration of the main
loop: CleanFieldDivergence + HydroMHDAdvance (compute_dt + 3 RK3 stages =
42 3D transforms).  The state is resident in HBM before timing starts.

Prints ONE JSON line.  value = FP64 steps/s; the FP32-bulk mode (FFTs and
products in FP32, state and RK3 in FP64) is reported beside it.  roofline is
measured live for the kernel with the largest share of the FP64 step, with HIP
events on the library's stream; cpu_baseline times the FP64 oracle on the host
cores for a bounded sample of the same workload.

Beside the timed line (not part of `value`):
  fp32_bulk.accuracy_vs_f64  FP32-bulk against FP64 after --accuracy-steps
                             steps (energy, cross helicity, spectra);
  taylor_green_64            64^3 HD Taylor-Green, 100 steps, GPU and oracle;
  mti_rkl2                   256^3 Boussinesq MTI with anisotropic conduction,
                             FP64 RKL2 on theta: steps/s and stages per cycle;
                             the same with FP32 bulk (state and RKL2 recurrence
                             FP64): steps/s and difference from FP64;
  hbi2d_512                  problems/hbi2d at 512x512x2 (Boussinesq MHD, RKL2
                             conduction), one box per precision with and without
                             the quasi-2D plane kernels; FP64 against the oracle;
  ensemble                   48 MHD boxes at 64^3 + 16 at 128^3 + 32 Boussinesq-MHD
                             boxes at 512x512x2 with RKL2 conduction, FP32 bulk,
                             one batch per shape (spk_ensemble): steps of every box
                             per second, algorithmic GB/s, every box against an
                             FP64 ensemble and one box per kind against the oracle;
  closure                    learned subgrid closure, shortened: a 512^3 forced DNS
                             spun up to t = 0.5, 6 filtered samples, one network
                             trained per seed (--closure-seeds), then the 128^3 run
                             with each closure, the plain 128^3 run and the DNS over
                             0.25 time units: shell-spectrum error up to kc/2 and
                             wall-time ratio (--no-closure skips it).
"""
import argparse
import importlib
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="oracle steps of the CPU baseline, also compared with the GPU path (accuracy_vs_oracle)")
    ap.add_argument("--precisions", default="f64,f32")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the FP32-bulk vs FP64 accuracy run (0: skip)")
    ap.add_argument("--mti-steps", type=int, default=10)
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the Taylor-Green 64^3 and MTI legs")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-acc-steps", type=int, default=10,
                    help="steps of the FP32-vs-FP64 ensemble accuracy check (0: skip)")
    ap.add_argument("--ensemble-split", type=int, default=-1,
                    help="batches per shape group (0: box by box; -1: autotuned)")
    ap.add_argument("--ensemble-group", default="all", choices=["all", "mhd64^3", "mhd128^3", "hbi512x512x2"],
                    help="run one shape group of the ensemble alone (--mode ensemble)")
    ap.add_argument("--closure-spinup", type=float, default=2.0)
    ap.add_argument("--closure-iters", type=int, default=3000)
    ap.add_argument("--closure-dns", type=int, default=512)
    ap.add_argument("--closure-samples", type=int, default=12)
    ap.add_argument("--closure-sample-dt", type=float, default=0.1)
    ap.add_argument("--closure-dT", type=float, default=0.5)
    ap.add_argument("--closure-seeds", default="7,8", help="network initializations, one trained closure each")
    ap.add_argument("--no-closure", dest="closure", action="store_false",
                    help="skip the closure leg of the default run (a shortened DNS: spin-up to t = 0.5, "
                         "6 samples 0.05 apart, 1500 training iterations per seed, 0.25 time units compared)")
    ap.add_argument("--mode", choices=["full", "passes", "kernels", "mti", "hbi2d", "ensemble", "closure",
                                       "closure-kernels"],
                    default="full",
                    help="passes: only the FFT pass kernels and a copy kernel; kernels: each pipeline "
                         "kernel of the 256^3 FP64 step twice, plus the copy kernel (both for counter runs); "
                         "mti: only the MTI leg (--mti-steps steps); hbi2d: only the problems/hbi2d leg")
    return ap.parse_args()


def mhd_params(n):
    import math
    return dict(nx=n, ny=n, nz=n, incompressible=1, mhd=1, lx=2 * math.pi, ly=2 * math.pi, lz=2 * math.pi,
                nu=1e-3, nu_m=1e-3, cfl=0.5, t_final=1e9, toutput_flow=1e9)


def time_steps(spk, torch, ctx, state, prec, n, steps, warmup):
    sol = spk.Solver(ctx, prec, **mhd_params(n))
    sol.set_fields(state)
    fused_used = sol.pipeline()
    sol.main_loop(warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    sol.main_loop(steps)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    t = sol.time()
    ev = sum(sol.energy(f) for f in range(3))
    em = sum(sol.energy(f) for f in range(3, 6))
    sol.destroy()
    return ms, t, ev + em, fused_used


def time_fft_passes(spk, torch, ctx, prec, n, reps=20):
    """Average duration of each FFT pass kernel, launched exactly as in the RK3
    step (8 product fields per launch), with HIP events on the context stream."""
    esz = 16 if prec == "f64" else 8
    nzh = n // 2 + 1
    nf = 8
    field_bytes = n * n * nzh * esz  # S_w
    plan = spk.FftPlan(ctx, n, n, n, prec, nf)
    buf = torch.empty(nf * field_bytes // 4, dtype=torch.float32, device="cuda")
    buf.normal_()
    p = buf.data_ptr()
    out = {}
    for name, axis, d in (("x_fwd", 0, -1), ("y_fwd", 1, -1), ("z_r2c", 2, -1), ("x_inv", 0, 1), ("y_inv", 1, 1),
                          ("z_c2r", 2, 1)):
        for _ in range(3):
            plan.one_pass(axis, d, p, p)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.one_pass(axis, d, p, p)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        algo = nf * 2 * field_bytes  # read + write every field once
        out[name] = dict(ms=ms, bytes=algo, gbs=algo / (ms * 1e-3) / 1e9)
    plan.destroy()
    del buf
    return out


KERNEL_NAMES = {
    "inv_x": "k_inv_x<double,256> (inverse x pass from the FP64 state, 6 fields)",
    "inv_y": "k_pass_strided<double,256,+1,1> (inverse y pass, 6 fields)",
    "zmid": "k_zmid<double,256,U|B> (z C2R of 6 fields, products, z R2C of 8)",
    "fwd_y": "k_pass_strided<double,256,-1,1> (forward y pass, 8 fields)",
    "fwd_x": "k_pass_strided<double,256,-1,0> (forward x pass, 8 fields)",
    "epilogue": "k_epilogue<double,MHD,stage 2> (RHS + RK3 stage update)",
    "clean_div": "k_clean_div (CleanFieldDivergence)",
    "fwdx_epilogue": "k_fwdx_epilogue<double,256,MHD,stage 2> (forward x pass + RHS + RK3 update + next inverse x)",
}


def kernel_table(spk, ctx, state, n, prec, fused, reps=10):
    """Live per-kernel average launch time (HIP events, spk_kernel_timings) and
    algorithmic HBM bytes per launch for the MHD step (6 state fields, 8 products);
    fused: the step runs the fused forward-x/epilogue kernel (autotune's choice)."""
    sol = spk.Solver(ctx, prec, **mhd_params(n))
    sol.set_fields(state)
    sol.main_loop(1)
    t = sol.kernel_timings(reps)
    sol.destroy()
    s64 = n * n * (n // 2 + 1) * 16
    sb = s64 if prec == "f64" else s64 // 2
    nin, nout, nf = 6, 8, 6
    algo = {"inv_x": nin * (s64 + sb), "inv_y": nin * 2 * sb, "zmid": (nin + nout) * sb, "fwd_y": nout * 2 * sb,
            "fwd_x": nout * 2 * sb, "epilogue": nout * sb + 4 * nf * s64, "clean_div": 2 * nf * s64,
            # stage 2: products in, U0 and the increment register in, U2 and the new increment out
            # (the register holds the bulk precision), next inverse x pass out
            "fwdx_epilogue": nout * sb + 2 * nf * s64 + 2 * nf * sb + nin * sb}
    if fused:
        # fused pipeline: forward x + epilogue + projection + next inverse x in one kernel
        per_step = {"inv_y": 3, "zmid": 3, "fwd_y": 3, "fwdx_epilogue": 3}
    else:
        per_step = {"inv_x": 3, "inv_y": 3, "zmid": 3, "fwd_y": 3, "fwd_x": 3, "epilogue": 3, "clean_div": 1}
    algo = {k: v for k, v in algo.items() if k in per_step}
    out = {}
    for k, b in algo.items():
        ms = t[k]
        out[k] = dict(ms=ms, bytes=b, gbs=b / (ms * 1e-3) / 1e9, per_step=per_step[k])
    tot = sum(v["ms"] * v["per_step"] for v in out.values())
    for v in out.values():
        v["share"] = v["ms"] * v["per_step"] / tot
    return out


def copy_gbs(torch, nbytes):
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    return 2 * a.nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(n, state, steps):
    """The FP64 oracle timed over `steps` steps of the same workload; returns the
    baseline record and the oracle's state and energy after those steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    sim = orc.Sim(**mhd_params(n))
    sim.set_fields(state)
    t0 = time.perf_counter()
    sim.main_loop(steps)
    dt = time.perf_counter() - t0
    cores = orc.lib().orc_num_threads()
    rec = dict(value=steps / dt, unit="steps/s", cores=cores, kind="port",
               sample="%d MHD RK3 step(s) at %d^3, FP64 oracle (C, OpenMP), same initial state" % (steps, n),
               seconds=dt)
    out = dict(fields=sim.fields.copy(), energy=sum(sim.energy(f) for f in range(6)), t=sim.time()["t"])
    return rec, out


def accuracy_vs_oracle(spk, ctx, state, n, steps, orc_out):
    """The GPU path (FP64 and FP32 bulk, the back end the benchmark runs) against
    the FP64 oracle after the same `steps` steps from the same state: full-state
    relative L2 and relative total-energy error."""
    import numpy as np
    res = dict(steps=steps)
    ref = orc_out["fields"]
    nref = float(np.linalg.norm(ref.ravel()))
    for prec in ("f64", "f32"):
        sol = spk.Solver(ctx, prec, **mhd_params(n))
        sol.set_fields(state)
        sol.main_loop(steps)
        got = sol.get_fields()
        e = sum(sol.energy(f) for f in range(6))
        res[prec] = dict(rel_l2_state=float(np.linalg.norm((got - ref).ravel())) / nref,
                         rel_energy_err=abs(e - orc_out["energy"]) / abs(orc_out["energy"]),
                         rel_time_err=abs(sol.time()["t"] - orc_out["t"]) / orc_out["t"])
        sol.destroy()
        del got
    return res


def rkl2_stages(dt, dt_sts):
    """RKL2 stage count of one cycle for dt / dt_sts (rkl.cu:466-470)."""
    import math
    if not (dt_sts > 0 and math.isfinite(dt_sts)):
        return 0
    r = dt / dt_sts
    return 1 + int(4.0 * (1.0 + 2.0 * r) / (1.0 + math.sqrt(9.0 + 16.0 * r)))


def fp32_accuracy(spk, ctx, state, n, steps):
    """FP32-bulk against FP64 after `steps` main-loop steps from the same state:
    relative errors of total energy ev+em, cross helicity <u.B> and the six
    shell spectra Kx..Mz (largest relative error over bins holding >= 1e-6
    of the spectrum's peak)."""
    out = {}
    for prec in ("f64", "f32"):
        sol = spk.Solver(ctx, prec, **mhd_params(n))
        sol.set_fields(state)
        sol.main_loop(steps)
        e = sum(sol.energy(f) for f in range(6))
        sol.complex2real()
        hc = sum(sol.correlation(i, 3 + i) for i in range(3))
        spec = [sol.spectrum(f, f) for f in range(6)]
        out[prec] = dict(e=e, hc=hc, spec=spec, t=sol.time()["t"])
        sol.destroy()
    a, b = out["f64"], out["f32"]
    serr = 0.0
    for sa, sb in zip(a["spec"], b["spec"]):
        keep = sa > 1e-6 * sa.max()
        serr = max(serr, float((abs(sb[keep] - sa[keep]) / sa[keep]).max()))
    return dict(steps=steps, rel_energy_err=abs(b["e"] - a["e"]) / abs(a["e"]),
                rel_cross_helicity_err=abs(b["hc"] - a["hc"]) / max(abs(a["hc"]), 1e-300),
                spectra_max_rel_err=serr, t_f64=a["t"], t_f32=b["t"])


def taylor_green_leg(spk, ini, torch, ctx, cpu):
    """64^3 HD Taylor-Green, nu = 0.01, CFL 0.5, 100 steps on the GPU (timed)
    and on the FP64 oracle; kinetic energy and dissipation 2 nu <w^2> compared."""
    import math
    n, steps = 64, 100
    p = dict(nx=n, ny=n, nz=n, incompressible=1, lx=2 * math.pi, ly=2 * math.pi, lz=2 * math.pi, nu=0.01,
             cfl=0.5, t_final=1e9, toutput_flow=1e9)
    u0 = ini.taylor_green(n)
    sol = spk.Solver(ctx, "f64", **p)
    sol.set_fields(u0)
    sol.main_loop(2)
    sol.set_fields(u0)
    sol.set_time(0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.main_loop(steps)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    ev = sum(sol.energy(f) for f in range(3))
    eps = 2 * 0.01 * sol.enstrophy(0)
    sol.destroy()
    res = dict(steps=steps, gpu_steps_per_s=steps / gpu_s, ev=ev, eps=eps)
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle as orc
        sim = orc.Sim(**p)
        sim.set_fields(u0)
        t0 = time.perf_counter()
        sim.main_loop(steps)
        cs = time.perf_counter() - t0
        ev_r = sum(sim.energy(f) for f in range(3))
        eps_r = 2 * 0.01 * sim.enstrophy(0)
        res.update(cpu_steps_per_s=steps / cs, cpu_cores=orc.lib().orc_num_threads(),
                   rel_err_ev=abs(ev - ev_r) / ev_r, rel_err_eps=abs(eps - eps_r) / eps_r)
    return res


def mti_leg(spk, ini, torch, ctx, n, steps, warmup, cpu):
    """256^3 Boussinesq MTI with anisotropic conduction, FP64, RKL2 on theta
    (see spooky-git_amd.initial.mti_params): steps/s with HIP events, and the
    RKL2 stage count of the run's dt / dt_sts; the same run with FP32 bulk
    (transforms and products in FP32, the RK3 state and the RKL2 recurrence in
    FP64): steps/s and each field group's relative L2 difference from the FP64
    state after the same steps; the FP64 oracle times one step."""
    import numpy as np
    p = ini.mti_params(n)
    state = ini.mti_white_noise((n, n, n), seed=7)
    out = {}
    for prec in ("f64", "f32"):
        sol = spk.Solver(ctx, prec, **p)
        sol.set_fields(state)
        sol.main_loop(warmup)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        sol.main_loop(steps)
        e1.record()
        torch.cuda.synchronize()
        out[prec] = dict(ms=e0.elapsed_time(e1) / steps, time=sol.time(), fields=sol.get_fields(), vars=sol.vars)
        sol.destroy()
    ms, tm, v = out["f64"]["ms"], out["f64"]["time"], out["f64"]["vars"]
    a, b = out["f32"]["fields"], out["f64"]["fields"]
    groups = {"u": slice(v["VX"], v["VX"] + 3), "B": slice(v["BX"], v["BX"] + 3), "theta": slice(v["TH"], v["TH"] + 1)}
    diff = {g: float(np.linalg.norm((a[sl] - b[sl]).ravel()) / max(np.linalg.norm(b[sl].ravel()), 1e-300))
            for g, sl in groups.items()}
    res = dict(grid=[n, n, n], steps=steps, ms_per_step=ms, steps_per_s=1000.0 / ms, dt=tm["dt"],
               dt_sts=tm["dt_sts"], rkl2_stages=rkl2_stages(tm["dt"], tm["dt_sts"]),
               fp32_bulk=dict(ms_per_step=out["f32"]["ms"], steps_per_s=1000.0 / out["f32"]["ms"],
                              ratio_vs_f64=ms / out["f32"]["ms"], steps_compared=warmup + steps,
                              rel_l2_vs_f64=diff),
               params={k: p[k] for k in ("nu", "nu_m", "nu_th", "reynolds_ani", "N2", "OmegaT2", "lx")},
               ic="u=0, B=(1e-4,0,0), theta white noise uniform in [-1e-4,1e-4], seed 7")
    del out, a, b
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle as orc
        sim = orc.Sim(**p)
        sim.set_fields(state)
        t0 = time.perf_counter()
        sim.main_loop(1)
        res.update(cpu_s_per_step=time.perf_counter() - t0, cpu_cores=orc.lib().orc_num_threads())
    return res


def hbi2d_leg(spk, ini, torch, ctx, steps, warmup, cpu):
    """problems/hbi2d at its own size (512 x 512 x 2, Boussinesq MHD with
    stratification along y and anisotropic conduction advanced by RKL2 on theta;
    spooky-git_amd.initial.hbi_params / hbi_noise): steps/s of one box in each
    precision with the quasi-2D plane kernels and without them (the separate y
    passes and z kernel), and the FP64 state against the oracle after 2 steps."""
    import numpy as np
    n = 512
    p = ini.hbi_params(n)
    state = ini.hbi_noise((n, n, 2), seed=11)
    res = dict(grid=[n, n, 2], steps=steps)
    for prec in ("f64", "f32"):
        for plane in (True, False):
            sol = spk.Solver(ctx, prec, **p)
            sol.set_plane(plane)
            sol.set_fields(state)
            sol.main_loop(warmup)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            sol.main_loop(steps)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            key = prec + ("_plane" if plane else "_three_kernels")
            res[key] = dict(ms_per_step=round(ms, 4), steps_per_s=round(1000.0 / ms, 1), plane=sol.plane())
            if plane and prec == "f64":
                tm = sol.time()
                res["rkl2_stages"] = rkl2_stages(tm["dt"], tm["dt_sts"])
            sol.destroy()
    res["plane_speedup"] = {pr: round(res[pr + "_three_kernels"]["ms_per_step"] / res[pr + "_plane"]["ms_per_step"], 3)
                            for pr in ("f64", "f32")}
    if cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle as orc
        sol = spk.Solver(ctx, "f64", **p)
        sol.set_fields(state)
        sol.main_loop(2)
        sim = orc.Sim(**p)
        sim.set_fields(state)
        t0 = time.perf_counter()
        sim.main_loop(2)
        res["cpu_s_per_step"] = round((time.perf_counter() - t0) / 2, 3)
        r = sim.fields
        res["f64_rel_l2_vs_oracle_2_steps"] = float(np.linalg.norm(sol.get_fields() - r) / np.linalg.norm(r))
        sol.destroy()
    return res


def ensemble_boxes(ini, seed=2026):
    """The batched-ensemble workload: 48 MHD boxes at 64^3, 16 at 128^3 (2 pi
    boxes, CFL 0.5) and 32 Boussinesq-MHD boxes at 512x512x2 with anisotropic
    conduction and RKL2 on theta (problems/hbi2d physics); per box nu, eta
    log-uniform in [1e-4, 1e-2] and chi_par uniform in [1e-3, 1e-1]."""
    import numpy as np
    rng = np.random.default_rng(seed)
    boxes, kinds = [], []
    for count, n in ((48, 64), (16, 128)):
        for _ in range(count):
            nu, eta = 10.0 ** rng.uniform(-4, -2, 2)
            boxes.append(dict(mhd_params(n), nu=float(nu), nu_m=float(eta)))
            kinds.append("mhd%d^3" % n)
    for _ in range(32):
        nu, eta = 10.0 ** rng.uniform(-4, -2, 2)
        chi = rng.uniform(1e-3, 1e-1)
        boxes.append(ini.hbi_params(512, nu=float(nu), eta=float(eta), chi=float(chi)))
        kinds.append("hbi512x512x2")
    return boxes, kinds


def box_state(ini, p, seed):
    shape = (p["nx"], p["ny"], p["nz"])
    if p.get("boussinesq"):
        return ini.hbi_noise(shape, seed=seed)
    return ini.mhd_random(shape, seed=seed)


def ensemble_bytes(p, bulk_bytes):
    """Algorithmic HBM bytes of one main-loop step of a box, counted as for the
    256^3 step: every 3D transform one sweep of its fields (2 S_bulk per field,
    read + write) and the FP64 RK3 state traffic (U read/write, increment
    register) of the fused back end; RKL2 stages as 6 transforms + 6 S_64 of
    state traffic + 3 S_bulk of real-space B per stage (taken from the box's own
    stage count)."""
    nzh = p["nz"] // 2 + 1
    ntc = p["nx"] * p["ny"] * nzh
    s64, sb = 16 * ntc, bulk_bytes * ntc
    nf = 3 + (3 if p.get("mhd") else 0) + (1 if p.get("boussinesq") else 0)
    nin = nf
    nout = 5 + (3 if p.get("mhd") else 0) + (3 if p.get("boussinesq") else 0)
    transforms = 3 * (nin + nout)
    return 2 * transforms * sb + 7 * nf * s64, 6 * 2 * sb + 6 * s64 + 3 * sb


def ensemble_leg(spk, ini, ctx, steps, acc_steps=10, split=-1, group="all"):
    """The ensemble in FP32 bulk.  Boxes of one shape run as one batch (one
    launch per kernel for all of them); the batch split is autotuned, then
    `steps` steps of every box are timed (host clock around run + sync) with the
    graph, without it, and box by box over 16 lanes (the unbatched launch).
    Accuracy: from the same initial states, an FP32-bulk and an FP64-bulk
    ensemble advance acc_steps steps and every box's state and energy are
    compared (max over the boxes of each kind); the FP64 ensemble's first box of
    each kind is checked against the FP64 oracle."""
    import numpy as np
    boxes, kinds = ensemble_boxes(ini)
    if group != "all":
        keep = [i for i, k in enumerate(kinds) if k == group]
        boxes, kinds = [boxes[i] for i in keep], [kinds[i] for i in keep]
    t0 = time.perf_counter()
    ens = spk.Ensemble(boxes, "f32")
    states = []
    for i, p in enumerate(boxes):
        states.append(box_state(ini, p, 5000 + i))
        ens.set_fields(i, states[-1])
    setup_s = time.perf_counter() - t0
    print("# ensemble: %d boxes created and initialized in %.1f s" % (len(boxes), setup_s), file=sys.stderr,
          flush=True)
    if split < 0:
        split, _ = ens.tune(probe=2)
    else:
        ens.set_split(split)
    layout = ens.layout()
    print("# ensemble: %s after tuning" % layout, file=sys.stderr, flush=True)

    def timed(n):
        ens.run(2)
        ens.sync()
        t1 = time.perf_counter()
        ens.run(n)
        ens.sync()
        return (time.perf_counter() - t1) * 1000.0 / n

    graph_ms = timed(steps)
    hbi = [i for i, k in enumerate(kinds) if k.startswith("hbi")]
    stages = [rkl2_stages(ens.time(i)["dt"], ens.time(i)["dt_sts"]) for i in hbi]
    ens.set_graph(False)
    eager_ms = timed(steps)
    ens.set_graph(True)
    box_ms = None
    if split != 0:
        ens.set_lanes(16)
        box_ms = timed(steps)
    print("# ensemble: %.2f ms per step of every box (%s, graph); %.2f without the graph; %s box by box over "
          "16 lanes (graph)" % (graph_ms, layout, eager_ms, box_ms), file=sys.stderr, flush=True)
    ens.destroy()
    # algorithmic bytes per step of every box (RKL2 stage count of each HBI box at the timed point)
    step_b = 0.0
    st_iter = iter(stages)
    for p in boxes:
        b, per_stage = ensemble_bytes(p, 4)
        step_b += b + (per_stage * next(st_iter) if p.get("supertimestepping") else 0.0)
    cells = sum(p["nx"] * p["ny"] * p["nz"] for p in boxes)

    # accuracy: FP32 bulk against FP64 bulk, box by box, and FP64 against the oracle
    if acc_steps <= 0:
        return dict(boxes={k: kinds.count(k) for k in dict.fromkeys(kinds)}, layout=layout,
                    ms_per_ensemble_step=round(graph_ms, 3), algorithmic_gb_per_ensemble_step=round(step_b / 1e9, 3),
                    achieved_gbs=round(step_b / (graph_ms * 1e-3) / 1e9, 1))
    e32 = spk.Ensemble(boxes, "f32")
    e64 = spk.Ensemble(boxes, "f64")
    for i, st in enumerate(states):
        e32.set_fields(i, st)
        e64.set_fields(i, st)
    e32.run(acc_steps)
    e64.run(acc_steps)
    acc = {}
    for i, k in enumerate(kinds):
        a, r = e32.get_fields(i), e64.get_fields(i)
        ea = sum(e32.boxes[i].energy(f) for f in range(e32.boxes[i].nf))
        er = sum(e64.boxes[i].energy(f) for f in range(e64.boxes[i].nf))
        d = acc.setdefault(k, dict(boxes=0, max_rel_l2_state=0.0, max_rel_energy_err=0.0))
        d["boxes"] += 1
        d["max_rel_l2_state"] = max(d["max_rel_l2_state"], float(np.linalg.norm(a - r) / np.linalg.norm(r)))
        d["max_rel_energy_err"] = max(d["max_rel_energy_err"], abs(ea - er) / abs(er))
    oracle = {}
    try:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import pyoracle as orc
        for k in dict.fromkeys(kinds):
            i = kinds.index(k)
            sim = orc.Sim(**boxes[i])
            sim.set_fields(states[i])
            sim.main_loop(acc_steps)
            r = sim.fields
            oracle[k] = dict(box=i, rel_l2_state_f64=float(np.linalg.norm(e64.get_fields(i) - r) / np.linalg.norm(r)))
    except OSError as ex:  # the oracle library is not built on this host
        oracle = {"skipped": str(ex)}
    e32.destroy()
    e64.destroy()
    return dict(boxes={k: kinds.count(k) for k in dict.fromkeys(kinds)}, bulk="f32", layout=layout, steps=steps,
                ms_per_ensemble_step=round(graph_ms, 3), box_steps_per_s=round(len(boxes) * 1000.0 / graph_ms, 1),
                cell_updates_per_s=cells * 1000.0 / graph_ms,
                algorithmic_gb_per_ensemble_step=round(step_b / 1e9, 3),
                achieved_gbs=round(step_b / (graph_ms * 1e-3) / 1e9, 1),
                frac_of_hbm_peak=round(step_b / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                eager_batched_ms_per_step=round(eager_ms, 3), box_by_box_16_lanes_ms_per_step=None if box_ms is None else round(box_ms, 3),
                rkl2_stages_hbi=[min(stages), max(stages)] if stages else None, setup_s=round(setup_s, 1),
                accuracy_f32_vs_f64=dict(steps=acc_steps, per_kind=acc), f64_vs_oracle=oracle,
                params="nu, eta log-uniform [1e-4, 1e-2]; chi_par uniform [1e-3, 1e-1] (hbi); seeds 5000+i")


def forced_modes(n, L=2 * 3.141592653589793, kf=2.0, f0=0.3, seed=11):
    """Large-scale forcing field on an n^3 grid: a solenoidal random field on the
    modes 0 < |k| <= kf (integer k in a 2 pi box), real-space rms f0, seeded so
    that every grid gets the same physical field.  Returns (indices, values)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kmax = int(kf)
    ks = [(a, b, c) for a in range(-kmax, kmax + 1) for b in range(-kmax, kmax + 1) for c in range(0, kmax + 1)
          if 0 < a * a + b * b + c * c <= kf * kf]
    nzh = n // 2 + 1
    F = np.zeros((3, n, n, nzh), np.complex128)
    for (a, b, c) in ks:
        v = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        k = np.array([a, b, c], float)
        v -= k * (k @ v) / (k @ k)
        F[:, a % n, b % n, c] = v
    r = np.stack([np.fft.irfftn(f, s=(n, n, n), axes=(0, 1, 2)) for f in F])  # Hermitian-consistent, real
    r *= f0 / np.sqrt(np.mean(np.sum(r ** 2, axis=0)))
    F = np.stack([np.fft.rfftn(x) for x in r])
    idx = np.array([((a % n) * n + (b % n)) * nzh + c for (a, b, c) in ks], dtype=np.int64)
    idx = np.unique(np.concatenate([idx, [((-a % n) * n + (-b % n)) * nzh for (a, b, c) in ks if c == 0]]))
    flat = F.reshape(3, -1)
    return idx, flat[:, idx].T.copy()


def shell_energy(sol):
    """E(k) of the velocity (sum of the three component spectra)."""
    v = sol.vars
    return sum(sol.spectrum(v["VX"] + i, v["VX"] + i) for i in range(3))


def closure_leg(spk, ini, ctx, torch, N=512, t_spin=1.0, nsamples=6, dt_sample=0.05, dT=0.5, iters=2000,
                lr=1e-3, seeds=(7,)):
    """Learned subgrid closure for forced HD turbulence: a 512^3 FP32-bulk DNS
    (nu = 2e-4, forcing at |k| <= 2, random-phase E(k) ~ k^4 exp(-2 (k/4)^2)
    start) gives filtered velocity / subgrid stress pairs on 128^3; the network is
    trained on them; then from the filtered DNS state the 128^3 run with the
    closure, the plain 128^3 run and the DNS advance by dT, and their shell
    spectra and wall times are compared (bins 1 .. kc/2, kc = the 128^3 cut).
    One network is trained and run per seed (its initial weights); the summary
    gives the mean and the worst over the seeds."""
    import numpy as np
    T = 2 * np.pi
    n = N // 4
    t0 = t_spin + (nsamples - 1) * dt_sample
    t_end = t0 + dT
    base = dict(incompressible=1, nu=2e-4, cfl=0.5, lx=T, ly=T, lz=T, toutput_flow=1e9)
    dns = spk.Solver(ctx, "f32", **dict(base, nx=N, ny=N, nz=N, t_final=t_end))
    dns.set_fields(ini.random_solenoidal((N, N, N), (T, T, T), seed=2027, k0=4.0))
    idx, val = forced_modes(N)
    dns.set_forcing(idx, val)
    res = dict(dns_grid=N, les_grid=n, nu=2e-4, forcing="solenoidal random field on 0<|k|<=2, rms 0.3, fixed",
               ic="random-phase solenoidal, E(k)~k^4 exp(-2(k/4)^2), rms 1, seed 2027")

    def run_until(sol, t, after=None):
        steps, last = 0, time.perf_counter()
        while sol.time()["t"] < t - 1e-12:
            sol.main_loop(1)
            if after:
                after(sol)
            steps += 1
            if time.perf_counter() - last > 30:
                last = time.perf_counter()
                print("# closure: t = %.4f after %d steps" % (sol.time()["t"], steps), file=sys.stderr, flush=True)
        return steps

    def clock():
        torch.cuda.synchronize()
        return time.perf_counter()

    tic = clock()
    res["spinup_steps"] = run_until(dns, t_spin)
    res["spinup_s"] = round(clock() - tic, 2)
    print("# closure: DNS %d^3 spun up to t=%.2f in %d steps (%.1f s)" % (N, t_spin, res["spinup_steps"],
          res["spinup_s"]), file=sys.stderr, flush=True)
    fs = n * n * (n + 2)
    samples = []
    for s in range(nsamples):
        if s:
            run_until(dns, t_spin + s * dt_sample)
        ub, tau = ctx.alloc(3 * fs * 8), ctx.alloc(6 * fs * 8)
        spk.sgs_stress(ctx, dns.d_fields.value, N, ub.ptr, tau.ptr, n)
        samples.append((ub, tau))
    ubh = [u.download(np.float64, (3, n, n, n + 2))[..., :n] for u, _ in samples]
    tah = [t.download(np.float64, (6, n, n, n + 2))[..., :n] for _, t in samples]
    u_rms = float(np.sqrt(np.mean([np.mean(np.sum(u ** 2, 0)) for u in ubh])))
    tau_rms = [float(np.sqrt(np.mean([np.mean(t[c] ** 2) for t in tah]))) for c in range(6)]
    les_p = dict(base, nx=n, ny=n, nz=n, t_final=t_end)
    uhat0 = np.stack([np.fft.rfftn(u) for u in ubh[-1]])
    lidx, lval = forced_modes(n)
    dtau = ctx.alloc(6 * fs * 8)

    def les_run(cl):
        """the 128^3 run from the filtered DNS state at t0, with the closure cl or plain"""
        les = spk.Solver(ctx, "f32", **les_p)
        les.set_fields(uhat0)
        les.set_time(t0)
        les.set_forcing(lidx, lval)
        tic = clock()
        steps = run_until(les, t_end, (lambda s_: cl.apply(s_)) if cl is not None else None)
        wall = clock() - tic
        E = shell_energy(les)
        les.destroy()
        return dict(steps=steps, wall_s=wall, E=E, finite=bool(np.all(np.isfinite(E))))

    runs = {"plain": les_run(None)}
    per_seed = []
    for seed in seeds:
        cl = spk.Closure(ctx, n)
        rng = np.random.default_rng(seed)
        p0 = np.zeros(cl.nparams, np.float32)
        for l, ((ws, bsl), cin) in enumerate(zip(spk.Closure.layer_slices(), spk.Closure.CIN)):
            w = rng.standard_normal((27, 32, cin)) / np.sqrt(27 * (3 if cin == 16 else cin))
            if cin == 16:
                w[..., 3:] = 0
            if l == 3:
                w[:, 6:, :] = 0
                w *= 0.1
            p0[ws] = w.reshape(-1)
        cl.set_params(p0)
        cl.set_scales(1.0 / u_rms, tau_rms)
        curve = []
        train_s = 0.0
        first = None
        for it in range(iters):
            rate = lr if it < 0.6 * iters else 0.3 * lr
            ub, tau = samples[it % (nsamples - 1)]  # the last sample is held out
            tic = clock()
            loss = cl.train_step(ub.ptr, tau.ptr, rate, want_loss=(it == 0))
            train_s += clock() - tic
            if it == 0:
                first = loss
            if (it + 1) % max(1, iters // 4) == 0 or it == iters - 1:
                tr = float(np.mean([cl.grads(u_.ptr, t_.ptr) for u_, t_ in samples[:-1]]))
                held = cl.grads(samples[-1][0].ptr, samples[-1][1].ptr)
                curve.append([it + 1, round(tr, 4), round(held, 4)])
                print("# closure seed %d: iteration %d: training loss %.4f, held-out %.4f" % (seed, it + 1, tr, held),
                      file=sys.stderr, flush=True)
        if seed == seeds[0]:
            # inference time on 128^3
            cl.forward(samples[-1][0].ptr, dtau.ptr)
            tic = clock()
            for _ in range(10):
                cl.forward(samples[-1][0].ptr, dtau.ptr)
            res["ms_per_inference"] = round((clock() - tic) * 100, 3)
        r = les_run(cl)
        runs["closure_seed%d" % seed] = r
        per_seed.append(dict(seed=seed, train_iters=iters, train_s=round(train_s, 2),
                             ms_per_train_iter=round(1000 * train_s / iters, 3), train_loss_first=first,
                             train_loss_last=curve[-1][1], heldout_loss=curve[-1][2], loss_curve=curve))
        cl.destroy()
    res.update(u_rms_filtered=u_rms, tau_rms=tau_rms, training=per_seed)
    tic = clock()
    dns_steps = run_until(dns, t_end)
    dns_wall = clock() - tic
    # filtered DNS spectrum at t_end on the 128^3 grid
    ub, tau = ctx.alloc(3 * fs * 8), ctx.alloc(6 * fs * 8)
    spk.sgs_stress(ctx, dns.d_fields.value, N, ub.ptr, tau.ptr, n)
    fl = spk.Solver(ctx, "f64", **les_p)
    fl.set_fields(np.stack([np.fft.rfftn(u) for u in ub.download(np.float64, (3, n, n, n + 2))[..., :n]]))
    E_dns = shell_energy(fl)
    fl.destroy()
    kc = int(2.0 / 3.0 * (n / 2 - 1))
    bins = np.arange(1, kc // 2 + 1)
    out = {}
    for name, r in runs.items():
        err = np.abs(r["E"][bins] / E_dns[bins] - 1.0) if r["finite"] else np.full(len(bins), np.inf)
        out[name] = dict(steps=r["steps"], wall_s=round(r["wall_s"], 3), max_rel_spectrum_err=float(np.max(err)),
                         mean_rel_spectrum_err=float(np.mean(err)), finite=r["finite"],
                         speedup_vs_dns=round(dns_wall / r["wall_s"], 1))
        print("# closure: %s %d^3: %d steps %.2f s, spectrum error max %.3f mean %.3f (bins 1..%d)" %
              (name, n, r["steps"], r["wall_s"], float(np.max(err)), float(np.mean(err)), kc // 2),
              file=sys.stderr, flush=True)
    cls = [out["closure_seed%d" % sd] for sd in seeds]
    summary = dict(seeds=list(seeds),
                   closure_max_err_mean_over_seeds=float(np.mean([c["max_rel_spectrum_err"] for c in cls])),
                   closure_max_err_worst_seed=float(np.max([c["max_rel_spectrum_err"] for c in cls])),
                   closure_mean_err_mean_over_seeds=float(np.mean([c["mean_rel_spectrum_err"] for c in cls])),
                   plain_max_err=out["plain"]["max_rel_spectrum_err"],
                   plain_mean_err=out["plain"]["mean_rel_spectrum_err"],
                   closure_speedup_vs_dns_min=float(np.min([c["speedup_vs_dns"] for c in cls])),
                   target="closure run within 10% of the filtered DNS spectrum up to kc/2 at >= 20x less wall time")
    print("# closure summary: %s" % json.dumps(summary), file=sys.stderr, flush=True)
    res.update(t0=t0, dT=dT, dns_steps_dT=dns_steps, dns_wall_s_dT=round(dns_wall, 2), compare_bins=[1, kc // 2],
               les=out, summary=summary, E_dns_filtered=[float(x) for x in E_dns[: kc + 1]],
               E_les_closure=[float(x) for x in runs["closure_seed%d" % seeds[0]]["E"][: kc + 1]],
               E_les_plain=[float(x) for x in runs["plain"]["E"][: kc + 1]])
    dns.destroy()
    return res


def closure_kernels(spk, ctx, torch, n=128, reps=20):
    """The closure's kernels alone on a 128^3 grid: inference and training
    iterations timed with HIP events, and the matrix-core rate of one inference
    (algorithmic flops: 2 P 27 (3*32 + 32*32 + 32*32 + 32*6))."""
    import numpy as np
    rng = np.random.default_rng(5)
    cl = spk.Closure(ctx, n)
    p = (rng.standard_normal(cl.nparams) * 0.05).astype(np.float32)
    cl.set_params(p)
    cl.set_scales(1.0, [1.0] * 6)
    fs = n * n * (n + 2)
    u = ctx.alloc(3 * fs * 8)
    u.upload(rng.standard_normal(3 * fs))
    tau = ctx.alloc(6 * fs * 8)
    tau.upload(rng.standard_normal(6 * fs) * 0.1)
    cl.forward(u.ptr, tau.ptr)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        cl.forward(u.ptr, tau.ptr)
    e1.record()
    torch.cuda.synchronize()
    inf_ms = e0.elapsed_time(e1) / reps
    cl.train_step(u.ptr, tau.ptr, 1e-4, want_loss=False)
    e0.record()
    for _ in range(reps):
        cl.train_step(u.ptr, tau.ptr, 1e-4, want_loss=False)
    e1.record()
    torch.cuda.synchronize()
    tr_ms = e0.elapsed_time(e1) / reps
    P = n ** 3
    flops = 2.0 * P * 27 * (3 * 32 + 32 * 32 + 32 * 32 + 32 * 6)
    cl.destroy()
    return dict(grid=n, inference_ms=round(inf_ms, 4), train_iter_ms=round(tr_ms, 4),
                inference_tflops=round(flops / inf_ms / 1e9, 1), inference_frac_of_bf16_peak=round(
                    flops / inf_ms / 1e9 / 2500.0, 4))


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: the solver has no multi-GPU path; its sharding is a design "
              "(slab decomposition with one all-to-all per 3D transform) (--gpus %d refused)" % a.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    spk = importlib.import_module("spooky-git_amd")
    ini = importlib.import_module("spooky-git_amd.initial")
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    ctx = spk.Ctx(0, stream.cuda_stream)
    n = a.n
    if a.mode == "passes":
        passes = time_fft_passes(spk, torch, ctx, "f64", n, reps=5)
        copy = copy_gbs(torch, n * n * (n // 2 + 1) * 16 * 8)
        print(json.dumps({"fft_pass": passes, "copy_gbs": copy, "copy_bytes": n * n * (n // 2 + 1) * 16 * 8}))
        return
    if a.mode == "mti":
        print(json.dumps(mti_leg(spk, ini, torch, ctx, n, a.mti_steps, 2, not a.no_cpu)))
        return
    if a.mode == "hbi2d":
        print(json.dumps(hbi2d_leg(spk, ini, torch, ctx, 40, 2, not a.no_cpu)))
        return
    if a.mode == "ensemble":
        print(json.dumps(ensemble_leg(spk, ini, ctx, a.ensemble_steps, a.ensemble_acc_steps, a.ensemble_split,
                                      a.ensemble_group)))
        return
    if a.mode == "closure-kernels":
        print(json.dumps(closure_kernels(spk, ctx, torch)))
        return
    if a.mode == "closure":
        print(json.dumps(closure_leg(spk, ini, ctx, torch, N=a.closure_dns, t_spin=a.closure_spinup,
                                     iters=a.closure_iters, nsamples=a.closure_samples,
                                     dt_sample=a.closure_sample_dt, dT=a.closure_dT,
                                     seeds=tuple(int(x) for x in a.closure_seeds.split(",")))))
        return
    state = ini.mhd_random((n, n, n), seed=2026)
    if a.mode == "kernels":
        sol = spk.Solver(ctx, "f64", **mhd_params(n))
        sol.set_fields(state)
        sol.main_loop(1)
        t = sol.kernel_timings(2)
        copy = copy_gbs(torch, n * n * (n // 2 + 1) * 16 * 8)
        print(json.dumps({"kernel_ms": t, "pipeline": sol.pipeline(), "copy_gbs": copy,
                          "copy_bytes": n * n * (n // 2 + 1) * 16 * 8}))
        return

    res = {}
    for prec in a.precisions.split(","):
        ms, t, e, fz = time_steps(spk, torch, ctx, state, prec, n, a.steps, a.warmup)
        res[prec] = dict(ms=ms, steps_per_s=1000.0 / ms, t=t["t"], dt=t["dt"], energy=e, fused=fz[0],
                         fused_points=fz[1])
        print("# %s: %.3f ms/step, %.1f steps/s, t=%.4g, E=%.6g, fused back end: %s (%d points/thread)" %
              (prec, ms, 1000.0 / ms, t["t"], e, fz[0], fz[1]), file=sys.stderr, flush=True)

    passes = time_fft_passes(spk, torch, ctx, "f64", n)
    for k, v in passes.items():
        print("# pass %s f64: %.3f ms, %.0f GB/s" % (k, v["ms"], v["gbs"]), file=sys.stderr, flush=True)
    copy = copy_gbs(torch, n * n * (n // 2 + 1) * 16 * 8)
    kern = kernel_table(spk, ctx, state, n, "f64", res["f64"]["fused"] if "f64" in res else False)
    for k, v in kern.items():
        print("# kernel %-9s f64: %8.3f ms/launch x%d per step, %6.0f GB/s, %4.1f%% of step" %
              (k, v["ms"], v["per_step"], v["gbs"], 100 * v["share"]), file=sys.stderr, flush=True)
    kern32 = None
    if "f32" in res:
        kern32 = kernel_table(spk, ctx, state, n, "f32", res["f32"]["fused"])
        for k, v in kern32.items():
            print("# kernel %-9s f32: %8.3f ms/launch x%d per step, %6.0f GB/s, %4.1f%% of step" %
                  (k, v["ms"], v["per_step"], v["gbs"], 100 * v["share"]), file=sys.stderr, flush=True)
    dom_name = max(kern, key=lambda k: kern[k]["ms"] * kern[k]["per_step"])
    dom = kern[dom_name]
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_f64_%d.json" % n)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f)["bytes_per_launch"].get(dom_name)
    roofline = dict(bound="hbm", achieved=round(dom["gbs"], 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(dom["gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel=KERNEL_NAMES[dom_name], algorithmic_bytes_per_launch=dom["bytes"],
                    avg_launch_ms=round(dom["ms"], 5), share_of_step=round(dom["share"], 4),
                    copy_gbs_measured=round(copy, 1))

    acc = fp32_accuracy(spk, ctx, state, n, a.accuracy_steps) if a.accuracy_steps > 0 else None
    if acc:
        print("# fp32 vs fp64 after %d steps: energy %.2e, cross helicity %.2e, spectra %.2e" %
              (acc["steps"], acc["rel_energy_err"], acc["rel_cross_helicity_err"], acc["spectra_max_rel_err"]),
              file=sys.stderr, flush=True)
    tg = taylor_green_leg(spk, ini, torch, ctx, not a.no_cpu) if a.extra else None
    if tg:
        print("# taylor-green 64^3: %s" % json.dumps(tg), file=sys.stderr, flush=True)
    mti = mti_leg(spk, ini, torch, ctx, n, a.mti_steps, 2, not a.no_cpu) if a.extra else None
    if mti:
        print("# mti %d^3: %s" % (n, json.dumps(mti)), file=sys.stderr, flush=True)
    q2d = hbi2d_leg(spk, ini, torch, ctx, 40, 2, not a.no_cpu) if a.extra else None
    if q2d:
        print("# hbi2d 512x512x2: %s" % json.dumps(q2d), file=sys.stderr, flush=True)
    ens = ensemble_leg(spk, ini, ctx, a.ensemble_steps, a.ensemble_acc_steps, a.ensemble_split) if a.extra and a.ensemble_steps > 0 else None
    closure = None
    if a.extra and a.closure:
        closure = closure_leg(spk, ini, ctx, torch, N=512, t_spin=0.5, nsamples=6, dt_sample=0.05, dT=0.25,
                              iters=1500, seeds=tuple(int(x) for x in a.closure_seeds.split(",")))
        closure["config"] = "DNS 512^3 spun up to t=0.5, 6 samples 0.05 apart, 1500 training iterations per seed, " \
                            "closure / plain 128^3 / DNS compared after 0.25 time units"

    cpu = None
    acc_orc = None
    if not a.no_cpu:
        cpu, orc_out = cpu_baseline(n, state, a.cpu_steps)
        print("# cpu oracle: %.3f s/step on %d threads" % (1.0 / cpu["value"], cpu["cores"]), file=sys.stderr,
              flush=True)
        acc_orc = accuracy_vs_oracle(spk, ctx, state, n, a.cpu_steps, orc_out)
        del orc_out
        print("# gpu vs oracle after %d steps: %s" % (a.cpu_steps, json.dumps(acc_orc)), file=sys.stderr, flush=True)

    f64 = res.get("f64") or next(iter(res.values()))
    line = {
        "metric": "MHD RK3 steps/s at 256^3 (FP64; FP32-bulk beside it)",
        "value": round(f64["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(f64["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "mhd_rk3_%d^3" % n, "grid": [n, n, n], "fields": 6, "nu": 1e-3, "eta": 1e-3,
                   "cfl": 0.5, "box": "2pi periodic",
                   "ic": "random-phase solenoidal u,B, E(k)~k^4 exp(-2(k/4)^2), rms 1 each, seed 2026",
                   "parallelism": "single GPU"},
        "pipeline_f64": {"fused_back_end": f64["fused"], "points_per_thread": f64["fused_points"]},
        "fp32_bulk": ({"steps_per_s": round(res["f32"]["steps_per_s"], 3), "ms_per_step": round(res["f32"]["ms"], 4),
                       "fused_back_end": res["f32"]["fused"], "points_per_thread": res["f32"]["fused_points"],
                       "ratio_vs_f64": round(res["f32"]["steps_per_s"] / f64["steps_per_s"], 3),
                       "rel_energy_diff_vs_f64_after_timed_run": abs(res["f32"]["energy"] - f64["energy"]) /
                       abs(f64["energy"]),
                       "accuracy_vs_f64": acc}
                      if "f32" in res and "f64" in res else None),
        "taylor_green_64": tg,
        "hbi2d_512": q2d,
        "mti_rkl2": mti,
        "ensemble": ens,
        "closure": closure,
        "fft_pass_gbs_f64": {k: round(v["gbs"], 1) for k, v in passes.items()},
        "kernels_f64": {k: {"ms": round(v["ms"], 4), "gbs": round(v["gbs"], 1), "share": round(v["share"], 4)}
                        for k, v in kern.items()},
        "kernels_f32": {k: {"ms": round(v["ms"], 4), "gbs": round(v["gbs"], 1), "share": round(v["share"], 4)}
                        for k, v in kern32.items()} if kern32 else None,
        "roofline": roofline,
        "accuracy_vs_oracle": acc_orc,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
