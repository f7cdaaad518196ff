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
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--precisions", default="f64,f32")
    ap.add_argument("--mode", choices=["full", "passes", "kernels"], default="full",
                    help="passes: only the FFT pass kernels and a copy kernel; kernels: each pipeline "
                         "kernel of the 256^3 FP64 step twice, plus the copy kernel (both for counter runs)")
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
            "fwdx_epilogue": nout * sb + 4 * nf * s64 + nin * sb}
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
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import pyoracle as orc

    sim = orc.Sim(**mhd_params(n))
    sim.set_fields(state)
    t0 = time.perf_counter()
    sim.main_loop(steps)
    dt = time.perf_counter() - t0
    cores = orc.lib().orc_num_threads()
    return dict(value=steps / dt, unit="steps/s", cores=cores, kind="port",
                sample="%d MHD RK3 step(s) at %d^3, FP64 oracle (C, OpenMP), same initial state" % (steps, n),
                seconds=dt)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py measures one GPU only: every box here has a single MI355X and this build has no "
              "multi-GPU path (--gpus %d refused)" % a.gpus, file=sys.stderr)
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

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(n, state, a.cpu_steps)
        print("# cpu oracle: %.3f s/step on %d threads" % (1.0 / cpu["value"], cpu["cores"]), file=sys.stderr)

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
                       "rel_energy_diff_vs_f64": abs(res["f32"]["energy"] - f64["energy"]) / abs(f64["energy"])}
                      if "f32" in res and "f64" in res else None),
        "fft_pass_gbs_f64": {k: round(v["gbs"], 1) for k, v in passes.items()},
        "kernels_f64": {k: {"ms": round(v["ms"], 4), "gbs": round(v["gbs"], 1), "share": round(v["share"], 4)}
                        for k, v in kern.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
