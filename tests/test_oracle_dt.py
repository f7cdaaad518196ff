"""Pins of the oracle's CFL step (orc_compute_dt, restating compute_dt.cu:15-309)
and of the stratification source in every direction (BoussinesqStrat,
cuda_kernels.cu:307-325), against what the mathematics fixes.

compute_dt: fields are single Fourier modes whose real-space maxima are known
exactly (sin(2 pi m y / L) reaches +-1 on a grid of 4m | N points), on boxes with
a different length and point count per direction, so a swapped direction, a
dropped 2/3, a wrong sts bucket or a wrong clamp branch changes the answer.
Expected values are written from the definitions:
  kmax_i = (2/3) (2 pi / L_i) (N_i / 2 - 1)          (wavevector.cu:70-96)
  dt_hyp = cfl / (sum_i kmax_i max|u_i| + sum_i kmax_i max|B_i|
                  + sqrt|N2| / safety_source + sqrt|OmegaT2| / safety_source
                  + |omega| / safety_source)
  dt_par = cfl_par / (sum of nu_f kmax_i^2 over explicitly stepped terms)
  dt_sts = cfl_sts / (the same sum over the sts variables)
  dt = min(dt_hyp, dt_par), capped at safety_sts dt_sts, then clamped to
  t_final and to the next output time; dt_sts <= dt with sts.
  heat equation: dt = cfl cfl_par / (nu_th sum_i kmax_i^2).

Stratification: for theta and u_s on a mode k perpendicular to e_s, with no
diffusion, the nonlinear terms vanish identically and the linear system
d/dt (u_s, theta) = A (u_s, theta), A = [[0, -1], [N2, 0]], is advanced by
Wray's RK3 with the exact amplification I + hA + (hA)^2/2 + (hA)^3/6 per step.
"""
import math

import numpy as np
import pytest

import pyoracle as orc

TWO_PI = 2 * math.pi
BIG_T = dict(t_final=1e9, toutput_flow=1e9)


def kmax(n, L):
    return (2.0 / 3.0) * (TWO_PI / L) * (n / 2 - 1)


def grid_coords(shape, L):
    """Sample positions of the reference grid x_i = -L/2 + i L/N (initflow.cu:32-34)."""
    return np.meshgrid(*[-l / 2 + np.arange(n) * l / n for n, l in zip(shape, L)], indexing="ij")


def spectral(r):
    return np.fft.rfftn(r)


def mode(shape, L, axis, m, amp, phase="sin"):
    """amp * sin(2 pi m x_axis / L_axis) (or cos) sampled on the grid; its
    grid maximum of |.| is amp when 4m divides N_axis."""
    X = grid_coords(shape, L)[axis]
    arg = TWO_PI * m * X / L[axis]
    return amp * (np.sin(arg) if phase == "sin" else np.cos(arg))


def make_state(shape, L, comps):
    """comps: list of (field, axis, m, amp); a solenoidal choice (axis != field
    direction) for velocity and B."""
    nf = max(c[0] for c in comps) + 1
    r = np.zeros((nf,) + shape)
    for f, axis, m, amp in comps:
        r[f] += mode(shape, L, axis, m, amp)
    return np.stack([spectral(x) for x in r])


SHAPE = (16, 32, 8)
L = (1.0, 2.0, 3.0)
KX, KY, KZ = (kmax(n, l) for n, l in zip(SHAPE, L))
K2 = KX ** 2 + KY ** 2 + KZ ** 2


def base(**kw):
    d = dict(nx=SHAPE[0], ny=SHAPE[1], nz=SHAPE[2], lx=L[0], ly=L[1], lz=L[2], incompressible=1, cfl=0.7,
             cfl_par=0.9, cfl_sts=1.3, safety_source=0.25, **BIG_T)
    d.update(kw)
    return d


def run_dt(params, state, t=0.0, t_lastsnap=0.0):
    sim = orc.Sim(**params)
    full = np.zeros((sim.nf,) + sim.shape_c, complex)
    full[: state.shape[0]] = state
    sim.set_fields(full)
    sim.set_time(t, 0, t_lastsnap)
    sim.compute_dt()
    return sim.time()


def close(a, b, rtol=1e-13):
    return abs(a - b) <= rtol * abs(b)


def test_wavevector_kmax_matches_definition():
    _, _, scal = orc.wavevector(**base())
    # scal: kxmax, kymax, kzmax after the 2/3 rescaling (wavevector.cu:94-96)
    assert close(scal[0], KX) and close(scal[1], KY) and close(scal[2], KZ)


def test_dt_hydro_single_modes():
    # u_x varies along y, u_y along z, u_z along x: divergence free; maxima 1.7, 0.4, 2.3
    a = (1.7, 0.4, 2.3)
    st = make_state(SHAPE, L, [(0, 1, 1, a[0]), (1, 2, 2, a[1]), (2, 0, 1, a[2])])
    nu = 3e-3
    t = run_dt(base(nu=nu), st)
    dt_hyp = 0.7 / (KX * a[0] + KY * a[1] + KZ * a[2])
    dt_par = 0.9 / (nu * K2)
    assert close(t["dt_hyp"], dt_hyp) and close(t["dt_par"], dt_par)
    assert close(t["dt"], min(dt_hyp, dt_par))
    # viscous limit: large nu makes dt_par the step
    t = run_dt(base(nu=5.0), st)
    assert close(t["dt"], 0.9 / (5.0 * K2))


def test_dt_mhd_boussinesq_aniso_rotation():
    a = (0.6, 1.1, 0.3)
    b = (0.9, 0.2, 1.4)
    comps = [(0, 1, 1, a[0]), (1, 2, 2, a[1]), (2, 0, 1, a[2]),
             (3, 2, 2, b[0]), (4, 0, 2, b[1]), (5, 1, 1, b[2])]
    st = make_state(SHAPE, L, comps)
    p = base(mhd=1, boussinesq=1, stratification=1, anisotropic_diffusion=1, N2=-0.49, OmegaT2=2.25,
             reynolds_ani=40.0, nu=2e-3, nu_m=7e-3, nu_th=11e-3, rotating=1, omega=0.8)
    t = run_dt(p, st)
    gam = (KX * a[0] + KY * a[1] + KZ * a[2]) + (KX * b[0] + KY * b[1] + KZ * b[2]) \
        + (0.7 + 1.5 + 0.8) / 0.25
    gpar = (2e-3 + 7e-3 + 1.0 / 40.0) * K2  # anisotropic conduction replaces nu_th
    assert close(t["dt_hyp"], 0.7 / gam)
    assert close(t["dt_par"], 0.9 / gpar)
    # without anisotropic conduction theta diffuses with nu_th and OmegaT2 drops out
    p2 = dict(p, anisotropic_diffusion=0)
    t2 = run_dt(p2, st)
    gam2 = gam - 1.5 / 0.25
    assert close(t2["dt_hyp"], 0.7 / gam2)
    assert close(t2["dt_par"], 0.9 / ((2e-3 + 7e-3 + 11e-3) * K2))


def test_dt_sts_buckets_and_cap():
    st = make_state(SHAPE, L, [(0, 1, 1, 0.05), (1, 2, 2, 0.05), (2, 0, 1, 0.05)])
    # theta (field 3) and B_y (field 4) as sts variables; nu_th dominates
    p = base(mhd=1, boussinesq=1, nu=1e-4, nu_m=2e-4, nu_th=0.05, supertimestepping=1,
             sts_mask=(1 << 6) | (1 << 4), safety_sts=7)
    t = run_dt(p, st)
    g_sts = 0.05 * K2 + 2e-4 * KY ** 2
    g_par = 1e-4 * K2 + 2e-4 * (KX ** 2 + KZ ** 2)
    dt_hyp = 0.7 / (0.05 * (KX + KY + KZ))
    assert close(t["dt_par"], 0.9 / g_par)
    uncapped = min(dt_hyp, 0.9 / g_par)
    dt_sts = 1.3 / g_sts
    assert uncapped > 7 * dt_sts  # the cap binds here
    assert close(t["dt"], 7 * dt_sts)
    assert close(t["dt_sts"], dt_sts)  # dt_sts < dt: unchanged
    # a tight toutput_flow clamp makes dt < dt_sts, which then shrinks to dt
    p3 = dict(p, toutput_flow=0.5 * dt_sts)
    t3 = run_dt(p3, st, t=0.2, t_lastsnap=0.2)
    assert close(t3["dt"], 0.5 * dt_sts)
    assert close(t3["dt_sts"], t3["dt"])


def test_dt_heat_equation_branch():
    n = (64, 2, 2)
    Lh = (1.0, 1.0, 1.0)
    p = dict(nx=n[0], ny=n[1], nz=n[2], lx=1.0, ly=1.0, lz=1.0, heat_equation=1, nu_th=1e-3, cfl=1.5,
             cfl_par=0.8, antialiasing=0, **BIG_T)
    st = spectral(mode(n, Lh, 0, 1, 1.0))[None]
    t = run_dt(p, st)
    km = [(TWO_PI / l) * (m / 2 - 1) for m, l in zip(n, Lh)]  # no 2/3 without antialiasing
    assert close(t["dt"], 1.5 * 0.8 / (1e-3 * sum(k * k for k in km)))
    assert t["dt_sts"] == math.inf  # no sts variable: cfl_par / 0


def test_dt_clamps_t_final_and_output():
    st = make_state(SHAPE, L, [(0, 1, 1, 1.0)])
    p = base(nu=1e-3, t_final=1.0, toutput_flow=0.3)
    dt_free = min(0.7 / (KX * 1.0), 0.9 / (1e-3 * K2))
    # next output time: t_lastsnap + toutput_flow
    t = run_dt(p, st, t=0.3 - 0.25 * dt_free, t_lastsnap=0.0)
    assert close(t["dt"], 0.25 * dt_free, 1e-9)
    # t_final takes precedence over the output clamp
    t = run_dt(p, st, t=1.0 - 0.5 * dt_free, t_lastsnap=0.9)
    assert close(t["dt"], 0.5 * dt_free, 1e-9)
    # neither binds
    t = run_dt(p, st, t=0.1, t_lastsnap=0.1)
    assert close(t["dt"], dt_free)


def rk3_matrix(h, A):
    hA = h * A
    return np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_stratification_direction_exact_rk3(sd):
    """A gravity wave on a mode perpendicular to e_s: (u_s, theta) follows the
    exact RK3 amplification of d/dt (u_s, theta) = [[0, -1], [N2, 0]] (u_s, theta)
    per step, with the dt the oracle chose."""
    shape = (8, 16, 8)
    Ls = (1.0, 1.5, 2.0)
    kdir = (sd + 1) % 3  # the mode varies along this axis, perpendicular to e_s
    N2 = 0.64
    p = dict(nx=shape[0], ny=shape[1], nz=shape[2], lx=Ls[0], ly=Ls[1], lz=Ls[2], incompressible=1,
             boussinesq=1, stratification=1, strat_direction=sd, N2=N2, nu=0.0, nu_th=0.0, cfl=0.4,
             safety_source=0.2, **BIG_T)
    sim = orc.Sim(**p)
    th0 = mode(shape, Ls, kdir, 1, 0.3, "cos")
    us0 = mode(shape, Ls, kdir, 1, 0.1, "sin")
    f = np.zeros((sim.nf,) + sim.shape_c, complex)
    f[sd] = spectral(us0)
    f[sim.vars["TH"]] = spectral(th0)
    sim.set_fields(f)
    A = np.array([[0.0, -1.0], [N2, 0.0]])
    ref = np.stack([f[sd], f[sim.vars["TH"]]]).reshape(2, -1)
    for step in range(6):
        sim.main_loop(1)
        h = sim.time()["dt"]
        ref = rk3_matrix(h, A) @ ref
    got = np.stack([sim.fields[sd], sim.fields[sim.vars["TH"]]]).reshape(2, -1)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-13, (sd, err)
    # the other velocity components stay zero
    for q in range(3):
        if q != sd:
            assert np.abs(sim.fields[q]).max() < 1e-12 * np.abs(ref).max()
