"""Pins the FP64 oracle's physics against what the mathematics fixes, independently
of the oracle's own formulas:

- wavevector / 2/3 mask / kmax / nbins: closed forms (wavevector.cu:50-110)
- CleanFieldDivergence: k.P(v) = 0 and P(P(v)) = P(v)
- nonlinear terms: the Basdevant stress divergence and the curl of u x B are
  recomputed in numpy in advective form, -(u.grad)u + (B.grad)B and
  (B.grad)u - (u.grad)B, with spectral derivatives; equal after mask + projection
- ideal MHD: the dealiased nonlinear terms conserve energy and cross helicity
- RK3: on an Alfven wave and on diffusion, one step is exactly the amplification
  1 + Z + Z^2/2 + Z^3/6 of the linear operator Z = A dt (Wray RK3)
- RKL2: b_j closed form; R(x) = e^{-x} + O(x^3); |R| <= 1 on the design interval
- end to end: Taylor-Green decay (check-taylor-green.py) and the MTI growth rate
  sigma = 0.9341413811120219 (check-mti-eigenmode.py:22), with the cubic of
  mti_disp_rel.py restated
"""
import math

import numpy as np
import pytest

import pyoracle as orc
from physics_cases import initial_state, relerr

TWO_PI = 2 * math.pi


def test_wavevector_closed_forms():
    kvec, mask, scal = orc.wavevector(nx=64, ny=64, nz=64, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI)
    # |k_i| <= 2/3 * 31 = 20.67 survives: 41 x 41 x 21 modes
    assert int(mask.sum()) == 41 * 41 * 21
    kxmax, kymax, kzmax, kmax, deltak, nbins = scal
    assert kxmax == pytest.approx(2.0 / 3.0 * 31)
    assert kmax == pytest.approx(math.sqrt(3) * 2.0 / 3.0 * 31)
    assert deltak == pytest.approx(1.0)
    assert nbins == int(math.sqrt(3) * 62.0 / 3.0 + 0.5)
    k = kvec.reshape(3, 64, 64, 33)
    assert k[0, 32, 0, 0] == -32.0 and k[0, 31, 0, 0] == 31.0 and k[1, 0, 63, 0] == -1.0
    assert k[2, 0, 0, 32] == 32.0
    # anisotropic box and with2d
    kvec, mask, scal = orc.wavevector(nx=8, ny=16, nz=4, lx=1.0, ly=2.0, lz=0.5, with2d=1)
    k = kvec.reshape(3, 8, 16, 3)
    assert np.all(k[2] == 0.0) and scal[2] == 0.0
    assert k[0, 1, 0, 0] == pytest.approx(TWO_PI) and k[1, 0, 1, 0] == pytest.approx(math.pi)


def test_projection_invariants():
    sim = orc.Sim(**dict(nx=16, ny=16, nz=16, incompressible=1, mhd=1, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI))
    rng = np.random.default_rng(4)
    f = np.fft.rfftn(rng.standard_normal((6, 16, 16, 16)), axes=(1, 2, 3))
    sim.set_fields(f)
    sim.clean_divergence()
    k = sim.kvec
    once = sim.fields.copy()
    for g in (0, 3):
        div = k[0] * once[g] + k[1] * once[g + 1] + k[2] * once[g + 2]
        assert np.abs(div).max() < 1e-12 * np.abs(once[g:g + 3]).max() * 16
    sim.clean_divergence()
    assert relerr(sim.fields, once) < 1e-15


def _grid_k(shape, L):
    nx, ny, nz = shape
    kx = 2 * np.pi / L[0] * np.fft.fftfreq(nx, 1.0 / nx)
    ky = 2 * np.pi / L[1] * np.fft.fftfreq(ny, 1.0 / ny)
    kz = 2 * np.pi / L[2] * np.arange(nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def _advective(a_hat, b_hat, shape, K):
    """(a . grad) b in real space from spectral a, b (numpy FFTs)."""
    a = [np.fft.irfftn(c, s=shape, axes=(0, 1, 2)) for c in a_hat]
    out = []
    for i in range(3):
        grad = [np.fft.irfftn(1j * K[j] * b_hat[i], s=shape, axes=(0, 1, 2)) for j in range(3)]
        out.append(sum(a[j] * grad[j] for j in range(3)))
    return out


def _mask_project(v_hat, K, mask, project=True):
    v = [c * mask for c in v_hat]
    if not project:
        return np.stack(v)
    K2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    ik2 = np.where(K2 > 0, 1.0 / np.where(K2 > 0, K2, 1.0), 1.0)
    div = K[0] * v[0] + K[1] * v[1] + K[2] * v[2]
    return np.stack([v[i] - K[i] * div * ik2 for i in range(3)])


@pytest.mark.parametrize("name", ["hd", "mhd"])
def test_nonlinear_terms_against_advective_form(name):
    from physics_cases import CASES
    p = dict(CASES[name]["params"], nu=0.0, nu_m=0.0)
    shape = (p["nx"], p["ny"], p["nz"])
    L = (p["lx"], p["ly"], p["lz"])
    u0 = initial_state(name, seed=3)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sim.complex2real()
    sim.compute_dfield(0)
    K = _grid_k(shape, L)
    _, mask, _ = orc.wavevector(**{k: p[k] for k in ("nx", "ny", "nz", "lx", "ly", "lz")})
    mask = mask.reshape(K[0].shape)
    n = np.prod(shape)
    u, b = u0[0:3], (u0[3:6] if name == "mhd" else None)
    nl = [-x for x in _advective(u, u, shape, K)]
    if b is not None:
        bb = _advective(b, b, shape, K)
        nl = [nl[i] + bb[i] for i in range(3)]
    du = _mask_project([np.fft.rfftn(x) for x in nl], K, mask)
    assert relerr(sim.dfields[0:3], du) < 1e-11
    if b is not None:
        ind = [x - y for x, y in zip(_advective(b, u, shape, K), _advective(u, b, shape, K))]
        db = _mask_project([np.fft.rfftn(x) for x in ind], K, mask, project=False)
        assert relerr(sim.dfields[3:6], db) < 1e-11
    assert n > 0


def test_ideal_mhd_conserves_energy_and_cross_helicity():
    from physics_cases import CASES
    p = dict(CASES["mhd"]["params"], nu=0.0, nu_m=0.0)
    u0 = initial_state("mhd", seed=9)
    sim = orc.Sim(**p)
    sim.set_fields(u0)
    sim.complex2real()
    sim.compute_dfield(0)
    d = sim.dfields
    w = np.full(u0.shape[1:], 2.0)
    w[:, :, 0] = 1.0  # half-space weights (kz = 0 counted once)
    dE = np.sum(w * np.real(np.conj(u0) * d))
    dHc = np.sum(w * np.real(np.conj(u0[0:3]) * d[3:6] + np.conj(u0[3:6]) * d[0:3]))
    scale = np.sqrt(np.sum(w * np.abs(u0) ** 2) * np.sum(w * np.abs(d) ** 2))
    assert abs(dE) < 1e-12 * scale
    assert abs(dHc) < 1e-12 * scale


def _rk3_amp(Z):
    I = np.eye(Z.shape[0])
    return I + Z + Z @ Z / 2 + Z @ Z @ Z / 6


def test_rk3_exact_amplification_alfven_wave():
    """Transverse Alfven wave along a uniform B0 z-hat: an exact nonlinear
    solution whose per-mode dynamics is d/dt (u, b) = [[-nu k^2, i k B0], [i k B0, -eta k^2]] (u, b)."""
    n = (8, 8, 16)
    L = 1.0
    nu = 1e-3
    p = dict(nx=n[0], ny=n[1], nz=n[2], incompressible=1, mhd=1, nu=nu, nu_m=nu, cfl=0.5, lx=L, ly=L, lz=L,
             t_final=1e9, toutput_flow=1e9)
    X, Y, Z = np.meshgrid(*[L * np.arange(m) / m - L / 2 for m in n], indexing="ij")
    k = TWO_PI * 2 / L
    A = 1e-3
    zero = 0 * X
    real = [A * np.cos(k * Z), zero, zero, A * 0.3 * np.sin(k * Z), zero, 1.0 + zero]
    state = np.stack([np.fft.rfftn(r) for r in real])
    sim = orc.Sim(**p)
    sim.set_fields(state)
    uk = state[0, 0, 0, 2]
    bk = state[3, 0, 0, 2]
    v = np.array([uk, bk])
    for _ in range(5):
        sim.main_loop(1)
        dt = sim.time()["dt"]
        M = np.array([[-nu * k * k, 1j * k * 1.0], [1j * k * 1.0, -nu * k * k]]) * dt
        v = _rk3_amp(M) @ v
    got = np.array([sim.fields[0, 0, 0, 2], sim.fields[3, 0, 0, 2]])
    assert np.abs(got - v).max() < 1e-13 * np.abs(v).max()
    # the other modes stay empty, the mean field untouched
    f = sim.fields.copy()
    f[0, 0, 0, 2] = 0
    f[3, 0, 0, 2] = 0
    f[5, 0, 0, 0] -= np.prod(n)
    assert np.abs(f).max() < 1e-12 * abs(uk)


def test_rk3_exact_amplification_heat_equation():
    """Heat equation (test_rk3_heatdiff.py restates the same RK3 in numpy):
    each mode is multiplied by 1 + z + z^2/2 + z^3/6, z = -chi k^2 dt, every step."""
    nx = 64
    chi = 1e-3
    p = dict(nx=nx, ny=2, nz=2, heat_equation=1, nu_th=chi, antialiasing=0, cfl=1.5, lx=1.0, ly=1.0, lz=1.0,
             t_final=1e9, toutput_flow=1e9)
    x = np.arange(nx) / nx - 0.5
    th = np.broadcast_to((np.exp(-(x / 0.1) ** 2))[:, None, None], (nx, 2, 2)).copy()
    c0 = np.fft.rfftn(th)[None]
    sim = orc.Sim(**p)
    sim.set_fields(c0)
    K = _grid_k((nx, 2, 2), (1.0, 1.0, 1.0))
    k2 = K[0] ** 2 + K[1] ** 2 + K[2] ** 2
    c = c0[0].copy()
    for _ in range(10):
        sim.main_loop(1)
        z = -chi * k2 * sim.time()["dt"]
        c = c * (1 + z + z * z / 2 + z ** 3 / 6)
    assert relerr(sim.fields[0], c) < 1e-13


def _rkl2_R(x, coef):
    """Scalar RKL2 amplification for y' = -y over a super-step with x = tau."""
    s = coef.shape[0]
    y0 = 1.0
    mut1 = coef[0, 2]
    yjm2 = y0
    yjm1 = y0 + mut1 * x * (-y0)
    dy0 = -y0
    for j in range(2, s + 1):
        mu, nu, mut, gam = coef[j - 1]
        yj = mu * yjm1 + nu * yjm2 + (1 - mu - nu) * y0 + mut * x * (-yjm1) + gam * x * dy0
        yjm2, yjm1 = yjm1, yj
    return yjm1


def test_rkl2_coefficients():
    # r = 1.6e5 needs about 800 stages (the 512^2 x 2 x 2 heat-diffusion-sts run)
    for r in (1.0, 7.3, 40.0, 100.0, 1.6e5):
        s, coef = orc.rkl2_coeffs(r, 1.0)
        # stage count: 4 r / dt_exp = s^2 + s - 2 solved for s (rkl.cu:473-477)
        assert s == 1 + int(4 * (1 + 2 * r) / (1 + math.sqrt(9 + 16 * r)))
        w1 = 4.0 / (s * s + s - 2)
        assert coef[0, 2] == pytest.approx(w1 / 3)
        # b_j = (j^2 + j - 2) / (2 j (j + 1)) for j >= 2 and b_0 = b_1 = 1/3 (Meyer 2014)
        b = [1 / 3, 1 / 3] + [(j * j + j - 2) / (2 * j * (j + 1)) for j in range(2, s + 1)]
        for j in range(2, s + 1):
            mu, nu, mut, gam = coef[j - 1]
            assert mu == pytest.approx((2 * j - 1) / j * b[j] / b[j - 1], rel=1e-14)
            assert nu == pytest.approx(-(j - 1) / j * b[j] / b[j - 2], rel=1e-14)
            assert mut == pytest.approx(w1 * mu, rel=1e-14)
            assert gam == pytest.approx(-(1 - b[j - 1]) * mut, rel=1e-14)
        # second order: R(x) - e^{-x} = O(x^3)
        e1 = abs(_rkl2_R(1e-2, coef) - math.exp(-1e-2))
        e2 = abs(_rkl2_R(2e-2, coef) - math.exp(-2e-2))
        assert e2 / e1 == pytest.approx(8.0, rel=0.1)
        # stable over the design interval x in [0, 2/w1] (explicit-Euler limit times (s^2+s-2)/4)
        xs = np.linspace(0, 2.0 / w1 * 0.999, 400)
        assert max(abs(_rkl2_R(x, coef)) for x in xs) <= 1.0 + 1e-12


def test_rkl2_heat_diffusion_second_order():
    """RKL2 on the heat equation (check-heat-diffusion-sts.py: L1 error vs
    irfftn(T0 exp(-chi k^2 t))); the error of one super-step shrinks as dt^3."""
    nx = 64
    errs = []
    for tfin in (0.02, 0.01):
        p = dict(nx=nx, ny=2, nz=2, heat_equation=1, nu_th=1.0, antialiasing=0, supertimestepping=1, sts_mask=1,
                 cfl_par=1.0, cfl_sts=1.0, safety_sts=50, lx=4.0, ly=4.0, lz=4.0, t_final=tfin, toutput_flow=1e9)
        x = 4.0 * np.arange(nx) / nx - 2.0
        th = np.broadcast_to(np.exp(-(x / 0.5) ** 2)[:, None, None], (nx, 2, 2)).copy()
        c0 = np.fft.rfftn(th)[None]
        sim = orc.Sim(**p)
        sim.set_fields(c0)
        # heat equation with theta in sts: current_dt = dt_par = inf, clamped to t_final
        sim.compute_dt()
        sim.rkl2_cycle()
        K = _grid_k((nx, 2, 2), (4.0, 4.0, 4.0))
        exact = c0[0] * np.exp(-(K[0] ** 2) * tfin)
        errs.append(relerr(sim.fields[0], exact))
    assert errs[0] / errs[1] > 6.0  # ~8 for a third-order local error


def test_taylor_green_decay():
    """2D Taylor-Green in (y, z), L = 1: an exact Navier-Stokes solution decaying
    as exp(-2 nu (2 pi)^2 t) per component (check-taylor-green.py:32-36)."""
    n = 16
    nu = 1.0 / 200
    p = dict(nx=n, ny=n, nz=n, incompressible=1, nu=nu, cfl=0.5, lx=1.0, ly=1.0, lz=1.0, t_final=0.5,
             toutput_flow=1e9)
    X, Y, Z = np.meshgrid(*[np.arange(n) / n - 0.5] * 3, indexing="ij")
    real = [0 * X, np.sin(TWO_PI * Y) * np.cos(TWO_PI * Z), -np.cos(TWO_PI * Y) * np.sin(TWO_PI * Z)]
    state = np.stack([np.fft.rfftn(r) for r in real])
    sim = orc.Sim(**p)
    sim.set_fields(state)
    while sim.time()["t"] < 0.5:
        sim.main_loop(1)
    t = sim.time()["t"]
    assert t == pytest.approx(0.5, abs=1e-12)
    exact = state * math.exp(-2 * nu * TWO_PI ** 2 * t)
    assert relerr(sim.fields, exact) < 1e-7


def mti_sigma():
    """Largest real root of the MTI cubic (mti_disp_rel.py:44-70), restated."""
    Pe, ReNu, ReEta, B0, N2 = 1e2, 1e5, 1e5, 1e-4, 0.1
    Pr, q, Lam = Pe / ReNu, Pe / ReEta, B0 ** 2 * Pe
    k2 = (TWO_PI * 12) ** 2 / Pe
    coeff = [1.0, k2 * (1 + q + Pr), N2 + k2 * Lam + (q + Pr) * k2 ** 2 + q * Pr * k2 ** 2,
             q * Pr * k2 ** 3 + N2 * q * k2 + Lam * k2 ** 2 - k2]
    r = np.roots(coeff)
    return float(np.max(np.real(r[np.abs(np.imag(r)) < 1e-12])))


def test_mti_cubic_root():
    assert mti_sigma() == pytest.approx(0.9341413811120219, rel=1e-12)


MTI_PARAMS = dict(nx=512, ny=4, nz=4, incompressible=1, boussinesq=1, stratification=1, strat_direction=2, mhd=1,
                  anisotropic_diffusion=1, nu=1e-5, nu_m=1e-5, nu_th=1e-6, reynolds_ani=100.0, N2=0.1, OmegaT2=1.0,
                  cfl=1.5, cfl_par=1.0, safety_source=0.2, lx=1.0, ly=1.0, lz=1.0, t_final=1e9, toutput_flow=0.1)


def mti_initial_state():
    """Eigenmode of check-mti-eigenmode.py:14-44, in the spectral layout."""
    nx, ny, nz = 512, 4, 4
    sigma, Pe, Reeta, B0, N2 = 0.9341413811120219, 100.0, 1e5, 1e-4, 0.1
    kp = TWO_PI * 12.0
    x = -0.5 + np.arange(nx) / nx
    X = np.broadcast_to(x[:, None, None], (nx, ny, nz))
    vz = -1e-5 * np.sin(kp * X)
    bz = -1e-5 * np.cos(kp * X) * B0 * kp / (sigma + kp * kp / Reeta)
    th = 1.0 / (sigma + kp * kp / Pe) * (N2 - kp * kp / Pe / (sigma + kp * kp / Reeta)) * vz
    zero = 0 * X
    real = [zero, zero, vz, B0 + zero, zero, bz, th]
    return np.stack([np.fft.rfftn(r) for r in real]), real


def test_mti_eigenmode_growth():
    state, real = mti_initial_state()
    sim = orc.Sim(**MTI_PARAMS)
    sim.set_fields(state)
    tend = 0.2
    while sim.time()["t"] < tend - 1e-12:
        sim.main_loop(1)
    t = sim.time()["t"]
    g = math.exp(0.9341413811120219 * t)
    sim.complex2real()
    rb = sim.realbuf
    expect = [real[0], real[1], real[2] * g, real[3], real[4], real[5] * g, real[6] * g]
    l1 = sum(np.sum(np.abs(rb[i] - expect[i])) for i in range(7)) / (512 * 16)
    assert l1 < 1e-7  # the reference's tolerance (check-mti-eigenmode.py:47)
    # and the growth itself is resolved: the vz amplitude grew by e^{sigma t}
    amp = np.abs(np.fft.rfftn(rb[2])[12, 0, 0]) / np.abs(np.fft.rfftn(real[2])[12, 0, 0])
    assert amp == pytest.approx(g, rel=1e-4)
