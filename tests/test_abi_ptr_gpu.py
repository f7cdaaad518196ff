"""The C ABI on caller-owned arrays (spk_*_ptr, spk_project, spk_rkl2_stage,
spk_aniso_flux, spk_maxabs): every entry runs on arrays allocated with spk_malloc
and gives the handle path's results bit for bit (the same kernels on the same
inputs), leaves the solver's own arrays untouched, and -- where the handle path
has no counterpart -- matches the FP64 oracle or numpy.  The reference methods
they replace take the arrays as arguments (physics.hpp:18-42,
timestepping.hpp:39-43, rkl.hpp:42, spooky_outputs.hpp:26-29)."""
import ctypes

import numpy as np
import pytest

import pyoracle as orc
from physics_cases import CASES, initial_state, relerr
from spkmod import spk

pytestmark = pytest.mark.gpu

L = spk.lib


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


class Pair:
    """A solver used through its own arrays (handle path) and a second one of
    the same configuration used only on caller arrays."""

    def __init__(self, ctx, name, prec="f64", seed=1):
        self.ctx = ctx
        self.p = CASES[name]["params"]
        self.h = spk.Solver(ctx, prec, **self.p)
        self.c = spk.Solver(ctx, prec, **self.p)
        self.nf, self.shape_c = self.h.nf, self.h.shape_c
        self.u0 = initial_state(name, seed)
        self.h.set_fields(self.u0)
        fb = 16 * int(np.prod(self.shape_c))
        self.F = ctx.alloc(fb * self.nf)
        self.F.upload(self.u0)
        self.D = ctx.alloc(fb * self.nf)
        self.D.upload(np.zeros((self.nf,) + self.shape_c, np.complex128))
        self.R = ctx.alloc(fb * self.nf)
        self.S = ctx.alloc(fb * self.nf)
        self.fb = fb

    def fields(self):
        return self.F.download(np.complex128, (self.nf,) + self.shape_c)

    def own_untouched(self):
        """the second solver's own state and dU are still all zero"""
        return (not np.any(self.c.get_fields())) and (not np.any(self.c.get_dfields()))


def rc(x):
    spk.check(x)


def test_compute_dt_and_rk3_bitwise(ctx):
    pr = Pair(ctx, "mhd")
    pr.h.compute_dt()
    pr.h.rk3()
    rc(L().spk_compute_dt_ptr(pr.c.h, pr.F.ptr, pr.R.ptr))
    rc(L().spk_rk3_ptr(pr.c.h, pr.F.ptr, None, pr.S.ptr))
    assert np.array_equal(pr.fields(), pr.h.get_fields())
    th, tc = pr.h.time(), pr.c.time()
    for k in ("t", "dt", "dt_hyp", "dt_par"):
        assert th[k] == tc[k], k
    assert pr.own_untouched()


def test_compute_dt_fills_real_buffer(ctx):
    pr = Pair(ctx, "mhd")
    rc(L().spk_compute_dt_ptr(pr.c.h, pr.F.ptr, pr.R.ptr))
    pr.h.complex2real()
    p = pr.p
    got = pr.R.download(np.float64, (pr.nf, p["nx"], p["ny"], 2 * (p["nz"] // 2 + 1)))[..., : p["nz"]]
    assert np.array_equal(got, pr.h.get_realbuf())
    # complex2real_ptr leaves its input unchanged
    rc(L().spk_complex2real_ptr(pr.c.h, pr.F.ptr, pr.R.ptr, pr.nf))
    assert np.array_equal(pr.fields(), pr.u0)


def test_compute_dfield_bitwise(ctx):
    pr = Pair(ctx, "mhd_bouss_aniso")
    pr.h.compute_dfield()
    rc(L().spk_compute_dfield_ptr(pr.c.h, pr.F.ptr, None, pr.D.ptr))
    got = pr.D.download(np.complex128, (pr.nf,) + pr.shape_c)
    assert np.array_equal(got, pr.h.get_dfields())
    assert pr.own_untouched()


def test_clean_divergence_and_project_bitwise(ctx):
    pr = Pair(ctx, "mhd")
    rng = np.random.default_rng(3)
    noisy = pr.u0 + 0.1 * (rng.standard_normal(pr.u0.shape) + 1j * rng.standard_normal(pr.u0.shape))
    pr.h.set_fields(noisy)
    pr.F.upload(noisy)
    pr.h.clean_divergence()
    rc(L().spk_clean_divergence_ptr(pr.c.h, pr.F.ptr))
    ref = pr.h.get_fields()
    assert np.array_equal(pr.fields(), ref)
    # spk_project on each vector (u at field 0, B at field 3) is the same operation
    pr.F.upload(noisy)
    rc(L().spk_project(pr.c.h, pr.F.ptr))
    rc(L().spk_project(pr.c.h, ctypes.c_void_p(pr.F.ptr.value + 3 * pr.fb)))
    assert np.array_equal(pr.fields(), ref)


def test_hydro_mhd_advance_with_rkl2_bitwise(ctx):
    pr = Pair(ctx, "mhd_bouss_rkl2")
    for _ in range(2):  # odd and even steps place the RKL2 cycle before and after RK3
        pr.h.advance()
        rc(L().spk_hydro_mhd_advance_ptr(pr.c.h, pr.F.ptr, pr.R.ptr, pr.S.ptr))
    assert np.array_equal(pr.fields(), pr.h.get_fields())
    assert pr.own_untouched()


def test_rkl2_cycle_and_stages(ctx):
    pr = Pair(ctx, "mhd_bouss_rkl2")
    pr.h.compute_dt()
    rc(L().spk_compute_dt_ptr(pr.c.h, pr.F.ptr, None))
    pr.h.rkl2_cycle()
    rc(L().spk_rkl2_cycle_ptr(pr.c.h, pr.F.ptr, None))
    ref = pr.h.get_fields()
    assert np.array_equal(pr.fields(), ref)
    # the same cycle stage by stage with the recurrence's coefficients
    t = pr.c.time()
    s, coef = orc.rkl2_coeffs(t["dt"], t["dt_sts"])
    assert s > 2
    pr.F.upload(pr.u0)
    nsts = 1
    fb = pr.fb
    uc0, uc1, du0 = ctx.alloc(fb * nsts), ctx.alloc(fb * nsts), ctx.alloc(fb * nsts)
    for j in range(1, s + 1):
        c = np.ascontiguousarray(coef[j - 1])
        rc(L().spk_rkl2_stage(pr.c.h, j, c.ctypes.data_as(ctypes.c_void_p), t["dt"], pr.F.ptr, uc0.ptr, uc1.ptr,
                              du0.ptr))
    # the cycle takes real B once per cycle, the stage entry per stage: rounding only
    assert relerr(pr.fields(), ref) < 1e-13
    assert pr.own_untouched()


def test_aniso_flux_against_oracle(ctx):
    """dtheta += i mask k . FFT[flux]: the oracle's RHS with conduction minus the
    same RHS with chi = 1/reynolds_ani ~ 0 is exactly that term."""
    name = "mhd_bouss_aniso"
    pr = Pair(ctx, name)
    p = pr.p
    th = pr.h.vars["TH"]
    # complex_dTheta is the theta slot of the dfields (physics.cu:214, 320)
    rc(L().spk_aniso_flux(pr.c.h, pr.F.ptr, None, ctypes.c_void_p(pr.D.ptr.value + th * pr.fb)))
    got = pr.D.download(np.complex128, (pr.nf,) + pr.shape_c)
    d = {}
    for tag, ra in (("on", p["reynolds_ani"]), ("off", 1e300)):
        sim = orc.Sim(**dict(p, reynolds_ani=ra))
        sim.set_fields(pr.u0)
        sim.complex2real()
        sim.compute_dfield(0)
        d[tag] = sim.dfields[th].copy()
    ref = d["on"] - d["off"]
    assert relerr(got[th], ref) < 1e-11
    others = [f for f in range(pr.nf) if f != th]
    assert not np.any(got[others])


def test_maxabs_and_measurements(ctx):
    pr = Pair(ctx, "mhd")
    p = pr.p
    rc(L().spk_complex2real_ptr(pr.c.h, pr.F.ptr, pr.R.ptr, pr.nf))
    real = pr.R.download(np.float64, (pr.nf, p["nx"], p["ny"], 2 * (p["nz"] // 2 + 1)))
    out = ctx.alloc(8)
    for f in range(pr.nf):
        rc(L().spk_maxabs(pr.c.h, ctypes.c_void_p(pr.R.ptr.value + f * pr.fb), out.ptr))
        assert out.download(np.float64, (1,))[0] == np.abs(real[f, ..., : p["nz"]]).max()
    pr.h.complex2real()
    v = ctypes.c_double()
    for f in range(pr.nf):
        rc(L().spk_energy_ptr(pr.c.h, ctypes.c_void_p(pr.F.ptr.value + f * pr.fb), ctypes.byref(v)))
        assert v.value == pr.h.energy(f)
    rc(L().spk_enstrophy_ptr(pr.c.h, ctypes.c_void_p(pr.F.ptr.value + 3 * pr.fb), ctypes.byref(v)))
    assert v.value == pr.h.enstrophy(3)
    rc(L().spk_correlation_ptr(pr.c.h, pr.R.ptr, ctypes.c_void_p(pr.R.ptr.value + 3 * pr.fb), ctypes.byref(v)))
    assert v.value == pr.h.correlation(0, 3)
    cap = 4096
    buf = (ctypes.c_double * cap)()
    nb = ctypes.c_int()
    rc(L().spk_spectrum_ptr(pr.c.h, pr.F.ptr, pr.F.ptr, buf, cap, ctypes.byref(nb)))
    ref = pr.h.spectrum(0, 0)
    got = np.array(buf[: nb.value])
    assert got.shape == ref.shape
    # shell sums use atomics: equal up to the order of additions
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_bad_pointers_are_refused(ctx):
    pr = Pair(ctx, "hd")
    assert L().spk_compute_dt_ptr(pr.c.h, None, None) == -1
    assert L().spk_rk3_ptr(pr.c.h, ctypes.c_void_p(pr.F.ptr.value + 8), None, None) == -1
    assert b"aligned" in L().spk_last_error()
    assert L().spk_aniso_flux(pr.c.h, pr.F.ptr, None, pr.D.ptr) == -4  # no conduction in this run
    assert L().spk_rkl2_stage(pr.c.h, 1, None, 0.1, pr.F.ptr, pr.F.ptr, pr.F.ptr, pr.F.ptr) == -4
