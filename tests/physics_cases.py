"""Shared run configurations and seeded initial states for the solver parity tests.

Each case is a parameter dict (Parameters fields, parameters.cu:77-330) plus a
function building a spectral initial state with the field order of
Parameters::popVariablesGrid (parameters.cu:568-640).
"""
import math

import numpy as np

from spkmod import spk

ini = __import__("importlib").import_module("spooky-git_amd.initial")

TWO_PI = 2 * math.pi
BIG_T = dict(t_final=1e9, toutput_flow=1e9)


def scalar_field(shape, L, seed, k0=3.0):
    """A smooth random scalar (dealiased, unit rms), spectral."""
    rng = np.random.default_rng(seed)
    KX, KY, KZ = ini.wavenumbers(*shape, *L)
    c = np.fft.rfftn(rng.standard_normal(shape)) * np.exp(-(KX ** 2 + KY ** 2 + KZ ** 2) / (k0 * 2 * np.pi / min(L)) ** 2)
    c *= ini.dealias_mask(*shape, *L)
    r = np.fft.irfftn(c, s=shape, axes=(0, 1, 2))
    return np.fft.rfftn(r / max(np.sqrt(np.mean(r ** 2)), 1e-300))


CASES = {
    "hd": dict(params=dict(nx=16, ny=16, nz=16, incompressible=1, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                           nu=0.01, cfl=0.5, **BIG_T), fields="u"),
    "mhd": dict(params=dict(nx=16, ny=8, nz=32, incompressible=1, mhd=1, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                            nu=0.01, nu_m=0.02, cfl=0.5, **BIG_T), fields="ub"),
    "hd_strat": dict(params=dict(nx=8, ny=16, nz=16, incompressible=1, boussinesq=1, stratification=1,
                                 strat_direction=2, N2=0.3, nu=0.01, nu_th=0.02, lx=1.0, ly=2.0, lz=1.5,
                                 cfl=0.5, **BIG_T), fields="ut"),
    "mhd_bouss_aniso": dict(params=dict(nx=16, ny=16, nz=16, incompressible=1, mhd=1, boussinesq=1,
                                        stratification=1, anisotropic_diffusion=1, N2=0.1, OmegaT2=1.0,
                                        reynolds_ani=50.0, nu=0.01, nu_m=0.01, nu_th=0.01, cfl=0.5,
                                        lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ubt", b0=1.0),
    "mhd_bouss_rkl2": dict(params=dict(nx=16, ny=8, nz=16, incompressible=1, mhd=1, boussinesq=1,
                                       stratification=1, anisotropic_diffusion=1, N2=0.1, OmegaT2=1.0,
                                       reynolds_ani=20.0, nu=0.01, nu_m=0.01, nu_th=0.01, cfl=0.5,
                                       supertimestepping=1, sts_mask=1 << 6, safety_sts=20, cfl_par=1.0,
                                       lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ubt", b0=1.0),
    "mhd_bouss_rkl2_x32": dict(params=dict(nx=32, ny=8, nz=16, incompressible=1, mhd=1, boussinesq=1,
                                           stratification=1, strat_direction=2, anisotropic_diffusion=1, N2=0.1,
                                           OmegaT2=1.0, reynolds_ani=30.0, nu=0.01, nu_m=0.01, nu_th=0.01,
                                           cfl=0.5, supertimestepping=1, sts_mask=1 << 6, safety_sts=20,
                                           cfl_par=1.0, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T),
                               fields="ubt", b0=1.0),
    "shear_mhd": dict(params=dict(nx=16, ny=16, nz=8, incompressible=1, mhd=1, shearing=1, rotating=1,
                                  shear=1.5, omega=1.0, nu=1e-3, nu_m=1e-3, cfl=0.5, lx=1.0, ly=1.5,
                                  lz=1.0, **BIG_T), fields="ub"),
    "quasi2d": dict(params=dict(nx=32, ny=16, nz=2, incompressible=1, mhd=1, nu=1e-3, nu_m=1e-3, cfl=0.5,
                                lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ub"),
    "mhd_x32": dict(params=dict(nx=32, ny=16, nz=16, incompressible=1, mhd=1, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                                nu=0.01, nu_m=0.02, cfl=0.5, **BIG_T), fields="ub"),
    "hd_x64": dict(params=dict(nx=64, ny=8, nz=16, incompressible=1, lx=TWO_PI, ly=1.0, lz=2.0, nu=0.005, cfl=0.5,
                               **BIG_T), fields="u"),
    "mhd_strat_x32": dict(params=dict(nx=32, ny=8, nz=16, incompressible=1, mhd=1, boussinesq=1, stratification=1,
                                      strat_direction=0, N2=-0.2, nu=0.01, nu_m=0.01, nu_th=0.02, cfl=0.5,
                                      lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ubt"),
    "hd_strat_x32": dict(params=dict(nx=32, ny=16, nz=8, incompressible=1, boussinesq=1, stratification=1,
                                     strat_direction=1, N2=0.5, nu=0.01, nu_th=0.01, cfl=0.5, lx=1.0, ly=1.0,
                                     lz=1.0, **BIG_T), fields="ut"),
    "hd_strat_sts_u_x32": dict(params=dict(nx=32, ny=16, nz=8, incompressible=1, boussinesq=1, stratification=1,
                                           strat_direction=1, N2=0.5, nu=0.02, nu_th=0.01, cfl=0.5,
                                           supertimestepping=1, sts_mask=0b111, safety_sts=10, cfl_par=1.0,
                                           lx=1.0, ly=1.0, lz=1.0, **BIG_T), fields="ut"),
    "mhd_x1024": dict(params=dict(nx=1024, ny=8, nz=8, incompressible=1, mhd=1, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                                  nu=0.01, nu_m=0.02, cfl=0.5, **BIG_T), fields="ub"),
    "mhd_bouss_rkl2_x512": dict(params=dict(nx=512, ny=8, nz=8, incompressible=1, mhd=1, boussinesq=1,
                                            stratification=1, strat_direction=2, anisotropic_diffusion=1, N2=0.1,
                                            OmegaT2=1.0, reynolds_ani=30.0, nu=0.01, nu_m=0.01, nu_th=0.01,
                                            cfl=0.5, supertimestepping=1, sts_mask=1 << 6, safety_sts=20,
                                            cfl_par=1.0, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T),
                                fields="ubt", b0=1.0),
    "hd_bouss_x32": dict(params=dict(nx=32, ny=16, nz=8, incompressible=1, boussinesq=1, nu=0.01, nu_th=0.01,
                                     cfl=0.5, lx=1.0, ly=1.0, lz=1.0, **BIG_T), fields="ut"),
    "mhd_bouss_x32": dict(params=dict(nx=32, ny=8, nz=16, incompressible=1, mhd=1, boussinesq=1, nu=0.01,
                                      nu_m=0.01, nu_th=0.02, cfl=0.5, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T),
                          fields="ubt"),
    # quasi-2D grids on which the plane kernels run (NZ = 2, NY = 64..512)
    "q2d_hd_y64": dict(params=dict(nx=32, ny=64, nz=2, incompressible=1, nu=0.01, cfl=0.5, lx=TWO_PI, ly=TWO_PI,
                                   lz=TWO_PI, **BIG_T), fields="u"),
    "q2d_bouss_y64": dict(params=dict(nx=64, ny=64, nz=2, incompressible=1, boussinesq=1, nu=0.01, nu_th=0.02,
                                      cfl=0.5, lx=1.0, ly=1.0, lz=1.0, **BIG_T), fields="ut"),
    "q2d_mhd_y256": dict(params=dict(nx=32, ny=256, nz=2, incompressible=1, mhd=1, nu=2e-3, nu_m=3e-3, cfl=0.5,
                                     lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ub"),
    "q2d_mhd_bouss_y128": dict(params=dict(nx=32, ny=128, nz=2, incompressible=1, mhd=1, boussinesq=1, nu=0.01,
                                           nu_m=0.01, nu_th=0.02, cfl=0.5, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                                           **BIG_T), fields="ubt"),
    "q2d_hbi_y512": dict(params=dict(nx=32, ny=512, nz=2, incompressible=1, mhd=1, boussinesq=1, stratification=1,
                                     strat_direction=1, anisotropic_diffusion=1, N2=0.2, OmegaT2=-1.0,
                                     reynolds_ani=20.0, nu=2e-3, nu_m=2e-3, nu_th=1e-3, cfl=0.5,
                                     supertimestepping=1, sts_mask=1 << 6, safety_sts=20, cfl_par=1.0,
                                     lx=1.0, ly=4.0, lz=1.0, **BIG_T), fields="ubt", b0=1.0),
    "q2d_mhd_y1024": dict(params=dict(nx=32, ny=1024, nz=2, incompressible=1, mhd=1, nu=2e-3, nu_m=3e-3, cfl=0.5,
                                      lx=TWO_PI, ly=TWO_PI, lz=TWO_PI, **BIG_T), fields="ub"),
    "heat": dict(params=dict(nx=64, ny=2, nz=2, heat_equation=1, nu_th=1e-3, antialiasing=0, cfl=1.5,
                             lx=1.0, ly=1.0, lz=1.0, **BIG_T), fields="t"),
}


def initial_state(name, seed=1):
    c = CASES[name]
    p = c["params"]
    shape = (p["nx"], p["ny"], p["nz"])
    L = (p.get("lx", 1.0), p.get("ly", 1.0), p.get("lz", 1.0))
    parts = []
    if "u" in c["fields"]:
        parts.append(0.5 * ini.random_solenoidal(shape, L, seed=seed, k0=2.0))
    if "b" in c["fields"]:
        b = 0.5 * ini.random_solenoidal(shape, L, seed=seed + 7, k0=2.0)
        if c.get("b0"):
            b[0, 0, 0, 0] += c["b0"] * np.prod(shape)  # uniform B_x, keeps B^2 away from 0
        parts.append(b)
    if "t" in c["fields"]:
        parts.append(0.3 * scalar_field(shape, L, seed + 13)[None])
    return np.concatenate(parts)


def relerr(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


# cases on which spk_main_loop takes the fused path (32 <= NX <= 1024, no shearing or in-RHS
# conduction): the stratified ones run the temperature and velocity groups in the same blocks,
# the super-time-stepped ones interleave RKL2 cycles (theta alone: its inverse x pass is made again
# after an even step's cycle; velocity: the next step starts from a fresh projection and transform)
FUSED_CASES = ["mhd_x32", "hd_x64", "hd_bouss_x32", "mhd_bouss_x32", "quasi2d", "mhd_strat_x32", "hd_strat_x32",
               "mhd_bouss_rkl2_x32", "hd_strat_sts_u_x32", "mhd_x1024", "mhd_bouss_rkl2_x512"]

# quasi-2D cases of the plane kernels: (RK3 stages, RKL2 stages) expected on them
PLANE_CASES = {"q2d_hd_y64": (True, False), "q2d_bouss_y64": (True, False), "q2d_mhd_y256": (True, False),
               "q2d_mhd_bouss_y128": (True, False), "q2d_hbi_y512": (True, True),
               "q2d_mhd_y1024": (False, False)}  # NY = 1024: the separate y passes and z kernel
