"""ctypes binding of the FP64 CPU oracle (oracle/spk_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker or the CPU baseline, never
as the thing being measured or shipped.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libspk_oracle.so")

_lib = None

PARAM_INT_FIELDS = [
    "nx", "ny", "nz", "incompressible", "heat_equation", "boussinesq", "stratification",
    "strat_direction", "mhd", "anisotropic_diffusion", "supertimestepping", "shearing",
    "rotating", "with2d", "antialiasing", "sts_mask", "safety_sts",
]
PARAM_DBL_FIELDS = [
    "lx", "ly", "lz", "nu", "nu_m", "nu_th", "reynolds_ani", "N2", "OmegaT2", "omega",
    "shear", "cfl", "cfl_par", "cfl_sts", "safety_source", "t_initial", "t_final",
    "toutput_flow",
]


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in PARAM_INT_FIELDS] + [(n, ctypes.c_double) for n in PARAM_DBL_FIELDS]


# Parameters defaults: parameters.cu:77-330 (reynolds* default 1, cfl 1.5, ...)
DEFAULTS = dict(
    nx=32, ny=32, nz=32, incompressible=0, heat_equation=0, boussinesq=0, stratification=0,
    strat_direction=2, mhd=0, anisotropic_diffusion=0, supertimestepping=0, shearing=0,
    rotating=0, with2d=0, antialiasing=1, sts_mask=0, safety_sts=1,
    lx=1.0, ly=1.0, lz=1.0, nu=1.0, nu_m=1.0, nu_th=1.0, reynolds_ani=1.0, N2=0.0,
    OmegaT2=0.0, omega=0.0, shear=0.0, cfl=1.5, cfl_par=1.0, cfl_sts=1.0,
    safety_source=0.2, t_initial=0.0, t_final=1.0, toutput_flow=1.0,
)


def make_params(cls=OrcParams, **kw):
    d = dict(DEFAULTS)
    for k in kw:
        if k not in d:
            raise KeyError("unknown parameter " + k)
    d.update(kw)
    p = cls()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.orc_params_size.restype = ctypes.c_size_t
        L.orc_num_threads.restype = ctypes.c_int
        for fn in ("orc_fft_r2c", "orc_fft_c2r"):
            getattr(L, fn).argtypes = [ctypes.c_int] * 3 + [vp, vp]
            getattr(L, fn).restype = None
        L.orc_fft1d.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_wavevector.argtypes = [ctypes.POINTER(OrcParams), vp, vp, vp]
        L.orc_sim_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_sim_create.restype = vp
        for fn in ("orc_sim_fields", "orc_sim_dfields", "orc_sim_realbuf", "orc_sim_kvec"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = dp
        L.orc_sim_nfields.argtypes = [vp]
        L.orc_sim_ntc.argtypes = [vp]
        L.orc_sim_ntc.restype = ctypes.c_size_t
        L.orc_sim_vars.argtypes = [vp, vp]
        L.orc_sim_get_time.argtypes = [vp, vp]
        L.orc_sim_set_time.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        for fn in ("orc_sim_destroy", "orc_clean_divergence", "orc_complex2real", "orc_compute_dt",
                   "orc_rk3", "orc_rkl2_cycle", "orc_hydro_mhd_advance", "orc_shift_time",
                   "orc_remap_all_fields"):
            getattr(L, fn).argtypes = [vp]
            getattr(L, fn).restype = None
        L.orc_compute_dfield.argtypes = [vp, ctypes.c_int]
        L.orc_main_loop.argtypes = [vp, ctypes.c_int]
        L.orc_shear_wavevector.argtypes = [vp, ctypes.c_double]
        L.orc_rkl2_coeffs.argtypes = [ctypes.c_double, ctypes.c_double, vp, ctypes.c_int]
        L.orc_rkl2_coeffs.restype = ctypes.c_int
        for fn in ("orc_energy", "orc_enstrophy"):
            getattr(L, fn).argtypes = [vp, ctypes.c_int]
            getattr(L, fn).restype = ctypes.c_double
        L.orc_correlation.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_correlation.restype = ctypes.c_double
        L.orc_spectrum.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int]
        L.orc_spectrum.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def padded_real(r):
    """[nx,ny,nz] real -> [nx,ny,2(nz//2+1)] padded layout (hdf5_io.cu:93-94)."""
    nx, ny, nz = r.shape
    out = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    out[:, :, :nz] = r
    return out


def r2c(r):
    """3D forward transform of a [nx,ny,nz] real array -> complex [nx,ny,nz//2+1]."""
    nx, ny, nz = r.shape
    buf = padded_real(np.asarray(r, dtype=np.float64))
    out = np.zeros((nx, ny, nz // 2 + 1), dtype=np.complex128)
    lib().orc_fft_r2c(nx, ny, nz, _ptr(buf), _ptr(out))
    return out


def c2r(c, nz):
    """3D inverse transform (1/N scaled) of complex [nx,ny,nz//2+1] -> real [nx,ny,nz]."""
    c = np.ascontiguousarray(c, dtype=np.complex128)
    nx, ny, _ = c.shape
    out = np.zeros((nx, ny, 2 * (nz // 2 + 1)))
    lib().orc_fft_c2r(nx, ny, nz, _ptr(c), _ptr(out))
    return out[:, :, :nz].copy()


def wavevector(**kw):
    p = make_params(**kw)
    ntc = p.nx * p.ny * (p.nz // 2 + 1)
    kvec = np.zeros(3 * ntc)
    mask = np.zeros(ntc)
    scal = np.zeros(6)
    lib().orc_wavevector(ctypes.byref(p), _ptr(kvec), _ptr(mask), _ptr(scal))
    return kvec.reshape(3, ntc), mask, scal


class Sim:
    """One oracle simulation (Fields + Physics + TimeStepping state)."""

    def __init__(self, **kw):
        self.params = make_params(**kw)
        self.h = lib().orc_sim_create(ctypes.byref(self.params))
        v = np.zeros(10, dtype=np.int32)
        lib().orc_sim_vars(self.h, _ptr(v))
        self.vars = dict(zip(["VX", "VY", "VZ", "BX", "BY", "BZ", "TH", "VEL", "MAG", "NUM_FIELDS"], v.tolist()))
        self.nf = self.vars["NUM_FIELDS"]
        self.ntc = lib().orc_sim_ntc(self.h)
        p = self.params
        self.shape_c = (p.nx, p.ny, p.nz // 2 + 1)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def _view(self, fn, n):
        ptr = getattr(lib(), fn)(self.h)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    @property
    def fields(self):
        """complex view [nf, nx, ny, nzh] (shares memory with the oracle)."""
        a = self._view("orc_sim_fields", 2 * self.ntc * self.nf)
        return a.view(np.complex128).reshape((self.nf,) + self.shape_c)

    @property
    def dfields(self):
        a = self._view("orc_sim_dfields", 2 * self.ntc * self.nf)
        return a.view(np.complex128).reshape((self.nf,) + self.shape_c)

    @property
    def realbuf(self):
        p = self.params
        a = self._view("orc_sim_realbuf", 2 * self.ntc * self.nf)
        return a.reshape(self.nf, p.nx, p.ny, 2 * (p.nz // 2 + 1))[..., : p.nz]

    @property
    def kvec(self):
        return self._view("orc_sim_kvec", 3 * self.ntc).reshape((3,) + self.shape_c)

    def set_fields(self, c):
        self.fields[...] = c

    def time(self):
        o = np.zeros(8)
        lib().orc_sim_get_time(self.h, _ptr(o))
        return dict(zip(["t", "dt", "tremap", "dt_hyp", "dt_par", "dt_sts", "t_lastsnap", "step"], o.tolist()))

    def set_time(self, t, step=0, t_lastsnap=0.0):
        lib().orc_sim_set_time(self.h, t, step, t_lastsnap)

    def clean_divergence(self):
        lib().orc_clean_divergence(self.h)

    def complex2real(self):
        lib().orc_complex2real(self.h)

    def compute_dt(self):
        lib().orc_compute_dt(self.h)
        return self.time()["dt"]

    def compute_dfield(self, stage_step=0):
        lib().orc_compute_dfield(self.h, stage_step)

    def rk3(self):
        lib().orc_rk3(self.h)

    def rkl2_cycle(self):
        lib().orc_rkl2_cycle(self.h)

    def advance(self):
        lib().orc_hydro_mhd_advance(self.h)

    def main_loop(self, nsteps):
        lib().orc_main_loop(self.h, nsteps)

    def energy(self, f):
        return lib().orc_energy(self.h, f)

    def enstrophy(self, f0):
        return lib().orc_enstrophy(self.h, f0)

    def correlation(self, fa, fb):
        return lib().orc_correlation(self.h, fa, fb)

    def spectrum(self, fa, fb, cap=4096):
        out = np.zeros(cap)
        nb = lib().orc_spectrum(self.h, fa, fb, _ptr(out), cap)
        return out[:nb].copy()


def rkl2_coeffs(dt, dt_sts):
    s = lib().orc_rkl2_coeffs(dt, dt_sts, None, 0)
    coef = np.zeros(4 * (s + 1))
    lib().orc_rkl2_coeffs(dt, dt_sts, _ptr(coef), s)
    return s, coef[: 4 * s].reshape(s, 4)
