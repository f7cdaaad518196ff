"""Python binding of libspk (the gfx950 SPOOKY solver) over its C ABI, include/spk.h.

Mirrors the reference's host objects for the hot path:
  FftPlan  -- init_plan / r2c_fft / c2r_fft / finish_cufft (cufft_routines.hpp:10-17)
  Solver   -- Fields + TimeStepping + RKLegendre device state and the calls
              CleanFieldDivergence, compute_dt, compute_dfield, RungeKutta3,
              compute_cycle_RKL2, HydroMHDAdvance and the main-loop body
              (symmetries.cu:114, compute_dt.cu:15, compute_dfield.cu:11,
              hydro_mhd_advance.cu:18-189, rkl.cu:437, supervisor.cu:169)
The library must be built (spooky-git_amd/build.py); nothing here falls back
to a CPU path: a missing library raises.
"""
import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libspk.so")

SPK_F64, SPK_F32 = 0, 1
_PREC = {"f64": SPK_F64, "fp64": SPK_F64, SPK_F64: SPK_F64, "f32": SPK_F32, "fp32": SPK_F32, SPK_F32: SPK_F32}

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

# Parameters defaults (parameters.cu:77-330)
DEFAULTS = dict(
    nx=32, ny=32, nz=32, incompressible=0, heat_equation=0, boussinesq=0, stratification=0,
    strat_direction=2, mhd=0, anisotropic_diffusion=0, supertimestepping=0, shearing=0,
    rotating=0, with2d=0, antialiasing=1, sts_mask=0, safety_sts=1,
    lx=1.0, ly=1.0, lz=1.0, nu=1.0, nu_m=1.0, nu_th=1.0, reynolds_ani=1.0, N2=0.0,
    OmegaT2=0.0, omega=0.0, shear=0.0, cfl=1.5, cfl_par=1.0, cfl_sts=1.0,
    safety_source=0.2, t_initial=0.0, t_final=1.0, toutput_flow=1.0,
)


class SpkParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in PARAM_INT_FIELDS] + [(n, ctypes.c_double) for n in PARAM_DBL_FIELDS]


def make_params(**kw):
    d = dict(DEFAULTS)
    for k in kw:
        if k not in d:
            raise KeyError("unknown parameter " + k)
    d.update(kw)
    p = SpkParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class SpkError(RuntimeError):
    pass


_lib = None

# (name, argtypes) of every exported entry point; restype is int unless noted
_VP, _I, _D, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
SIGNATURES = {
    "spk_abi_version": [],
    "spk_last_error": [],
    "spk_params_size": [],
    "spk_ctx_create": [_I, _VP, _VP],
    "spk_ctx_destroy": [_VP],
    "spk_ctx_sync": [_VP],
    "spk_malloc": [_VP, _SZ, _VP],
    "spk_free": [_VP, _VP],
    "spk_memcpy_h2d": [_VP, _VP, _VP, _SZ],
    "spk_memcpy_d2h": [_VP, _VP, _VP, _SZ],
    "spk_fft_plan_create": [_VP, _I, _I, _I, _I, _I, _VP],
    "spk_fft_plan_destroy": [_VP],
    "spk_fft_r2c": [_VP, _VP, _VP],
    "spk_fft_c2r": [_VP, _VP, _VP],
    "spk_fft_pass": [_VP, _I, _I, _VP, _VP],
    "spk_solver_create": [_VP, _VP, _I, _VP],
    "spk_solver_destroy": [_VP],
    "spk_solver_info": [_VP, _VP, _VP],
    "spk_solver_buffers": [_VP, _VP, _VP, _VP],
    "spk_solver_get_time": [_VP, _VP],
    "spk_solver_set_time": [_VP, _D, _I, _D],
    "spk_clean_divergence": [_VP],
    "spk_complex2real": [_VP],
    "spk_compute_dt": [_VP],
    "spk_compute_dfield": [_VP],
    "spk_rk3": [_VP],
    "spk_rkl2_cycle": [_VP],
    "spk_hydro_mhd_advance": [_VP],
    "spk_main_loop": [_VP, _I],
    "spk_kernel_timings": [_VP, _I, _VP, _I, _VP],
    "spk_solver_set_fused": [_VP, _I],
    "spk_solver_set_fused_points": [_VP, _I],
    "spk_solver_set_plane": [_VP, _I],
    "spk_main_loop_resume": [_VP, _I],
    "spk_solver_plane": [_VP, _VP, _VP],
    "spk_solver_pipeline": [_VP, _VP, _VP],
    "spk_energy": [_VP, _I, _VP],
    "spk_enstrophy": [_VP, _I, _VP],
    "spk_correlation": [_VP, _I, _I, _VP],
    "spk_spectrum": [_VP, _I, _I, _VP, _I, _VP],
    "spk_timevar": [_VP, ctypes.c_char_p, _VP],
    "spk_spectrum_bins": [_VP, _VP, _VP],
    "spk_spectrum_named": [_VP, ctypes.c_char_p, _VP, _I, _VP],
    "spk_user_timevar": [_VP, ctypes.c_char_p, _VP],
    "spk_snapshot_real": [_VP],
    "spk_solver_set_rkl_device": [_VP, _I],
    "spk_solver_rkl_plan": [_VP, _VP, _VP, _VP],
    "spk_ensemble_create": [_I, _I, _VP, _I, _I, _VP],
    "spk_ensemble_destroy": [_VP],
    "spk_ensemble_size": [_VP, _VP, _VP],
    "spk_ensemble_box": [_VP, _I, _VP],
    "spk_ensemble_run": [_VP, _I],
    "spk_ensemble_sync": [_VP],
    "spk_ensemble_set_graph": [_VP, _I],
    "spk_ensemble_set_lanes": [_VP, _I],
    "spk_ensemble_tune": [_VP, _I, _VP, _VP],
    "spk_ensemble_set_split": [_VP, _I],
    "spk_ensemble_layout": [_VP, _VP, _VP, _VP],
    "spk_closure_num_params": [],
    "spk_closure_create": [_VP, _I, _VP],
    "spk_closure_destroy": [_VP],
    "spk_closure_set_params": [_VP, _VP],
    "spk_closure_get_params": [_VP, _VP],
    "spk_closure_set_scales": [_VP, _D, _VP],
    "spk_closure_forward": [_VP, _VP, _VP],
    "spk_closure_grads": [_VP, _VP, _VP, _VP],
    "spk_closure_get_grads": [_VP, _VP],
    "spk_closure_train_step": [_VP, _VP, _VP, _D, _VP],
    "spk_closure_apply": [_VP, _VP],
    "spk_sgs_stress": [_VP, _VP, _I, _VP, _VP, _I],
    "spk_solver_set_forcing": [_VP, _I, _VP, _VP],
    "spk_compute_dt_ptr": [_VP, _VP, _VP],
    "spk_compute_dfield_ptr": [_VP, _VP, _VP, _VP],
    "spk_rk3_ptr": [_VP, _VP, _VP, _VP],
    "spk_hydro_mhd_advance_ptr": [_VP, _VP, _VP, _VP],
    "spk_clean_divergence_ptr": [_VP, _VP],
    "spk_project": [_VP, _VP],
    "spk_complex2real_ptr": [_VP, _VP, _VP, _I],
    "spk_rkl2_cycle_ptr": [_VP, _VP, _VP],
    "spk_rkl2_stage": [_VP, _I, _VP, _D, _VP, _VP, _VP, _VP],
    "spk_aniso_flux": [_VP, _VP, _VP, _VP],
    "spk_maxabs": [_VP, _VP, _VP],
    "spk_energy_ptr": [_VP, _VP, _VP],
    "spk_energy_noalloc": [_VP, _I, _I, _I, _VP, _VP],
    "spk_enstrophy_ptr": [_VP, _VP, _VP],
    "spk_correlation_ptr": [_VP, _VP, _VP, _VP],
    "spk_spectrum_ptr": [_VP, _VP, _VP, _VP, _I, _VP],
}


def lib():
    """Loads libspk.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SpkError("libspk.so not built: run `python spooky-git_amd/build.py` (%s)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            if not hasattr(L, name):  # reported by the ABI test; a call raises AttributeError
                continue
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.spk_last_error.restype = ctypes.c_char_p
        L.spk_params_size.restype = ctypes.c_size_t
        L.spk_closure_num_params.restype = ctypes.c_size_t
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise SpkError("spk error %d: %s" % (rc, lib().spk_last_error().decode()))


def _out_ptr():
    return ctypes.c_void_p()


class Ctx:
    """Device + stream (spk_ctx).  stream: an int hipStream_t handle or None."""

    def __init__(self, device=0, stream=None):
        h = ctypes.c_void_p()
        check(lib().spk_ctx_create(device, ctypes.c_void_p(stream or 0), ctypes.byref(h)))
        self.h = h

    def sync(self):
        check(lib().spk_ctx_sync(self.h))

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def close(self):
        if self.h:
            lib().spk_ctx_destroy(self.h)
            self.h = None


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().spk_malloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes, (a.nbytes, self.nbytes)
        check(lib().spk_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def download(self, dtype, shape):
        a = np.empty(shape, dtype=dtype)
        assert a.nbytes <= self.nbytes
        check(lib().spk_memcpy_d2h(self.ctx.h, a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes))
        return a

    def free(self):
        if self.ptr:
            lib().spk_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def download_ptr(ctx, ptr, dtype, shape):
    a = np.empty(shape, dtype=dtype)
    check(lib().spk_memcpy_d2h(ctx.h, a.ctypes.data_as(ctypes.c_void_p), ptr, a.nbytes))
    return a


def upload_ptr(ctx, ptr, arr):
    a = np.ascontiguousarray(arr)
    check(lib().spk_memcpy_h2d(ctx.h, ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))


class FftPlan:
    """3D R2C/C2R plan on the reference's padded in-place layout."""

    def __init__(self, ctx, nx, ny, nz, precision="f64", batch=1):
        self.ctx = ctx
        self.shape = (nx, ny, nz)
        self.prec = _PREC[precision]
        self.batch = batch
        h = ctypes.c_void_p()
        check(lib().spk_fft_plan_create(ctx.h, nx, ny, nz, self.prec, batch, ctypes.byref(h)))
        self.h = h

    def r2c(self, d_in, d_out):
        check(lib().spk_fft_r2c(self.h, d_in, d_out))

    def c2r(self, d_in, d_out):
        check(lib().spk_fft_c2r(self.h, d_in, d_out))

    def one_pass(self, axis, direction, d_in, d_out):
        check(lib().spk_fft_pass(self.h, axis, direction, d_in, d_out))

    def destroy(self):
        if self.h:
            lib().spk_fft_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


VAR_NAMES = ["VX", "VY", "VZ", "BX", "BY", "BZ", "TH", "VEL", "MAG", "NUM_FIELDS"]


class Solver:
    """Device state and time stepping of one run (Fields + TimeStepping + RKLegendre)."""

    def __init__(self, ctx, precision="f64", _handle=None, _params=None, **params):
        self.ctx = ctx
        self.params = _params if _params is not None else make_params(**params)
        self.prec = _PREC[precision]
        self.owned = _handle is None
        if _handle is None:
            h = ctypes.c_void_p()
            check(lib().spk_solver_create(ctx.h, ctypes.byref(self.params), self.prec, ctypes.byref(h)))
        else:
            h = _handle
        self.h = h
        v = (ctypes.c_int * 10)()
        ntc = ctypes.c_size_t()
        check(lib().spk_solver_info(h, v, ctypes.byref(ntc)))
        self.vars = dict(zip(VAR_NAMES, list(v)))
        self.nf = self.vars["NUM_FIELDS"]
        self.ntc = ntc.value
        p = self.params
        self.shape_c = (p.nx, p.ny, p.nz // 2 + 1)
        f, d, r = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().spk_solver_buffers(h, ctypes.byref(f), ctypes.byref(d), ctypes.byref(r)))
        self.d_fields, self.d_dfields, self.d_realbuf = f, d, r

    def destroy(self):
        if self.h and self.owned:
            lib().spk_solver_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # state transfer
    def set_fields(self, c):
        c = np.ascontiguousarray(c, dtype=np.complex128).reshape((self.nf,) + self.shape_c)
        # the buffer query also tells the library that the state is about to change
        # (the next main_loop_resume then starts from a fresh inverse x pass)
        f, d, r = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().spk_solver_buffers(self.h, ctypes.byref(f), ctypes.byref(d), ctypes.byref(r)))
        upload_ptr(self.ctx, self.d_fields, c)

    def get_fields(self):
        return download_ptr(self.ctx, self.d_fields, np.complex128, (self.nf,) + self.shape_c)

    def get_mode(self, field, i, j, k):
        """Complex value of spectral mode (i, j, k) of one state field (synchronizes)."""
        nx, ny, nzh = self.shape_c
        off = ((field * nx + i) * ny + j) * nzh + k
        a = download_ptr(self.ctx, ctypes.c_void_p(self.d_fields.value + 16 * off), np.complex128, (1,))
        return complex(a[0])

    def get_dfields(self):
        return download_ptr(self.ctx, self.d_dfields, np.complex128, (self.nf,) + self.shape_c)

    def get_realbuf(self):
        p = self.params
        a = download_ptr(self.ctx, self.d_realbuf, np.float64, (self.nf, p.nx, p.ny, 2 * (p.nz // 2 + 1)))
        return a[..., : p.nz]

    def time(self):
        o = (ctypes.c_double * 8)()
        check(lib().spk_solver_get_time(self.h, o))
        return dict(zip(["t", "dt", "tremap", "dt_hyp", "dt_par", "dt_sts", "t_lastsnap", "step"], list(o)))

    def set_time(self, t, step=0, t_lastsnap=0.0):
        check(lib().spk_solver_set_time(self.h, t, step, t_lastsnap))

    # hot path
    def clean_divergence(self):
        check(lib().spk_clean_divergence(self.h))

    def complex2real(self):
        check(lib().spk_complex2real(self.h))

    def compute_dt(self):
        check(lib().spk_compute_dt(self.h))
        return self.time()["dt"]

    def compute_dfield(self):
        check(lib().spk_compute_dfield(self.h))

    def rk3(self):
        check(lib().spk_rk3(self.h))

    def rkl2_cycle(self):
        check(lib().spk_rkl2_cycle(self.h))

    def advance(self):
        check(lib().spk_hydro_mhd_advance(self.h))

    def main_loop(self, nsteps, resume=False):
        """nsteps iterations of the main loop.  resume: continue the previous call,
        the state unchanged since (spk_main_loop_resume): the first step reuses the
        inverse x pass the last fused step left, where there is one."""
        check((lib().spk_main_loop_resume if resume else lib().spk_main_loop)(self.h, nsteps))

    KERNEL_CLASSES = ["inv_x", "inv_y", "zmid", "fwd_y", "fwd_x", "epilogue", "clean_div", "dt", "fwdx_epilogue"]

    def set_fused(self, mode):
        """0: unfused kernels; 1: autotuned choice (default); 2: fused kernel wherever allowed."""
        check(lib().spk_solver_set_fused(self.h, int(mode)))

    def set_fused_points(self, points):
        """Points per thread of the fused kernel's x line: 4, 8, or 0 for the autotuned tile."""
        check(lib().spk_solver_set_fused_points(self.h, int(points)))

    def set_plane(self, on):
        """Quasi-2D grids: the plane kernels (y passes and products in one kernel) on (default) or off."""
        check(lib().spk_solver_set_plane(self.h, 1 if on else 0))

    def plane(self):
        """(plane kernel in the RK3 stages, plane kernel in the RKL2 stages) of main_loop."""
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().spk_solver_plane(self.h, ctypes.byref(a), ctypes.byref(b)))
        return bool(a.value), bool(b.value)

    def pipeline(self):
        """(fused back end used by main_loop, points per thread of its x line)."""
        f, pts = ctypes.c_int(), ctypes.c_int()
        check(lib().spk_solver_pipeline(self.h, ctypes.byref(f), ctypes.byref(pts)))
        return bool(f.value), pts.value

    def kernel_timings(self, reps=10):
        """Average ms per launch of each pipeline kernel class (None where it does not run)."""
        out = (ctypes.c_double * 9)()
        n = ctypes.c_int()
        check(lib().spk_kernel_timings(self.h, reps, out, 9, ctypes.byref(n)))
        return {k: (v if v >= 0 else None) for k, v in zip(self.KERNEL_CLASSES, list(out))}

    # measurement
    def energy(self, f):
        o = ctypes.c_double()
        check(lib().spk_energy(self.h, f, ctypes.byref(o)))
        return o.value

    def enstrophy(self, f0):
        o = ctypes.c_double()
        check(lib().spk_enstrophy(self.h, f0, ctypes.byref(o)))
        return o.value

    def correlation(self, fa, fb):
        o = ctypes.c_double()
        check(lib().spk_correlation(self.h, fa, fb, ctypes.byref(o)))
        return o.value

    def timevar(self, name):
        """One timevar quantity by the reference's name (ev, Kx, hv, thetaB, dissgradT, ...);
        the real buffer must hold the current state (complex2real)."""
        o = ctypes.c_double()
        check(lib().spk_timevar(self.h, name.encode(), ctypes.byref(o)))
        return o.value

    def spectrum(self, fa, fb, cap=4096):
        out = np.zeros(cap)
        nb = ctypes.c_int()
        check(lib().spk_spectrum(self.h, fa, fb, out.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(nb)))
        return out[: nb.value].copy()

    def spectrum_bins(self):
        """(nbins, deltak) of the shell spectra."""
        nb, dk = ctypes.c_int(), ctypes.c_double()
        check(lib().spk_spectrum_bins(self.h, ctypes.byref(nb), ctypes.byref(dk)))
        return nb.value, dk.value

    def spectrum_named(self, name):
        """One spectrum of the spectrum file by the reference's name (Kx, bxbz,
        advec_x, emfpower_y, th_aniso_diss, ...); transfer spectra read the real
        buffer, which must hold the current state (complex2real)."""
        nb = self.spectrum_bins()[0]
        out = np.zeros(max(nb, 1))
        check(lib().spk_spectrum_named(self.h, name.encode(), out.ctypes.data_as(ctypes.c_void_p), nb, None))
        return out[:nb].copy()

    def user_timevar(self, name):
        """One quantity of the user timevar file (kpartheta, uservar1, uservar2)."""
        o = ctypes.c_double()
        check(lib().spk_user_timevar(self.h, name.encode(), ctypes.byref(o)))
        return o.value

    def set_rkl_device(self, on=True):
        """RKL2 cycles planned on the device (no host synchronization per cycle)."""
        check(lib().spk_solver_set_rkl_device(self.h, 1 if on else 0))

    def rkl_plan(self):
        """(stages of the last device plan, overflow flag, stages launched per cycle)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().spk_solver_rkl_plan(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_forcing(self, index, values):
        """Fixed forcing F on the given modes (index: int64 positions in a field;
        values: complex [nmodes, 3]); each step adds dt F."""
        idx = np.ascontiguousarray(index, dtype=np.int64)
        val = np.ascontiguousarray(np.asarray(values, dtype=np.complex128).reshape(len(idx), 3))
        check(lib().spk_solver_set_forcing(self.h, len(idx), idx.ctypes.data_as(ctypes.c_void_p),
                                           val.ctypes.data_as(ctypes.c_void_p)))

    def snapshot_real(self):
        """The real fields a snapshot holds into the real buffer (unsheared in a shearing box)."""
        check(lib().spk_snapshot_real(self.h))


class Ensemble:
    """Many independent boxes advanced together (spk_ensemble): one solver per
    box; boxes of one shape advanced as a batch (one launch per kernel for all
    of them), batches on their own HIP streams, two steps of every box per graph
    launch.  boxes: a list of parameter dicts (as Solver's keyword arguments)."""

    def __init__(self, boxes, precision="f32", lanes=0, device=0):
        self.params = [make_params(**b) for b in boxes]
        arr = (SpkParams * len(self.params))(*self.params)
        self.prec = _PREC[precision]
        h = ctypes.c_void_p()
        check(lib().spk_ensemble_create(device, len(self.params), arr, self.prec, lanes, ctypes.byref(h)))
        self.h = h
        # host transfers go through the null stream after an ensemble sync
        self.ctx = Ctx(device)
        self.boxes = []
        for i, p in enumerate(self.params):
            b = ctypes.c_void_p()
            check(lib().spk_ensemble_box(h, i, ctypes.byref(b)))
            self.boxes.append(Solver(self.ctx, precision, _handle=b, _params=p))

    def __len__(self):
        return len(self.boxes)

    def lanes(self):
        n, l = ctypes.c_int(), ctypes.c_int()
        check(lib().spk_ensemble_size(self.h, ctypes.byref(n), ctypes.byref(l)))
        return l.value

    def run(self, nsteps):
        check(lib().spk_ensemble_run(self.h, nsteps))

    def sync(self):
        check(lib().spk_ensemble_sync(self.h))

    def set_graph(self, on):
        check(lib().spk_ensemble_set_graph(self.h, 1 if on else 0))

    def set_lanes(self, lanes):
        """box-by-box launches over `lanes` streams (leaves batching)"""
        check(lib().spk_ensemble_set_lanes(self.h, lanes))

    def set_split(self, split):
        """split >= 1: each shape group advanced in `split` batches; 0: box by box"""
        check(lib().spk_ensemble_set_split(self.h, split))

    def layout(self):
        g, s, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().spk_ensemble_layout(self.h, ctypes.byref(g), ctypes.byref(s), ctypes.byref(b)))
        return dict(groups=g.value, split=s.value, batches=b.value, lanes=self.lanes())

    def tune(self, probe=2):
        """batched: the best split of 1, 2, 4; box by box: the best lane count"""
        best, ms = ctypes.c_int(), ctypes.c_double()
        check(lib().spk_ensemble_tune(self.h, probe, ctypes.byref(best), ctypes.byref(ms)))
        return best.value, ms.value

    def set_fields(self, i, c):
        self.sync()
        self.boxes[i].set_fields(c)

    def get_fields(self, i):
        self.sync()
        return self.boxes[i].get_fields()

    def time(self, i):
        self.sync()
        return self.boxes[i].time()

    def destroy(self):
        if self.h:
            for b in self.boxes:
                b.h = None
            lib().spk_ensemble_destroy(self.h)
            self.h = None
            self.ctx.close()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Closure:
    """Learned subgrid-stress closure (spk_closure): a 4-layer 3D conv net on a
    periodic n^3 grid, BF16 on the matrix cores.  Velocities and stresses are
    device buffers in the padded real FP64 layout [3 or 6][n][n][n + 2]."""

    CIN = (16, 32, 32, 32)

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, n
        h = ctypes.c_void_p()
        check(lib().spk_closure_create(ctx.h, n, ctypes.byref(h)))
        self.h = h
        self.nparams = int(lib().spk_closure_num_params())

    @classmethod
    def layer_slices(cls):
        """(W, b) slices of each layer in the flat parameter vector."""
        out, o = [], 0
        for cin in cls.CIN:
            nw = 27 * 32 * cin
            out.append((slice(o, o + nw), slice(o + nw, o + nw + 32)))
            o += nw + 32
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(lib().spk_closure_set_params(self.h, p.ctypes.data_as(ctypes.c_void_p)))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        check(lib().spk_closure_get_params(self.h, p.ctypes.data_as(ctypes.c_void_p)))
        return p

    def get_grads(self):
        g = np.empty(self.nparams, np.float32)
        check(lib().spk_closure_get_grads(self.h, g.ctypes.data_as(ctypes.c_void_p)))
        return g

    def set_scales(self, in_scale, out_scale):
        o = (ctypes.c_double * 6)(*[float(x) for x in out_scale])
        check(lib().spk_closure_set_scales(self.h, float(in_scale), o))

    def forward(self, d_u, d_tau):
        check(lib().spk_closure_forward(self.h, d_u, d_tau))

    def grads(self, d_u, d_tau):
        loss = ctypes.c_double()
        check(lib().spk_closure_grads(self.h, d_u, d_tau, ctypes.byref(loss)))
        return loss.value

    def train_step(self, d_u, d_tau, lr, want_loss=True):
        loss = ctypes.c_double()
        check(lib().spk_closure_train_step(self.h, d_u, d_tau, float(lr), ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def apply(self, sol):
        """SGS term of the solver's state after its step (same stream as the solver)."""
        check(lib().spk_closure_apply(self.h, sol.h))

    def destroy(self):
        if self.h:
            lib().spk_closure_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def sgs_stress(ctx, d_U_hi, N, d_ubar, d_tau, n):
    """Filtered velocity and subgrid stress on n^3 from a fine N^3 spectral velocity (spk_sgs_stress)."""
    check(lib().spk_sgs_stress(ctx.h, d_U_hi, N, d_ubar, d_tau, n))


def padded_real(r):
    nx, ny, nz = r.shape[-3:]
    out = np.zeros(r.shape[:-1] + (2 * (nz // 2 + 1),))
    out[..., :nz] = r
    return out
