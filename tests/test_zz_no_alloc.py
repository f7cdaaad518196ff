"""spk_energy_noalloc, the entry point that allocates nothing: computeEnergy
(spooky_output_functions.cu:23-55) of a spectral field with its reduction
partials in the library's static scratch array.  Checked bit for bit against
spk_energy_ptr (same kernels, same partial count, same order) on two calls in a
row, the second on a smaller field, and against the definition in numpy.  The
first field needs every one of the 1024 partial slots, the second two, so the
second call runs over the first call's leftovers."""
import ctypes

import numpy as np
import pytest

from spkmod import spk

pytestmark = pytest.mark.gpu


def energy_numpy(a, npts):
    """computeEnergy: (sum |a|^2 - 1/2 sum_{kz=0} |a|^2) / N^2"""
    w = np.ones(a.shape)
    w[..., 0] = 0.5
    return float(np.sum(w * (a.real ** 2 + a.imag ** 2)) / float(npts) ** 2)


def field(shape, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    c = (nx, ny, nz // 2 + 1)
    return (rng.standard_normal(c) + 1j * rng.standard_normal(c)) * 1e3


def test_energy_noalloc_matches_energy_ptr():
    L = spk.lib()
    ctx = spk.Ctx(0)
    out = ctx.alloc(8)
    try:
        for shape, seed in (((128, 64, 64), 11), ((16, 8, 4), 12)):
            nx, ny, nz = shape
            print("spk_energy_noalloc: grid %s seed %d" % (shape, seed), flush=True)
            a = field(shape, seed)
            F = ctx.alloc(a.nbytes)
            F.upload(a)
            sol = spk.Solver(ctx, "f64", nx=nx, ny=ny, nz=nz, incompressible=1)
            ref = ctypes.c_double()
            spk.check(L.spk_energy_ptr(sol.h, F.ptr, ctypes.byref(ref)))
            spk.check(L.spk_energy_noalloc(ctx.h, nx, ny, nz, F.ptr, out.ptr))
            ctx.sync()
            got = float(out.download(np.float64, (1,))[0])
            sol.destroy()
            F.free()
            assert got == ref.value, (shape, got, ref.value)
            want = energy_numpy(a, nx * ny * nz)
            assert abs(got - want) <= 1e-13 * abs(want), (shape, got, want)
    finally:
        out.free()
        ctx.close()


def test_energy_noalloc_rejects_bad_arguments():
    L = spk.lib()
    ctx = spk.Ctx(0)
    out = ctx.alloc(16)
    F = ctx.alloc(16 * 4 * 4 * 3)
    try:
        base = out.ptr.value
        assert L.spk_energy_noalloc(ctx.h, 4, 4, 4, None, out.ptr) == -1
        assert L.spk_energy_noalloc(ctx.h, 4, 4, 4, F.ptr, None) == -1
        assert L.spk_energy_noalloc(ctx.h, 4, 4, 4, F.ptr, ctypes.c_void_p(base + 4)) == -1
        assert L.spk_energy_noalloc(ctx.h, 4, 4, 4, ctypes.c_void_p(F.ptr.value + 8), out.ptr) == -1
        assert L.spk_energy_noalloc(ctx.h, 0, 4, 4, F.ptr, out.ptr) == -1
        assert b"spk_energy_noalloc" in L.spk_last_error()
    finally:
        F.free()
        out.free()
        ctx.close()
