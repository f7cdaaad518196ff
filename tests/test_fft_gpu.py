"""GPU parity of the HIP 3D R2C/C2R (spk_fft_*) against numpy golden vectors and
the FP64 oracle.  Tolerances: FP64 rel-L2 <= 1e-13, FP32 rel-L2 <= 2e-6."""
import os

import numpy as np
import pytest

import pyoracle as orc
from spkmod import spk

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
SHAPES = sorted({k.split("_", 1)[1] for k in GOLD.files})
TOL = {"f64": 1e-13, "f32": 2e-6}
RDT = {"f64": np.float64, "f32": np.float32}
CDT = {"f64": np.complex128, "f32": np.complex64}


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


def relerr(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


def gpu_r2c(ctx, r, prec, inplace, batch=1):
    nx, ny, nz = r.shape[-3:]
    nzh = nz // 2 + 1
    plan = spk.FftPlan(ctx, nx, ny, nz, prec, batch)
    buf = spk.padded_real(r).astype(RDT[prec])
    d_in = ctx.alloc(buf.nbytes)
    d_in.upload(buf)
    d_out = d_in if inplace else ctx.alloc(buf.nbytes)
    plan.r2c(d_in.ptr, d_out.ptr)
    shape = r.shape[:-1] + (nzh,)
    return d_out.download(CDT[prec], shape).astype(np.complex128)


def gpu_c2r(ctx, c, nz, prec, inplace, batch=1):
    nx, ny = c.shape[-3:-1]
    plan = spk.FftPlan(ctx, nx, ny, nz, prec, batch)
    a = np.ascontiguousarray(c.astype(CDT[prec]))
    d_in = ctx.alloc(a.nbytes)
    d_in.upload(a)
    d_out = d_in if inplace else ctx.alloc(a.nbytes)
    plan.c2r(d_in.ptr, d_out.ptr)
    out = d_out.download(RDT[prec], c.shape[:-1] + (2 * (nz // 2 + 1),)).astype(np.float64)
    if not inplace:  # input untouched
        back = d_in.download(CDT[prec], c.shape)
        assert np.array_equal(back, a)
    return out


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("tag", SHAPES)
def test_r2c_golden(ctx, tag, prec):
    r = GOLD["real_" + tag]
    for inplace in (True, False):
        got = gpu_r2c(ctx, r, prec, inplace)
        assert relerr(got, GOLD["spec_" + tag]) < TOL[prec], (tag, prec, inplace)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("tag", SHAPES)
def test_c2r_golden(ctx, tag, prec):
    c = GOLD["cspec_" + tag]
    ref = GOLD["creal_" + tag]
    nz = ref.shape[2]
    for inplace in (True, False):
        got = gpu_c2r(ctx, c, nz, prec, inplace)
        assert relerr(got[..., :nz], ref) < TOL[prec], (tag, prec, inplace)
        assert np.all(got[..., nz:] == 0.0)  # padding written as zero


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (128, 32, 256), (512, 512, 2), (2048, 2, 2), (16, 2048, 4), (4, 4, 1024)])
def test_against_oracle_and_roundtrip(ctx, shape, prec):
    rng = np.random.default_rng(11)
    r = rng.standard_normal(shape)
    c_gpu = gpu_r2c(ctx, r, prec, inplace=True)
    c_orc = orc.r2c(r)
    assert relerr(c_gpu, c_orc) < TOL[prec]
    back = gpu_c2r(ctx, c_orc, shape[2], prec, inplace=False)[..., : shape[2]]
    assert relerr(back, r) < TOL[prec]


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_batch_of_fields(ctx, prec):
    rng = np.random.default_rng(5)
    r = rng.standard_normal((3, 16, 8, 32))
    got = gpu_r2c(ctx, r, prec, inplace=True, batch=3)
    for f in range(3):
        assert relerr(got[f], orc.r2c(r[f])) < TOL[prec]


def test_roundtrip_256_cubed(ctx):
    """Full-size size-independent property: c2r(r2c(x)) == x (reference pin:
    rel-L2 < 1e-12 at 512^3, src/tests/tests.cu:121)."""
    n = 256
    rng = np.random.default_rng(3)
    r = rng.standard_normal((n, n, n))
    plan = spk.FftPlan(ctx, n, n, n, "f64", 1)
    buf = spk.padded_real(r)
    d = ctx.alloc(buf.nbytes)
    d.upload(buf)
    for _ in range(5):
        plan.r2c(d.ptr, d.ptr)
        plan.c2r(d.ptr, d.ptr)
    out = d.download(np.float64, buf.shape)[..., :n]
    assert relerr(out, r) < 1e-13
