"""Pins the oracle's 3D R2C/C2R against numpy golden vectors, a brute-force DFT
and closed forms.  Convention: forward unnormalized, inverse scaled by 1/N
(cufft_routines.cu:14-40)."""
import os

import numpy as np
import pytest

import pyoracle as orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
SHAPES = sorted({k.split("_", 1)[1] for k in GOLD.files})


def relerr(a, b):
    return np.linalg.norm(np.ravel(a - b)) / max(np.linalg.norm(np.ravel(b)), 1e-300)


@pytest.mark.parametrize("tag", SHAPES)
def test_r2c_matches_numpy_golden(tag):
    r = GOLD["real_" + tag]
    assert relerr(orc.r2c(r), GOLD["spec_" + tag]) < 1e-13


@pytest.mark.parametrize("tag", SHAPES)
def test_c2r_matches_numpy_golden(tag):
    c = GOLD["cspec_" + tag]
    nz = GOLD["creal_" + tag].shape[2]
    assert relerr(orc.c2r(c, nz), GOLD["creal_" + tag]) < 1e-13


def brute_rdft(r):
    nx, ny, nz = r.shape
    out = np.zeros((nx, ny, nz // 2 + 1), dtype=complex)
    x = np.arange(nx)[:, None, None]
    y = np.arange(ny)[None, :, None]
    z = np.arange(nz)[None, None, :]
    for a in range(nx):
        for b in range(ny):
            for c in range(nz // 2 + 1):
                ph = np.exp(-2j * np.pi * (a * x / nx + b * y / ny + c * z / nz))
                out[a, b, c] = np.sum(r * ph)
    return out


@pytest.mark.parametrize("shape", [(2, 2, 2), (4, 4, 2), (4, 2, 8), (8, 4, 4)])
def test_r2c_matches_brute_force(shape):
    r = np.random.default_rng(1).standard_normal(shape)
    assert relerr(orc.r2c(r), brute_rdft(r)) < 1e-13


def test_single_mode_closed_form():
    nx, ny, nz = 16, 8, 32
    x = np.arange(nx)[:, None, None] / nx
    z = np.arange(nz)[None, None, :] / nz
    r = np.broadcast_to(np.cos(2 * np.pi * (3 * x + 5 * z)) + 2.0, (nx, ny, nz))
    c = orc.r2c(r)
    n = nx * ny * nz
    expect = np.zeros_like(c)
    expect[0, 0, 0] = 2.0 * n
    expect[3, 0, 5] = n / 2
    expect[nx - 3, 0, 5] = 0.0  # (-3,+5) is not a mode of cos(3x+5z)
    assert np.abs(c - expect).max() < 1e-9 * n


def test_round_trip_identity_and_parseval():
    r = np.random.default_rng(7).standard_normal((32, 16, 64))
    c = orc.r2c(r)
    assert relerr(orc.c2r(c, 64), r) < 1e-14
    # Parseval on the half spectrum: kz=0 and kz=nz/2 planes count once, others twice
    w = np.full(c.shape, 2.0)
    w[:, :, 0] = 1.0
    w[:, :, -1] = 1.0
    assert abs(np.sum(w * np.abs(c) ** 2) / r.size - np.sum(r ** 2)) < 1e-10 * np.sum(r ** 2)


def test_c2r_ignores_imaginary_dc_and_nyquist():
    rng = np.random.default_rng(3)
    r = rng.standard_normal((8, 8, 8))
    c = orc.r2c(r)
    c2 = c.copy()
    c2[0, 0, 0] += 5j
    c2[0, 0, 4] += 3j
    assert relerr(orc.c2r(c2, 8), r) < 1e-14
