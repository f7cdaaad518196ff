"""Writes the FFT golden vectors in tests/golden/fft_golden.npz.

The vectors come from numpy.fft.rfftn / irfftn, which share the reference's
transform convention: forward unnormalized, inverse scaled by 1/(NX NY NZ)
(cufft_routines.cu:14-40, scaleKernel at cufft_routines.cu:30).  Inputs are
seeded, so re-running this script reproduces the committed file bit for bit.

Layout of each entry (shape nx x ny x nz):
  real_<tag>   : float64 [nx, ny, nz]       real input
  spec_<tag>   : complex128 [nx, ny, nz//2+1] rfftn(real)
  cspec_<tag>  : complex128 [nx, ny, nz//2+1] Hermitian-consistent spectrum
  creal_<tag>  : float64 [nx, ny, nz]       irfftn(cspec) * 1 (numpy's 1/N included)
"""
import os

import numpy as np

SHAPES = [(2, 2, 2), (4, 4, 2), (8, 8, 8), (16, 8, 4), (4, 8, 16), (32, 32, 32), (1024, 2, 2), (64, 4, 4)]


def tag(shape):
    return "x".join(str(s) for s in shape)


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for shp in SHAPES:
        nx, ny, nz = shp
        r = rng.standard_normal(shp)
        out["real_" + tag(shp)] = r
        out["spec_" + tag(shp)] = np.fft.rfftn(r)
        r2 = rng.standard_normal(shp)
        c = np.fft.rfftn(r2)  # a spectrum that is exactly the transform of a real field
        out["cspec_" + tag(shp)] = c
        out["creal_" + tag(shp)] = np.fft.irfftn(c, s=shp, axes=(0, 1, 2))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "arrays")


if __name__ == "__main__":
    main()
