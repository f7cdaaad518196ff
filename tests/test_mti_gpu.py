"""3D Boussinesq MTI with anisotropic conduction at 256^3 on the GPU (FP64,
RKL2 on theta): gravity along -z, dT/dz = -1 (N2 = 0.1, OmegaT2 = 1, the
mapping of tests/mti-eigenmode/spooky.cfg, pinned on the oracle by the
512x4x4 eigenmode test), B = 1e-4 along x, chi = 1e-2, nu = eta = 1e-4,
seeded white-noise theta of amplitude 1e-4.

In the linear phase every mode (n, 0, 0) with k along B evolves on its own;
its growing eigenmode dominates after a few e-folds of the gap to the other
two roots (about 2 here), so the growth of |theta(n, 0, 0)| fitted over
t in [5, 8] must equal the largest root of the MTI cubic
(mti_disp_rel.py:44-70) for that kx.  On the oracle at 64^3 (dt ten times
larger) the same fit gives n = 2 within 0.2%; a window starting at t = 2.5
still carries the decaying roots (+2.7% for n = 2)."""
import importlib
import math

import numpy as np
import pytest

from spkmod import spk

ini = importlib.import_module("spooky-git_amd.initial")
pytestmark = pytest.mark.gpu

N = 256
MODES = (2, 4, 8)


def test_mti_256_linear_growth_rate():
    ctx = spk.Ctx(0)
    try:
        p = ini.mti_params(N)
        sol = spk.Solver(ctx, "f64", **p)
        sol.set_fields(ini.mti_white_noise((N, N, N), seed=7))
        th = sol.vars["TH"]
        ts, amps = [], []
        while True:
            t = sol.time()["t"]
            ts.append(t)
            amps.append([abs(sol.get_mode(th, n, 0, 0)) for n in MODES])
            if t >= 8.0:
                break
            sol.main_loop(10)
        tm = sol.time()
        ts, la = np.array(ts), np.log(np.array(amps))
        sel = ts >= 5.0
        assert sel.sum() >= 10
        # RKL2 really sub-cycles here: dt is several times the conduction limit
        assert tm["dt"] > 3 * tm["dt_sts"]
        for c, n in enumerate(MODES):
            slope = np.polyfit(ts[sel], la[sel, c], 1)[0]
            sigma = ini.mti_growth_rate(2 * math.pi * n, 1e-2, 1e-4, 1e-4, 1e-4, 0.1)
            assert slope == pytest.approx(sigma, rel=1e-2), (n, slope, sigma)
        sol.destroy()
    finally:
        ctx.close()
