"""CPU checks of the numpy restatements that the GPU tests compare against
(they pin the references themselves, independently of the GPU code):

- the closure's convolution, its transpose and its weight contraction are
  adjoint / derivative of one another (exact identities of linear maps);
- the float64 back-propagation of the closure restatement matches finite
  differences of its loss;
- GELU' is the derivative of GELU;
- the transfer terms of the spectrum file conserve what the dealiased
  equations conserve (numpy only);
- the shell spectrum of a single Fourier mode is its energy in one shell."""
import numpy as np
import pytest

import test_closure_gpu as tc
import test_diag_gpu as td


def test_conv_transpose_and_weight_contraction_are_adjoint():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((6, 6, 6, 5))
    W = rng.standard_normal((27, 4, 5))
    D = rng.standard_normal((6, 6, 6, 4))
    lhs = np.sum(tc.conv(X, W) * D)
    assert np.sum(X * tc.conv_t(D, W)) == pytest.approx(lhs, rel=1e-12)
    # L(W) = <conv(X, W), D> is linear in W with gradient conv_dw(D, X)
    g = tc.conv_dw(D, X)
    for (t, co, ci) in ((0, 0, 0), (13, 2, 3), (26, 3, 4), (21, 1, 2)):
        E = np.zeros_like(W)
        E[t, co, ci] = 1.0
        assert np.sum(tc.conv(X, W + E) * D) - lhs == pytest.approx(g[t, co, ci], rel=1e-9, abs=1e-9)


def test_taps_shift_the_input():
    """Tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) reads the input at (x + dx, y + dy, z + dz)."""
    X = np.random.default_rng(1).standard_normal((5, 6, 7, 1))
    W = np.zeros((27, 1, 1))
    W[9 * 2 + 3 * 0 + 1, 0, 0] = 1.0  # (dx, dy, dz) = (1, -1, 0)
    Y = tc.conv(X, W)
    assert np.array_equal(Y[1, 2, 3], X[2, 1, 3])


def test_closure_backprop_matches_finite_differences():
    n = 8
    rng = np.random.default_rng(2)
    p = tc.random_params(rng).astype(np.float64)
    u = rng.standard_normal((3, n, n, n))
    tgt = rng.standard_normal((n, n, n, 6)) * 0.3
    _, grads, loss = tc.reference(p, u, 1.0, tgt, exact=True)
    sl = tc.spk.Closure.layer_slices()
    h = 1e-5
    for l, (ws, bsl) in enumerate(sl):
        cin = tc.spk.Closure.CIN[l]
        for flat in (5, 1234, 27 * 32 * min(cin, 3) - 7):
            k = ws.start + flat % (ws.stop - ws.start)
            t, co, ci = np.unravel_index(k - ws.start, (27, 32, cin))
            if (l == 0 and ci >= 3) or (l == 3 and co >= 6):
                continue
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (tc.reference(pp, u, 1.0, tgt, exact=True)[2] - tc.reference(pm, u, 1.0, tgt, exact=True)[2]) / (2 * h)
            assert grads[l][0][t, co, ci] == pytest.approx(fd, rel=1e-5, abs=1e-9), (l, t, co, ci)
        kb = bsl.start + 1
        pp, pm = p.copy(), p.copy()
        pp[kb] += h
        pm[kb] -= h
        fd = (tc.reference(pp, u, 1.0, tgt, exact=True)[2] - tc.reference(pm, u, 1.0, tgt, exact=True)[2]) / (2 * h)
        assert grads[l][1][1] == pytest.approx(fd, rel=1e-5, abs=1e-9), l


def test_gelu_derivative():
    z = np.linspace(-5, 5, 41)
    h = 1e-6
    fd = (tc.gelu(z + h) - tc.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(tc.gelu_d(z) - fd)) < 1e-8


def test_transfer_terms_conserve_on_cpu():
    td.test_transfer_terms_conserve(None)


def test_shell_spectrum_of_one_mode():
    """cos(3 x) on the 16^3, 2 pi box: all of <u^2> = 1/2 sits in shell 3 (the
    shell spectrum sums Re(a b*) over the whole spectrum, so its bins add up to
    <u v>, here <u^2>)."""
    X, Y, Z = td.ini.grid_coords(16, 16, 16, td.L, td.L, td.L)
    c = np.fft.rfftn(np.cos(3 * X))
    s = td.shell(c, c, 8)
    assert s[3] == pytest.approx(0.5, rel=1e-12) and np.sum(np.abs(np.delete(s, 3))) < 1e-12
