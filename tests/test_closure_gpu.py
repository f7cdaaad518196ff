"""Learned subgrid-stress closure (spk_closure) on the GPU against numpy.

The numpy restatement rounds to BF16 where the GPU stores BF16 (the input, the
weights, every hidden activation, the stored pre-activations and the
back-propagated gradients) and computes everything else in float64; the GPU
accumulates in FP32 on the matrix cores.  Tolerances are relative L2 over a
whole tensor.  Known answers that do not go through the restatement: biases
alone give a constant stress; a one-tap network is a shift of the input by
(dx, dy, dz) composed with GELU; the output-bias gradient is the mean error."""
import numpy as np
import pytest
from scipy.special import erf

from spkmod import spk

pytestmark = pytest.mark.gpu

N = 32


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


def bf(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def taps():
    for t in range(27):
        yield t, (t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1)


def conv(X, W):
    """Y[p][co] = sum_t sum_ci X[p + d_t][ci] W[t][co][ci] on the periodic grid."""
    Y = np.zeros(X.shape[:3] + (W.shape[1],))
    for t, d in taps():
        Y += np.roll(X, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2)) @ W[t].T
    return Y


def conv_t(D, W):
    """dX[q][ci] = sum_t sum_co D[q - d_t][co] W[t][co][ci]."""
    X = np.zeros(D.shape[:3] + (W.shape[2],))
    for t, d in taps():
        X += np.roll(D, shift=d, axis=(0, 1, 2)) @ W[t]
    return X


def conv_dw(D, X):
    """dW[t][co][ci] = sum_p D[p][co] X[p + d_t][ci]."""
    P = D.shape[0] * D.shape[1] * D.shape[2]
    out = np.zeros((27, D.shape[3], X.shape[3]))
    for t, d in taps():
        Xs = np.roll(X, shift=(-d[0], -d[1], -d[2]), axis=(0, 1, 2))
        out[t] = D.reshape(P, -1).T @ Xs.reshape(P, -1)
    return out


def unpack(params):
    Ws, bs = [], []
    for (ws, bsl), cin in zip(spk.Closure.layer_slices(), spk.Closure.CIN):
        Ws.append(params[ws].astype(np.float64).reshape(27, 32, cin))
        bs.append(params[bsl].astype(np.float64))
    return Ws, bs


def reference(params, u, in_scale, tgt=None, exact=False):
    """numpy forward (and backward when tgt is given); u: [3][n][n][n].  exact:
    float64 throughout (no BF16 rounding), for checks of the restatement itself."""
    global bf
    rnd = bf
    if exact:
        bf = lambda x: np.asarray(x, dtype=np.float64)  # noqa: E731
    try:
        return _reference(params, u, in_scale, tgt)
    finally:
        bf = rnd


def _reference(params, u, in_scale, tgt=None):
    n = u.shape[-1]
    Ws, bs = unpack(params)
    Wb = [bf(w) for w in Ws]
    x0 = np.zeros((n, n, n, 16))
    x0[..., :3] = bf(np.moveaxis(u.astype(np.float32) * np.float32(in_scale), 0, -1))
    acts, zs = [x0], []
    for l in range(3):
        z = conv(acts[-1], Wb[l]) + bs[l]
        zs.append(bf(z))
        acts.append(bf(gelu(z)))
    out = conv(acts[-1], Wb[3])[..., :6] + bs[3][:6]
    if tgt is None:
        return out
    P = n ** 3
    d = np.zeros((n, n, n, 32))
    d[..., :6] = bf((2.0 / (6 * P)) * (out - tgt))
    grads = [None] * 4
    for l in (3, 2, 1, 0):
        gw = conv_dw(d, acts[l])
        grads[l] = (gw, d.reshape(P, 32).sum(0))
        if l > 0:
            d = bf(conv_t(d, Wb[l]) * gelu_d(zs[l - 1]))
    loss = np.mean((out - tgt) ** 2)
    return out, grads, loss


def padded(fields):
    f = np.asarray(fields)
    out = np.zeros(f.shape[:-1] + (f.shape[-1] + 2,))
    out[..., : f.shape[-1]] = f
    return out


def random_params(rng, scale=1.0):
    p = np.zeros(sum(27 * 32 * c + 32 for c in spk.Closure.CIN), np.float32)
    for (ws, bsl), cin in zip(spk.Closure.layer_slices(), spk.Closure.CIN):
        w = rng.standard_normal((27, 32, cin)) * scale / np.sqrt(27 * 3 if cin == 16 else 27 * cin)
        if cin == 16:
            w[..., 3:] = 0.0
        p[ws] = w.reshape(-1)
        p[bsl] = 0.1 * rng.standard_normal(32)
    ws, bsl = spk.Closure.layer_slices()[3]
    w = p[ws].reshape(27, 32, 32)
    w[:, 6:, :] = 0.0
    p[ws] = w.reshape(-1)
    p[bsl][6:] = 0.0
    return p


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def device_fields(ctx, f):
    buf = ctx.alloc(f.nbytes)
    buf.upload(f)
    return buf


@pytest.mark.parametrize("n", [N, 128])
def test_forward_matches_numpy_and_known_answers(ctx, n):
    """At 32^3 and at the closure's own 128^3 grid."""
    N = n
    rng = np.random.default_rng(1)
    cl = spk.Closure(ctx, N)
    assert cl.nparams == sum(27 * 32 * c + 32 for c in spk.Closure.CIN)
    u = rng.standard_normal((3, N, N, N))
    du = device_fields(ctx, padded(u))
    dtau = ctx.alloc(6 * N * N * (N + 2) * 8)
    scales = [1.0, 2.0, 0.5, 1.5, 3.0, 0.25]
    cl.set_scales(0.7, scales)
    # biases only: a constant stress
    p = np.zeros(cl.nparams, np.float32)
    p[spk.Closure.layer_slices()[3][1]][:6] = [1, 2, 3, 4, 5, 6]
    cl.set_params(p)
    cl.forward(du.ptr, dtau.ptr)
    tau = dtau.download(np.float64, (6, N, N, N + 2))
    for c in range(6):
        assert np.all(tau[c, ..., :N] == (c + 1) * scales[c]) and np.all(tau[c, ..., N:] == 0)
    # one tap per layer: tau_c = g(g(g(u_c(x + 1, y, z - 1)))) for c < 3 (layer 1 shifts by
    # (+1, 0, -1): tap 9 * 2 + 3 * 1 + 0 = 21; layers 2-4 take the centre tap 13)
    p = np.zeros(cl.nparams, np.float32)
    sl = spk.Closure.layer_slices()
    for l, (ws, _) in enumerate(sl):
        cin = spk.Closure.CIN[l]
        w = np.zeros((27, 32, cin), np.float32)
        for c in range(3):
            w[21 if l == 0 else 13, c, c] = 1.0
        p[ws] = w.reshape(-1)
    cl.set_params(p)
    cl.set_scales(1.0, [1.0] * 6)
    cl.forward(du.ptr, dtau.ptr)
    tau = dtau.download(np.float64, (6, N, N, N + 2))[..., :N]
    a = bf(np.roll(u.astype(np.float32), shift=(-1, 0, 1), axis=(1, 2, 3)))
    for _ in range(3):
        a = bf(gelu(a))
    assert rel(tau[:3], a) < 1e-3 and np.all(tau[3:] == 0)
    # random network against the numpy restatement: the whole grid at 32^3; at
    # 128^3 a 24^3 block that wraps around the periodic boundary, from a 32^3
    # crop of the input (each of the 4 layers consumes one halo point)
    p = random_params(rng)
    cl.set_params(p)
    cl.set_scales(0.7, scales)
    cl.forward(du.ptr, dtau.ptr)
    tau = dtau.download(np.float64, (6, N, N, N + 2))[..., :N]
    sc = np.array(scales)[:, None, None, None]
    if N <= 32:
        ref = np.moveaxis(reference(p, u, 0.7), -1, 0) * sc
        assert rel(tau, ref) < 1e-2, rel(tau, ref)
    else:
        m, h = 24, 4
        idx = [(np.arange(-h, m + h) + s0) % N for s0 in (N - 10, 50, N - 3)]
        crop = u[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]]
        ref = np.moveaxis(reference(p, crop, 0.7), -1, 0)[:, h:h + m, h:h + m, h:h + m] * sc
        inner = [ix[h:h + m] for ix in idx]
        got = tau[:, inner[0]][:, :, inner[1]][:, :, :, inner[2]]
        assert rel(got, ref) < 1e-2, rel(got, ref)
    cl.destroy()


def test_gradients_match_numpy_backprop(ctx):
    rng = np.random.default_rng(2)
    cl = spk.Closure(ctx, N)
    p = random_params(rng)
    cl.set_params(p)
    cl.set_scales(1.0, [1.0] * 6)
    u = rng.standard_normal((3, N, N, N))
    tgt = rng.standard_normal((6, N, N, N)) * 0.3
    du = device_fields(ctx, padded(u))
    dt = device_fields(ctx, padded(tgt))
    loss = cl.grads(du.ptr, dt.ptr)
    g = cl.get_grads()
    out, grads, loss_ref = reference(p, u, 1.0, np.moveaxis(tgt, 0, -1))
    assert loss == pytest.approx(loss_ref, rel=5e-3)
    for l, (ws, bsl) in enumerate(spk.Closure.layer_slices()):
        cin = spk.Closure.CIN[l]
        gw_ref = grads[l][0][:, :, :cin] if cin == 16 else grads[l][0]
        gw = g[ws].reshape(27, 32, cin)
        if l == 3:  # output channels >= 6 have no gradient
            assert np.all(gw[:, 6:, :] == 0)
        assert rel(gw, gw_ref) < 2e-2, (l, rel(gw, gw_ref))
        assert rel(g[bsl], grads[l][1]) < 2e-2, (l, rel(g[bsl], grads[l][1]))
    # output-layer bias: dL/db_c = 2 mean_p (out_c - tgt_c) / 6, independent of the restatement
    b4 = g[spk.Closure.layer_slices()[3][1]][:6]
    direct = 2.0 * np.mean(out - np.moveaxis(tgt, 0, -1), axis=(0, 1, 2)) / 6.0
    assert np.max(np.abs(b4 - direct)) < 1e-2 * np.max(np.abs(direct))
    cl.destroy()


def test_training_fits_a_teacher(ctx):
    """A student network trained with Adam on a teacher network's output lowers
    the loss below half in 150 steps (0.40 of the first on the MI355X)."""
    rng = np.random.default_rng(3)
    teacher = spk.Closure(ctx, N)
    teacher.set_params(random_params(rng, 1.5))
    teacher.set_scales(1.0, [1.0] * 6)
    u = rng.standard_normal((3, N, N, N))
    du = device_fields(ctx, padded(u))
    dt = ctx.alloc(6 * N * N * (N + 2) * 8)
    teacher.forward(du.ptr, dt.ptr)
    ctx.sync()
    student = spk.Closure(ctx, N)
    student.set_params(random_params(rng, 1.0))
    student.set_scales(1.0, [1.0] * 6)
    first = student.train_step(du.ptr, dt.ptr, 3e-3)
    for _ in range(148):
        student.train_step(du.ptr, dt.ptr, 3e-3, want_loss=False)
    last = student.train_step(du.ptr, dt.ptr, 3e-3)
    print("teacher fit: loss %.4g -> %.4g" % (first, last))
    assert np.isfinite(last) and last < 0.5 * first, (first, last)
    teacher.destroy()
    student.destroy()


ini = __import__("importlib").import_module("spooky-git_amd.initial")
TWO_PI = 2 * np.pi


def hd_params(n, nu=0.01):
    return dict(nx=n, ny=n, nz=n, incompressible=1, nu=nu, cfl=0.5, lx=TWO_PI, ly=TWO_PI, lz=TWO_PI,
                t_final=1e9, toutput_flow=1e9)


def test_forcing_single_shear_mode(ctx):
    """U += dt F after each step: on a single shear mode (no nonlinear term) the
    state follows x_{n+1} = R(-nu k^2 dt_n) x_n + dt_n F, R(z) = 1 + z + z^2/2 + z^3/6
    (the RK3 amplification), with the run's own dt_n."""
    n, nu = 32, 0.05
    sol = spk.Solver(ctx, "f64", **hd_params(n, nu))
    nzh = n // 2 + 1
    i1 = (0 * n + 1) * nzh          # (kx, ky, kz) = (0, 1, 0)
    i2 = (0 * n + (n - 1)) * nzh    # (0, -1, 0), its Hermitian partner on kz = 0
    c = (0.3 + 0.2j) * n ** 3
    sol.set_forcing([i1, i2], [[c, 0, 0], [np.conj(c), 0, 0]])
    x = 0.0 + 0.0j
    for _ in range(20):
        sol.main_loop(1)
        dt = sol.time()["dt"]
        z = -nu * dt
        x = (1 + z + z * z / 2 + z ** 3 / 6) * x + dt * c
    U = sol.get_fields()
    assert abs(U[0].reshape(-1)[i1] - x) < 1e-12 * abs(x)
    assert abs(U[0].reshape(-1)[i2] - np.conj(x)) < 1e-12 * abs(x)
    U[0].reshape(-1)[[i1, i2]] = 0
    assert np.max(np.abs(U)) < 1e-12 * abs(x)
    with pytest.raises(RuntimeError):
        sol.set_forcing([n ** 3], [[1, 0, 0]])
    sol.destroy()


def test_sgs_stress_two_modes(ctx):
    """u_x = cos(3 y) + cos(12 y) on 64^3, coarse grid 32^3 (cut kc = 10): ubar_x =
    cos(3 y) and tau_xx = F(u_x^2) - F(ubar_x^2) = 1/2 + cos(9 y); the other
    components vanish."""
    N, n = 64, 32
    X, Y, Z = ini.grid_coords(N, N, N, TWO_PI, TWO_PI, TWO_PI)
    u = np.stack([np.cos(3 * Y) + np.cos(12 * Y), 0 * Y, 0 * Y])
    U = np.stack([np.fft.rfftn(f) for f in u])
    dU = device_fields(ctx, U)
    dub = ctx.alloc(3 * n * n * (n + 2) * 8)
    dtau = ctx.alloc(6 * n * n * (n + 2) * 8)
    spk.sgs_stress(ctx, dU.ptr, N, dub.ptr, dtau.ptr, n)
    ub = dub.download(np.float64, (3, n, n, n + 2))[..., :n]
    tau = dtau.download(np.float64, (6, n, n, n + 2))[..., :n]
    x, y, z = ini.grid_coords(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    assert np.max(np.abs(ub[0] - np.cos(3 * y))) < 1e-12 and np.max(np.abs(ub[1:])) < 1e-12
    assert np.max(np.abs(tau[0] - (0.5 + np.cos(9 * y)))) < 1e-12
    assert np.max(np.abs(tau[1:])) < 1e-12


def test_closure_apply_adds_the_projected_divergence(ctx):
    n = 32
    rng = np.random.default_rng(4)
    p = hd_params(n)
    sol = spk.Solver(ctx, "f64", **p)
    sol.set_fields(0.5 * ini.random_solenoidal((n, n, n), (TWO_PI,) * 3, seed=9, k0=3.0))
    sol.main_loop(1)
    cl = spk.Closure(ctx, n)
    cl.set_params(random_params(rng))
    scales = [0.1, 0.05, 0.05, 0.1, 0.05, 0.1]
    cl.set_scales(2.0, scales)
    before = sol.get_fields()
    dt = sol.time()["dt"]
    # the stress the network gives on this state
    sol.complex2real()
    dtau = ctx.alloc(6 * n * n * (n + 2) * 8)
    cl.forward(sol.d_realbuf.value, dtau.ptr)
    tau = dtau.download(np.float64, (6, n, n, n + 2))[..., :n]
    cl.apply(sol)
    after = sol.get_fields()
    T = [np.fft.rfftn(t) for t in tau]
    KX, KY, KZ = ini.wavenumbers(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    m = ini.dealias_mask(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    d = [KX * T[0] + KY * T[1] + KZ * T[2], KX * T[1] + KY * T[3] + KZ * T[4], KX * T[2] + KY * T[4] + KZ * T[5]]
    g = [-1j * m * di for di in d]
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    kg = KX * g[0] + KY * g[1] + KZ * g[2]
    with np.errstate(invalid="ignore", divide="ignore"):
        g = [np.where(K2 > 0, gi - k * kg / np.where(K2 > 0, K2, 1), gi) for gi, k in zip(g, (KX, KY, KZ))]
    expect = before + dt * np.stack(g)
    assert np.max(np.abs(after - expect)) < 1e-10 * np.max(np.abs(after - before))
    assert np.max(np.abs(after - before)) > 0
    cl.destroy()
    sol.destroy()
