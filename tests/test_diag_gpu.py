"""Timevar quantities (spk_timevar: WriteTimevarOutput of
spooky_output_timevar.cu with the functions of spooky_output_functions.cu) on
the GPU, against direct numpy computations of their definitions from the same
state (numpy FFTs, the reference's grid and normalization), and against closed
forms:
  hv, hm  : <A . V> with A = i k x V / k^2 (k = 0 -> 1), the Helicity kernel;
            an ABC flow (curl u = u at |k| = 1) gives <u . u> = A^2 + B^2 + C^2.
  pv      : <curl u . grad theta>
  dissgradT (isotropic) : sum_n energy(i k_n theta) = <|grad theta|^2> / 2 for
            dealiased fields without Nyquist content; a single mode gives k^2/4.
  dissgradT (anisotropic, with mhd) : -<theta div_m(B (B.grad theta) / B^2)>
  fluxbbgradT : <theta div_m(B B_s / B^2)>
  thetaB, bz2, phiB : <|b_z|>, <b_z^2>, <acos(|B_x| / |B_xy|)> of b = B / |B|;
            a uniform B along z gives 1, 1; along (1, 1, 0) gives 0, 0, pi/4.
div_m is the divergence with the 2/3 mask."""
import numpy as np
import pytest

from physics_cases import TWO_PI, relerr
from spkmod import spk

pytestmark = pytest.mark.gpu
ini = __import__("importlib").import_module("spooky-git_amd.initial")


@pytest.fixture(scope="module")
def ctx():
    c = spk.Ctx(0)
    yield c
    c.close()


N = 16
L = TWO_PI
P_MTI = dict(nx=N, ny=N, nz=N, incompressible=1, mhd=1, boussinesq=1, stratification=1, anisotropic_diffusion=1,
             N2=0.1, OmegaT2=1.0, reynolds_ani=50.0, nu=0.01, nu_m=0.01, nu_th=0.01, lx=L, ly=L, lz=L,
             t_final=1e9, toutput_flow=1e9)


def kvecs(n=N):
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    return np.meshgrid(kx, kx, kz, indexing="ij")


def state():
    u = ini.random_solenoidal((N, N, N), (L, L, L), seed=11)
    b = ini.random_solenoidal((N, N, N), (L, L, L), seed=12)
    b[0, 0, 0, 0] += 0.7 * N ** 3  # a mean field keeps |B| away from 0
    th = ini.random_solenoidal((N, N, N), (L, L, L), seed=13)[0]
    return np.concatenate([u, b, th[None]])


def irf(c):
    return np.fft.irfftn(c, s=(N, N, N), axes=(0, 1, 2))


def rf(r):
    return np.fft.rfftn(r, axes=(0, 1, 2))


def test_timevar_quantities_match_numpy(ctx):
    st = state()
    sol = spk.Solver(ctx, "f64", **P_MTI)
    sol.set_fields(st)
    sol.complex2real()
    KX, KY, KZ = kvecs()
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    ik2 = np.where(K2 > 0, 1.0 / np.where(K2 > 0, K2, 1), 1.0)
    m = ini.dealias_mask(N, N, N, L, L, L)
    r = [irf(c) for c in st]
    u, B, th = st[0:3], st[3:6], st[6]
    ur, Br, thr = r[0:3], r[3:6], r[6]

    def curl(V):
        return [1j * (KY * V[2] - KZ * V[1]), 1j * (KZ * V[0] - KX * V[2]), 1j * (KX * V[1] - KY * V[0])]

    def divm(F):
        return 1j * m * (KX * F[0] + KY * F[1] + KZ * F[2])

    grad = [1j * KX * th, 1j * KY * th, 1j * KZ * th]
    gr = [irf(g) for g in grad]
    Bm = np.sqrt(Br[0] ** 2 + Br[1] ** 2 + Br[2] ** 2)
    expect = {
        "hv": sum(np.mean(irf(c * ik2) * v) for c, v in zip(curl(u), ur)),
        "hm": sum(np.mean(irf(c * ik2) * v) for c, v in zip(curl(B), Br)),
        "pv": sum(np.mean(irf(c) * g) for c, g in zip(curl(u), gr)),
        "dissgradT": -np.mean(thr * irf(divm([rf(Br[q] * sum(Br[i] * gr[i] for i in range(3)) / Bm ** 2)
                                              for q in range(3)]))),
        "fluxbbgradT": np.mean(thr * irf(divm([rf(Br[q] * Br[2] / Bm ** 2) for q in range(3)]))),
        "thetaB": np.mean(np.abs(Br[2] / (1e-16 + Bm))),
        "bz2": np.mean((Br[2] / (1e-16 + Bm)) ** 2),
        "phiB": np.mean(np.arccos((np.abs(Br[0]) / (1e-16 + np.sqrt(Br[0] ** 2 + Br[1] ** 2))).astype(np.float32))),
        "hc": sum(np.mean(ur[q] * Br[q]) for q in range(3)),
        "thvz": np.mean(thr * ur[2]),
        "vxvy": np.mean(ur[0] * ur[1]),
    }
    for name, ref in expect.items():
        got = sol.timevar(name)
        tol = 2e-6 if name == "phiB" else 1e-10  # phiB: single-precision acos, as the reference
        assert abs(got - ref) <= tol * max(abs(ref), 1e-3), (name, got, ref)
    # the energies follow computeEnergy (kz = 0 plane weighted 1/2)
    w = np.ones(st.shape[1:])
    w[:, :, 0] = 0.5
    for name, f in (("Kx", 0), ("My", 4), ("et", 6)):
        assert sol.timevar(name) == pytest.approx(np.sum(w * np.abs(st[f]) ** 2) / N ** 6, rel=1e-12)
    assert sol.timevar("ev") == pytest.approx(sum(sol.timevar(k) for k in ("Kx", "Ky", "Kz")), rel=1e-13)
    with pytest.raises(RuntimeError):
        sol.timevar("divbbz")
    sol.destroy()


def test_closed_forms(ctx):
    X, Y, Z = ini.grid_coords(N, N, N, L, L, L)
    p = dict(P_MTI, anisotropic_diffusion=0, stratification=0)
    sol = spk.Solver(ctx, "f64", **p)
    # ABC flow with A = B = C = 1 (curl u = u): helicity <u . u> = 3
    u = [np.sin(Z) + np.cos(Y), np.sin(X) + np.cos(Z), np.sin(Y) + np.cos(X)]
    th = np.sin(3 * X)
    for b, (tb, b2, ph) in (((0, 0, 1.0), (1.0, 1.0, None)), ((1.0, 1.0, 0.0), (0.0, 0.0, np.pi / 4))):
        fields = u + [0 * X + b[0], 0 * X + b[1], 0 * X + b[2]] + [th]
        sol.set_fields(np.stack([rf(f) for f in fields]))
        sol.complex2real()
        assert sol.timevar("hv") == pytest.approx(3.0, rel=1e-12)
        assert sol.timevar("thetaB") == pytest.approx(tb, abs=1e-12)
        assert sol.timevar("bz2") == pytest.approx(b2, abs=1e-12)
        if ph is not None:
            assert sol.timevar("phiB") == pytest.approx(ph, rel=1e-6)
        # isotropic dissipation of sin(3x): k^2 / 4
        assert sol.timevar("dissgradT") == pytest.approx(9.0 / 4.0, rel=1e-12)
    sol.destroy()
