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


# ---- spectrum file (spk_spectrum_named: WriteSpectrumOutput, spooky_output_spectrum.cu) ----

def shell(a, b, nb):
    """Spectrum1d (cuda_kernels_generic.cu:441-471) in numpy for the 2 pi box:
    Re(a b*) / N^2 per half-space mode (k_z > 0 twice), summed over shells
    round(|k| / deltak), deltak = 1, bins below nbins kept."""
    KX, KY, KZ = kvecs()
    pw = np.real(a * np.conj(b)) / float(N ** 3) ** 2
    pw[:, :, 1:] *= 2.0
    bins = (np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2) + 0.5).astype(int)
    keep = bins < nb
    return np.bincount(bins[keep], weights=pw[keep], minlength=nb)[:nb]


def transfer_terms(st):
    """The nonlinear terms WriteSpectrumOutput pairs with a state field, from
    their definitions (products in real space, derivatives spectral, 2/3 mask m)."""
    KX, KY, KZ = kvecs()
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    ik2 = np.where(K2 > 0, 1.0 / np.where(K2 > 0, K2, 1), 1.0)
    m = ini.dealias_mask(N, N, N, L, L, L)
    k = (KX, KY, KZ)
    r = [irf(c) for c in st]
    ur, Br, thr = r[0:3], r[3:6], r[6]

    def advect(a):  # -i m d_j (a_i a_j)
        return [-1j * m * sum(k[j] * rf(a[i] * a[j]) for j in range(3)) for i in range(3)]

    def curlm(F):
        return [1j * m * (KY * F[2] - KZ * F[1]), 1j * m * (KZ * F[0] - KX * F[2]), 1j * m * (KX * F[1] - KY * F[0])]

    def divm(F):
        return 1j * m * (KX * F[0] + KY * F[1] + KZ * F[2])

    emf = [rf(ur[1] * Br[2] - ur[2] * Br[1]), rf(ur[2] * Br[0] - ur[0] * Br[2]), rf(ur[0] * Br[1] - ur[1] * Br[0])]
    B = st[3:6]
    hel = [1j * (KY * B[2] - KZ * B[1]) * ik2, 1j * (KZ * B[0] - KX * B[2]) * ik2, 1j * (KX * B[1] - KY * B[0]) * ik2]
    Bm2 = Br[0] ** 2 + Br[1] ** 2 + Br[2] ** 2
    gr = [irf(1j * kk * st[6]) for kk in k]
    bg = sum(Br[i] * gr[i] for i in range(3))
    return dict(advec=advect(ur), lorentz=advect(Br), emfpower=curlm(emf),
                th_advec=-1j * m * sum(k[j] * rf(ur[j] * thr) for j in range(3)), helicity=hel,
                aniso_diss=divm([rf(Br[q] * bg / Bm2) for q in range(3)]), bybz=rf(Br[1] * Br[2]),
                aniso_inj=divm([rf(Br[q] * Br[2] / Bm2) for q in range(3)]))


def test_transfer_terms_conserve(ctx):
    """Pins the numpy definitions above: with solenoidal u, B the dealiased
    advection conserves kinetic energy and theta variance, and the lorentz term
    as written there, <u_i (-d_j (B_i B_j))>, equals the induction transfer
    <B_i curl(u x B)_i> component by component (their difference is
    <u_j d_j B_i^2 / 2> = 0): magnetic energy gained is kinetic energy lost."""
    st = state()
    T = transfer_terms(st)
    w = np.full(st.shape[1:], 2.0)
    w[:, :, 0] = 1.0

    def tot(a, b):
        return float(np.sum(w * np.real(a * np.conj(b))))

    scale = np.sqrt(tot(st[0], st[0]) * tot(T["advec"][0], T["advec"][0]))
    assert abs(sum(tot(st[i], T["advec"][i]) for i in range(3))) < 1e-12 * scale
    for i in range(3):
        lz = tot(st[i], T["lorentz"][i])
        assert abs(lz - tot(st[3 + i], T["emfpower"][i])) < 1e-12 * abs(lz)
    assert abs(tot(st[6], T["th_advec"])) < 1e-12 * np.sqrt(tot(st[6], st[6]) * tot(T["th_advec"], T["th_advec"]))


def test_spectra_match_numpy(ctx):
    st = state()
    sol = spk.Solver(ctx, "f64", **P_MTI)
    sol.set_fields(st)
    sol.complex2real()
    nb, dk = sol.spectrum_bins()
    assert dk == pytest.approx(1.0, rel=1e-15) and nb == int(np.sqrt(3) * (2.0 / 3.0) * (N // 2 - 1) + 0.5)
    T = transfer_terms(st)
    ex = {"Kx": (st[0], st[0]), "Ky": (st[1], st[1]), "Mz": (st[5], st[5]), "Eth": (st[6], st[6]),
          "vxvy": (st[0], st[1]), "vyvz": (st[1], st[2]), "bybz": (st[4], st[5]),
          "bxbz": (st[3], st[4]),  # B_x with B_y, as the reference pairs them
          "thvy": (st[6], st[1]), "th_advec": (st[6], T["th_advec"]),
          "th_aniso_diss": (st[6], T["bybz"] + T["aniso_diss"]), "th_aniso_inj": (st[6], T["aniso_inj"])}
    for i, c in enumerate("xyz"):
        ex["advec_" + c] = (st[i], T["advec"][i])
        ex["lorentz_" + c] = (st[i], T["lorentz"][i])
        ex["emfpower_" + c] = (st[3 + i], T["emfpower"][i])
        ex["helicity_" + c] = (st[3 + i], T["helicity"][i])
    got = {}
    for name, (a, b) in ex.items():
        ref = shell(a, b, nb)
        got[name] = sol.spectrum_named(name)
        assert got[name].shape == (nb,)
        assert np.abs(got[name] - ref).max() <= 1e-11 * np.abs(ref).max(), name
    # the bin sums are the timevar quantities of the same terms (every retained
    # mode falls in a bin here): helicity -> hm, injection -> fluxbbgradT, and the
    # anisotropic dissipation without the B_y B_z term -> -dissgradT; the lorentz
    # and emfpower sums agree component by component
    for c in "xyz":
        assert got["lorentz_" + c].sum() == pytest.approx(got["emfpower_" + c].sum(), rel=1e-10)
    assert abs(sum(got["advec_" + c].sum() for c in "xyz")) < 1e-12 * np.abs(got["advec_x"]).max()
    assert sum(got["helicity_" + c].sum() for c in "xyz") == pytest.approx(sol.timevar("hm"), rel=1e-10)
    assert got["th_aniso_inj"].sum() == pytest.approx(sol.timevar("fluxbbgradT"), rel=1e-10)
    diss = got["th_aniso_diss"].sum() - shell(st[6], T["bybz"], nb).sum()
    assert diss == pytest.approx(-sol.timevar("dissgradT"), rel=1e-9)
    for bad in ("Ex", "advec_w", "uservar1"):
        with pytest.raises(RuntimeError):
            sol.spectrum_named(bad)
    sol.destroy()
    # without the modules a name is not computed (the reference writes no line)
    sol = spk.Solver(ctx, "f64", nx=N, ny=N, nz=N, incompressible=1, lx=L, ly=L, lz=L, t_final=1e9,
                     toutput_flow=1e9)
    sol.set_fields(st[0:3])
    sol.complex2real()
    assert np.abs(sol.spectrum_named("advec_y") - shell(st[1], T["advec"][1], nb)).max() <= 1e-11 * np.abs(
        shell(st[1], T["advec"][1], nb)).max()
    for bad in ("Mx", "emfpower_x", "th_advec", "Eth", "helicity_z"):
        with pytest.raises(RuntimeError):
            sol.spectrum_named(bad)
    sol.destroy()


def test_user_kpartheta(ctx):
    """computekpartheta (user_output_functions.cu:31-58): <(b . grad theta)^2> /
    (2 E_theta) with b = B / (1e-16 + |B|); for theta = sin(3x) and B along x it
    is k_par^2 = 9, for B along y it is 0."""
    st = state()
    sol = spk.Solver(ctx, "f64", **P_MTI)
    sol.set_fields(st)
    sol.complex2real()
    KX, KY, KZ = kvecs()
    r = [irf(c) for c in st]
    Br, thr = r[3:6], r[6]
    gr = [irf(1j * kk * st[6]) for kk in (KX, KY, KZ)]
    bg = sum(Br[i] * gr[i] for i in range(3)) / (1e-16 + np.sqrt(sum(b ** 2 for b in Br)))
    ref = np.mean(bg ** 2) / np.mean(thr ** 2)
    assert sol.user_timevar("kpartheta") == pytest.approx(ref, rel=1e-11)
    assert sol.user_timevar("uservar1") == 0.0 and sol.user_timevar("uservar2") == 0.0
    X, Y, Z = ini.grid_coords(N, N, N, L, L, L)
    for b, kp in (((1.0, 0.0, 0.0), 9.0), ((0.0, 1.0, 0.0), 0.0)):
        fields = [0 * X] * 3 + [0 * X + b[0], 0 * X + b[1], 0 * X + b[2], np.sin(3 * X)]
        sol.set_fields(np.stack([rf(f) for f in fields]))
        sol.complex2real()
        assert sol.user_timevar("kpartheta") == pytest.approx(kp, abs=1e-12)
    with pytest.raises(RuntimeError):
        sol.user_timevar("divbbz")
    sol.destroy()


def test_snapshot_real_unshears(ctx):
    """UnshearOutput (unshear_output.cu:21-54): in a shearing box the snapshot
    holds every field shifted by -c/2 along y, c = Lx S fmod(t, 2 Ly / (S Lx));
    with Lx = 1, Ly = 1.5, S = 1.5, NY = 16 and t = 2.25 the shift is two cells."""
    from physics_cases import CASES
    p = dict(CASES["shear_mhd"]["params"])
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    sol = spk.Solver(ctx, "f64", **p)
    Ls = (p["lx"], p["ly"], p["lz"])
    st = np.concatenate([ini.random_solenoidal((nx, ny, nz), Ls, seed=5, k0=2.0),
                         ini.random_solenoidal((nx, ny, nz), Ls, seed=6, k0=2.0)])
    sol.set_fields(st)
    real = np.stack([np.fft.irfftn(c, s=(nx, ny, nz), axes=(0, 1, 2)) for c in st])
    sol.set_time(2.25)
    sol.snapshot_real()
    assert np.max(np.abs(sol.get_realbuf() - np.roll(real, 2, axis=2))) < 1e-13 * np.max(np.abs(real))
    assert np.max(np.abs(sol.get_fields() - st)) == 0.0  # the state is untouched
    sol.set_time(2.0)  # fmod -> 0: no shift
    sol.snapshot_real()
    assert np.max(np.abs(sol.get_realbuf() - real)) < 1e-13 * np.max(np.abs(real))
    sol.destroy()
    # without shearing the snapshot is the plain inverse transform
    q = dict(p, shearing=0, rotating=0)
    sol = spk.Solver(ctx, "f64", **q)
    sol.set_fields(st)
    sol.set_time(2.25)
    sol.snapshot_real()
    assert np.max(np.abs(sol.get_realbuf() - real)) < 1e-13 * np.max(np.abs(real))
    sol.destroy()
