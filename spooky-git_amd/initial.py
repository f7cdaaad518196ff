"""Initial conditions in the solver's spectral layout (numpy; host side).

The spectral convention is the reference's: a field's spectrum is the
unnormalized 3D R2C transform of its real values (numpy rfftn), shape
[NX, NY, NZ/2+1], with the real grid x_i = -L/2 + i L/N (initflow.cu:32-34).
"""
import numpy as np


def grid_coords(nx, ny, nz, lx, ly, lz):
    x = -lx / 2 + lx * np.arange(nx) / nx
    y = -ly / 2 + ly * np.arange(ny) / ny
    z = -lz / 2 + lz * np.arange(nz) / nz
    return np.meshgrid(x, y, z, indexing="ij")


def wavenumbers(nx, ny, nz, lx, ly, lz):
    kx = 2 * np.pi / lx * np.fft.fftfreq(nx, 1.0 / nx)
    ky = 2 * np.pi / ly * np.fft.fftfreq(ny, 1.0 / ny)
    kz = 2 * np.pi / lz * np.arange(nz // 2 + 1)
    return np.meshgrid(kx, ky, kz, indexing="ij")


def dealias_mask(nx, ny, nz, lx, ly, lz):
    """The 2/3 rule of wavevector.cu:79-93: zero where |k_i| > 2/3 * 2pi/L_i (N_i/2 - 1)."""
    KX, KY, KZ = wavenumbers(nx, ny, nz, lx, ly, lz)
    m = np.ones(KX.shape)
    for K, n, L in ((KX, nx, lx), (KY, ny, ly), (KZ, nz, lz)):
        m[np.abs(K) > 2.0 / 3.0 * (2 * np.pi / L * (n // 2 - 1))] = 0.0
    return m


def taylor_green(n, L=2 * np.pi, with_b=False):
    """u = (sin x cos y cos z, -cos x sin y cos z, 0) on an n^3, L-periodic box;
    returns the spectral state [3 or 6, n, n, n//2+1]."""
    X, Y, Z = grid_coords(n, n, n, L, L, L)
    k = 2 * np.pi / L
    u = [np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z), -np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z), 0 * X]
    fields = u + ([0 * X, 0 * X, 0 * X] if with_b else [])
    return np.stack([np.fft.rfftn(f) for f in fields])


def random_solenoidal(shape, L=(2 * np.pi,) * 3, seed=0, k0=4.0, rms=1.0, ncomp=3):
    """Random-phase, solenoidal, dealiased vector field with shell spectrum
    E(k) ~ k^4 exp(-2 (k/k0)^2), normalized to sqrt(<|u|^2>) = rms.
    Returns the spectral components [ncomp=3, NX, NY, NZ/2+1] (complex128)."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    KX, KY, KZ = wavenumbers(nx, ny, nz, *L)
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    Kabs = np.sqrt(K2)
    c = [np.fft.rfftn(rng.standard_normal(shape)) for _ in range(ncomp)]
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where(K2 > 0, np.sqrt(Kabs ** 4 * np.exp(-2 * (Kabs / k0) ** 2) / (4 * np.pi * np.maximum(K2, 1e-300))), 0.0)
        ik2 = np.where(K2 > 0, 1.0 / np.maximum(K2, 1e-300), 0.0)
    c = [a * amp for a in c]
    # solenoidal projection
    div = KX * c[0] + KY * c[1] + KZ * c[2]
    c = [c[0] - KX * div * ik2, c[1] - KY * div * ik2, c[2] - KZ * div * ik2]
    # zero the Nyquist planes and apply the 2/3 mask
    m = dealias_mask(nx, ny, nz, *L)
    m[nx // 2, :, :] = 0.0
    m[:, ny // 2, :] = 0.0
    m[:, :, nz // 2] = 0.0
    c = [a * m for a in c]
    # Hermitian-consistent: go through real space once
    r = [np.fft.irfftn(a, s=shape, axes=(0, 1, 2)) for a in c]
    cur = np.sqrt(np.mean(sum(x ** 2 for x in r)))
    r = [x * (rms / cur) for x in r]
    return np.stack([np.fft.rfftn(x) for x in r])


def mhd_random(shape, L=(2 * np.pi,) * 3, seed=0, k0=4.0):
    """u and B independent random-phase solenoidal fields, rms 1 each (state [6, ...])."""
    u = random_solenoidal(shape, L, seed=seed, k0=k0)
    b = random_solenoidal(shape, L, seed=seed + 1000003, k0=k0)
    return np.concatenate([u, b])


# 3D Boussinesq MTI with anisotropic (Braginskii) conduction: gravity along -z,
# background dT/dz = -1 (N2, OmegaT2 as in the reference's mti-eigenmode test,
# tests/mti-eigenmode/spooky.cfg), weak uniform B along x.
def mti_params(n=256, chi=1e-2, nu=1e-4, eta=1e-4, nu_th=1e-6, N2=0.1, OmegaT2=1.0, L=1.0, cfl=1.5,
               safety_sts=100):
    """Solver parameters of the MTI run: theta (field 6) is the sts variable,
    advanced by RKL2; everything else by RK3."""
    return dict(nx=n, ny=n, nz=n, incompressible=1, boussinesq=1, stratification=1, strat_direction=2, mhd=1,
                anisotropic_diffusion=1, supertimestepping=1, sts_mask=1 << 6, safety_sts=safety_sts,
                nu=nu, nu_m=eta, nu_th=nu_th, reynolds_ani=1.0 / chi, N2=N2, OmegaT2=OmegaT2,
                cfl=cfl, cfl_par=1.0, cfl_sts=1.0, safety_source=0.2, lx=L, ly=L, lz=L,
                t_final=1e9, toutput_flow=1e9)


def mti_white_noise(shape, L=(1.0, 1.0, 1.0), seed=0, amp=1e-4, b0=1e-4):
    """State [7, ...]: u = 0, B = (b0, 0, 0), theta = seeded real-space white
    noise uniform in [-amp, amp], Nyquist planes zeroed and 2/3 mask applied."""
    rng = np.random.default_rng(seed)
    th = np.fft.rfftn(rng.uniform(-amp, amp, size=shape))
    m = dealias_mask(*shape, *L)
    m[shape[0] // 2, :, :] = 0.0
    m[:, shape[1] // 2, :] = 0.0
    m[:, :, shape[2] // 2] = 0.0
    th = np.fft.rfftn(np.fft.irfftn(th * m, s=shape, axes=(0, 1, 2)))
    out = np.zeros((7,) + th.shape, dtype=np.complex128)
    out[3, 0, 0, 0] = b0 * np.prod(shape)
    out[6] = th
    return out


def mti_growth_rate(kx, chi, nu, eta, b0, N2):
    """Growth rate of the MTI eigenmode with wavevector along B (x), OmegaT2 = 1:
    largest real root of the cubic of tests/mti-eigenmode/mti_disp_rel.py:44-70,
    written in physical units (k2 = chi kx^2, q = eta/chi, Pr = nu/chi,
    Lambda = b0^2/chi)."""
    k2 = chi * kx * kx
    q, pr, lam = eta / chi, nu / chi, b0 * b0 / chi
    coeff = [1.0, k2 * (1 + q + pr), N2 + k2 * lam + (q + pr) * k2 ** 2 + q * pr * k2 ** 2,
             q * pr * k2 ** 3 + N2 * q * k2 + lam * k2 ** 2 - k2]
    r = np.roots(coeff)
    real = np.real(r[np.abs(np.imag(r)) < 1e-9 * np.maximum(1.0, np.abs(r))])
    return float(np.max(real)) if real.size else float("nan")


def hbi_noise(shape, L=(1.0, 1.0, 1.0), seed=0, amp=0.01, b0=1e-5):
    """State [7, ...] of the 2D heat-flux-buoyancy problem (problems/hbi2d,
    its initial-condition script): u_x, u_y with uniform random amplitudes
    amp * U(-1/2, 1/2) * sqrt(27/8 N) and uniform random phases at every
    scale, u_z = 0, B = (0, b0, 0), theta = 0."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    csh = (nx, ny, nz // 2 + 1)
    fact = np.sqrt(27.0 / 8.0 * nx * ny * nz)

    def noise():
        c = amp * rng.uniform(-0.5, 0.5, csh) * np.exp(2j * np.pi * rng.uniform(-0.5, 0.5, csh)) * fact
        return np.fft.rfftn(np.fft.irfftn(c, s=shape, axes=(0, 1, 2)))  # Hermitian-consistent

    out = np.zeros((7,) + csh, dtype=np.complex128)
    out[0], out[1] = noise(), noise()
    out[4, 0, 0, 0] = b0 * nx * ny * nz
    return out


def hbi_params(n=512, nu=1.0 / 12800.0, eta=1.0 / 12800.0, chi=1.0 / 1600.0, N2=0.2, OmegaT2=-1.0):
    """Solver parameters of problems/hbi2d/spooky.cfg (gravity and stratification
    along y, B along y, RKL2 on theta), on an n x n x 2 box of unit size."""
    return dict(nx=n, ny=n, nz=2, incompressible=1, boussinesq=1, stratification=1, strat_direction=1, mhd=1,
                anisotropic_diffusion=1, supertimestepping=1, sts_mask=1 << 6, safety_sts=100, nu=nu, nu_m=eta,
                nu_th=1e-6, reynolds_ani=1.0 / chi, N2=N2, OmegaT2=OmegaT2, cfl=1.5, cfl_par=1.0, cfl_sts=1.0,
                safety_source=0.2, lx=1.0, ly=1.0, lz=1.0, t_final=1e9, toutput_flow=1e9)
