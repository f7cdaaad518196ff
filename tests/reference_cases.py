"""Initial conditions and closed-form solutions of the reference's end-to-end
tests (tests/*/check-*.py), restated in numpy on the reference's grid
x_i = -L/2 + i L/N.  Each returns real-space fields {name: [NX, NY, NZ]}."""
import numpy as np

TWO_PI = 2 * np.pi


def grid(n, L=(1.0, 1.0, 1.0)):
    ax = [-Li / 2 + Li * np.arange(ni) / ni for ni, Li in zip(n, L)]
    return np.meshgrid(*ax, indexing="ij")


def taylor_green_yz(t, nu=1.0 / 200.0, n=(64, 64, 64)):
    """check-taylor-green.py:14-33 and :118-121: v decays as exp(-2 nu t (2 pi)^2)."""
    X, Y, Z = grid(n)
    g = np.exp(-2.0 * nu * t * TWO_PI ** 2)
    return {"vx": 0 * X, "vy": np.sin(TWO_PI * Y) * np.cos(TWO_PI * Z) * g,
            "vz": -np.cos(TWO_PI * Y) * np.sin(TWO_PI * Z) * g}


def alfven_wave(t, n=(64, 64, 64)):
    """check-alfven-wave.py:14-38 and :113-123 (its closed form multiplies the
    real initial fields by exp(-i omega_A t) and keeps the real part)."""
    X, Y, Z = grid(n)
    a0 = np.real(1e-5 * np.exp(1j * TWO_PI * Z - 1j * np.pi / 2.0))
    w = TWO_PI
    c = np.cos(w * t)
    zero = 0 * X
    return {"vx": a0 * c, "vy": zero, "vz": zero, "bx": -a0 * c, "by": zero, "bz": zero + 1.0}


MTI_SIGMA = 0.9341413811120219  # check-mti-eigenmode.py:22


def mti_eigenmode(t, n=(512, 4, 4)):
    """check-mti-eigenmode.py:14-44 and :95-108: vz, bz, theta grow as exp(sigma t)."""
    X, Y, Z = grid(n)
    Pe, Reeta, B0, N2 = 100.0, 1e5, 1e-4, 0.1
    kp = TWO_PI * 12.0
    g = np.exp(MTI_SIGMA * t)
    vz = -1e-5 * np.sin(kp * X)
    bz = -1e-5 * np.cos(kp * X) * B0 * kp / (MTI_SIGMA + kp * kp / Reeta)
    th = 1.0 / (MTI_SIGMA + kp * kp / Pe) * (N2 - kp * kp / Pe / (MTI_SIGMA + kp * kp / Reeta)) * vz
    zero = 0 * X
    return {"vx": zero, "vy": zero, "vz": vz * g, "bx": zero + B0, "by": zero, "bz": bz * g, "th": th * g}


def l1(a, b, names):
    """Sum over fields of sum |a - b|, divided by the number of points (the
    reference's L1 error)."""
    n = a[names[0]].size
    return sum(float(np.sum(np.abs(a[k] - b[k]))) for k in names) / n


def advected_mhd_vortex(t, n=(512, 512, 2)):
    """check-advected-mhd-vortex.py:14-46: a Gaussian MHD vortex (sigma = 200)
    carried by the uniform flow (0.05, 0.05); ideal closed form (the run has
    nu = eta = 1e-4, hence the reference's tolerance of 1e-4)."""
    X, Y, Z = grid(n)
    u0 = v0 = 0.05
    sigma = 200.0
    Xt, Yt = X - u0 * t, Y - v0 * t
    g = np.exp((1.0 - sigma * (Xt ** 2 + Yt ** 2)) / 2.0) / TWO_PI
    zero = 0 * X
    return {"vx": v0 - g * Yt, "vy": u0 + g * Xt, "vz": zero, "bx": -g * Yt, "by": g * Xt, "bz": zero}


def sheared(w, t, shear, L):
    """w(x, y + shear x t) on the periodic grid, by a y-spectral phase
    exp(i ky x shear t) (the analytic shearing of check-shearing-wave.py:60-63)."""
    n = w.shape
    X = grid(n, L)[0]
    ky = TWO_PI * np.fft.fftfreq(n[1], d=L[1] / n[1])[None, :, None]
    return np.real(np.fft.ifft(np.fft.fft(w, axis=1) * np.exp(1j * ky * X * shear * t), axis=1))


def shearing_wave(t, n=(64, 64, 64)):
    """check-shearing-wave.py:14-45 and :104-113: vz = cos(ky0 y) sheared by the
    background flow, decaying viscously as exp(-nu (ky0^2 t + S^2 ky0^2 t^3 / 3))
    (kx0 = kz0 = 0), nu = 1/2000, S = 1.5, L = (1, 1.5, 1)."""
    L = (1.0, 1.5, 1.0)
    X, Y, Z = grid(n, L)
    shear, nu, ky0 = 1.5, 1.0 / 2000.0, TWO_PI / L[1]
    decay = np.exp(-nu * (ky0 ** 2 * t + shear ** 2 * ky0 ** 2 * t ** 3 / 3.0))
    zero = 0 * X
    return {"vx": zero, "vy": zero, "vz": sheared(np.cos(ky0 * Y), t, shear, L) * decay}


def mri_eigenmode(t, n=(64, 64, 64)):
    """check-mri-eigenmode.py:14-74 and :118-135: the viscous-resistive MRI
    channel mode (kx = 2 pi, kz = 4 pi, B0z = 0.1, Omega = 1, S = 1.5,
    nu = eta = 1/3000) growing as exp(sigma t); bz includes the mean B0z."""
    L = (1.0, 1.5, 1.0)
    X, Y, Z = grid(n, L)
    S, Om, B0z, re = 1.5, 1.0, 0.1, 3000.0
    kZ, kX = 2.0 * TWO_PI, TWO_PI
    gamma2 = kZ ** 2 / (kZ ** 2 + kX ** 2)
    kappa2 = 2 * Om * (2 * Om - S)
    wA2 = B0z ** 2 * kZ ** 2
    w2 = wA2 + 0.5 * kappa2 * gamma2 * (1.0 - np.sqrt(1.0 + 16.0 * wA2 * Om ** 2 / (kappa2 ** 2 * gamma2)))
    s_nu = np.sqrt(-w2)
    sigma = s_nu - (kX ** 2 + kZ ** 2) / re
    s_eta = sigma + (kX ** 2 + kZ ** 2) / re
    a = 1e-4
    ph = kX * X + kZ * Z
    vx = a * np.cos(ph)
    vy = vx * (S * wA2 / s_eta ** 2 - (2 * Om - S)) / (s_nu + wA2 / s_eta)
    vz = -(kX / kZ) * a * np.cos(ph)
    bx = -a * np.sin(ph) * B0z * kZ / s_eta
    by = bx * (-S / s_eta + (S * wA2 / s_eta ** 2 - (2 * Om - S)) / (s_nu + wA2 / s_eta))
    bzp = B0z * (kX / kZ) * a * np.sin(ph) * kZ / s_eta
    g = np.exp(sigma * t)
    out = {k: sheared(v, t, S, L) * g for k, v in (("vx", vx), ("vy", vy), ("vz", vz), ("bx", bx), ("by", by),
                                                      ("bz", bzp))}
    out["bz"] = out["bz"] + B0z
    return out
