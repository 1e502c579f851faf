"""Seeded synthetic inputs for the LenseFlow workloads (host-side numpy).

Gaussian random fields follow the recipe of the reference's simulation code
(simulate.py:70-113): a diagonal covariance C = C_l(|l|) / Delta^2 on the rfft
half plane with C[0,0] = 0, Fourier coefficients sqrt(C/2) (n1 + i n2), and the
map irfft2(F, norm="ortho").  The |l| grid is gen_ell_grid (util.py:163-179).

The lensing potential uses C_l^phiphi ~ (l+20)^-4 exp(-(l/500)^2), rescaled
so the RMS deflection |grad phi| is 2.5 arcmin.  Without the exp(-(l/500)^2)
damping this spectrum, taken up to the pixel Nyquist, folds the map
(det(I + grad grad phi) <= 0 on a large fraction of pixels); with it the
minimum determinant stays around +0.2..+0.35.
"""
from __future__ import annotations

import math

import numpy as np

ARCMIN = math.pi / 180.0 / 60.0


def pix_width(theta_pix_arcmin: float) -> float:
    return theta_pix_arcmin * ARCMIN


def ell_grid(N: int, theta_pix_arcmin: float) -> np.ndarray:
    """|l| on the [N, N//2+1] half plane (util.py:163-179)."""
    d = pix_width(theta_pix_arcmin)
    lx = 2 * np.pi * np.fft.rfftfreq(N, d)
    ly = 2 * np.pi * np.fft.fftfreq(N, d)
    LX, LY = np.meshgrid(lx, ly)
    return np.sqrt(LX ** 2 + LY ** 2)


def grf_fourier(N, theta_pix_arcmin, cl_fn, rng) -> np.ndarray:
    """Fourier coefficients of a GRF (simulate.py:100-113), before the ortho irfft2."""
    ell = ell_grid(N, theta_pix_arcmin)
    with np.errstate(divide="ignore", invalid="ignore"):
        C = np.where(ell > 0, cl_fn(np.where(ell > 0, ell, 1.0)), 0.0)
    C = C / pix_width(theta_pix_arcmin) ** 2
    C[0, 0] = 0.0
    shape = (N, N // 2 + 1)
    return np.sqrt(C / 2) * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


def grf_map(N, theta_pix_arcmin, cl_fn, rng) -> np.ndarray:
    return np.fft.irfft2(grf_fourier(N, theta_pix_arcmin, cl_fn, rng), norm="ortho")


def cl_tt(ell):
    """C_l^TT = 1e3 (l+30)^-2 exp(-(l/2500)^2) uK^2."""
    return 1e3 * (ell + 30.0) ** -2 * np.exp(-(ell / 2500.0) ** 2)


def cl_ee(ell):
    """C_l^EE = 30 (l+50)^-2 exp(-(l/2500)^2) uK^2."""
    return 30.0 * (ell + 50.0) ** -2 * np.exp(-(ell / 2500.0) ** 2)


def cl_phiphi(ell, damping_ell=500.0):
    """C_l^phiphi ~ (l+20)^-4 exp(-(l/damping)^2) (unnormalised)."""
    return (ell + 20.0) ** -4 * np.exp(-(ell / damping_ell) ** 2)


def deflection_rms(phi_map: np.ndarray, theta_pix_arcmin: float) -> float:
    N = phi_map.shape[-1]
    d = pix_width(theta_pix_arcmin)
    kx = 2 * np.pi * np.fft.fftfreq(N, d)[:, None]
    ky = 2 * np.pi * np.fft.rfftfreq(N, d)[None, :]
    ky = ky.copy()
    ky[0, -1] = -ky[0, -1]
    F = np.fft.rfft2(phi_map)
    gx = np.fft.irfft2(1j * kx * F)
    gy = np.fft.irfft2(1j * ky * F)
    return float(np.sqrt(np.mean(gx ** 2 + gy ** 2)))


def phi_map(N, theta_pix_arcmin, seed, rms_arcmin=2.5, damping_ell=500.0) -> np.ndarray:
    """Lensing potential (MAP, float64) with RMS deflection `rms_arcmin`."""
    rng = np.random.default_rng(seed)
    phi = grf_map(N, theta_pix_arcmin, lambda l: cl_phiphi(l, damping_ell), rng)
    return phi * (rms_arcmin * ARCMIN / deflection_rms(phi, theta_pix_arcmin))


def min_det_magnification(phi: np.ndarray, theta_pix_arcmin: float) -> float:
    """min over pixels and t in [0,1] of det(I + t Hess phi) (closed form in t)."""
    N = phi.shape[-1]
    d = pix_width(theta_pix_arcmin)
    kx = 2 * np.pi * np.fft.fftfreq(N, d)[:, None]
    ky = 2 * np.pi * np.fft.rfftfreq(N, d)[None, :].copy()
    ky[0, -1] = -ky[0, -1]
    F = np.fft.rfft2(phi)
    hxx = np.fft.irfft2(-kx * kx * F)
    hxy = np.fft.irfft2(-kx * ky * F)
    hyy = np.fft.irfft2(-ky * ky * F)
    tr, dh = hxx + hyy, hxx * hyy - hxy ** 2
    m = np.minimum(1.0, 1.0 + tr + dh)
    with np.errstate(divide="ignore", invalid="ignore"):
        ts = np.where(dh > 0, -tr / (2 * dh), -1.0)
    inner = (ts > 0) & (ts < 1)
    m = np.where(inner, np.minimum(m, 1.0 - tr ** 2 / (4 * np.where(inner, dh, 1.0))), m)
    return float(m.min())


def t_map(N, theta_pix_arcmin, seed) -> np.ndarray:
    """Temperature map [1, N, N] float64."""
    rng = np.random.default_rng(seed)
    return grf_map(N, theta_pix_arcmin, cl_tt, rng)[None]


def eb_to_qu_fourier(E, B, N, theta_pix_arcmin):
    """(util.py:140-149) on numpy arrays."""
    d = pix_width(theta_pix_arcmin)
    lx = 2 * np.pi * np.fft.rfftfreq(N, d)
    ly = 2 * np.pi * np.fft.fftfreq(N, d)
    LX, LY = np.meshgrid(lx, ly)
    LX = LX.copy()
    LX[: N // 2 + 1, -1] *= -1
    a = np.arctan2(LX, LY)
    c2, s2 = np.cos(2 * a), np.sin(2 * a)
    return -E * c2 + B * s2, -E * s2 - B * c2


def qu_map(N, theta_pix_arcmin, seed) -> np.ndarray:
    """Q/U maps [2, N, N] float64 from an E-only GRF (B = 0)."""
    rng = np.random.default_rng(seed)
    E = grf_fourier(N, theta_pix_arcmin, cl_ee, rng)
    Q, U = eb_to_qu_fourier(E, np.zeros_like(E), N, theta_pix_arcmin)
    return np.fft.irfft2(np.stack([Q, U]), norm="ortho")


def lenseflow_inputs(config: int, seed: int = 0):
    """(f [C,N,N] float64, phi [N,N] float64, theta_pix, nsteps) of a workload.

    config 0: T, N=128, 3 arcmin; config 1: QU, N=256, 2 arcmin; both n = 7."""
    if config == 0:
        N, th = 128, 3.0
        return t_map(N, th, seed), phi_map(N, th, seed + 1), th, 7
    if config == 1:
        N, th = 256, 2.0
        return qu_map(N, th, seed), phi_map(N, th, seed + 1), th, 7
    raise ValueError(config)
