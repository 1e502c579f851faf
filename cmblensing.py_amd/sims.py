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


# ---------------------------------------------------------------------------
# Wiener-filter / likelihood problems
# ---------------------------------------------------------------------------
def covariance_diag(N, theta_pix_arcmin, cl_fn) -> np.ndarray:
    """Fourier-diagonal covariance C_l(|l|) / Delta^2 on the half plane with the
    (0,0) mode zeroed (simulate.py:70-97 with rescale)."""
    ell = ell_grid(N, theta_pix_arcmin)
    with np.errstate(divide="ignore", invalid="ignore"):
        C = np.where(ell > 0, cl_fn(np.where(ell > 0, ell, 1.0)), 0.0)
    C = C / pix_width(theta_pix_arcmin) ** 2
    C[0, 0] = 0.0
    return C


def noise_cl(uk_arcmin, ell_knee=None, alpha_knee=3.0):
    """N_l = (uK-arcmin in rad)^2 (1 + (l_knee/l)^alpha) (simulate.py:54-65 form);
    the stated level is used directly for T and for each of Q, U."""
    w = (uk_arcmin * ARCMIN) ** 2

    def fn(ell):
        if ell_knee is None:
            return w * np.ones_like(ell)
        return w * (1.0 + (ell_knee / ell) ** alpha_knee)
    return fn


def beam_diag(N, theta_pix_arcmin, fwhm_arcmin) -> np.ndarray:
    """b_l = exp(-l^2 sigma^2 / 2), sigma = FWHM / sqrt(8 ln 2) (simulate.py:50-51,135)."""
    ell = ell_grid(N, theta_pix_arcmin)
    sig = fwhm_arcmin * ARCMIN / math.sqrt(8 * math.log(2))
    return np.exp(-0.5 * (ell * sig) ** 2)


def lowpass_mask_diag(N, theta_pix_arcmin, ell_cut=3000.0, delta_ell=50.0) -> np.ndarray:
    """Fourier-diagonal low-pass 'mask' like the reference's (simulate.py:117-130):
    1 below ell_cut - delta_ell, a cosine taper to 0 at ell_cut, 1 at l = 0."""
    ell = ell_grid(N, theta_pix_arcmin)
    x = np.clip((ell - (ell_cut - delta_ell)) / delta_ell, 0.0, 1.0)
    m = 0.5 * (1 + np.cos(np.pi * x))
    m[0, 0] = 1.0
    return m


def square_mask(N, frac_area=0.8, taper_px=8) -> np.ndarray:
    """Centred square covering `frac_area` of the map with a cosine taper."""
    side = math.sqrt(frac_area) * N
    c = (N - 1) / 2.0
    d = np.abs(np.arange(N) - c)
    inner = side / 2 - taper_px / 2

    def prof(v):
        t = np.clip((v - inner) / taper_px, 0.0, 1.0)
        return 0.5 * (1 + np.cos(np.pi * t))
    return np.outer(prof(d), prof(d))


HOLES_PER_SQ_ARCMIN = 200 / (512 * 1.5) ** 2   # 200 holes on a 512^2 x 1.5' map


def point_source_mask(N, theta_pix_arcmin, nholes=None, radius_arcmin=5.0, seed=4) -> np.ndarray:
    """1 everywhere except `nholes` discs of `radius_arcmin` at uniform random
    centres (periodic distances).  By default the hole count scales with the
    map's area at the density of 200 holes on a 512^2 map of 1.5 arcmin pixels
    (about 2.6 % of the area masked at any size), at least one hole."""
    if nholes is None:
        nholes = max(1, int(round(HOLES_PER_SQ_ARCMIN * (N * theta_pix_arcmin) ** 2)))
    rng = np.random.default_rng(seed)
    m = np.ones((N, N))
    r_px = radius_arcmin / theta_pix_arcmin
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    for cx, cy in rng.uniform(0, N, size=(nholes, 2)):
        dx = np.abs(ii - cx)
        dy = np.abs(jj - cy)
        dx = np.minimum(dx, N - dx)
        dy = np.minimum(dy, N - dy)
        m[dx * dx + dy * dy <= r_px * r_px] = 0.0
    return m


def wf_problem(N, theta_pix_arcmin, pol=True, seed=0, mask="fourier", noise_uk=5.0,
               ell_knee=None, beam_fwhm=1.0):
    """A lensed Wiener-filter problem as numpy arrays (EB / T FOURIER fields,
    per-mode operators [C, N, N//2+1]):

      f (truth), phi (MAP), Cf, Cn, B, and either M (Fourier-diagonal) with
      d = M B L f + n in Fourier, or mask_pix with d = m map(B L f) + n as maps.
    The lensing of the truth uses the FP64 numpy FFTs of this module's
    helpers and 7 RK4 steps of the same LenseFlow (imported lazily)."""
    rng = np.random.default_rng(seed)
    th = theta_pix_arcmin
    C = 2 if pol else 1
    cl = cl_ee if pol else cl_tt
    Cf1 = covariance_diag(N, th, cl)
    Cn1 = covariance_diag(N, th, noise_cl(noise_uk, ell_knee))
    # The monopole gets the variance of the fundamental mode: with C[0,0] = 0
    # the nonzero-safe pinv (util.py:45-47) leaves it without prior, noise
    # weight or preconditioner, and a pixel mask couples it to every mode, so
    # PCG could never reduce the residual there.
    for A in (Cf1, Cn1):
        A[0, 0] = A[0, 1]
    Cn = np.stack([Cn1] * C)
    Cf = np.stack([Cf1, 1e-3 * Cf1])[:C] if pol else Cf1[None]   # small B-mode power
    B = np.stack([beam_diag(N, th, beam_fwhm)] * C)
    shape = (C, N, N // 2 + 1)
    f = np.sqrt(Cf / 2) * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    f = np.fft.rfft2(np.fft.irfft2(f))           # Hermitian-consistent half plane
    phi = phi_map(N, th, seed + 1)
    return dict(N=N, theta_pix=th, C=C, f=f, phi=phi, Cf=Cf, Cn=Cn, B=B, mask_kind=mask,
                rng_seed=seed)


def wf_data(prob, lensed_qu_maps, seed=None):
    """Attach data to a problem given the lensed T or Q/U maps [C, N, N] of its
    truth: Fourier mask -> d = M B L f + n (EB Fourier); pixel mask ->
    d = m map(B L f) + n (maps)."""
    N, th, C = prob["N"], prob["theta_pix"], prob["C"]
    rng = np.random.default_rng(prob["rng_seed"] + 7 if seed is None else seed)
    shape = (C, N, N // 2 + 1)
    nF = np.sqrt(prob["Cn"] / 2) * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    nF = np.fft.rfft2(np.fft.irfft2(nF))        # the transform of a real noise map
    LF = np.fft.rfft2(lensed_qu_maps)
    if C == 2:
        e, b = _qu2eb_np(LF[0], LF[1], N, th)
        LF = np.stack([e, b])
    if prob["mask_kind"] == "fourier":
        M = np.stack([lowpass_mask_diag(N, th, ell_cut=0.8 * np.pi / pix_width(th))] * C)
        prob["M"] = M
        prob["mask_pix"] = None
        prob["d"] = M * prob["B"] * LF + nF
    else:
        m = square_mask(N) if prob["mask_kind"] == "square" else point_source_mask(N, th)
        prob["M"] = None
        prob["mask_pix"] = m
        BLF = prob["B"] * LF
        if C == 2:
            q, u = eb_to_qu_fourier(BLF[0], BLF[1], N, th)
            BLF = np.stack([q, u])
            ne, nb = eb_to_qu_fourier(nF[0], nF[1], N, th)
            nF = np.stack([ne, nb])
        prob["d"] = m * np.fft.irfft2(BLF) + np.fft.irfft2(nF)
    return prob


def _qu2eb_np(q, u, N, th):
    d = pix_width(th)
    lx = 2 * np.pi * np.fft.rfftfreq(N, d)
    ly = 2 * np.pi * np.fft.fftfreq(N, d)
    LX, LY = np.meshgrid(lx, ly)
    LX = LX.copy()
    LX[: N // 2 + 1, -1] *= -1
    a = np.arctan2(LX, LY)
    c2, s2 = np.cos(2 * a), np.sin(2 * a)
    return -q * c2 - u * s2, q * s2 - u * c2


def phi_map_and_cov(N, theta_pix_arcmin, seed, rms_arcmin=2.5, damping_ell=500.0):
    """phi (MAP) and the Fourier-diagonal covariance of its rfft2 coefficients
    (same normalisation as the reference's fixtures: <|rfft2(phi)_l|^2> per
    mode), with the monopole given the fundamental mode's variance."""
    rng = np.random.default_rng(seed)
    cl = lambda l: cl_phiphi(l, damping_ell)
    F = grf_fourier(N, theta_pix_arcmin, cl, rng)
    phi = np.fft.irfft2(F, norm="ortho")
    s = rms_arcmin * ARCMIN / deflection_rms(phi, theta_pix_arcmin)
    C = covariance_diag(N, theta_pix_arcmin, cl) * (s * N) ** 2
    C[0, 0] = C[0, 1]
    # floor the damped tail: below ~1e-10 of the peak the variance is smaller
    # than the FFT rounding of phi itself, and dot(phi, C^-1 phi) would be
    # dominated by that rounding
    C = np.maximum(C, 1e-10 * C.max())
    return phi * s, C


# ---------------------------------------------------------------------------
# CAMB-layout spectra for the simulation code (no CAMB here)
# ---------------------------------------------------------------------------
def camb_like_spectra(lmax_prime=5000, phi_rms_arcmin=2.5, phi_damping=1000.0, r=0.2):
    """Analytic stand-ins for CAMB's output in its layout (rows l = 0..lmax_prime-1,
    D_l = l(l+1) C_l / 2pi in muK^2; columns TT, EE, BB, TE; lens potential
    [l(l+1)]^2 C_l^phiphi / 2pi): (unlensed_scalar, tensor, total, lens_potential).

    Scalar TT/EE follow cl_tt / cl_ee, TE = 0.5 sqrt(TT EE) cos(l / 300)
    (changes sign: the reference then interpolates linearly), scalar BB = 0;
    tensors carry r times a low-l bump with a power-law tail; 'total' adds a lensing-like B-mode and
    10 % extra TT/EE power.  C_l^phiphi ~ (l+20)^-4 exp(-(l/phi_damping)^2),
    scaled so that sum_l l^3 C_l / 2pi = (phi_rms_arcmin)^2."""
    ell = np.arange(lmax_prime, dtype=np.float64)
    l2 = np.where(ell >= 2, ell, 2.0)
    to_d = l2 * (l2 + 1) / (2 * np.pi)
    tt, ee = cl_tt(l2), cl_ee(l2)
    te = 0.5 * np.sqrt(tt * ee) * np.cos(l2 / 300.0)
    zero = np.zeros_like(l2)
    scal = np.stack([tt, ee, zero, te], axis=1) * to_d[:, None]
    # a power-law tail (as CAMB's tensor spectra have): a Gaussian cut-off would
    # leave C_l^BB ~ 1e-77 and pinv(Cf) overflowing in the B-mode prior
    bump = r * 1e-2 * (l2 + 10.0) ** -2 / (1.0 + (l2 / 150.0) ** 4)
    ten = np.stack([bump, 0.1 * bump, 0.1 * bump, 0.05 * bump], axis=1) * to_d[:, None]
    lensb = 2e-6 * np.ones_like(l2) * np.exp(-(l2 / 3000.0) ** 2)
    tot = np.stack([1.1 * tt, 1.1 * ee, lensb, te], axis=1) * to_d[:, None] + ten
    cpp = (l2 + 20.0) ** -4 * np.exp(-(l2 / phi_damping) ** 2)
    msk = ell >= 2
    amp = (phi_rms_arcmin * ARCMIN) ** 2 / np.sum((l2 ** 3 * cpp)[msk] / (2 * np.pi))
    lens = amp * cpp * (l2 * (l2 + 1)) ** 2 / (2 * np.pi)
    for a in (scal, ten, tot):
        a[~msk] = 0.0
    lens[~msk] = 0.0
    return scal, ten, tot, lens
