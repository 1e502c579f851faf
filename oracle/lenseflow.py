"""FP64 numpy restatement of CMBLensing.py's LenseFlow operator (oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Conventions follow the reference exactly:
  * MAP arrays [..., N, N] float64; FOURIER arrays [..., N, N//2+1] complex128
    (numpy rfft2: R2C along axis -1, C2C along axis -2).
  * KX = 2 pi fftfreq(N, d) on axis -2, KY = 2 pi rfftfreq(N, d) on axis -1 with
    its last (Nyquist) entry negated                      (util.py:103-107).
  * forward FFT unnormalised, inverse 1/N^2, C2R ignores the imaginary part of
    self-conjugate bins (pocketfft == XLA/DUCC == FFTW convention).
"""
from __future__ import annotations

import math

import numpy as np

FORWARD_LENSE = +1   # constants.py:8
INVERSE_LENSE = -1   # constants.py:7
ARCMIN_PER_DEGREE = 60  # constants.py:6
DET_EPS = 1e-12      # lense_flow.py:408

_rfft2 = np.fft.rfft2
_irfft2 = np.fft.irfft2


def pix_width_from_theta(theta_pix_arcmin: float) -> float:
    """Pixel width in radians (constants.py:20: deg2rad(theta/60))."""
    return float(np.deg2rad(theta_pix_arcmin / ARCMIN_PER_DEGREE))


def fourier_weights(N: int) -> np.ndarray:
    """[1, 2, ..., 2, 1] along the rfft half axis (util.py:90-96)."""
    w = 2.0 * np.ones(N // 2 + 1)
    w[0] = 1.0
    w[-1] = 1.0
    return w


def primal_dot(a, b, N: int) -> float:
    """Re sum conj(a) b w / N^2 over the last two axes (util.py:100-101)."""
    w = fourier_weights(N)
    return np.real(np.sum(np.conj(a) * b * w * (1.0 / (N * N)), axis=(-2, -1)))


def field_dot(a, b, N: int) -> float:
    """Field dot: sum of primal_dot over components (fields.py:136-142).

    ``a`` and ``b`` are FOURIER arrays [..., C, N, N//2+1]; the component axis
    is -3 and is summed too."""
    return np.sum(primal_dot(a, b, N), axis=-1)


def k_meshgrid(N: int, pix_width: float):
    """(KX, KY) on the [N, N//2+1] half plane, KY Nyquist negated (util.py:103-107)."""
    kx = 2 * np.pi * np.fft.fftfreq(N, pix_width)
    ky = 2 * np.pi * np.fft.rfftfreq(N, pix_width)
    ky[-1] = -ky[-1]
    return np.meshgrid(kx, ky, indexing="ij")


def primal_derivatives(f, pix_width):
    """(Fx, Fy, Fxx, Fxy, Fyy) of a real map (util.py:110-120)."""
    N = f.shape[-1]
    KX, KY = k_meshgrid(N, pix_width)
    F = _rfft2(f)
    return (_irfft2(1j * KX * F), _irfft2(1j * KY * F), _irfft2(-(KX ** 2) * F),
            _irfft2(-(KX * KY) * F), _irfft2(-(KY ** 2) * F))


def primal_gradient(f, pix_width):
    """(Fx, Fy) only -- what XLA keeps of util.py:110-120 after dead-code elimination."""
    N = f.shape[-1]
    KX, KY = k_meshgrid(N, pix_width)
    F = _rfft2(f)
    return _irfft2(1j * KX * F), _irfft2(1j * KY * F)


def primal_derivatives_from_fourier(F, pix_width):
    """Derivatives from an rfft2 array (util.py:125-130)."""
    N = F.shape[-2]
    KX, KY = k_meshgrid(N, pix_width)
    return (_irfft2(1j * KX * F), _irfft2(1j * KY * F), _irfft2(-(KX ** 2) * F),
            _irfft2(-(KX * KY) * F), _irfft2(-(KY ** 2) * F))


def primal_derivatives_to_fourier(f, pix_width):
    """Derivative multipliers applied to rfft2(f), kept in Fourier (util.py:133-137)."""
    N = f.shape[-1]
    KX, KY = k_meshgrid(N, pix_width)
    F = _rfft2(f)
    return (1j * KX * F), (1j * KY * F), (-(KX ** 2) * F), (-(KX * KY) * F), (-(KY ** 2) * F)


def phi_derivatives(phi, pix_width):
    """MAP phi (square) or FOURIER phi (rectangular) -> 5 derivative maps
    (lense_flow.py:297-302)."""
    if phi.shape[-2] != phi.shape[-1]:
        return primal_derivatives_from_fourier(phi, pix_width)
    return primal_derivatives(phi, pix_width)


# --------------------------------------------------------------------------
# EB <-> QU (util.py:140-172)
# --------------------------------------------------------------------------
def _l_angle(nside: int, theta_pix: float):
    d = math.radians(theta_pix / ARCMIN_PER_DEGREE)                # util.py:165
    lx = 2 * np.pi * np.fft.rfftfreq(nside, d)                     # util.py:167
    ly = 2 * np.pi * np.fft.fftfreq(nside, d)                      # util.py:169
    lx, ly = np.meshgrid(lx, ly)                                   # util.py:171 ("xy")
    lx = lx.copy()
    lx[: nside // 2 + 1, -1] *= -1                                 # util.py:143
    return np.arctan2(lx, ly)                                      # util.py:145


def eb2qu(e, b, nside: int, theta_pix: float):
    """(util.py:140-149)"""
    a = _l_angle(nside, theta_pix)
    c2, s2 = np.cos(2 * a), np.sin(2 * a)
    return -e * c2 + b * s2, -e * s2 - b * c2


def qu2eb(q, u, nside: int, theta_pix: float):
    """(util.py:152-161)"""
    a = _l_angle(nside, theta_pix)
    c2, s2 = np.cos(2 * a), np.sin(2 * a)
    return -q * c2 - u * s2, q * s2 - u * c2


# --------------------------------------------------------------------------
# RK4 (util.py:184-199)
# --------------------------------------------------------------------------
def rk4_solve(F, y0, t0, t1, nsteps, args=()):
    """Classic fixed-step RK4 over tuples of arrays (util.py:184-199)."""
    single = not isinstance(y0, tuple)
    y = (y0,) if single else y0
    Fw = (lambda t, yy, a: (F(t, yy[0], a),)) if single else F
    h = (t1 - t0) / nsteps
    h_half = h / 2.0
    for i in range(nsteps):
        t = t0 + i * h
        k1 = Fw(t, y, args)
        k2 = Fw(t + h_half, tuple(yi + h_half * ki for yi, ki in zip(y, k1)), args)
        k3 = Fw(t + h_half, tuple(yi + h_half * ki for yi, ki in zip(y, k2)), args)
        k4 = Fw(t + h, tuple(yi + h * ki for yi, ki in zip(y, k3)), args)
        y = tuple(yi + h * (a + 2 * b + 2 * c + d) / 6
                  for yi, a, b, c, d in zip(y, k1, k2, k3, k4))
    return y[0] if single else y


# --------------------------------------------------------------------------
# LenseFlow rates (lense_flow.py:335-415)
# --------------------------------------------------------------------------
def m_matrix(t, phi_xx, phi_xy, phi_yy):
    """M = I + t Hess(phi) (lense_flow.py:397-405)."""
    return 1 + t * phi_xx, t * phi_xy, 1 + t * phi_yy


def inverse_matrix(m_xx, m_xy, m_yy, eps=DET_EPS):
    """Symmetric 2x2 inverse with det + eps (lense_flow.py:407-415)."""
    det = m_xx * m_yy - m_xy * m_xy
    inv_det = 1.0 / (det + eps)
    return m_yy * inv_det, -m_xy * inv_det, m_xx * inv_det


def p_vector(phi_x, phi_y, phi_xx, phi_xy, phi_yy, t):
    """p = M(t)^-1 grad phi (lense_flow.py:372-381)."""
    m_xx, m_xy, m_yy = m_matrix(t, phi_xx, phi_xy, phi_yy)
    ixx, ixy, iyy = inverse_matrix(m_xx, m_xy, m_yy)
    return phi_x * ixx + phi_y * ixy, phi_x * ixy + phi_y * iyy


def standard_term(f, derivs, t, pix_width):
    """grad phi . M(t)^-1 grad f (lense_flow.py:364-370, 383-395)."""
    phi_x, phi_y, phi_xx, phi_xy, phi_yy = derivs
    fx, fy = primal_gradient(f, pix_width)
    m_xx, m_xy, m_yy = m_matrix(t, phi_xx, phi_xy, phi_yy)
    ixx, ixy, iyy = inverse_matrix(m_xx, m_xy, m_yy)
    v_x = fx * ixx + fy * ixy
    v_y = fy * iyy + fx * ixy
    return phi_x * v_x + phi_y * v_y


def adjoint_term(f, derivs, t, pix_width):
    """d_x(p_x f) + d_y(p_y f) (lense_flow.py:354-362)."""
    p_x, p_y = p_vector(*derivs, t)
    dx, _ = primal_gradient(p_x * f, pix_width)
    _, dy = primal_gradient(p_y * f, pix_width)
    return dx + dy


def lense_times(direction: int, adjoint: bool):
    """(t0, t1) from the truth table of lense_flow.py:304-322."""
    if (direction == FORWARD_LENSE) != bool(adjoint):
        return 0.0, 1.0
    return 1.0, 0.0


def primal_lense_flow(f, phi, pix_width, n=10, direction=FORWARD_LENSE, adjoint=False,
                      derivs=None):
    """LenseFlow of MAP array(s) f [..., N, N] (lense_flow.py:295-333).

    ``phi`` is a MAP [N, N] or FOURIER [N, N//2+1] array; ``derivs`` may pass the
    five precomputed phi derivatives instead."""
    if derivs is None:
        derivs = phi_derivatives(phi, pix_width)
    t0, t1 = lense_times(direction, adjoint)
    rate = adjoint_term if adjoint else standard_term
    return rk4_solve(lambda t, y, a: rate(y, derivs, t, pix_width),
                     np.asarray(f, dtype=np.float64), t0, t1, n)


def lense_flow(f, phi, pix_width, n=10, direction=FORWARD_LENSE, adjoint=False):
    """Apply LenseFlow to every component of f [..., C, N, N] (lense_flow.py:289-292).

    The reference recomputes the phi derivatives per component; the value is the
    same, so they are computed once here."""
    return primal_lense_flow(f, phi, pix_width, n, direction, adjoint)


# --------------------------------------------------------------------------
# Gradient ODE (lense_flow.py:10-126, 156-252)
# --------------------------------------------------------------------------
def delta_phi_rate(fs, dfs, derivs, time, pix_width, phi_fourier):
    """d(delta phi)/dt (lense_flow.py:195-252). fs, dfs: lists of MAP arrays."""
    phi_x, phi_y, phi_xx, phi_xy, phi_yy = derivs
    fdf_x = 0.0
    fdf_y = 0.0
    for f, df in zip(fs, dfs):                                    # :201-213
        fx, fy = primal_gradient(f, pix_width)
        fdf_x = fdf_x + df * fx
        fdf_y = fdf_y + df * fy
    m_xx, m_xy, m_yy = m_matrix(time, phi_xx, phi_xy, phi_yy)     # :218-219
    ixx, ixy, iyy = inverse_matrix(m_xx, m_xy, m_yy)
    mx = ixx * fdf_x + ixy * fdf_y                                # :221-222
    my = ixy * fdf_x + iyy * fdf_y
    p_x, p_y = p_vector(*derivs, time)                            # :225
    w_xx, w_xy, w_yx, w_yy = p_x * mx, p_x * my, p_y * mx, p_y * my   # :226-229
    if phi_fourier:                                               # :233-239
        d = primal_derivatives_to_fourier
    else:                                                         # :240-248
        d = primal_derivatives
    div = d(mx, pix_width)[0] + d(my, pix_width)[1]
    lap = (d(time * w_xx, pix_width)[2] + d(time * w_xy, pix_width)[3]
           + d(time * w_yx, pix_width)[3] + d(time * w_yy, pix_width)[4])
    return lap + div                                              # :249-251


def lensing_operator_gradients(phi, fields, cotangents, direction, n, pix_width):
    """Joint RK4 of (f, delta f, delta phi) (lense_flow.py:53-126).

    ``fields``/``cotangents``: lists of MAP arrays (the components present; the
    reference zero-pads T/QU to TQU, whose zero components contribute exactly 0).
    Returns (fields', delta_fields, delta_phi); delta_phi is FOURIER-shaped when
    phi is FOURIER (lense_flow.py:84-92)."""
    if direction == INVERSE_LENSE:                                # :62-74
        t0, t1 = 1.0, 0.0
    else:
        t0, t1 = 0.0, 1.0
    phi_fourier = phi.shape[-2] != phi.shape[-1]
    derivs = phi_derivatives(phi, pix_width)
    nc = len(fields)
    if phi_fourier:
        dphi0 = np.zeros(phi.shape, dtype=np.complex128)
    else:
        dphi0 = np.zeros(phi.shape, dtype=np.float64)
    y0 = tuple(np.asarray(f, np.float64) for f in fields) + \
        tuple(np.asarray(c, np.float64) for c in cotangents) + (dphi0,)

    def step(time, y, _):                                         # :156-191
        fs, dfs = list(y[:nc]), list(y[nc:2 * nc])
        rf = tuple(standard_term(f, derivs, time, pix_width) for f in fs)
        rdf = tuple(adjoint_term(df, derivs, time, pix_width) for df in dfs)
        return rf + rdf + (delta_phi_rate(fs, dfs, derivs, time, pix_width, phi_fourier),)

    out = rk4_solve(step, y0, t0, t1, n)
    return list(out[:nc]), list(out[nc:2 * nc]), out[2 * nc]


def undo_inner_product(m, N):
    """fields.py:146-154: MAP arrays -> irfft2(conj(rfft2(m)/w) N^2); FOURIER as is."""
    if m.shape[-2] != m.shape[-1]:
        return m
    return _irfft2(np.conj(_rfft2(m) / fourier_weights(N)) * N ** 2)


def lense_flow_vjp(phi, f, cotangent, pix_width, n=10, direction=FORWARD_LENSE):
    """Backward rule of lense_flow_wrapper (lense_flow.py:23-46) for MAP fields
    f, cotangent [C, N, N]: integrate from the transformed field with -direction.
    Returns (delta_f [C,N,N], delta_phi) after undo_inner_product.  The reference
    also swaps Q<->U in delta_f for spin-2 fields (lense_flow.py:140-152); that
    output is never consumed and is not reproduced here."""
    N = f.shape[-1]
    Lf = lense_flow(f, phi, pix_width, n, direction, False)
    _, dfs, dphi = lensing_operator_gradients(phi, list(Lf), list(cotangent), -direction, n,
                                              pix_width)
    dfs = np.stack([undo_inner_product(d, N) for d in dfs])
    return dfs, undo_inner_product(dphi, N)
