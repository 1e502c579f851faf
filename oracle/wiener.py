"""FP64 numpy restatement of CMBLensing.py's gradf_logpdf, logpdf and the
preconditioned-CG Wiener filter (oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/ and bench.py's cpu_baseline leg.

Fields are FOURIER arrays [C, N, N//2+1] complex128 with C=1 (T, FlatS0),
C=2 (E, B of a FlatS2 in the EB parametrisation) or C=3 (T, E, B of a
FlatS02).  Diagonal operators (DiagonalScalar / DiagonalEB,
matrix_operators.py:45-90) are real arrays of the same shape, applied per
mode; BlockTEB operators (matrix_operators.py:92-148) are ``TEB`` objects.
``theta_pix`` is in arcmin.
"""
from __future__ import annotations

import numpy as np

from . import lenseflow as lf


def _recip(m):
    """util.py:45-47: 1/m where m != 0, else 0."""
    m = np.asarray(m)
    out = np.zeros_like(m)
    nz = m != 0
    out[nz] = 1.0 / m[nz]
    return out


def _log_det1(m, w):
    """util.py:39-42: sign(prod) * sum log|m| w over the nonzero modes."""
    val = np.sum(np.where(m != 0, np.log(np.abs(np.where(m != 0, m, 1))), 0) * w)
    sign = np.prod(np.sign(np.where(m != 0, m, 1)))
    return sign * val


class TEB:
    """BlockTEB operator (matrix_operators.py:92-148): per mode
    [[TT, TE, 0], [ET, EE, 0], [0, 0, BB]] acting on (T, E, B)."""
    __array_ufunc__ = None          # numpy defers array * TEB to __rmul__

    def __init__(self, TT, TE, ET, EE, BB):
        self.TT, self.TE, self.ET, self.EE, self.BB = (np.asarray(a) for a in (TT, TE, ET, EE, BB))

    def __mul__(self, o):
        if isinstance(o, TEB):                                      # :115-121
            return TEB(self.TT * o.TT + self.TE * o.ET, self.TT * o.TE + self.TE * o.EE,
                       self.ET * o.TT + self.EE * o.ET, self.ET * o.TE + self.EE * o.EE,
                       self.BB * o.BB)
        if isinstance(o, np.ndarray) and o.ndim >= 3 and o.shape[-3] == 3:   # :123-127
            t, e, b = o[..., 0, :, :], o[..., 1, :, :], o[..., 2, :, :]
            return np.stack([self.TT * t + self.TE * e, self.ET * t + self.EE * e, self.BB * b],
                            axis=-3)
        return TEB(*(m * o for m in (self.TT, self.TE, self.ET, self.EE, self.BB)))  # :129-133

    def __rmul__(self, o):
        return TEB(*(o * m for m in (self.TT, self.TE, self.ET, self.EE, self.BB)))

    def __add__(self, o):
        return TEB(self.TT + o.TT, self.TE + o.TE, self.ET + o.ET, self.EE + o.EE,
                   self.BB + o.BB)

    def pinv(self):
        """Schur-complement inverse of the TE block (util.py:59-69), reciprocal BB."""
        a, b, c, d = self.TT, self.TE, self.ET, self.EE
        rd = _recip(d)
        s = _recip(a - b * rd * c)
        return TEB(s, -s * b * rd, -rd * c * s, rd + rd * c * s * b * rd, _recip(self.BB))

    def log_det(self, N):
        """util.py:51-53 for the TE block plus the BB block (matrix_operators.py:193-200)."""
        w = lf.fourier_weights(N)
        return (_log_det1(self.TT, w) + _log_det1(self.EE - self.ET * _recip(self.TT) * self.TE, w)
                + _log_det1(self.BB, w))


def pinv(op):
    """Nonzero-safe reciprocal (util.py:45-47; matrix_operators.py:208-219), or
    the block inverse of a TEB operator (matrix_operators.py:222-234)."""
    if isinstance(op, TEB):
        return op.pinv()
    return _recip(op)


def log_det(op, N):
    """Weighted log-determinant of a diagonal operator (util.py:39-42;
    matrix_operators.py:179-190: DiagonalEB sums both blocks) or a TEB block
    operator."""
    if isinstance(op, TEB):
        return op.log_det(N)
    w = lf.fourier_weights(N)
    total = 0.0
    for m in np.asarray(op).reshape(-1, *np.shape(op)[-2:]):
        total = total + _log_det1(m, w)
    return total


def eb2qu_field(f, N, theta_pix):
    """fields.py:182-199: T passes through, (E,B) -> (Q,U)."""
    if f.shape[0] == 1:
        return f
    q, u = lf.eb2qu(f[-2], f[-1], N, theta_pix)
    return np.stack(([f[0]] if f.shape[0] == 3 else []) + [q, u])


def qu2eb_field(f, N, theta_pix):
    """fields.py:209-226."""
    if f.shape[0] == 1:
        return f
    e, b = lf.qu2eb(f[-2], f[-1], N, theta_pix)
    return np.stack(([f[0]] if f.shape[0] == 3 else []) + [e, b])


def to_map(f):
    """fields.py:167-172 (irfft2 of every component)."""
    return np.fft.irfft2(f)


def to_fourier(f):
    """fields.py:158-163."""
    return np.fft.rfft2(f)


def phi_map(phi):
    """map(phi) for a FOURIER phi (gradients.py:17); MAP phi passes through."""
    return phi if phi.shape[-2] == phi.shape[-1] else np.fft.irfft2(phi)


def gradf_logpdf(f, phi, d, Cf, Cn, M, B, theta_pix, n=10, mask_pix=None):
    """d logpdf / d f = L^T B M Cn^-1 (d - M B L f) - Cf^-1 f (gradients.py:12-26).

    With ``mask_pix`` (a real [N, N] map; an extension, the reference's mask is
    Fourier-diagonal) M acts on the T or Q/U maps, ``d`` is a MAP array
    [C, N, N] and ``M`` is ignored:
        g = B qu2eb(fourier(m map(eb2qu(Cn^-1 qu2eb(fourier(d - m map(eb2qu(B L f))))))))"""
    N = f.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    fq = to_map(eb2qu_field(f, N, theta_pix))                       # :16
    phim = phi_map(phi)                                             # :17
    lensed = qu2eb_field(to_fourier(lf.lense_flow(fq, phim, pw, n)), N, theta_pix)  # :18
    if mask_pix is None:
        g = B * M * pinv(Cn) * (d - M * B * lensed)                 # :20
    else:
        q = to_map(eb2qu_field(B * lensed, N, theta_pix))
        r = d - mask_pix * q
        s_ = to_map(eb2qu_field(pinv(Cn) * qu2eb_field(to_fourier(r), N, theta_pix), N, theta_pix))
        g = B * qu2eb_field(to_fourier(mask_pix * s_), N, theta_pix)
    g = to_map(eb2qu_field(g, N, theta_pix))                        # :22
    g = lf.lense_flow(g, phim, pw, n, adjoint=True)                 # :23
    return (qu2eb_field(to_fourier(g), N, theta_pix)
            - pinv(Cf) * qu2eb_field(to_fourier(fq), N, theta_pix))  # :25


def data_residual(lensed, d, M, B, N, theta_pix, mask_pix=None):
    """d - M B L f in EB Fourier (statistics.py:159); with a pixel mask the
    residual is formed on the maps, d - m map(eb2qu(B L f)), and returned in EB
    Fourier."""
    if mask_pix is None:
        return d - M * B * lensed
    q = to_map(eb2qu_field(B * lensed, N, theta_pix))
    return qu2eb_field(to_fourier(d - mask_pix * q), N, theta_pix)


def logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n=10, mask_pix=None):
    """statistics.py:138-165 (note the argument order differs from gradf_logpdf)."""
    N = f.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    logdets = log_det(Cphi, N) + log_det(Cn, N) + log_det(Cf, N)    # :143-145
    field_product = lf.field_dot(f, pinv(Cf) * f, N)                # :148
    phi_product = lf.primal_dot(phi, pinv(Cphi) * phi, N)           # :149 / :185-186
    fq = to_map(eb2qu_field(f, N, theta_pix))                       # :155
    lensed = qu2eb_field(to_fourier(lf.lense_flow(fq, phi, pw, n)), N, theta_pix)  # :156
    diff = data_residual(lensed, d, M, B, N, theta_pix, mask_pix)   # :159
    data_product = lf.field_dot(diff, pinv(Cn) * diff, N)           # :160
    return -np.real(data_product + field_product + phi_product + logdets) / 2   # :164-165


def preconditioner(Cf, B, M, Cn, mask_pix=None):
    """P = Cf^-1 + B M Cn^-1 M B (wiener_filter.py:35-43); with a pixel mask M
    is replaced by its mean square."""
    if mask_pix is not None:
        return pinv(Cf) + B * B * pinv(Cn) * float(np.mean(mask_pix ** 2))
    return pinv(Cf) + B * M * pinv(Cn) * M * B


def conjugate_gradient(x0, d, phi, M, B, Cn, Cf, b, P, theta_pix, maxiter=500, tol=1e-1,
                       n=10, history=None, mask_pix=None, rtol=None):
    """PCG exactly as wiener_filter.py:51-105 (absolute stop on dot(r, P^-1 r));
    ``rtol`` adds the relative stop sqrt(dot(r,r)/dot(b,b)) < rtol."""
    N = x0.shape[-2]

    def A(x):                                                       # :45-49, :57-59
        dz = 0 * x if mask_pix is None else np.zeros(x.shape[:-1] + (N,))
        return gradf_logpdf(x, phi, dz, Cf, Cn, M, B, theta_pix, n, mask_pix)

    bb = lf.field_dot(b, b, N)

    def rel_ok(r):
        return rtol is not None and np.sqrt(lf.field_dot(r, r, N) / bb) < rtol

    Pinv = pinv(P)                                                  # :63
    r = b - A(x0)                                                   # :61-62
    z = Pinv * r                                                    # :65
    p = z
    res = lf.field_dot(r, z, N)                                     # :69
    res_curr = res
    x = x0
    k = 0
    while res_curr >= tol and k < maxiter and not rel_ok(r):        # :75-77
        Ap = A(p)                                                   # :82
        alpha = res / lf.field_dot(p, Ap, N)                        # :84
        x = x + alpha * p                                           # :86
        r = r - alpha * Ap                                          # :88
        z = Pinv * r                                                # :90
        res_curr = lf.field_dot(r, z, N)                            # :92
        p = z + (res_curr / res) * p                                # :94-95
        res = res_curr
        k += 1
        if history is not None:
            history.append(res_curr)
    return x


def wiener_filter(x0, phi, d, Cf, Cn, M, B, theta_pix, maxiter=500, tol=1e-1, n=10,
                  history=None, mask_pix=None, rtol=None):
    """wiener_filter.py:10-26 (b = -gradf(0, d); P as above; PCG)."""
    b = -gradf_logpdf(0 * x0, phi, d, Cf, Cn, M, B, theta_pix, n, mask_pix)   # :28-33
    P = preconditioner(Cf, B, M, Cn, mask_pix)
    return conjugate_gradient(x0, d, phi, M, B, Cn, Cf, b, P, theta_pix, maxiter, tol, n,
                              history, mask_pix, rtol)


def logpdf_lensed_cotangent(f, phi, d, Cn, M, B, theta_pix, n=10, mask_pix=None):
    """(L f, d logpdf / d(L f)) on the maps: with r = d - M B L f,
    the cotangent is map(eb2qu(B M^T Cn^-1 r)) (Fourier mask) or
    m map(eb2qu(Cn^-1 r)) projected through B (pixel mask)."""
    N = f.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    fq = to_map(eb2qu_field(f, N, theta_pix))
    Lf = lf.lense_flow(fq, phi, pw, n)
    lensed = qu2eb_field(to_fourier(Lf), N, theta_pix)
    r = data_residual(lensed, d, M, B, N, theta_pix, mask_pix)
    if mask_pix is None:
        g = B * M * pinv(Cn) * r
    else:
        s_ = to_map(eb2qu_field(pinv(Cn) * r, N, theta_pix))
        g = B * qu2eb_field(to_fourier(mask_pix * s_), N, theta_pix)
    return Lf, to_map(eb2qu_field(g, N, theta_pix))


def grad_phi_logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n=10, mask_pix=None):
    """jax.grad(logpdf, argnums=1) for a FOURIER phi as the reference defines it
    (gradients.py:29-37): the gradient-ODE delta phi accumulated in Fourier
    (lense_flow.py:23-46, undo_inner_product leaves FOURIER arrays alone)
    minus the prior term 2 Cphi^-1 phi / 2 (statistics.py:196-208)."""
    pw = lf.pix_width_from_theta(theta_pix)
    Lf, cot = logpdf_lensed_cotangent(f, phi, d, Cn, M, B, theta_pix, n, mask_pix)
    _, _, dphi = lf.lensing_operator_gradients(phi, list(Lf), list(cot), -1, n, pw)
    return dphi - pinv(Cphi) * phi
