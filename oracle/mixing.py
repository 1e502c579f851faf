"""FP64 numpy restatement of CMBLensing.py's mixed parametrisation and the
gradient the HMC phi sampler follows (oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/ (and never by the package).

Fields and operators as in ``oracle.wiener`` (FOURIER arrays [C, N, N//2+1];
diagonal operators as arrays, BlockTEB as ``wiener.TEB``).  phi is a FOURIER
half plane (the reference differentiates with respect to it).
"""
from __future__ import annotations

import numpy as np

from . import lenseflow as lf
from . import wiener as w


def mix(f, phi, D, G, theta_pix, n=10):
    """mixing.py:7-20: f' = qu2eb(fourier(L(phi) map(eb2qu(D f)))), phi' = G phi."""
    N = f.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    x = w.to_map(w.eb2qu_field(D * f, N, theta_pix))
    mixed = w.qu2eb_field(w.to_fourier(lf.lense_flow(x, w.phi_map(phi), pw, n)), N, theta_pix)
    return mixed, G * phi


def unmix(fm, phim, D, G, theta_pix, n=10):
    """mixing.py:23-37: phi = G^-1 phi', f = D^-1 qu2eb(fourier(L^-1(phi) map(eb2qu f')))."""
    N = fm.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    phi = w.pinv(G) * phim
    x = w.to_map(w.eb2qu_field(fm, N, theta_pix))
    y = lf.lense_flow(x, w.phi_map(phi), pw, n, direction=lf.INVERSE_LENSE)
    return w.pinv(D) * w.qu2eb_field(w.to_fourier(y), N, theta_pix), phi


def mixed_logpdf(fm, phim, d, Cn, Cphi, Cf, M, B, G, D, theta_pix, n=10):
    """statistics.py:168-181 (mixing_g before mixing_d, as the reference)."""
    N = fm.shape[-2]
    f, phi = unmix(fm, phim, D, G, theta_pix, n)
    return (w.logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n)
            - np.real(w.log_det(G, N)) - np.real(w.log_det(D, N)))


def mixing_jacobian_phi(fm, phi, gf, D, theta_pix, n=10):
    """gradients.py:69-85: the pullback of phi -> map(D^-1 qu2eb(fourier(L^-1(phi)
    map(eb2qu f')))) applied to map(grad_f).  The linear part's transpose on
    the maps is map(eb2qu(D^-1 fourier(.))); the lensing part is the gradient
    ODE integrated from the unlensed field forward (lense_flow.py:23-46),
    whose Fourier accumulation is what the reference returns for a FOURIER
    phi."""
    N = fm.shape[-2]
    pw = lf.pix_width_from_theta(theta_pix)
    x = w.to_map(w.eb2qu_field(fm, N, theta_pix))
    y = lf.lense_flow(x, w.phi_map(phi), pw, n, direction=lf.INVERSE_LENSE)
    cot = w.to_map(w.eb2qu_field(w.pinv(D) * w.to_fourier(w.to_map(gf)), N, theta_pix))
    _, _, dphi = lf.lensing_operator_gradients(phi, list(y), list(cot), lf.FORWARD_LENSE, n, pw)
    return dphi


def mixed_grad_phi_logpdf(fm, phim, d, Cn, Cphi, Cf, M, B, D, G, theta_pix, n=10):
    """gradients.py:40-65: G^-1 (grad_phi + J^T grad_f) at the unmixed point."""
    f, phi = unmix(fm, phim, D, G, theta_pix, n)
    gphi = w.grad_phi_logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n)
    gf = w.gradf_logpdf(f, phi, d, Cf, Cn, M, B, theta_pix, n)
    return w.pinv(G) * (gphi + mixing_jacobian_phi(fm, phi, gf, D, theta_pix, n))


def symplectic_integrate(x0, p0, fm, d, Cn, Cphi, Cf, M, B, D, G, mass, theta_pix, n=10,
                         num_steps=30, step_size=0.01):
    """sampling_ar.py:105-140 (leapfrog on the mixed phi; mass is a diagonal
    array).  Returns (delta_h, x, p)."""
    N = fm.shape[-2]
    minv = w.pinv(mass)

    def grad(x):
        return mixed_grad_phi_logpdf(fm, x, d, Cn, Cphi, Cf, M, B, D, G, theta_pix, n)

    def hamiltonian(x, p):
        lp = mixed_logpdf(fm, x, d, Cn, Cphi, Cf, M, B, G, D, theta_pix, n)
        return lp - lf.primal_dot(p, minv * p, N) / 2

    g = grad(x0)
    x, p = x0, p0
    for _ in range(num_steps):
        g_prev = g
        x = x - step_size * minv * (p - 0.5 * step_size * g)
        g = grad(x)
        p = p - 0.5 * step_size * (g + g_prev)
    return hamiltonian(x, p) - hamiltonian(x0, p0), x, p
