"""FP64 numpy restatement of CMBLensing.py's MAP-joint estimator
(map_joint.py), composed of the oracle's Wiener filter, mixing and logpdf.

TEST INFRASTRUCTURE ONLY -- imported by tests/ (and never by the package).

Fields are FOURIER arrays [C, N, N//2+1]; diagonal operators arrays of the
same shape (BlockTEB as wiener.TEB); phi a FOURIER [N, N//2+1] array.
"""
from __future__ import annotations

import numpy as np

from . import lenseflow as lf
from . import mixing as mx
from . import wiener as w


def phi_gradient_hessian(Cphi, QE):
    """map_joint.py:71-76"""
    return w.pinv(Cphi) + w.pinv(QE)


def logpdf_at_phi_new(alpha, g, fm, phim, d, B, M, Cf, Cphi, Cn, D, G, H, theta_pix, n=10):
    """map_joint.py:78-95"""
    test = phim + alpha * (H * g)
    f, phi = mx.unmix(fm, test, D, G, theta_pix, n)
    return -w.logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n)


def line_search(alpha, phim, H, g, fm, D, G, d, Cn, Cphi, Cf, M, B, theta_pix, n=10, num=10):
    """map_joint.py:97-114: (best alpha, the values at linspace(0, alpha, num))."""
    alphas = np.linspace(0.0, alpha, num)
    vals = np.array([logpdf_at_phi_new(a, g, fm, phim, d, B, M, Cf, Cphi, Cn, D, G, H,
                                       theta_pix, n) for a in alphas])
    return alphas[int(np.argmin(vals))], vals


def map_joint(d, Cf, Cn, Cphi, QE, M, B, D, theta_pix, num_steps=30, step="constant", n=10,
              maxiter=500, tol=1e-1, alpha_max=1.0, history=None):
    """map_joint.py:19-67 with G = 1 and step "constant" (alpha = 1), "grid"
    (line_search over [0, alpha_max]) or "bfgs" (scipy, as the reference).
    history, if given, receives logpdf(f, phi) after every Wiener filter."""
    N = d.shape[-2]
    G = np.ones((N, N // 2 + 1))
    f = 0 * d
    phi = np.zeros((N, N // 2 + 1), dtype=np.complex128)
    H = w.pinv(phi_gradient_hessian(Cphi, QE))
    alpha = 1.0
    for _ in range(num_steps):
        f = w.wiener_filter(f, phi, d, Cf, Cn, M, B, theta_pix, maxiter=maxiter, tol=tol, n=n)
        if history is not None:
            history.append(float(w.logpdf(f, phi, d, Cn, Cphi, Cf, M, B, theta_pix, n)))
        fm, phim = mx.mix(f, phi, D, G, theta_pix, n)
        g = mx.mixed_grad_phi_logpdf(fm, phim, d, Cn, Cphi, Cf, M, B, D, G, theta_pix, n)
        if step == "grid":
            alpha, _ = line_search(alpha_max, phim, H, g, fm, D, G, d, Cn, Cphi, Cf, M, B,
                                   theta_pix, n)
        elif step == "bfgs":
            from scipy.optimize import minimize
            res = minimize(lambda x: logpdf_at_phi_new(float(x[0]), g, fm, phim, d, B, M, Cf, Cphi,
                                                       Cn, D, G, H, theta_pix, n), 1,
                           method="BFGS")
            alpha = float(res.x[0])
        phi = phi + alpha * (H * g)
    return f, phi
