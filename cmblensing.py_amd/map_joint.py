"""MAP-joint estimation of (f, phi) on the GPU (map_joint.py of the reference).

Each of num_steps iterations (map_joint.py:19-67):
  1. f <- wiener_filter(f, phi, data, Cf, Cn, M, B)            (PCG, libcmbl.so)
  2. (f', phi') <- mix(f, phi, D, G = 1)                        (LenseFlow)
  3. g <- mixed_grad_phi_logpdf(f', phi', ...)                  (gradient ODE)
  4. alpha <- argmin_a -logpdf(unmix(f', phi' + a H g))         (line search)
  5. phi <- phi + alpha H g,   H = (Cphi^-1 + QE^-1)^-1

The step length follows the reference: scipy's BFGS from alpha = 1 on the
host, one batched GPU evaluation (unmix + logpdf) per function call, or a
constant alpha = 1.  ``step="grid"`` is the reference's line_search
(map_joint.py:97-114: the best of 10 alphas in [0, alpha]) evaluated as ONE
batched LenseFlow pair over every alpha, and over every trial when the data
set carries a trial axis (independent MAP-joint runs advance together).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .constants import DEFAULT_NSTEPS
from .fields import FlatS0, zero_scalar_field_like
from .matrix_operators import get_scalar_identity, pinv
from .mixing import mix, mixed_grad_phi_logpdf, unmix
from .statistics import logpdf
from .wiener_filter import wiener_filter


def phi_gradient_hessian(phi_covariance, quadratic_estimate):
    """map_joint.py:71-76: Cphi^-1 + QE^-1."""
    return pinv(phi_covariance) + pinv(quadratic_estimate)


def logpdf_at_phi_new(alpha, grad_phi, mixed_field, phi_predict, data, beam, mask,
                      field_covariance, phi_covariance, noise_covariance, mixing_d, mixing_g,
                      hessian, n: int = DEFAULT_NSTEPS):
    """map_joint.py:78-95: -logpdf at the unmixed (f', phi' + alpha H g)."""
    test_phi = phi_predict + alpha * (hessian * grad_phi)
    field, phi = unmix(mixed_field, test_phi, mixing_d, mixing_g, n)
    return -1 * logpdf(field, phi, data, noise_covariance, phi_covariance, field_covariance, mask,
                       beam, n=n)


def _lead(field) -> tuple:
    m = getattr(field, field._matrix_names()[0])
    return tuple(m.shape[:-2])


def _tile(field, T: int, K: int):
    """Matrices [T?, ..., N, M] -> [T*K, ..., N, M] (each trial repeated K times)."""
    def rep(m):
        if T == 0:
            return m.unsqueeze(0).expand(K, *m.shape).contiguous()
        return m.repeat_interleave(K, dim=0)
    return field.replace(**{nm: rep(getattr(field, nm)) for nm in field._matrix_names()})


def line_search(alpha, mixed_phi, hessian, mixed_grad_phi, mixed_field, mixing_d, mixing_g, data,
                noise_covariance, phi_covariance, field_covariance, mask, beam,
                n: int = DEFAULT_NSTEPS, num: int = 10, return_values: bool = False):
    """map_joint.py:97-114: the alpha in linspace(0, alpha, num) minimising
    -logpdf(unmix(f', phi' + alpha H g)), all alphas (and all trials of a
    batched data set) in one batched evaluation.  Returns alpha (a float, or a
    [T] tensor for T trials) and with return_values the [T?, num] values."""
    dev = mixed_phi.scalar_matrix.device
    alphas = torch.linspace(0.0, float(alpha), num, dtype=torch.float64, device=dev)
    step = (hessian * mixed_grad_phi).scalar_matrix
    phi0 = mixed_phi.scalar_matrix
    T = phi0.shape[0] if phi0.dim() == 3 else 0
    if T:
        test = phi0[:, None] + alphas[None, :, None, None] * step[:, None]
        test = test.reshape(T * num, *phi0.shape[-2:])
    else:
        test = phi0[None] + alphas[:, None, None] * step[None]
    lead = _lead(mixed_field)
    Tf = lead[0] if lead else 0
    mf = _tile(mixed_field, Tf, num)
    d = _tile(data, Tf, num) if _lead(data) else data
    field, phi = unmix(mf, mixed_phi.replace(scalar_matrix=test), mixing_d, mixing_g, n)
    vals = -1 * logpdf(field, phi, d, noise_covariance, phi_covariance, field_covariance, mask,
                       beam, n=n)
    vals = torch.as_tensor(vals).reshape(T, num) if T else torch.as_tensor(vals).reshape(num)
    best = alphas[torch.argmin(vals, dim=-1)]
    out = best if T else float(best)
    return (out, vals) if return_values else out


def map_joint(data_set, num_steps: int = 30, constant_step: bool = False, *,
              step: Optional[str] = None, n: int = DEFAULT_NSTEPS, maxiter: int = 500,
              tol: float = 1e-1, alpha_max: float = 1.0, history: Optional[list] = None):
    """map_joint.py:19-67: returns (f, phi) after num_steps iterations.

    step: "bfgs" (the reference's default: scipy BFGS from alpha = 1),
    "constant" (alpha = 1; also selected by constant_step=True) or "grid" (the
    batched line search over linspace(0, alpha_max, 10)).  A data set with a
    trial axis needs "constant" or "grid".  maxiter / tol go to every Wiener
    filter (the reference's defaults 500 / 0.1).  history, if given, receives
    per step a dict with alpha and the -logpdf value the line search reached."""
    if step is None:
        step = "constant" if constant_step else "bfgs"
    if step not in ("bfgs", "constant", "grid"):
        raise ValueError("step must be 'bfgs', 'constant' or 'grid'")
    noise_covariance = data_set.noise_covariance
    mixing_d = data_set.mixing_d
    mixing_g = get_scalar_identity(mixing_d)          # MAP-joint is invariant to G
    field_covariance = data_set.field_covariance
    phi_covariance = data_set.phi_covariance
    quadratic_estimate = data_set.quadratic_estimate
    mask, beam, data = data_set.mask, data_set.beam, data_set.data
    trials = _lead(data)
    if trials and step == "bfgs":
        raise ValueError("a data set with a trial axis needs step='constant' or 'grid'")

    field_predict = 0 * data
    phi_predict = zero_scalar_field_like(data)
    if trials:
        phi_predict = phi_predict.replace(
            scalar_matrix=phi_predict.scalar_matrix.expand(*trials, -1, -1).contiguous())
    hessian = pinv(phi_gradient_hessian(phi_covariance, quadratic_estimate))
    alpha = 1.0
    for _ in range(num_steps):
        field_predict = wiener_filter(field_predict, phi_predict, data, field_covariance,
                                      noise_covariance, mask, beam, n=n, maxiter=maxiter, tol=tol)
        mixed_field, mixed_phi = mix(field_predict, phi_predict, mixing_d, mixing_g, n)
        g = mixed_grad_phi_logpdf(mixed_field, mixed_phi, data, noise_covariance, phi_covariance,
                                  field_covariance, mask, beam, mixing_d, mixing_g, n=n)
        value = None
        if step == "bfgs":
            from scipy.optimize import minimize

            def f(x):
                return float(logpdf_at_phi_new(float(x[0]), g, mixed_field, mixed_phi, data, beam,
                                               mask, field_covariance, phi_covariance,
                                               noise_covariance, mixing_d, mixing_g, hessian, n))
            res = minimize(f, 1, method="BFGS")
            alpha, value = float(res.x[0]), float(res.fun)
        elif step == "grid":
            alpha, vals = line_search(alpha_max, mixed_phi, hessian, g, mixed_field, mixing_d,
                                      mixing_g, data, noise_covariance, phi_covariance,
                                      field_covariance, mask, beam, n=n, return_values=True)
            value = vals.min(dim=-1).values
        hg = (hessian * g).scalar_matrix
        a = alpha[:, None, None] if isinstance(alpha, torch.Tensor) else alpha
        phi_predict = phi_predict.replace(scalar_matrix=phi_predict.scalar_matrix + a * hg)
        if history is not None:
            history.append({"alpha": alpha, "neg_logpdf": value})
    return field_predict, phi_predict
