"""Mixed parametrisation on the GPU (mixing.py and gradients.py:40-85 of the
reference): the change of variables the MAP-joint and sampling algorithms
move phi in,

    f' = qu2eb(fourier(L(phi) map(eb2qu(D f)))),   phi' = G phi,

its inverse, the mixed logpdf (statistics.py:168-181) and the gradient of the
mixed logpdf with respect to phi' at fixed f':

    G^-1 (grad_phi logpdf + J^T grad_f logpdf),
    J = d/dphi [ map(D^-1 qu2eb(fourier(L^-1(phi) map(eb2qu f')))) ].

J^T is applied without materialising J: the linear part's transpose on the
maps is map(eb2qu(D^-1 fourier(.))) and the lensing part is the hand-written
gradient ODE (``cmbl_lenseflow_vjp``) integrated from the unlensed field, whose
Fourier delta-phi accumulation is what the reference returns for a FOURIER
phi.  Every LenseFlow, gradient ODE and gradf runs in libcmbl.so; fields may
carry a leading batch axis (chains), phi then one per chain.
"""
from __future__ import annotations

import torch

from .constants import DEFAULT_NSTEPS, FORWARD_LENSE, INVERSE_LENSE
from .fields import FlatS0, eb2qu, fourier, map as to_map, qu2eb
from .lense_flow import _phi_matrix, lense_flow, lense_flow_matrices, lense_flow_vjp_matrices
from .matrix_operators import log_det, pinv


def mix(field, phi, mixing_d, mixing_g, n: int = DEFAULT_NSTEPS):
    """mixing.py:7-20"""
    df = to_map(eb2qu(mixing_d * field))
    mixed_field = qu2eb(fourier(lense_flow(df, to_map(phi), n=n)))
    return mixed_field, mixing_g * phi


def unmix(mixed_field, mixed_phi, mixing_d, mixing_g, n: int = DEFAULT_NSTEPS):
    """mixing.py:23-37"""
    phi = pinv(mixing_g) * mixed_phi
    x = to_map(eb2qu(mixed_field))
    field = qu2eb(fourier(lense_flow(x, to_map(phi), n=n, direction=INVERSE_LENSE)))
    return pinv(mixing_d) * field, phi


def mixed_logpdf(mixed_field, mixed_phi, data, noise_covariance, phi_covariance,
                 field_covariance, mask, beam, mixing_g, mixing_d, n: int = DEFAULT_NSTEPS):
    """statistics.py:168-181 (mixing_g before mixing_d, as the reference)."""
    from .statistics import logpdf
    field, phi = unmix(mixed_field, mixed_phi, mixing_d, mixing_g, n)
    lp = logpdf(field, phi, data, noise_covariance, phi_covariance, field_covariance, mask, beam,
                n=n)
    return lp - torch.real(torch.as_tensor(log_det(mixing_g))) - \
        torch.real(torch.as_tensor(log_det(mixing_d)))


def _stack(field):
    names = field._matrix_names()
    m0 = getattr(field, names[0])
    N = m0.shape[-1]
    return torch.stack([getattr(field, nm).reshape(-1, N, N) for nm in names], dim=1), m0.shape[:-2]


def mixing_jacobian_phi_component(mixed_field, phi, grad_f, mixing_d, n: int = DEFAULT_NSTEPS):
    """gradients.py:69-85: J^T map(grad_f) as a FOURIER phi array (one per
    batch element when the fields are batched)."""
    x, lead = _stack(to_map(eb2qu(mixed_field)))
    phi_m = _phi_matrix(phi)
    y = lense_flow_matrices(x, phi_m, mixed_field.theta_pix, n, INVERSE_LENSE, False)
    cot, _ = _stack(to_map(eb2qu(pinv(mixing_d) * fourier(to_map(grad_f)))))
    _, d_phi = lense_flow_vjp_matrices(y, cot, phi_m, mixed_field.theta_pix, n, INVERSE_LENSE)
    N = x.shape[-1]
    return d_phi.reshape(*lead, N, N // 2 + 1) if len(lead) else d_phi[0]


def mixed_grad_phi_logpdf(mixed_field, mixed_phi, data, noise_covariance, phi_covariance,
                          field_covariance, mask, beam, mixing_d, mixing_g,
                          n: int = DEFAULT_NSTEPS) -> FlatS0:
    """gradients.py:40-65: G^-1 (grad_phi + J^T grad_f) at the unmixed point."""
    from .statistics import grad_logpdf
    field, phi = unmix(mixed_field, mixed_phi, mixing_d, mixing_g, n)
    # both gradients from one forward LenseFlow and one gradient ODE
    gf, gphi = grad_logpdf(field, phi, data, noise_covariance, phi_covariance, field_covariance,
                           mask, beam, n=n)
    # its inverse LenseFlow is not the unmixing's: that one lenses by map(phi)
    # (mixing.py:32), this one by the FOURIER phi (gradients.py:69-85, which
    # keeps the imaginary parts of the DC and Nyquist columns), and the two
    # differ once phi has such parts (a momentum step gives it them)
    chain = mixing_jacobian_phi_component(mixed_field, phi, gf, mixing_d, n)
    out = gphi.replace(scalar_matrix=gphi.scalar_matrix + chain.to(gphi.scalar_matrix.dtype))
    return pinv(mixing_g) * out
