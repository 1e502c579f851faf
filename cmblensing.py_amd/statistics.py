"""logpdf and its phi gradient on the GPU (statistics.py:138-213 and
gradients.py:29-37 of the reference).

logpdf = -1/2 Re[ dot(r, Cn^-1 r) + dot(f, Cf^-1 f) + dot(phi, Cphi^-1 phi)
                  + logdet Cn + logdet Cphi + logdet Cf ],   r = d - M B L f

logpdf is built like the reference's: the field is lensed by
lense_flow_wrapper (the hand-written gradient-ODE backward behind a
torch.autograd.Function, MAP or FOURIER phi) and the phi prior goes through
phi_dot_wrapper (its custom VJP 2 Cphi^-1 phi, statistics.py:183-213), so
torch.autograd.grad(logpdf, (f, phi)) gives d logpdf / d(f, phi) in the
reference's convention (Riesz representers under dot for FOURIER matrices):
the same values as gradf_logpdf and grad_phi_logpdf.

grad_phi_logpdf follows the reference's definition for a FOURIER phi: the
gradient-ODE delta phi accumulated in Fourier (the hand-written backward,
``cmbl_lenseflow_vjp``) minus Cphi^-1 phi (the custom VJP of the prior term,
statistics.py:196-208).  Both accept a PixelMask (the residual is then formed
on the maps).
"""
from __future__ import annotations

import torch

from . import _lib
from .constants import Basis, DEFAULT_NSTEPS, Parametrization
from .fields import (Field, FlatS0, FlatS02, dot, eb2qu, fourier, map as to_map, qu2eb)
from .fields import _needs_grad, _primal_dot_raw, _real_dtype
from .lense_flow import (_phi_matrix, lense_flow_matrices, lense_flow_vjp_matrices,
                         lense_flow_wrapper, phi_handles_for)
from .matrix_operators import PixelMask, log_det, pinv


def _lensed_maps(field: Field, phi, n: int):
    """(stacked unlensed QU/T maps, stacked lensed maps, group names)."""
    fq = to_map(eb2qu(field))
    names = fq._matrix_names()
    m0 = getattr(fq, names[0])
    N = m0.shape[-1]
    x = torch.stack([getattr(fq, nm).reshape(-1, N, N) for nm in names], dim=1)
    y = lense_flow_matrices(x, _phi_matrix(phi), field.theta_pix, n, 1, False)
    return fq, x, y, names, m0.shape[:-2]


def _residual(lensed_eb: Field, data: Field, mask, beam) -> Field:
    """r = d - M B L f (statistics.py:159) in EB Fourier."""
    if isinstance(mask, PixelMask):
        q = to_map(eb2qu(beam * lensed_eb))
        m = mask.matrix.to(dtype=getattr(q, q._matrix_names()[0]).dtype)
        return qu2eb(fourier(data - q * m))
    return data - mask * (beam * lensed_eb)


class _PhiDotFn(torch.autograd.Function):
    """dot(phi, Cphi^-1 phi) with the custom VJP of statistics.py:183-213:
    the gradient with respect to phi is 2 Cphi^-1 phi times the cotangent (the
    covariance is not differentiated)."""

    @staticmethod
    def forward(ctx, phi_m, cinv, nside):
        ctx.save_for_backward(phi_m, cinv)
        return _primal_dot_raw(phi_m, cinv * phi_m, nside)

    @staticmethod
    def backward(ctx, g):
        phi_m, cinv = ctx.saved_tensors
        gphi = (2 * g[..., None, None] * (cinv * phi_m)).to(phi_m.dtype).sum_to_size(phi_m.shape)
        return gphi, None, None


def phi_dot_wrapper(phi, phi_covariance) -> torch.Tensor:
    """dot(phi, pinv(Cphi) * phi) (statistics.py:183-186) with its custom VJP."""
    if not _needs_grad(phi.scalar_matrix):
        return dot(phi, pinv(phi_covariance) * phi)
    pm = phi.scalar_matrix
    cinv = pinv(phi_covariance).scalar_matrix.to(device=pm.device, dtype=_real_dtype(pm))
    return _PhiDotFn.apply(pm, cinv, phi.nside)


def logpdf(field, phi, data, noise_covariance, phi_covariance, field_covariance, mask, beam,
           n: int = DEFAULT_NSTEPS):
    """statistics.py:138-165 (same argument order as the reference).  phi may be
    MAP or FOURIER (the reference keeps it FOURIER, statistics.py:152-156);
    differentiable in field and phi with torch.autograd."""
    if not isinstance(phi, Field):
        phi = FlatS0(scalar_matrix=phi, theta_pix=field.theta_pix)
    logdets = log_det(phi_covariance) + log_det(noise_covariance) + log_det(field_covariance)
    field_product = dot(field, pinv(field_covariance) * field)
    phi_product = phi_dot_wrapper(phi, phi_covariance)
    lensed = qu2eb(fourier(lense_flow_wrapper(to_map(eb2qu(field)), phi, n)))
    r = _residual(lensed, data, mask, beam)
    data_product = dot(r, pinv(noise_covariance) * r)
    return -(data_product + field_product + phi_product + logdets) / 2


def _lensed_and_cotangent(f: Field, phi_m: torch.Tensor, data, noise_covariance, mask, beam,
                          n: int):
    """Stacked lensed maps y = L f [B, C, N, N] and the cotangent of the data
    term c = map(eb2qu(B M^T Cn^-1 (d - M B L f))), the lensing and the whole
    residual chain in one native call (cmbl_lensed_cotangent, T or EB
    fields), with the phi handles of the given phi basis; and the lead
    (batch) shape."""
    from .gradients import WfOperators, stack_matrices, _ncomp
    C = _ncomp(f)
    X, lead = stack_matrices(f)
    B, N = X.shape[0], X.shape[-2]
    rdt = torch.float32 if X.dtype == torch.complex64 else torch.float64
    plan = _lib.get_plan(N, f.theta_pix, rdt, B * C, X.device)
    ops = WfOperators(None, noise_covariance, mask, beam, rdt, X.device)
    D, _ = stack_matrices(data)
    D = D.to(X.dtype if not ops.pixel else rdt)
    if D.shape[0] == 1 and B > 1:
        D = D.expand(B, *D.shape[1:])
    D = D.contiguous()
    phis = phi_handles_for(plan, phi_m, B, n)
    y = torch.empty((B, C, N, N), dtype=rdt, device=X.device)
    c = torch.empty_like(y)
    _lib.lensed_cotangent(plan, phis, torch.view_as_real(X),
                          torch.view_as_real(D) if D.is_complex() else D, ops.struct(False),
                          y, c, B, C, n)
    return y, c, lead


def grad_phi_logpdf(unlensed_field, phi, data, noise_covariance, phi_covariance,
                    field_covariance, mask, beam, n: int = DEFAULT_NSTEPS) -> FlatS0:
    """d logpdf / d phi for a FOURIER phi (gradients.py:29-37 with
    lense_flow.py:10-51 and statistics.py:183-213)."""
    phi_m = _phi_matrix(phi)
    if phi_m.shape[-1] == phi_m.shape[-2]:
        raise ValueError("grad_phi_logpdf takes phi in the FOURIER basis, like the reference")
    f = unlensed_field
    if not isinstance(f, FlatS02):
        y, c, lead = _lensed_and_cotangent(f, phi_m, data, noise_covariance, mask, beam, n)
        N = y.shape[-1]
    else:
        fq, x, y, names, lead = _lensed_maps(f, phi, n)
        N = x.shape[-1]
        lensed = qu2eb(fourier(fq.replace(**{nm: y[:, c].reshape(*lead, N, N)
                                             for c, nm in enumerate(names)})))
        r = _residual(lensed, data, mask, beam)
        if isinstance(mask, PixelMask):
            s_ = to_map(eb2qu(pinv(noise_covariance) * r))
            m = mask.matrix.to(dtype=getattr(s_, s_._matrix_names()[0]).dtype)
            g = beam * qu2eb(fourier(s_ * m))
        else:
            g = beam * (mask * (pinv(noise_covariance) * r))
        cot = to_map(eb2qu(g))
        c = torch.stack([getattr(cot, nm).reshape(-1, N, N) for nm in names], dim=1)
    _, d_phi = lense_flow_vjp_matrices(y, c, phi_m, f.theta_pix, n, 1)
    return _phi_gradient(d_phi, phi, phi_m, phi_covariance, lead, N, f.theta_pix)


def _phi_gradient(d_phi, phi, phi_m, phi_covariance, lead, N, theta_pix) -> FlatS0:
    """The phi gradient from the gradient ODE's Fourier delta-phi: one per
    batch element (each chain's own phi) or a single one, minus the prior's
    Cphi^-1 phi (statistics.py:196-208)."""
    d_phi = d_phi.reshape(*lead, N, N // 2 + 1) if len(lead) else d_phi[0]
    prior = pinv(phi_covariance) * (phi if isinstance(phi, Field) else FlatS0(scalar_matrix=phi_m))
    out = d_phi.to(phi_m.dtype) - prior.scalar_matrix
    return FlatS0(scalar_matrix=out, theta_pix=theta_pix, basis=Basis.FOURIER,
                  parametrization=Parametrization.T)


def grad_logpdf(unlensed_field, phi, data, noise_covariance, phi_covariance, field_covariance,
                mask, beam, n: int = DEFAULT_NSTEPS):
    """Both gradients of logpdf, (d/d f, d/d phi), as gradf_logpdf and
    grad_phi_logpdf return them (EB or T FOURIER; FOURIER phi), from one
    forward LenseFlow and one gradient ODE -- what one reverse pass of the
    reference's jax.grad(logpdf, argnums=(0, 1)) computes (gradients.py:12-37,
    statistics.py:138-213).  The gradient ODE's delta f is L^T applied to the
    data term's cotangent, the adjoint LenseFlow gradf_logpdf runs, so
    d/d f = qu2eb(fourier(delta f)) - Cf^-1 f (cmbl_gradf_final) needs no
    second forward and no adjoint LenseFlow.  T + EB fields take the two
    separate calls."""
    from .gradients import gradf_logpdf, per_mode, stack_matrices, unstack_matrices
    f = unlensed_field
    phi_m = _phi_matrix(phi)
    if phi_m.shape[-1] == phi_m.shape[-2]:
        raise ValueError("grad_logpdf takes phi in the FOURIER basis, like the reference")
    if isinstance(f, FlatS02):
        return (gradf_logpdf(f, phi, data, field_covariance, noise_covariance, mask, beam, n=n),
                grad_phi_logpdf(f, phi, data, noise_covariance, phi_covariance, field_covariance,
                                mask, beam, n=n))
    y, c, lead = _lensed_and_cotangent(f, phi_m, data, noise_covariance, mask, beam, n)
    B, C, N = y.shape[0], y.shape[1], y.shape[-1]
    d_field, d_phi = lense_flow_vjp_matrices(y, c, phi_m, f.theta_pix, n, 1)
    X, lead_f = stack_matrices(f)
    plan = _lib.get_plan(N, f.theta_pix, y.dtype, B * C, X.device)
    cfinv = per_mode(pinv(field_covariance), y.dtype, X.device)
    out = torch.empty((B, C, N, N // 2 + 1), dtype=X.dtype, device=X.device)
    _lib.gradf_final(plan, d_field, torch.view_as_real(X), cfinv, torch.view_as_real(out), B, C)
    return (unstack_matrices(f, out, lead_f),
            _phi_gradient(d_phi, phi, phi_m, phi_covariance, lead, N, f.theta_pix))
