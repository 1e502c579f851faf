"""LenseFlow on the GPU: the drop-in for lense_flow (lense_flow.py:289-415).

``lense_flow(field, phi, n=10, direction=FORWARD_LENSE, adjoint=False)`` keeps
the reference's signature, defaults and truth table (lense_flow.py:317-322):

    direction  adjoint   integrates   rate
       +1       False      t 0 -> 1   grad phi . M(t)^-1 grad f     (forward)
       -1       False      t 1 -> 0   same                          (inverse)
       +1       True       t 1 -> 0   div(M(t)^-1 grad phi  f)      (adjoint)
       -1       True       t 0 -> 1   same                          (inverse adjoint)

with M(t) = I + t Hess(phi) and classic RK4 over n steps (util.py:184-199).
All arithmetic runs in libcmbl.so: the phi derivatives and the p(t) table in
float64 once per (phi, n) (cached), then two fused kernels per RK4 stage.
Fields may carry a leading batch axis; phi is either shared ([N, N] or
[N, N//2+1]) or one per batch element ([B, N, N] / [B, N, N//2+1]).
"""
from __future__ import annotations

import torch

from . import _lib
from .constants import DEFAULT_NSTEPS, FORWARD_LENSE, INVERSE_LENSE
from .fields import Field, FlatS0, FlatS02, FlatS2


def _phi_matrix(phi) -> torch.Tensor:
    return phi.scalar_matrix if isinstance(phi, Field) else phi


def phi_handles_for(plan, phi_m: torch.Tensor, batch: int, n: int):
    """Prepared phi per batch element: a shared phi ([N, N] / [N, N//2+1]) or
    one per element (a leading batch axis of size `batch`, or of size 1)."""
    if phi_m.dim() == 2:
        return [_lib.get_phi(plan, phi_m, n)] * batch
    lead = phi_m.shape[:-2]
    count = 1
    for d in lead:
        count *= d
    if count == 1:
        return [_lib.get_phi(plan, phi_m, n, index=(0,) * len(lead))] * batch
    if count != batch or len(lead) != 1:
        raise ValueError(f"phi batch {tuple(lead)} does not match field batch {batch}")
    # one batched preparation for the whole batch (the cache holds every handle
    # of the call; an entry the cache drops stays valid while a caller holds it)
    return _lib.get_phi_batch(plan, phi_m, n)


def lense_flow_matrices(x: torch.Tensor, phi_m: torch.Tensor, theta_pix: float, n: int,
                        direction: int, adjoint: bool) -> torch.Tensor:
    """LenseFlow of stacked MAP matrices x [B, C, N, N] (C = 1: T, 2: Q,U, 3:
    T,Q,U, lensed as a T group and a Q,U group)."""
    if direction not in (FORWARD_LENSE, INVERSE_LENSE):
        raise ValueError("direction must be FORWARD_LENSE (+1) or INVERSE_LENSE (-1)")
    if x.dim() != 4 or x.shape[1] not in (1, 2, 3) or x.shape[-1] != x.shape[-2]:
        raise ValueError(f"expected MAP matrices [B, C(1|2|3), N, N], got {tuple(x.shape)}")
    if x.shape[1] == 3:
        return torch.cat([lense_flow_matrices(x[:, :1].contiguous(), phi_m, theta_pix, n,
                                              direction, adjoint),
                          lense_flow_matrices(x[:, 1:].contiguous(), phi_m, theta_pix, n,
                                              direction, adjoint)], dim=1)
    if not x.is_cuda:
        raise _lib.CmblError("LenseFlow runs on the GPU: move the field to a HIP device")
    B, C, N, _ = x.shape
    x = x.contiguous()
    plan = _lib.get_plan(N, theta_pix, x.dtype, B * C, x.device)
    phis = phi_handles_for(plan, phi_m, B, n)
    out = torch.empty_like(x)
    _lib.lenseflow(plan, phis, x, out, B, C, n, direction, adjoint)
    return out


def _apply(mats, lead, N, phi_m, theta_pix, n, direction, adjoint):
    x = torch.stack([m.reshape(-1, N, N) for m in mats], dim=1)
    y = lense_flow_matrices(x, phi_m, theta_pix, n, direction, adjoint)
    return [y[:, c].reshape(*lead, N, N) for c in range(len(mats))]


def lense_flow(field: Field, phi, n: int = DEFAULT_NSTEPS, direction: int = FORWARD_LENSE,
               adjoint: bool = False) -> Field:
    """Lens every matrix of a MAP-basis field (T and/or Q,U) by phi
    (lense_flow.py:289-292).  Like the reference, EB fields must be converted
    to QU (eb2qu) and to MAP (map) first."""
    phi_m = _phi_matrix(phi)
    names = field._matrix_names()
    m0 = getattr(field, names[0])
    if m0.shape[-1] != m0.shape[-2]:
        raise ValueError("lense_flow needs MAP-basis matrices (apply map() first)")
    N = m0.shape[-1]
    lead = m0.shape[:-2]
    updates = {}
    if isinstance(field, (FlatS0, FlatS02)):
        (t,) = _apply([field.scalar_matrix], lead, N, phi_m, field.theta_pix, n, direction, adjoint)
        updates["scalar_matrix"] = t
    if isinstance(field, (FlatS2, FlatS02)):
        q, u = _apply([field.polar_matrix_1, field.polar_matrix_2], lead, N, phi_m,
                      field.theta_pix, n, direction, adjoint)
        updates["polar_matrix_1"] = q
        updates["polar_matrix_2"] = u
    if not updates:
        raise TypeError(f"Unsupported type: {type(field).__name__}")
    return field.replace(**updates)


# --------------------------------------------------------------------------
# differentiation (lense_flow.py:10-126, 156-252)
# --------------------------------------------------------------------------
def lense_flow_vjp_matrices(lensed: torch.Tensor, cot: torch.Tensor, phi_m: torch.Tensor,
                            theta_pix: float, n: int, direction: int):
    """Raw backward of LenseFlow on stacked MAP matrices [B, C, N, N]:
    returns (d_field [B, C, N, N], d_phi_fourier [B, N, N//2+1]) -- the joint
    RK4 of (f, delta f, delta phi) run from the lensed field with -direction.
    d_field = L^T cot exactly; irfft2(d_phi_fourier) is the pixel gradient of
    <cot, L f> with respect to a MAP phi."""
    if direction not in (FORWARD_LENSE, INVERSE_LENSE):
        raise ValueError("direction must be FORWARD_LENSE (+1) or INVERSE_LENSE (-1)")
    if lensed.shape[1] == 3:
        # T and Q,U groups; the delta-phi rate sums over all components
        dT, pT = lense_flow_vjp_matrices(lensed[:, :1], cot[:, :1], phi_m, theta_pix, n,
                                         direction)
        dP, pP = lense_flow_vjp_matrices(lensed[:, 1:], cot[:, 1:], phi_m, theta_pix, n,
                                         direction)
        return torch.cat([dT, dP], dim=1), pT + pP
    B, C, N, _ = lensed.shape
    lensed = lensed.contiguous()
    cot = cot.to(lensed.dtype).contiguous()
    plan = _lib.get_plan(N, theta_pix, lensed.dtype, B * C, lensed.device)
    phis = phi_handles_for(plan, phi_m, B, n)
    d_field = torch.empty_like(lensed)
    cdt = torch.complex64 if lensed.dtype == torch.float32 else torch.complex128
    d_phi = torch.empty((B, N, N // 2 + 1), dtype=cdt, device=lensed.device)
    _lib.lenseflow_vjp(plan, phis, lensed, cot, d_field, torch.view_as_real(d_phi), B, C, n,
                       direction)
    return d_field, d_phi


def _irfft2_t(X: torch.Tensor, theta_pix: float) -> torch.Tensor:
    from .fields import _irfft2
    return _irfft2(X, theta_pix)


class LenseFlowFunction(torch.autograd.Function):
    """y = LenseFlow(x; phi) on MAP matrices x [B, C, N, N] with phi MAP
    ([N, N], [1, N, N] or [B, N, N]) or FOURIER (the same with N//2+1
    complex columns); backward through the hand-written gradient ODE
    (lense_flow.py:23-46: states recomputed from y, not stored).  Gradients
    follow the reference's custom VJP: dL/dx = L^T dL/dy on the maps; for a
    FOURIER phi the delta-phi rate accumulated in Fourier (lense_flow.py:84-92,
    which keeps the anti-Hermitian Nyquist content; undo_inner_product leaves
    FOURIER arrays alone, fields.py:146-154), for a MAP phi its irfft2, the
    pixel gradient.  A phi shared by the batch gets the sum over the batch."""

    @staticmethod
    def forward(ctx, x, phi_m, theta_pix, n, direction):
        y = lense_flow_matrices(x, phi_m.detach(), theta_pix, n, direction, False)
        ctx.save_for_backward(y, phi_m)
        ctx.cfg = (theta_pix, n, direction)
        return y

    @staticmethod
    def backward(ctx, gy):
        y, phi_m = ctx.saved_tensors
        theta_pix, n, direction = ctx.cfg
        d_field, d_phi = lense_flow_vjp_matrices(y, gy, phi_m.detach(), theta_pix, n, direction)
        g_phi = None
        if ctx.needs_input_grad[1]:
            fourier_phi = phi_m.shape[-1] != phi_m.shape[-2]
            g = d_phi if fourier_phi else _irfft2_t(d_phi, theta_pix)   # [B, N, *]
            g = g.to(phi_m.dtype)
            if phi_m.dim() == 2 or phi_m[..., 0, 0].numel() == 1:
                g = g.sum(0)
            g_phi = g.reshape(phi_m.shape)
        return d_field if ctx.needs_input_grad[0] else None, g_phi, None, None, None


def lense_flow_wrapper(field: Field, phi, n: int = DEFAULT_NSTEPS, direction: int = FORWARD_LENSE,
                       adjoint: bool = False) -> Field:
    """Differentiable lense_flow (lense_flow.py:10-51): torch autograd flows to
    the field matrices and to phi (MAP or FOURIER) through the gradient ODE."""
    if adjoint:
        raise NotImplementedError("the gradient ODE differentiates the non-adjoint LenseFlow "
                                  "(as lense_flow.py:23-46)")
    phi_m = _phi_matrix(phi)
    names = field._matrix_names()
    m0 = getattr(field, names[0])
    N = m0.shape[-1]
    lead = m0.shape[:-2]
    up = {}
    groups = []
    if isinstance(field, (FlatS0, FlatS02)):
        groups.append(["scalar_matrix"])
    if isinstance(field, (FlatS2, FlatS02)):
        groups.append(["polar_matrix_1", "polar_matrix_2"])
    if not groups:
        raise TypeError(f"Unsupported type: {type(field).__name__}")
    for g in groups:
        x = torch.stack([getattr(field, nm).reshape(-1, N, N) for nm in g], dim=1)
        y = LenseFlowFunction.apply(x, phi_m, field.theta_pix, n, direction)
        for c, nm in enumerate(g):
            up[nm] = y[:, c].reshape(*lead, N, N)
    return field.replace(**up)


def lense_flow_vjp(field: Field, phi, cotangent: Field, n: int = DEFAULT_NSTEPS,
                   direction: int = FORWARD_LENSE, reference_convention: bool = False):
    """(delta field, delta phi) of <cotangent, lense_flow(field, phi)>, with phi
    MAP or FOURIER (lense_flow.py:23-46).  The delta phi is returned in phi's
    basis: the raw Fourier accumulation for a FOURIER phi (as the reference
    keeps it), its irfft2 (the pixel gradient) for a MAP phi.
    reference_convention=True applies undo_inner_product (fields.py:146-154) to
    both, exactly like lense_flow_backwards does."""
    from .fields import undo_inner_product
    phi_m = _phi_matrix(phi)
    names = field._matrix_names()
    m0 = getattr(field, names[0])
    N = m0.shape[-1]
    lead = m0.shape[:-2]
    fourier_phi = phi_m.shape[-1] != phi_m.shape[-2]
    up = {}
    dphi_total = None
    groups = []
    if isinstance(field, (FlatS0, FlatS02)):
        groups.append(["scalar_matrix"])
    if isinstance(field, (FlatS2, FlatS02)):
        groups.append(["polar_matrix_1", "polar_matrix_2"])
    for g in groups:
        x = torch.stack([getattr(field, nm).reshape(-1, N, N) for nm in g], dim=1)
        c = torch.stack([getattr(cotangent, nm).reshape(-1, N, N) for nm in g], dim=1)
        y = lense_flow_matrices(x, phi_m, field.theta_pix, n, direction, False)
        d_field, d_phi = lense_flow_vjp_matrices(y, c, phi_m, field.theta_pix, n, direction)
        for k, nm in enumerate(g):
            up[nm] = d_field[:, k].reshape(*lead, N, N)
        dphi_total = d_phi if dphi_total is None else dphi_total + d_phi
    dphi = dphi_total.reshape(*lead, N, N // 2 + 1) if len(lead) else dphi_total[0]
    if not fourier_phi:
        dphi = _irfft2_t(dphi, field.theta_pix)
    dfield = field.replace(**up)
    dphi_field = FlatS0(scalar_matrix=dphi, theta_pix=field.theta_pix)
    if reference_convention:
        dfield = undo_inner_product(dfield)
        dphi_field = undo_inner_product(dphi_field)
    return dfield, dphi_field
