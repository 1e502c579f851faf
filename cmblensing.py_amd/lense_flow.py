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


def _phi_list(phi_m: torch.Tensor, N: int, batch: int):
    """Split phi into per-batch-element matrices (shared phi -> same tensor)."""
    if phi_m.dim() == 2:
        return [phi_m] * batch
    flat = phi_m.reshape(-1, *phi_m.shape[-2:])
    if flat.shape[0] == 1:
        return [flat[0]] * batch
    if flat.shape[0] != batch:
        raise ValueError(f"phi batch {flat.shape[0]} does not match field batch {batch}")
    return [flat[i] for i in range(batch)]


def lense_flow_matrices(x: torch.Tensor, phi_m: torch.Tensor, theta_pix: float, n: int,
                        direction: int, adjoint: bool) -> torch.Tensor:
    """LenseFlow of stacked MAP matrices x [B, C, N, N] (C = 1 or 2)."""
    if direction not in (FORWARD_LENSE, INVERSE_LENSE):
        raise ValueError("direction must be FORWARD_LENSE (+1) or INVERSE_LENSE (-1)")
    if x.dim() != 4 or x.shape[1] not in (1, 2) or x.shape[-1] != x.shape[-2]:
        raise ValueError(f"expected MAP matrices [B, C(1|2), N, N], got {tuple(x.shape)}")
    if not x.is_cuda:
        raise _lib.CmblError("LenseFlow runs on the GPU: move the field to a HIP device")
    B, C, N, _ = x.shape
    x = x.contiguous()
    plan = _lib.get_plan(N, theta_pix, x.dtype, B * C, x.device)
    phis = [_lib.get_phi(plan, p, n) for p in _phi_list(phi_m, N, B)]
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
