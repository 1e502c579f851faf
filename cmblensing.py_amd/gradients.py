"""gradf_logpdf on the GPU (gradients.py:12-26 of the reference).

    gradf = L^T B^T M^T Cn^-1 (d - M B L f) - Cf^-1 f

for EB (FlatS2) or T (FlatS0) fields in the FOURIER basis, optionally with a
leading batch axis.  The whole chain -- EB->QU, irfft2, forward LenseFlow,
rfft2, the data term, the adjoint LenseFlow and the prior term -- runs in
libcmbl.so (``cmbl_gradf_logpdf``).  Like the reference (gradients.py:17) the
lensing potential is taken to the MAP basis first.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import _lib
from .constants import DEFAULT_NSTEPS
from .fields import Field, FlatS0, FlatS02, FlatS2, map as to_map
from .matrix_operators import (BlockTEB, DiagonalEB, DiagonalScalar, PixelMask, pinv)


def _ncomp(field: Field) -> int:
    if isinstance(field, FlatS02):
        return 3
    if isinstance(field, FlatS2):
        return 2
    if isinstance(field, FlatS0):
        return 1
    raise TypeError(f"Unsupported type: {type(field).__name__}")


def stack_matrices(field: Field) -> torch.Tensor:
    """[B, C, ...] stack of a field's matrices (B = 1 when unbatched)."""
    mats = [getattr(field, n) for n in field._matrix_names()]
    lead = mats[0].shape[:-2]
    return torch.stack([m.reshape(-1, *m.shape[-2:]) for m in mats], dim=1).contiguous(), lead


def unstack_matrices(field: Field, X: torch.Tensor, lead) -> Field:
    names = field._matrix_names()
    return field.replace(**{n: X[:, c].reshape(*lead, *X.shape[-2:]) for c, n in enumerate(names)})


def per_mode(op, dtype, device) -> torch.Tensor:
    """[C, N, N//2+1] planes of a diagonal operator, or the five BlockTEB planes
    [TT, TE, ET, EE, BB]."""
    if isinstance(op, (DiagonalScalar, DiagonalEB, BlockTEB)):
        return op.per_mode().to(device=device, dtype=dtype).contiguous()
    raise TypeError(f"Unsupported operator: {type(op).__name__}")


class WfOperators:
    """Per-mode operator arrays in the plan dtype, built once per set of
    operators: mb = M B, bmcn = B M Cn^-1, cfinv = Cf^-1 (Fourier-diagonal
    mask), or beam, cninv and a pixel mask.  BlockTEB products are formed in
    the reference's order (gradients.py:20: (B M) Cn^-1 and M B) before they are
    reduced to planes."""

    def __init__(self, field_covariance, noise_covariance, mask, beam, dtype, device):
        self.dtype, self.device = dtype, device
        self.pixel = isinstance(mask, PixelMask)
        pm = lambda op: per_mode(op, dtype, device)
        self.cfinv = pm(pinv(field_covariance)) if field_covariance is not None else None
        if any(isinstance(o, BlockTEB) for o in (noise_covariance, mask, beam)):
            if self.pixel:
                raise NotImplementedError("T+EB fields take BlockTEB Fourier operators, not a "
                                          "PixelMask")
            self.mb = pm(mask * beam)
            self.bmcn = pm(beam * mask * pinv(noise_covariance))
            self.beam = self.cninv = self.mask = None
            return
        if self.pixel:
            self.beam = pm(beam)
            self.cninv = pm(pinv(noise_covariance))
            self.mask = mask.matrix.to(device=device, dtype=dtype).contiguous()
            self.mb = self.bmcn = None
        else:
            b, m = pm(beam), pm(mask)
            self.mb = (m * b).contiguous()
            self.bmcn = (b * m * pm(pinv(noise_covariance))).contiguous()
            self.beam = self.cninv = self.mask = None

    def struct(self, with_prior: bool = True) -> "_lib.WfOps":
        p = lambda t: None if t is None else t.data_ptr()
        return _lib.WfOps(p(self.mb), p(self.bmcn), p(self.cfinv) if with_prior else None,
                          p(self.beam), p(self.cninv), p(self.mask))


def phi_handles(plan, phi, batch: int, n: int):
    """Prepared phi handles for each batch element (phi in MAP basis, as
    gradients.py:17 does)."""
    from .lense_flow import phi_handles_for
    pm = phi.scalar_matrix if isinstance(phi, Field) else phi
    if pm.shape[-1] != pm.shape[-2]:
        pm = _map_cache(pm, plan.theta_pix)
    return phi_handles_for(plan, pm, batch, n)


_map_cache_d: dict = {}


def _map_cache(pF: torch.Tensor, theta_pix: float) -> torch.Tensor:
    """map(phi) of a FOURIER phi, kept per phi object (and version) so the
    prepared handles of the MAP phi stay valid across calls."""
    import weakref
    key = id(pF)
    ent = _map_cache_d.get(key)
    if ent is not None and ent[0]() is pF and ent[1] == pF._version:
        return ent[2]
    pm = to_map(FlatS0(scalar_matrix=pF, theta_pix=theta_pix)).scalar_matrix
    _map_cache_d[key] = (weakref.ref(pF), pF._version, pm)
    if len(_map_cache_d) > 64:
        _map_cache_d.pop(next(iter(_map_cache_d)))
    return pm


def gradf_apply(plan, phis, X: Optional[torch.Tensor], D: Optional[torch.Tensor], ops: WfOperators,
                B: int, C: int, n: int, with_prior: bool = True) -> torch.Tensor:
    """Raw entry: X (EB Fourier stack [B, C, N, N/2+1] in the plan's complex
    dtype, or None for f = 0), D (data stack or None for d = 0)."""
    N = plan.N
    cdt = torch.complex64 if plan.dtype == torch.float32 else torch.complex128
    out = torch.empty((B, C, N, N // 2 + 1), dtype=cdt, device=plan.device)
    s = ops.struct(with_prior)
    _lib.gradf_logpdf(plan, phis, None if X is None else torch.view_as_real(X),
                      None if D is None else (torch.view_as_real(D) if D.is_complex() else D),
                      s, torch.view_as_real(out), B, C, n)
    return out


def gradf_logpdf(field: Field, phi, data: Field, field_covariance, noise_covariance, mask, beam,
                 n: int = DEFAULT_NSTEPS) -> Field:
    """d logpdf / d f in the EB FOURIER basis (gradients.py:12-26).  With a
    PixelMask, `data` is a MAP-basis field of T or Q/U maps."""
    C = _ncomp(field)
    X, lead = stack_matrices(field)
    B = X.shape[0]
    N = X.shape[-2]
    rdt = torch.float32 if X.dtype == torch.complex64 else torch.float64
    plan = _lib.get_plan(N, field.theta_pix, rdt, B * C, X.device)
    ops = WfOperators(field_covariance, noise_covariance, mask, beam, rdt, X.device)
    D, _ = stack_matrices(data)
    D = D.to(X.dtype if not ops.pixel else rdt)
    if D.shape[0] == 1 and B > 1:
        D = D.expand(B, *D.shape[1:]).contiguous()
    out = gradf_apply(plan, phi_handles(plan, phi, B, n), X, D, ops, B, C, n)
    return unstack_matrices(field, out, lead)
