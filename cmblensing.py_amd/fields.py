"""Field containers over device tensors, mirroring fields.py of the reference.

FlatS0 (T), FlatS2 (two polarisation matrices), FlatS02 (T + polarisation)
keep the reference's attribute names (scalar_matrix, polar_matrix_1/2,
nside, theta_pix, pix_width, basis, parametrization, fourier_weights) and its
element-wise arithmetic (fields.py:15-116).  Matrices are torch tensors on a
HIP device, MAP [..., N, N] real or FOURIER [..., N, N//2+1] complex; an
optional leading batch axis is carried through every operation.

Differences from the reference, all additive:
  * nside defaults to the matrix size (the reference defaults to 256 and
    silently mis-normalises other sizes); theta_pix still defaults to 2.5.
  * map/fourier run the library's HIP rfft2/irfft2 passes.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field as dc_field
from typing import Optional

import numpy as np
import torch

from . import _lib
from .constants import ARCMIN_PER_DEGREE, Basis, Parametrization, THETA_PIX_DEFAULT


def _real_dtype(t: torch.Tensor) -> torch.dtype:
    return {torch.complex64: torch.float32, torch.complex128: torch.float64}.get(t.dtype, t.dtype)


def _complex_dtype(t: torch.Tensor) -> torch.dtype:
    return {torch.float32: torch.complex64, torch.float64: torch.complex128}[t.dtype]


_WEIGHTS = {}


def fourier_weights_for(N: int, device=None, dtype=torch.float64) -> torch.Tensor:
    """[1, 2, ..., 2, 1] along the half axis (util.py:90-96).  One tensor per
    (N, device, dtype) is built and shared: it is read-only (every consumer
    reads it; none writes it), and building it anew launched three small
    kernels for each dot product and operator of a sampler's inner loop."""
    dev = torch.device(device) if device is not None else torch.device("cpu")
    if dev.type == "cuda" and dev.index is None:  # "cuda" means the current device
        dev = torch.device("cuda", torch.cuda.current_device())
    key = (N, str(dev), dtype)
    w = _WEIGHTS.get(key)
    if w is None:
        w = torch.full((N // 2 + 1,), 2.0, dtype=dtype, device=dev)
        w[0] = 1.0
        w[-1] = 1.0
        _WEIGHTS[key] = w
    return w


@dataclass
class Field:
    """Base container (fields.py:15-73)."""
    fourier_weights: Optional[torch.Tensor] = None
    nside: Optional[int] = None
    theta_pix: float = THETA_PIX_DEFAULT
    pix_width: Optional[float] = None
    basis: int = Basis.FOURIER

    def _matrix_names(self):
        raise NotImplementedError

    def _first(self) -> torch.Tensor:
        return getattr(self, self._matrix_names()[0])

    def __post_init__(self):
        m = self._first()
        if m is not None:
            if self.nside is None:
                self.nside = int(m.shape[-2])
            if m.shape[-1] == m.shape[-2] and self.basis == Basis.FOURIER and \
                    not torch.is_complex(m):
                self.basis = Basis.MAP
        if self.pix_width is None:
            self.pix_width = math.radians(self.theta_pix / ARCMIN_PER_DEGREE)

    def replace(self, **updates):
        return dataclasses.replace(self, **updates)

    def _binop(self, other, fn):
        if isinstance(other, self.__class__):
            up = {n: fn(getattr(self, n), getattr(other, n)) for n in self._matrix_names()}
        else:
            up = {n: fn(getattr(self, n), other) for n in self._matrix_names()}
        return self.replace(**up)

    def __add__(self, o):
        return self._binop(o, lambda a, b: a + b)

    def __radd__(self, o):
        return self.__add__(o)

    def __sub__(self, o):
        return self._binop(o, lambda a, b: a - b)

    def __rsub__(self, o):
        return self._binop(o, lambda a, b: b - a)

    def __mul__(self, o):
        from .matrix_operators import FieldOperator
        if isinstance(o, FieldOperator):
            return o.__mul__(self)
        return self._binop(o, lambda a, b: a * b)

    def __rmul__(self, o):
        return self.__mul__(o)

    def __truediv__(self, o):
        return self._binop(o, lambda a, b: a / b)

    def __rtruediv__(self, o):
        return self._binop(o, lambda a, b: b / a)

    def __neg__(self):
        return self._binop(-1, lambda a, b: a * b)


@dataclass
class FlatS0(Field):
    """Temperature-only field (fields.py:80-87)."""
    scalar_matrix: Optional[torch.Tensor] = None
    parametrization: int = Parametrization.T

    def _matrix_names(self):
        return ["scalar_matrix"]


@dataclass
class FlatS2(Field):
    """Polarisation-only field (fields.py:93-101); EB by default, QU after eb2qu."""
    polar_matrix_1: Optional[torch.Tensor] = None
    polar_matrix_2: Optional[torch.Tensor] = None
    parametrization: int = Parametrization.EB

    def _matrix_names(self):
        return ["polar_matrix_1", "polar_matrix_2"]


@dataclass
class FlatS02(Field):
    """Temperature + polarisation field (fields.py:107-116)."""
    scalar_matrix: Optional[torch.Tensor] = None
    polar_matrix_1: Optional[torch.Tensor] = None
    polar_matrix_2: Optional[torch.Tensor] = None
    parametrization: int = Parametrization.EB

    def _matrix_names(self):
        return ["scalar_matrix", "polar_matrix_1", "polar_matrix_2"]


def zero_scalar_field_like(field: Field) -> FlatS0:
    """fields.py:123-133: a FOURIER FlatS0 of zeros (complex128) on the field's
    grid and device."""
    N = field.nside
    dev = field._first().device
    return FlatS0(fourier_weights=fourier_weights_for(N, dev), nside=N, theta_pix=field.theta_pix,
                  pix_width=field.pix_width, basis=Basis.FOURIER,
                  scalar_matrix=torch.zeros((N, N // 2 + 1), dtype=torch.complex128, device=dev))


# --------------------------------------------------------------------------
# basis changes (fields.py:156-172) through the library's HIP FFT passes
# --------------------------------------------------------------------------
# The basis changes transform up to this many maps per launch (the shared
# plan of the geometry grows to hold them; larger stacks go in chunks).
FFT_BATCH_MAPS = 64


def _needs_grad(*ts) -> bool:
    return torch.is_grad_enabled() and any(t.requires_grad for t in ts)


# --------------------------------------------------------------------------
# autograd convention
# --------------------------------------------------------------------------
# Gradients of FOURIER matrices are Riesz representers under `dot` (the
# reference's convention: its phi-prior VJP returns 2 Cphi^-1 phi,
# statistics.py:196-208, and its LenseFlow VJP the Fourier-accumulated delta
# phi, lense_flow.py:84-92): g with d logpdf = dot(g, delta) for every
# perturbation delta.  Gradients of MAP matrices are the plain pixel
# gradients (d logpdf = sum_x g delta).  Under this pair of inner products
# the adjoint of rfft2 is irfft2 and the adjoint of irfft2 is rfft2, and every
# per-mode or per-pixel operation (operators, EB<->QU rotations, masks) is
# its own torch adjoint; so with fourier, map and dot defined below,
# torch.autograd.grad(logpdf, (f, phi)) returns gradf_logpdf and
# grad_phi_logpdf.

class _RFFT2Fn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, theta_pix):
        ctx.theta_pix = theta_pix
        return _rfft2_raw(x, theta_pix)

    @staticmethod
    def backward(ctx, G):
        return _irfft2_raw(G, ctx.theta_pix), None


class _IRFFT2Fn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, theta_pix):
        ctx.theta_pix = theta_pix
        return _irfft2_raw(X, theta_pix)

    @staticmethod
    def backward(ctx, g):
        return _rfft2_raw(g, ctx.theta_pix), None


def _rfft2(x: torch.Tensor, theta_pix: float) -> torch.Tensor:
    if _needs_grad(x):
        return _RFFT2Fn.apply(x, theta_pix)
    return _rfft2_raw(x, theta_pix)


def _irfft2(X: torch.Tensor, theta_pix: float) -> torch.Tensor:
    if _needs_grad(X):
        return _IRFFT2Fn.apply(X, theta_pix)
    return _irfft2_raw(X, theta_pix)


def _rfft2_raw(x: torch.Tensor, theta_pix: float) -> torch.Tensor:
    N = x.shape[-1]
    if x.shape[-2] != N:
        raise ValueError("rfft2 expects square MAP matrices")
    xc = x.contiguous()
    nmaps = xc.numel() // (N * N)
    out = torch.empty(x.shape[:-1] + (N // 2 + 1,), dtype=_complex_dtype(xc), device=x.device)
    plan = _lib.get_plan(N, theta_pix, xc.dtype, min(nmaps, FFT_BATCH_MAPS), x.device)
    _lib.rfft2(plan, xc, torch.view_as_real(out), nmaps)
    return out


def _irfft2_raw(X: torch.Tensor, theta_pix: float) -> torch.Tensor:
    N = X.shape[-2]
    if X.shape[-1] != N // 2 + 1:
        raise ValueError("irfft2 expects FOURIER [.., N, N/2+1] matrices")
    Xc = X.contiguous()
    nmaps = Xc.numel() // (N * (N // 2 + 1))
    rdt = _real_dtype(Xc)
    out = torch.empty(X.shape[:-1] + (N,), dtype=rdt, device=X.device)
    plan = _lib.get_plan(N, theta_pix, rdt, min(nmaps, FFT_BATCH_MAPS), X.device)
    _lib.irfft2(plan, torch.view_as_real(Xc), out, nmaps)
    return out


def fourier(field: Field) -> Field:
    """rfft2 of every matrix (fields.py:158-163); no basis guard, like the reference."""
    up = {n: _rfft2(getattr(field, n), field.theta_pix) for n in field._matrix_names()}
    return field.replace(basis=Basis.FOURIER, **up)


def map(field: Field) -> Field:  # noqa: A001 - reference name
    """irfft2 of every matrix (fields.py:167-172)."""
    up = {n: _irfft2(getattr(field, n), field.theta_pix) for n in field._matrix_names()}
    return field.replace(basis=Basis.MAP, **up)


# --------------------------------------------------------------------------
# EB <-> QU (util.py:140-172; fields.py:176-226)
# --------------------------------------------------------------------------
_rot_cache: dict = {}


def rotation_2psi(nside: int, theta_pix: float, device, dtype=torch.float64):
    """(cos 2psi, sin 2psi) on the [N, N//2+1] half plane, psi = arctan2(lx, ly)
    with lx the half axis and its Nyquist column negated for rows <= N/2."""
    key = (nside, round(float(theta_pix), 12), str(device), dtype)
    if key not in _rot_cache:
        d = math.radians(theta_pix / ARCMIN_PER_DEGREE)
        lx = 2 * np.pi * np.fft.rfftfreq(nside, d)
        ly = 2 * np.pi * np.fft.fftfreq(nside, d)
        LX, LY = np.meshgrid(lx, ly)
        LX = LX.copy()
        LX[: nside // 2 + 1, -1] *= -1
        a = np.arctan2(LX, LY)
        _rot_cache[key] = (torch.as_tensor(np.cos(2 * a), dtype=dtype, device=device),
                           torch.as_tensor(np.sin(2 * a), dtype=dtype, device=device))
    return _rot_cache[key]


def primal_eb2qu(e, b, nside, theta_pix):
    c2, s2 = rotation_2psi(nside, theta_pix, e.device, _real_dtype(e))
    return -e * c2 + b * s2, -e * s2 - b * c2


def primal_qu2eb(q, u, nside, theta_pix):
    c2, s2 = rotation_2psi(nside, theta_pix, q.device, _real_dtype(q))
    return -q * c2 - u * s2, q * s2 - u * c2


def eb2qu(field: Field) -> Field:
    if isinstance(field, FlatS0):
        return field
    q, u = primal_eb2qu(field.polar_matrix_1, field.polar_matrix_2, field.nside, field.theta_pix)
    return field.replace(polar_matrix_1=q, polar_matrix_2=u, parametrization=Parametrization.QU)


def qu2eb(field: Field) -> Field:
    if isinstance(field, FlatS0):
        return field
    e, b = primal_qu2eb(field.polar_matrix_1, field.polar_matrix_2, field.nside, field.theta_pix)
    return field.replace(polar_matrix_1=e, polar_matrix_2=b, parametrization=Parametrization.EB)


# --------------------------------------------------------------------------
# inner products (util.py:90-101; fields.py:136-154)
# --------------------------------------------------------------------------
def _primal_dot_raw(a: torch.Tensor, b: torch.Tensor, nside: int) -> torch.Tensor:
    w = fourier_weights_for(nside, a.device)
    prod = torch.conj(a.to(torch.complex128)) * b.to(torch.complex128)
    return torch.sum(prod.real * w, dim=(-2, -1)) / float(nside * nside)


class _DotFn(torch.autograd.Function):
    """dot with Riesz-representer gradients: d dot(a, b) = dot(b, da) + dot(a, db)."""

    @staticmethod
    def forward(ctx, a, b, nside):
        ctx.save_for_backward(a, b)
        return _primal_dot_raw(a, b, nside)

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        g = g[..., None, None]
        ga = (g * b).to(a.dtype).sum_to_size(a.shape) if ctx.needs_input_grad[0] else None
        gb = (g * a).to(b.dtype).sum_to_size(b.shape) if ctx.needs_input_grad[1] else None
        return ga, gb, None


def primal_dot(a: torch.Tensor, b: torch.Tensor, nside: int) -> torch.Tensor:
    """Re sum conj(a) b w / N^2 over the last two axes, accumulated in float64."""
    if _needs_grad(a, b):
        return _DotFn.apply(a, b, nside)
    return _primal_dot_raw(a, b, nside)


def dot(field_1: Field, field_2: Field) -> torch.Tensor:
    total = 0
    for n in field_1._matrix_names():
        total = total + primal_dot(getattr(field_1, n), getattr(field_2, n), field_1.nside)
    return total


def undo_inner_product(field: Field) -> Field:
    """fields.py:146-154: MAP matrices -> irfft2(conj(rfft2(m)/w) N^2)."""
    def undo(m):
        if m.shape[-1] != m.shape[-2]:
            return m
        N = field.nside
        F = _rfft2(m, field.theta_pix)
        w = fourier_weights_for(N, m.device, _real_dtype(F))
        return _irfft2(torch.conj(F / w) * (N * N), field.theta_pix)
    return field.replace(**{n: undo(getattr(field, n)) for n in field._matrix_names()})
