"""Diagonal field operators, mirroring matrix_operators.py of the reference.

DiagonalScalar (T), DiagonalEB (EE, BB blocks) and BlockTEB (TT, TE, ET, EE,
BB) hold real per-mode arrays [N, N//2+1] on the device and multiply fields
mode by mode (matrix_operators.py:45-148).  pinv is the nonzero-safe
reciprocal (Schur complement for TEB, util.py:45-69) and log_det the weighted
log-determinant (util.py:39-53).  This operator algebra runs once per solve on
small arrays; the per-iteration work of the Wiener filter is in libcmbl.so.

PixelMask is an addition: a real-space mask (configurations with point-source
holes or an apodised square) that the reference does not have -- its mask is
Fourier-diagonal.  gradf_logpdf / wiener_filter accept either.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass
from typing import Optional

import torch

from .constants import ARCMIN_PER_DEGREE, THETA_PIX_DEFAULT
from .fields import FlatS0, FlatS02, FlatS2, fourier_weights_for


def _like(m: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """Operator plane m in the precision of the field matrix x it multiplies
    (float32 for float32 / complex64 fields): an operator applied to a field
    keeps the field's dtype, so a float32 field stays float32."""
    rdt = {torch.complex64: torch.float32, torch.complex128: torch.float64}.get(x.dtype, x.dtype)
    return m if m.dtype == rdt or not m.is_floating_point() else m.to(rdt)


def reciprocal_matrix(m: torch.Tensor) -> torch.Tensor:
    """util.py:45-47"""
    return torch.where(m != 0, 1.0 / torch.where(m != 0, m, torch.ones_like(m)),
                       torch.zeros_like(m))


def primal_log_det(m: torch.Tensor, weights: torch.Tensor):
    """util.py:39-42 (weights broadcast along the half axis)."""
    nz = m != 0
    safe = torch.where(nz, m, torch.ones_like(m))
    val = torch.sum(torch.where(nz, torch.log(torch.abs(safe)), torch.zeros_like(safe)) * weights)
    sign = torch.prod(torch.sign(safe))
    return sign * val


@dataclass
class FieldOperator:
    fourier_weights: Optional[torch.Tensor] = None
    nside: Optional[int] = None
    theta_pix: float = THETA_PIX_DEFAULT
    pix_width: Optional[float] = None

    def _matrix_names(self):
        raise NotImplementedError

    def __post_init__(self):
        m = getattr(self, self._matrix_names()[0])
        if m is not None and self.nside is None:
            self.nside = int(m.shape[-2])
        if self.pix_width is None:
            self.pix_width = math.radians(self.theta_pix / ARCMIN_PER_DEGREE)

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)

    def _binop(self, other, fn):
        if isinstance(other, self.__class__):
            up = {n: fn(getattr(self, n), getattr(other, n)) for n in self._matrix_names()}
        else:
            up = {n: fn(getattr(self, n), other) for n in self._matrix_names()}
        return self.replace(**up)

    def __add__(self, o):
        return self._binop(o, lambda a, b: a + b)

    def __sub__(self, o):
        return self._binop(o, lambda a, b: a - b)

    def __rsub__(self, o):
        return self._binop(o, lambda a, b: b - a)

    def weights(self, device):
        return fourier_weights_for(self.nside, device)


@dataclass
class DiagonalScalar(FieldOperator):
    """matrix_operators.py:45-63"""
    scalar_matrix: Optional[torch.Tensor] = None

    def _matrix_names(self):
        return ["scalar_matrix"]

    def __mul__(self, other):
        if isinstance(other, DiagonalScalar):
            return self.replace(scalar_matrix=self.scalar_matrix * other.scalar_matrix)
        if isinstance(other, FlatS0):
            x = other.scalar_matrix
            return other.replace(scalar_matrix=_like(self.scalar_matrix, x) * x)
        return self.replace(scalar_matrix=self.scalar_matrix * other)

    def __rmul__(self, other):
        return self.__mul__(other)

    def per_mode(self) -> torch.Tensor:
        """[1, N, N//2+1] stack used by the native Wiener filter."""
        return self.scalar_matrix.reshape(1, *self.scalar_matrix.shape[-2:])


@dataclass
class DiagonalEB(FieldOperator):
    """matrix_operators.py:65-90: [EE 0; 0 BB]"""
    matrix_EE: Optional[torch.Tensor] = None
    matrix_BB: Optional[torch.Tensor] = None

    def _matrix_names(self):
        return ["matrix_EE", "matrix_BB"]

    def __mul__(self, other):
        if isinstance(other, DiagonalEB):
            return self.replace(matrix_EE=self.matrix_EE * other.matrix_EE,
                                matrix_BB=self.matrix_BB * other.matrix_BB)
        if isinstance(other, FlatS2):
            q, u = other.polar_matrix_1, other.polar_matrix_2
            return other.replace(polar_matrix_1=_like(self.matrix_EE, q) * q,
                                 polar_matrix_2=_like(self.matrix_BB, u) * u)
        return self.replace(matrix_EE=self.matrix_EE * other, matrix_BB=self.matrix_BB * other)

    def __rmul__(self, other):
        return self.__mul__(other)

    def per_mode(self) -> torch.Tensor:
        return torch.stack([self.matrix_EE, self.matrix_BB])


@dataclass
class BlockTEB(FieldOperator):
    """matrix_operators.py:92-148: [TT TE 0; ET EE 0; 0 0 BB]"""
    matrix_TT: Optional[torch.Tensor] = None
    matrix_TE: Optional[torch.Tensor] = None
    matrix_ET: Optional[torch.Tensor] = None
    matrix_EE: Optional[torch.Tensor] = None
    matrix_BB: Optional[torch.Tensor] = None

    def _matrix_names(self):
        return ["matrix_TT", "matrix_TE", "matrix_ET", "matrix_EE", "matrix_BB"]

    def __mul__(self, o):
        if isinstance(o, BlockTEB):
            return self.replace(matrix_TT=self.matrix_TT * o.matrix_TT + self.matrix_TE * o.matrix_ET,
                                matrix_ET=self.matrix_ET * o.matrix_TT + self.matrix_EE * o.matrix_ET,
                                matrix_TE=self.matrix_TT * o.matrix_TE + self.matrix_TE * o.matrix_EE,
                                matrix_EE=self.matrix_ET * o.matrix_TE + self.matrix_EE * o.matrix_EE,
                                matrix_BB=self.matrix_BB * o.matrix_BB)
        if isinstance(o, FlatS02):
            t, e, b = o.scalar_matrix, o.polar_matrix_1, o.polar_matrix_2
            L = lambda m: _like(m, t)
            return o.replace(
                scalar_matrix=L(self.matrix_TT) * t + L(self.matrix_TE) * e,
                polar_matrix_1=L(self.matrix_ET) * t + L(self.matrix_EE) * e,
                polar_matrix_2=L(self.matrix_BB) * b)
        return self.replace(**{n: getattr(self, n) * o for n in self._matrix_names()})

    def __rmul__(self, o):
        if isinstance(o, BlockTEB):
            return o.__mul__(self)
        return self.replace(**{n: getattr(self, n) * o for n in self._matrix_names()})

    def per_mode(self) -> torch.Tensor:
        """[5, N, N//2+1] planes TT, TE, ET, EE, BB used by the native Wiener
        filter."""
        return torch.stack([self.matrix_TT, self.matrix_TE, self.matrix_ET, self.matrix_EE,
                            self.matrix_BB])


@dataclass
class PixelMask:
    """Real-space mask applied to T or Q/U maps (an addition to the reference,
    whose mask is Fourier-diagonal).  ``matrix`` is a real [N, N] tensor."""
    matrix: torch.Tensor = None

    def mean_square(self) -> float:
        return float(torch.mean(self.matrix.double() ** 2))


def pinv(op):
    """matrix_operators.py:203-234"""
    if isinstance(op, DiagonalScalar):
        return op.replace(scalar_matrix=reciprocal_matrix(op.scalar_matrix))
    if isinstance(op, DiagonalEB):
        return op.replace(matrix_EE=reciprocal_matrix(op.matrix_EE),
                          matrix_BB=reciprocal_matrix(op.matrix_BB))
    if isinstance(op, BlockTEB):
        a, b, c, d = op.matrix_TT, op.matrix_TE, op.matrix_ET, op.matrix_EE
        rd = reciprocal_matrix(d)
        s = reciprocal_matrix(a - b * rd * c)
        return op.replace(matrix_TT=s, matrix_TE=-s * b * rd, matrix_ET=-rd * c * s,
                          matrix_EE=rd + rd * c * s * b * rd,
                          matrix_BB=reciprocal_matrix(op.matrix_BB))
    raise TypeError(f"Unsupported type: {type(op).__name__}")


def log_det(op):
    """matrix_operators.py:173-200"""
    if isinstance(op, DiagonalScalar):
        return primal_log_det(op.scalar_matrix, op.weights(op.scalar_matrix.device))
    if isinstance(op, DiagonalEB):
        w = op.weights(op.matrix_EE.device)
        return primal_log_det(op.matrix_EE, w) + primal_log_det(op.matrix_BB, w)
    if isinstance(op, BlockTEB):
        w = op.weights(op.matrix_TT.device)
        a, b, c, d = op.matrix_TT, op.matrix_TE, op.matrix_ET, op.matrix_EE
        return (primal_log_det(a, w) + primal_log_det(d - c * reciprocal_matrix(a) * b, w)
                + primal_log_det(op.matrix_BB, w))
    raise TypeError(f"Unsupported type: {type(op).__name__}")


def get_identity_like(op):
    """matrix_operators.py:151-155"""
    return op.replace(**{n: torch.ones_like(getattr(op, n)) for n in op._matrix_names()})


def get_scalar_identity(op) -> DiagonalScalar:
    """matrix_operators.py:158-170: a DiagonalScalar of ones on op's grid.
    The reference's ones are complex128; a real identity multiplies alike."""
    first = getattr(op, op._matrix_names()[0])
    N = op.nside
    return DiagonalScalar(fourier_weights=fourier_weights_for(N, first.device), nside=N,
                          theta_pix=op.theta_pix, pix_width=op.pix_width,
                          scalar_matrix=torch.ones((N, N // 2 + 1), dtype=torch.float64,
                                                   device=first.device))
