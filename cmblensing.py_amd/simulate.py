"""Simulated data sets on the GPU: simulate.py of the reference without CAMB.

The reference's load_sim (simulate.py:617-743) runs CAMB, turns its spectra
into Fourier-diagonal covariances, draws phi, the unlensed T/E/B fields and
white noise, lenses the fields with 10 RK4 steps and forms the data
d = M B L(phi) f + n, plus the mixing matrices and the quadratic-estimate
noise.  Here the spectra come in as the dict ``extract_all_cls`` returns
(the same keys as simulate.py:442-451), so CAMB stays outside; everything
else follows the reference:

  * covariances, mask, beam, noise spectra and the D/G mixing matrices are
    built once per configuration as float64 device arrays (torch);
  * the random draws run in libcmbl.so (``cmbl_grf_fourier``): one launch per
    field kind over all trials, Fourier coefficients straight from a
    counter-based generator, with the Hermitian projection of the
    reference's irfft2 -> rfft2 round trip applied per mode;
  * the fields are lensed in one batched LenseFlow (T and Q/U of every trial,
    each trial with its own phi) and the data model M B L f + n is one more
    launch of the same kernel, the noise drawn in place;
  * the quadratic-estimate normalisation runs its real-space products through
    the library's batched FFTs.

Random streams: trial t of master seed s draws kind k (0 phi, 1 T, 2 E, 3 B,
4..6 noise T, E, B) from Philox4x64-10 with key {s, 0} and stream
(k << 32) + t.  The reference's jax.random streams are not reproduced (the
draws have the same distribution and per-mode construction).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass
from itertools import combinations
from typing import Optional

import numpy as np
import torch

from . import _lib
from .constants import ARCMIN_PER_DEGREE, Basis, Parametrization
from .fields import (FlatS0, FlatS02, FlatS2, _irfft2, _rfft2, fourier_weights_for, primal_eb2qu,
                     primal_qu2eb)
from .lense_flow import lense_flow_matrices
from .matrix_operators import BlockTEB, DiagonalEB, DiagonalScalar, reciprocal_matrix

BEAM_TRANSFER_SCALAR = 8 * math.log(2)    # constants.py:12
NPHI_FAC = 2                              # constants.py:5
KIND_PHI, KIND_T, KIND_E, KIND_B, KIND_NT, KIND_NE, KIND_NB = range(7)
_F64 = torch.float64


def _dev(device):
    return torch.device(device) if device is not None else torch.device("cuda")


def _t(x, device):
    return torch.as_tensor(x, dtype=_F64, device=device)


# --------------------------------------------------------------------------
# grids, interpolation, spectra
# --------------------------------------------------------------------------
def gen_ell_grid(nside: int, theta_pix: float, device=None):
    """util.py:163-181: |l| on the [N, N//2+1] half plane, and the pixel width."""
    d = math.radians(theta_pix / ARCMIN_PER_DEGREE)
    lx = 2 * np.pi * np.fft.rfftfreq(nside, d)
    ly = 2 * np.pi * np.fft.fftfreq(nside, d)
    lx, ly = np.meshgrid(lx, ly)
    return _t(np.sqrt(lx ** 2 + ly ** 2), _dev(device)), d


def interp(x, xp, fp, left=None, right=None):
    """jnp.interp on tensors, including left/right = "extrapolate" (the end
    segments continued linearly)."""
    i = torch.clamp(torch.searchsorted(xp, x, right=True), 1, xp.numel() - 1)
    df = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = x - xp[i - 1]
    f = torch.where(dx == 0, fp[i], fp[i - 1] + (delta / torch.where(dx == 0, 1.0, dx)) * df)
    if not (isinstance(left, str) and left == "extrapolate"):
        f = torch.where(x < xp[0], fp[0] if left is None else torch.as_tensor(left, dtype=f.dtype,
                                                                              device=f.device), f)
    if not (isinstance(right, str) and right == "extrapolate"):
        f = torch.where(x > xp[-1], fp[-1] if right is None else torch.as_tensor(right, dtype=f.dtype,
                                                                                 device=f.device), f)
    return f


def dl2cl(dl_xx, lmax: int, lmax_prime: int, is_phi: bool = False, device=None):
    """simulate.py:21-45: D_l (CAMB layout, from l = 0) -> C_l on l = 2..lmax-1,
    log-log interpolated (linear with zero fill if any C_l <= 0)."""
    dev = _dev(device)
    ell = torch.arange(2, lmax, dtype=_F64, device=dev)
    ellp = torch.arange(2, lmax_prime, dtype=_F64, device=dev)
    dl = _t(dl_xx, dev)
    cl = dl[2:] * 2 * math.pi / ellp ** 4 if is_phi else dl[2:] * 2 * math.pi / (ellp * (ellp + 1))
    if bool(torch.all(cl > 0)):
        return torch.exp(interp(torch.log(ell), torch.log(ellp), torch.log(cl), "extrapolate",
                                "extrapolate"))
    return interp(ell, ellp, cl, 0.0, 0.0)


CAMB_COLS = {"TT": 0, "EE": 1, "BB": 2, "TE": 3}


def extract_all_cls(unlensed_scalar, tensor, total, lens_potential, lmax: int, lmax_prime: int,
                    device=None):
    """simulate.py:442-451 on CAMB-layout arrays ([lmax_prime, 4] D_l in muK^2
    with columns TT, EE, BB, TE; lens_potential [lmax_prime])."""
    cls = {}
    for st, col in CAMB_COLS.items():
        cls[f"scalar_{st}"] = dl2cl(np.asarray(unlensed_scalar)[:, col], lmax, lmax_prime,
                                    device=device)
        cls[f"tensor_{st}"] = dl2cl(np.asarray(tensor)[:, col], lmax, lmax_prime, device=device)
    for st in ("TT", "TE", "EE", "BB"):
        cls[f"total_{st}"] = dl2cl(np.asarray(total)[:, CAMB_COLS[st]], lmax, lmax_prime,
                                   device=device)
    cls["phi"] = dl2cl(lens_potential, lmax, lmax_prime, is_phi=True, device=device)
    return cls


def beam_cls(beam_fwhm: float, ell: torch.Tensor) -> torch.Tensor:
    """simulate.py:48-49 (the squared beam transfer b_l^2)."""
    sig2 = math.radians(beam_fwhm / ARCMIN_PER_DEGREE) ** 2
    return torch.exp(-ell.to(_F64) ** 2 * sig2 / BEAM_TRANSFER_SCALAR)


def noise_cls(lmax_prime: int, uk_arcmin_t: float, beam_fwhm: float = 0, l_knee: float = 100,
              alpha_knee: float = 3, device=None):
    """simulate.py:52-65: (TT, TE, EE, BB) on l = 2..lmax_prime-1; P = 2 T."""
    ellp = torch.arange(2, lmax_prime, dtype=_F64, device=_dev(device))
    bls = beam_cls(beam_fwhm, ellp)
    nls = 1 + (l_knee / ellp) ** alpha_knee
    tt = torch.nan_to_num(math.radians(uk_arcmin_t / ARCMIN_PER_DEGREE) ** 2 * nls / bls,
                          nan=0.0, posinf=0.0, neginf=0.0)
    return tt, torch.zeros_like(tt), 2 * tt, 2 * tt


def covar_matrix_from_cls(nside: int, pix_width: float, ell_grid: torch.Tensor, ells, cls,
                          origin_value=None, rescale: bool = True,
                          use_linear_interpolation: bool = False, fill_value=None):
    """simulate.py:70-97: C_l interpolated onto the |l| grid (log-log when every
    C_l > 0 and not forced linear), (0,0) set to origin_value, / pix_width^2."""
    dev = ell_grid.device
    ells, cls = _t(ells, dev), _t(cls, dev)
    lg = ell_grid.flatten()
    if bool(torch.all(cls > 0)) and not use_linear_interpolation:
        ls = torch.where(lg == 0, torch.full_like(lg, 1e-10), lg)
        res = torch.exp(interp(torch.log(ls), torch.log(ells), torch.log(cls), "extrapolate",
                               "extrapolate"))
    else:
        fv = fill_value if fill_value is not None else 0.0
        res = interp(lg, ells, cls, fv, fv)
    res = res.reshape(nside, nside // 2 + 1).clone()
    if origin_value is not None:
        res[0, 0] = origin_value
    if rescale:
        res = res / pix_width ** 2
    return res


def cos_ramp_up(length: int, device=None) -> torch.Tensor:
    """simulate.py:117-118"""
    return (torch.cos(torch.linspace(math.pi, 0, length, dtype=_F64, device=_dev(device))) + 1) / 2


def cos_ramp_down(length: int, device=None) -> torch.Tensor:
    """simulate.py:120-121"""
    return 1 - cos_ramp_up(length, device)


def low_pass(l_cutoff: int, delta_l: int = 50, device=None) -> torch.Tensor:
    """simulate.py:123-124"""
    dev = _dev(device)
    return torch.cat([torch.ones(l_cutoff - delta_l + 1, dtype=_F64, device=dev),
                      cos_ramp_down(delta_l, dev)])


def get_mask(l_cutoff: int, nside: int, pix_width: float, ell_grid: torch.Tensor):
    """simulate.py:126-130: the Fourier-diagonal low-pass 'mask'."""
    screen = low_pass(l_cutoff, device=ell_grid.device)
    ell = torch.arange(2, screen.numel(), dtype=_F64, device=ell_grid.device)
    return covar_matrix_from_cls(nside, pix_width, ell_grid, ell, screen[2:], origin_value=1,
                                 rescale=False)


def get_beam(nside: int, pix_width: float, ell_grid: torch.Tensor, lmax_prime: int,
             beam_fwhm: float = 0):
    """simulate.py:133-138: b_l linearly interpolated (1 beyond the table, 0 at l = 0)."""
    ellp = torch.arange(2, lmax_prime, dtype=_F64, device=ell_grid.device)
    b = torch.sqrt(beam_cls(beam_fwhm, ellp))
    return covar_matrix_from_cls(nside, pix_width, ell_grid, ellp, b, origin_value=0,
                                 rescale=False, use_linear_interpolation=True, fill_value=1.0)


# --------------------------------------------------------------------------
# block algebra (util.py:10-69) and mixing matrices (simulate.py:142-192)
# --------------------------------------------------------------------------
def teb_matrix_mult(tt1, te1, et1, ee1, bb1, tt2, te2, et2, ee2, bb2):
    """util.py:10-20: [[TT, TE], [ET, EE]] (+) BB block products per mode."""
    return (tt1 * tt2 + te1 * et2, tt1 * te2 + te1 * ee2, et1 * tt2 + ee1 * et2,
            et1 * te2 + ee1 * ee2, bb1 * bb2)


def block_matrix_sqrt(a, b, c, d):
    """util.py:25-35: principal square root of [[a, b], [c, d]] per mode."""
    s = torch.sqrt(a * d - b * c)
    t = reciprocal_matrix(torch.sqrt(a + (d + 2 * s)))
    return t * (a + s), t * b, t * c, t * (d + s)


def invert_block_matrix(a, b, c, d):
    """util.py:59-69 (Schur complement)."""
    rd = reciprocal_matrix(d)
    s = reciprocal_matrix(a - b * rd * c)
    return s, -s * b * rd, -rd * c * s, rd + rd * c * s * b * rd


def get_d_matrix(cf_tt, cf_te, cf_ee, cf_bb, cn_tt, cn_te, cn_ee, cn_bb):
    """simulate.py:142-166: D = sqrt((Cf + (5')^2 + 2 Cn) Cf^-1) per block."""
    pre = math.radians(5 / ARCMIN_PER_DEGREE) ** 2
    inv = invert_block_matrix(cf_tt, cf_te, cf_te, cf_ee)
    itt, ite, iet, iee, ibb = [torch.nan_to_num(x, nan=0.0, posinf=0.0, neginf=0.0)
                               for x in (*inv, reciprocal_matrix(cf_bb))]
    s_tt = cf_tt + pre + 2 * cn_tt
    s_te = cf_te + 2 * cn_te
    s_ee = cf_ee + pre + 2 * cn_ee
    s_bb = cf_bb + pre + 2 * cn_bb
    d_tt, d_te, d_et, d_ee, d_bb = teb_matrix_mult(s_tt, s_te, s_te, s_ee, s_bb,
                                                   itt, ite, iet, iee, ibb)
    # the reference roots [[TT, ET], [ET, EE]] (ET on both off-diagonals)
    d_tt, d_te, _, d_ee = block_matrix_sqrt(d_tt, d_et, d_et, d_ee)
    return d_tt, d_te, d_ee, torch.sqrt(d_bb)


def get_g_matrix(cphi_fid, nphi, a_phi_fid, a_phi=1):
    """simulate.py:180-185"""
    g0 = torch.sqrt(1 + 2 * nphi * reciprocal_matrix(cphi_fid))
    g = torch.sqrt(1 + 2 * nphi * reciprocal_matrix((a_phi / a_phi_fid) * cphi_fid))
    return reciprocal_matrix(g0) * g


# --------------------------------------------------------------------------
# quadratic-estimate normalisation (simulate.py:196-376)
# --------------------------------------------------------------------------
def get_indices(length: int):
    """simulate.py:198-213: every 0/1 tuple of `length`, grouped by the number
    of zeros, the zero positions in itertools.combinations order."""
    rows = []
    for n0 in range(length + 1):
        for zeros in combinations(range(length), n0):
            p = [1] * length
            for i in zeros:
                p[i] = 0
            rows.append(tuple(p))
    return rows


def levi_civita_3d():
    """simulate.py:263-266"""
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    return (i - j) * (j - k) * (k - i) / 2


def _kij(N: int, pix_width: float, device):
    kx = 2 * np.pi * np.fft.fftfreq(N, pix_width)
    ky = 2 * np.pi * np.fft.rfftfreq(N, pix_width)
    KX, KY = np.meshgrid(kx, ky, indexing="ij")
    return _t(KX, device), _t(KY, device)


def get_specific_derivative(field, pix_width: float, t1: int, t2: int):
    """simulate.py:216-253 for two derivative indices (0 = x, 1 = y)."""
    KX, KY = _kij(field.shape[-2], pix_width, field.device)
    if t1 == t2:
        return -(KX ** 2 if t1 == 0 else KY ** 2) * field
    return -(KX * KY) * field


def qe_leg(field, pix_width: float, encapsulated, isolated):
    """simulate.py:269-303: (i KX)^nx (i KY)^ny field / |K|^(len isolated),
    nan -> 0; the identity (as complex) without indices."""
    nd = len(isolated)
    sq = list(encapsulated) + list(isolated)
    ny = int(sum(sq))
    nx = len(sq) - ny
    field = torch.nan_to_num(field, nan=0.0)
    if nx == 0 and ny == 0:
        return field.to(torch.complex128)
    KX, KY = _kij(field.shape[-2], pix_width, field.device)
    fac = (1j ** (nx + ny)) * KX ** nx * KY ** ny / torch.sqrt((KX ** 2 + KY ** 2) ** nd)
    return torch.nan_to_num(fac * field, nan=0.0).to(torch.complex128)


def _irfft2_batch(legs, theta_pix):
    return _irfft2(torch.stack(legs).contiguous(), theta_pix)


def scalar_quadratic_estimate(cn_tt, cf_tt, cfl_tt, m_tt, b_tt, pix_width: float):
    """simulate.py:306-316, 350-359: 1 / sum_ij |d_i d_j rfft2(internal_ij)|."""
    theta = math.degrees(pix_width) * ARCMIN_PER_DEGREE
    tf = m_tt * b_tt
    sigma = tf ** 2 * cfl_tt + cn_tt
    a = tf ** 2 * cf_tt ** 2 / sigma
    b = tf ** 2 / sigma
    c = tf ** 2 * cf_tt / sigma
    idx = get_indices(2)
    legs = []
    for i, j in idx:
        legs += [qe_leg(a, pix_width, [i, j], []), qe_leg(b, pix_width, [], []),
                 qe_leg(c, pix_width, [i], []), qe_leg(c, pix_width, [j], [])]
    m = _irfft2_batch(legs, theta).reshape(len(idx), 4, *cf_tt.shape[:-1], cf_tt.shape[-2])
    internal = m[:, 0] * m[:, 1] + m[:, 2] * m[:, 3]
    F = _rfft2(internal.contiguous(), theta)
    qe = torch.zeros(cf_tt.shape, dtype=_F64, device=cf_tt.device)
    for k, (i, j) in enumerate(idx):
        qe = qe + torch.abs(get_specific_derivative(F[k], pix_width, i, j))
    return torch.nan_to_num(1 / qe, nan=0.0, posinf=0.0, neginf=0.0)


def polar_quadratic_estimate(cf_ee, cf_bb, cfl_ee, cfl_bb, cn_ee, cn_bb, mask_ee, mask_bb,
                             beam_ee, beam_bb, pix_width: float):
    """simulate.py:319-347, 362-376 (the terms whose Levi-Civita weight is zero
    add 0 times a finite map and are skipped).  Returns a real array; the
    reference's is complex with zero imaginary part."""
    theta = math.degrees(pix_width) * ARCMIN_PER_DEGREE
    eps = levi_civita_3d()
    tf2e = (mask_ee * beam_ee) ** 2
    tf2b = (mask_bb * beam_bb) ** 2
    se = tf2e * cfl_ee + cn_ee
    sb = tf2b * cfl_bb + cn_bb
    e2, e1, e0 = tf2e * cf_ee ** 2 / se, tf2e * cf_ee / se, tf2e / se
    b2, b1, b0 = tf2b * cf_bb ** 2 / sb, tf2b * cf_bb / sb, tf2b / sb
    pw = pix_width
    combos = [(k, l, m, n, p, q) for k, l, m, n, p, q in get_indices(6)
              if eps[m, p, 2] * eps[n, q, 2] != 0]
    qe = torch.zeros(cf_ee.shape, dtype=_F64, device=cf_ee.device)
    for i, j in get_indices(2):
        legs, wgts = [], []
        for k, l, m, n, p, q in combos:
            A, B = [k, l, m, n], [k, l, p, q]
            legs += [qe_leg(e2, pw, [i, j], A), qe_leg(b0, pw, [], B),
                     qe_leg(e1, pw, [i], A), qe_leg(b1, pw, [j], B),
                     qe_leg(e0, pw, [], A), qe_leg(b2, pw, [i, j], B)]
            wgts.append(4 * eps[m, p, 2] * eps[n, q, 2])
        mm = _irfft2_batch(legs, theta).reshape(len(combos), 6, *cf_ee.shape[:-1],
                                                cf_ee.shape[-2])
        terms = mm[:, 0] * mm[:, 1] - 2 * mm[:, 2] * mm[:, 3] + mm[:, 4] * mm[:, 5]
        w = torch.as_tensor(wgts, dtype=_F64, device=cf_ee.device).reshape(-1, 1, 1)
        internal = torch.sum(w * terms, dim=0)
        qe = qe + torch.abs(get_specific_derivative(_rfft2(internal.contiguous(), theta), pw, i, j))
    return torch.nan_to_num(1 / qe, nan=0.0, posinf=0.0, neginf=0.0)


# --------------------------------------------------------------------------
# random fields (libcmbl.so)
# --------------------------------------------------------------------------
def grf_fourier(covs, ntrials: int, seed: int, stream0: int, theta_pix: float,
                dtype=torch.float64, mul=None, add=None) -> torch.Tensor:
    """[ntrials, C, N, N//2+1] complex Fourier coefficients of Gaussian random
    maps with covariances covs [C, N, N//2+1] (float64): rfft2 of the
    reference's field_from_covar map (simulate.py:100-113) for each trial and
    component, drawn on the device.  With `add` ([ntrials, C, N, N//2+1], the
    output dtype) the result is mul * add + draw (the data model; mul
    [C, N, N//2+1] float64 or None for 1)."""
    covs = covs.to(_F64).contiguous()
    C, N = covs.shape[0], covs.shape[1]
    if covs.shape != (C, N, N // 2 + 1) or not 1 <= C <= 3:
        raise ValueError(f"covariances must be [C(1..3), N, N//2+1], got {tuple(covs.shape)}")
    if not covs.is_cuda:
        raise _lib.CmblError("the random draws run on the GPU: move the covariances to a HIP device")
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    out = torch.empty((ntrials, C, N, N // 2 + 1), dtype=cdt, device=covs.device)
    if mul is not None:
        mul = mul.to(_F64).contiguous()
        if mul.shape != covs.shape:
            raise ValueError("mul must have the covariances' shape")
    if add is not None:
        add = add.to(cdt).contiguous()
        if add.shape != out.shape:
            raise ValueError(f"add must be {tuple(out.shape)}")
    plan = _lib.get_plan(N, theta_pix, dtype, 1, covs.device)
    _lib.grf_fourier(plan, covs, mul, None if add is None else torch.view_as_real(add),
                     torch.view_as_real(out), ntrials, C, seed, stream0)
    return out


def field_from_covar_single_key(nside: int, covar_matrix, seed: int, stream: int = 0,
                                theta_pix: float = 2.5):
    """simulate.py:106-113: one Gaussian random MAP [N, N] (irfft2(F, "ortho"))."""
    F = grf_fourier(covar_matrix[None], 1, seed, stream, theta_pix)[0, 0]
    return _irfft2(F.contiguous(), theta_pix)


def lens_fields(fields: torch.Tensor, phi_f: torch.Tensor, nside: int, theta_pix: float,
                n: int = 10) -> torch.Tensor:
    """simulate.py:455-469 on FOURIER [B, C, N, N//2+1] fields (C = 1: T; 2: E, B;
    3: T, E, B) with a FOURIER phi ([N, N//2+1] or one per trial): E/B go
    through Q/U maps, everything is lensed in one batched LenseFlow."""
    B, C = fields.shape[:2]
    X = fields
    if C >= 2:
        q, u = primal_eb2qu(fields[:, -2], fields[:, -1], nside, theta_pix)
        X = torch.stack(([fields[:, 0]] if C == 3 else []) + [q, u], dim=1)
    x = _irfft2(X.contiguous(), theta_pix)
    y = lense_flow_matrices(x, phi_f, theta_pix, n, 1, False)
    Y = _rfft2(y, theta_pix)
    if C >= 2:
        e, b = primal_qu2eb(Y[:, -2], Y[:, -1], nside, theta_pix)
        Y = torch.stack(([Y[:, 0]] if C == 3 else []) + [e, b], dim=1)
    return Y


# --------------------------------------------------------------------------
# data sets (dataset.py)
# --------------------------------------------------------------------------
@dataclass
class DataSet:
    """dataset.py:15-22"""
    fid_r: float = 0.2
    fid_a_phi: float = 1.0
    fourier_weights: Optional[torch.Tensor] = None
    nside: int = 256
    theta_pix: float = 2.5
    pix_width: float = 0.0

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)


@dataclass
class DataSetT(DataSet):
    """dataset.py:24-45 (DiagonalScalar operators, FlatS0 fields)."""
    noise_covariance: Optional[DiagonalScalar] = None
    mixing_d: Optional[DiagonalScalar] = None
    mixing_g: Optional[DiagonalScalar] = None
    field_covariance: Optional[DiagonalScalar] = None
    lensed_field_covariance: Optional[DiagonalScalar] = None
    scalar_field_covariance: Optional[DiagonalScalar] = None
    tensor_field_covariance: Optional[DiagonalScalar] = None
    phi_covariance: Optional[DiagonalScalar] = None
    unscaled_phi_covariance: Optional[DiagonalScalar] = None
    mask: Optional[DiagonalScalar] = None
    beam: Optional[DiagonalScalar] = None
    quadratic_estimate: Optional[DiagonalScalar] = None
    data: Optional[FlatS0] = None
    unlensed_field: Optional[FlatS0] = None
    lensed_field: Optional[FlatS0] = None
    phi: Optional[FlatS0] = None


@dataclass
class DataSetEB(DataSet):
    """dataset.py:47-68 (DiagonalEB operators, FlatS2 fields)."""
    noise_covariance: Optional[DiagonalEB] = None
    mixing_d: Optional[DiagonalEB] = None
    mixing_g: Optional[DiagonalScalar] = None
    field_covariance: Optional[DiagonalEB] = None
    lensed_field_covariance: Optional[DiagonalEB] = None
    scalar_field_covariance: Optional[DiagonalEB] = None
    tensor_field_covariance: Optional[DiagonalEB] = None
    phi_covariance: Optional[DiagonalScalar] = None
    unscaled_phi_covariance: Optional[DiagonalScalar] = None
    mask: Optional[DiagonalEB] = None
    beam: Optional[DiagonalEB] = None
    quadratic_estimate: Optional[DiagonalScalar] = None
    data: Optional[FlatS2] = None
    unlensed_field: Optional[FlatS2] = None
    lensed_field: Optional[FlatS2] = None
    phi: Optional[FlatS0] = None


@dataclass
class DataSetTEB(DataSet):
    """dataset.py:70-91 (BlockTEB operators, FlatS02 fields)."""
    noise_covariance: Optional[BlockTEB] = None
    mixing_d: Optional[BlockTEB] = None
    mixing_g: Optional[DiagonalScalar] = None
    field_covariance: Optional[BlockTEB] = None
    lensed_field_covariance: Optional[BlockTEB] = None
    scalar_field_covariance: Optional[BlockTEB] = None
    tensor_field_covariance: Optional[BlockTEB] = None
    phi_covariance: Optional[DiagonalScalar] = None
    unscaled_phi_covariance: Optional[DiagonalScalar] = None
    mask: Optional[BlockTEB] = None
    beam: Optional[BlockTEB] = None
    quadratic_estimate: Optional[DiagonalScalar] = None
    data: Optional[FlatS02] = None
    unlensed_field: Optional[FlatS02] = None
    lensed_field: Optional[FlatS02] = None
    phi: Optional[FlatS0] = None


@dataclass
class SimConfig:
    """The per-configuration float64 arrays of a simulation (everything that
    does not depend on the seed)."""
    nside: int
    theta_pix: float
    pix_width: float
    pol: str
    cphi: torch.Tensor
    unscaled_cphi: torch.Tensor
    cf: dict
    cf_scalar: dict
    cf_tensor: dict
    cfl: dict
    cn: dict
    mask: torch.Tensor
    beam: torch.Tensor
    d: tuple
    qe: Optional[torch.Tensor]
    g: Optional[torch.Tensor]
    r: float
    a_phi: float


def sim_config(nside: int, theta_pix: float, pol: str, cls: dict, lmax: int = 17_000,
               uk_arcmin_t: float = 3, r: float = 0.2, nphi_fac: float = NPHI_FAC,
               a_phi: float = 1, with_qe: bool = True, device=None) -> SimConfig:
    """The seed-independent part of load_sim (simulate.py:625-727): phi and
    field covariances, lensed covariances, mask, beam, noise covariances, the
    D matrix and (with_qe) the quadratic-estimate noise and G matrix."""
    if pol not in ("I", "P", "IP"):
        raise ValueError("pol must be 'I', 'P' or 'IP'")
    dev = _dev(device)
    lmax_prime = min(lmax, 5000)
    ell_grid, pw = gen_ell_grid(nside, theta_pix, dev)
    ells = torch.arange(2, lmax, dtype=_F64, device=dev)
    cl = {k: _t(v, dev) for k, v in cls.items()}
    unscaled = covar_matrix_from_cls(nside, pw, ell_grid, ells, cl["phi"], origin_value=0)
    cphi = a_phi * unscaled
    comps = ("TT", "TE", "EE", "BB")
    cfs = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cl[f"scalar_{c}"], origin_value=0)
           for c in comps}
    cft = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cl[f"tensor_{c}"], origin_value=0)
           for c in comps}
    cf = {c: cfs[c] + cft[c] for c in comps}
    cfl = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cl[f"total_{c}"], origin_value=0)
           for c in comps}
    mask = get_mask(3000, nside, pw, ell_grid)
    beam = get_beam(nside, pw, ell_grid, lmax_prime)
    ellp = torch.arange(2, lmax_prime, dtype=_F64, device=dev)
    cn = {c: covar_matrix_from_cls(nside, pw, ell_grid, ellp, v, origin_value=0)
          for c, v in zip(comps, noise_cls(lmax_prime, uk_arcmin_t, device=dev))}
    d = get_d_matrix(cf["TT"], cf["TE"], cf["EE"], cf["BB"], cn["TT"], cn["TE"], cn["EE"],
                     cn["BB"])
    qe = g = None
    if with_qe:
        if pol == "I":
            qe = scalar_quadratic_estimate(cn["TT"], cf["TT"], cfl["TT"], mask, beam, pw)
        else:
            qe = polar_quadratic_estimate(cf["EE"], cf["BB"], cfl["EE"], cfl["BB"], cn["EE"],
                                          cn["BB"], mask, mask, beam, beam, pw)
        qe = qe / nphi_fac
        g = get_g_matrix(cphi, qe, a_phi, a_phi=a_phi)
    return SimConfig(nside, theta_pix, pw, pol, cphi, unscaled, cf, cfs, cft, cfl, cn, mask, beam,
                     d, qe, g, r, a_phi)


def _keys(pol):
    return {"I": ("TT",), "P": ("EE", "BB"), "IP": ("TT", "EE", "BB")}[pol]


def simulate(cfg: SimConfig, master_seed: int, ntrials: int = 1, dtype=torch.float64,
             n: int = 10):
    """Draw, lens and observe `ntrials` trials (batch axis first): returns the
    FOURIER arrays (phi [B, N, N//2+1], field, lensed, data [B, C, N, N//2+1])
    in `dtype`'s complex type.  phi is drawn in float64 and lensing runs in
    `dtype` (float32 or float64)."""
    keys = _keys(cfg.pol)
    th = cfg.theta_pix
    k0 = KIND_E if cfg.pol == "P" else KIND_T
    phi = grf_fourier(cfg.cphi[None], ntrials, master_seed, KIND_PHI << 32, th)[:, 0]
    field = grf_fourier(torch.stack([cfg.cf[k] for k in keys]), ntrials, master_seed, k0 << 32,
                        th, dtype)
    lensed = lens_fields(field, phi, cfg.nside, th, n)
    mb = (cfg.mask * cfg.beam)[None].expand(len(keys), -1, -1)
    data = grf_fourier(torch.stack([cfg.cn[k] for k in keys]), ntrials, master_seed,
                       (k0 + 3) << 32, th, dtype, mul=mb, add=lensed)
    return phi, field, lensed, data


def _scalar_op(cfg, m, w):
    return DiagonalScalar(fourier_weights=w, nside=cfg.nside, theta_pix=cfg.theta_pix,
                          pix_width=cfg.pix_width, scalar_matrix=m)


def build_dataset(cfg: SimConfig, phi, field, lensed, data):
    """A DataSetT / DataSetEB / DataSetTEB like load_sim's (simulate.py:473-614);
    fields carry the trial axis first when more than one trial was drawn."""
    N, th, pw = cfg.nside, cfg.theta_pix, cfg.pix_width
    w = fourier_weights_for(N, cfg.cphi.device)
    squeeze = phi.shape[0] == 1
    sq = (lambda x: x[0]) if squeeze else (lambda x: x)
    base = dict(fid_r=cfg.r, fid_a_phi=cfg.a_phi, fourier_weights=w, nside=N, theta_pix=th,
                pix_width=pw)
    fkw = dict(fourier_weights=w, nside=N, theta_pix=th, pix_width=pw, basis=Basis.FOURIER)
    phi_f = FlatS0(scalar_matrix=sq(phi), parametrization=Parametrization.T, **fkw)
    qe = _scalar_op(cfg, cfg.qe, w) if cfg.qe is not None else None
    g = _scalar_op(cfg, cfg.g, w) if cfg.g is not None else None
    cphi = _scalar_op(cfg, cfg.cphi, w)
    uc = _scalar_op(cfg, cfg.unscaled_cphi, w)
    d_tt, d_te, d_ee, d_bb = cfg.d
    if cfg.pol == "I":
        op = lambda m: _scalar_op(cfg, m, w)
        fld = lambda x: FlatS0(scalar_matrix=sq(x[:, 0]), parametrization=Parametrization.T, **fkw)
        return DataSetT(
            noise_covariance=op(cfg.cn["TT"]), field_covariance=op(cfg.cf["TT"]),
            lensed_field_covariance=op(cfg.cfl["TT"]),
            scalar_field_covariance=op(cfg.cf_scalar["TT"]),
            tensor_field_covariance=op(cfg.cf_tensor["TT"]), mixing_d=op(d_tt), mixing_g=g,
            phi_covariance=cphi, unscaled_phi_covariance=uc, mask=op(cfg.mask),
            beam=op(cfg.beam), quadratic_estimate=qe, data=fld(data),
            unlensed_field=fld(field), lensed_field=fld(lensed), phi=phi_f, **base)
    okw = dict(fourier_weights=w, nside=N, theta_pix=th, pix_width=pw)
    if cfg.pol == "P":
        op = lambda c: DiagonalEB(matrix_EE=c["EE"], matrix_BB=c["BB"], **okw)
        fld = lambda x: FlatS2(polar_matrix_1=sq(x[:, 0]), polar_matrix_2=sq(x[:, 1]), **fkw)
        return DataSetEB(
            noise_covariance=op(cfg.cn), field_covariance=op(cfg.cf),
            lensed_field_covariance=op(cfg.cfl), scalar_field_covariance=op(cfg.cf_scalar),
            tensor_field_covariance=op(cfg.cf_tensor),
            mixing_d=DiagonalEB(matrix_EE=d_ee, matrix_BB=d_bb, **okw), mixing_g=g,
            phi_covariance=cphi, unscaled_phi_covariance=uc,
            mask=DiagonalEB(matrix_EE=cfg.mask, matrix_BB=cfg.mask, **okw),
            beam=DiagonalEB(matrix_EE=cfg.beam, matrix_BB=cfg.beam, **okw),
            quadratic_estimate=qe, data=fld(data), unlensed_field=fld(field),
            lensed_field=fld(lensed), phi=phi_f, **base)
    zero = torch.zeros_like(cfg.mask)
    op = lambda c: BlockTEB(matrix_TT=c["TT"], matrix_TE=c["TE"], matrix_ET=c["TE"],
                            matrix_EE=c["EE"], matrix_BB=c["BB"], **okw)
    fld = lambda x: FlatS02(scalar_matrix=sq(x[:, 0]), polar_matrix_1=sq(x[:, 1]),
                            polar_matrix_2=sq(x[:, 2]), **fkw)
    return DataSetTEB(
        noise_covariance=op(cfg.cn), field_covariance=op(cfg.cf), lensed_field_covariance=op(cfg.cfl),
        scalar_field_covariance=op(cfg.cf_scalar), tensor_field_covariance=op(cfg.cf_tensor),
        mixing_d=BlockTEB(matrix_TT=d_tt, matrix_TE=d_te, matrix_ET=d_te, matrix_EE=d_ee,
                          matrix_BB=d_bb, **okw),
        mixing_g=g, phi_covariance=cphi, unscaled_phi_covariance=uc,
        mask=BlockTEB(matrix_TT=cfg.mask, matrix_TE=zero, matrix_ET=zero, matrix_EE=cfg.mask,
                      matrix_BB=cfg.mask, **okw),
        beam=BlockTEB(matrix_TT=cfg.beam, matrix_TE=zero, matrix_ET=zero, matrix_EE=cfg.beam,
                      matrix_BB=cfg.beam, **okw),
        quadratic_estimate=qe, data=fld(data), unlensed_field=fld(field), lensed_field=fld(lensed),
        phi=phi_f, **base)


def load_sim_from_cls(nside: int, theta_pix: float, pol: str, master_seed: int, cls: dict,
                      uk_arcmin_t: float = 3, lmax: int = 17_000, r: float = 0.2,
                      nphi_fac: float = NPHI_FAC, a_phi: float = 1, dtype=torch.float64,
                      n: int = 10, device=None):
    """load_sim (simulate.py:617-743) with the CAMB spectra given as `cls`
    (the dict of extract_all_cls): one data set of trial 0 of master_seed."""
    cfg = sim_config(nside, theta_pix, pol, cls, lmax, uk_arcmin_t, r, nphi_fac, a_phi,
                     device=device)
    return build_dataset(cfg, *simulate(cfg, master_seed, 1, dtype, n))


def batch_simulated_trials(num_trials: int, nside: int, theta_pix: float, cls: dict,
                           uk_arcmin_t: float = 10, lmax: int = 17_000, pol: str = "I",
                           master_seed: int = 0, dtype=torch.float64, n: int = 10, device=None):
    """simulate.py:745-752: num_trials data sets in one batch (trial axis first).
    The reference vmaps load_sim over random seeds; here the trials are the
    streams 0..num_trials-1 of one master seed."""
    cfg = sim_config(nside, theta_pix, pol, cls, lmax, uk_arcmin_t, device=device)
    return build_dataset(cfg, *simulate(cfg, master_seed, num_trials, dtype, n))
