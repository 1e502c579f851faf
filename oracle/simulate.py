"""FP64 numpy restatement of CMBLensing.py's simulation code (simulate.py) --
the covariance, instrument, noise, mixing-matrix and quadratic-estimate
construction and the lensed data model -- with the CAMB call left out: the
spectra come in as the dict _extract_all_cls returns (simulate.py:442-451).

TEST INFRASTRUCTURE ONLY -- imported by tests/ (and never by the package).

Random draws: the reference uses jax.random (simulate.py:100-113), which is
not reproducible here.  The library draws from Philox4x64-10 with key
{seed, 0} and counter {mode, stream, 0, 0}; ``philox4x64_10`` restates that
generator (Salmon et al., SC'11) and is pinned to numpy's own Philox bit
generator, and the map of a draw follows the reference's literal route
(irfft2(F, norm="ortho"), then rfft2).
"""
from __future__ import annotations

import math
from itertools import combinations

import numpy as np

from . import lenseflow as lf
from . import wiener as w

ARCMIN_PER_DEGREE = 60
BEAM_TRANSFER_SCALAR = 8 * np.log(2)          # constants.py:12
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_M32 = np.uint64(0xFFFFFFFF)

# stream kinds of one trial (the library's convention; the reference's key
# order is phi, T, E, B, noise T, E, B: simulate.py:648-696)
KIND_PHI, KIND_T, KIND_E, KIND_B, KIND_NT, KIND_NE, KIND_NB = range(7)


# --------------------------------------------------------------------------
# Philox4x64-10 and Box-Muller
# --------------------------------------------------------------------------
def _mulhilo(a, b):
    """64x64 -> (hi, lo) on uint64 arrays through 32-bit halves."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.uint64(b)
    s32 = np.uint64(32)
    al, ah = a & _M32, a >> s32
    bl, bh = b & _M32, b >> s32
    ll, lh, hl, hh = al * bl, al * bh, ah * bl, ah * bh
    mid = (ll >> s32) + (lh & _M32) + (hl & _M32)
    hi = hh + (lh >> s32) + (hl >> s32) + (mid >> s32)
    with np.errstate(over="ignore"):
        lo = a * b
    return hi, lo


def philox4x64_10(c0, c1, c2, c3, k0, k1=0):
    """Philox4x64-10 on uint64 counter arrays (Salmon, Moraes, Dror, Shaw,
    "Parallel random numbers: as easy as 1, 2, 3", SC'11)."""
    M0, M1 = 0xD2E7470EE14C6C93, 0xCA5A826395121157
    W0, W1 = 0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B
    c = [np.asarray(x, dtype=np.uint64) for x in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c = [x.copy() for x in c]
    ka, kb = int(k0) & 0xFFFFFFFFFFFFFFFF, int(k1) & 0xFFFFFFFFFFFFFFFF
    for r in range(10):
        if r:
            ka = (ka + W0) & 0xFFFFFFFFFFFFFFFF
            kb = (kb + W1) & 0xFFFFFFFFFFFFFFFF
        hi0, lo0 = _mulhilo(c[0], M0)
        hi1, lo1 = _mulhilo(c[2], M1)
        c = [hi1 ^ c[1] ^ np.uint64(ka), lo1, hi0 ^ c[3] ^ np.uint64(kb), lo0]
    return c


def normals(N, seed, sid):
    """Standard normal pairs (a, b) [N, N//2+1] of one map: counter {mode, sid}."""
    H1 = N // 2 + 1
    mode = np.arange(N * H1, dtype=np.uint64)
    x = philox4x64_10(mode, np.uint64(sid), 0, 0, seed, 0)
    u = ((x[0] >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    v = (x[1] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u))
    return (r * np.cos(2 * np.pi * v)).reshape(N, H1), (r * np.sin(2 * np.pi * v)).reshape(N, H1)


def grf_fourier(cov, ntrials, seed, stream0):
    """[ntrials, C, N, N//2+1]: rfft2(irfft2(sqrt(cov[c]/2)(a + i b), norm="ortho"))
    per trial and component (simulate.py:100-113, then the rfft2 of
    simulate.py:700-705); stream of (t, c) = stream0 + (c << 32) + t."""
    cov = np.asarray(cov, dtype=np.float64)
    C, N = cov.shape[0], cov.shape[1]
    out = np.empty((ntrials, C, N, N // 2 + 1), dtype=np.complex128)
    for t in range(ntrials):
        for c in range(C):
            a, b = normals(N, seed, stream0 + (c << 32) + t)
            F = np.sqrt(cov[c] / 2) * (a + 1j * b)
            out[t, c] = np.fft.rfft2(np.fft.irfft2(F, norm="ortho"))
    return out


# --------------------------------------------------------------------------
# grids, interpolation, spectra (simulate.py:21-138; util.py:163-181)
# --------------------------------------------------------------------------
def gen_ell_grid(nside, theta_pix):
    """util.py:163-181: |l| on the [N, N//2+1] half plane and the pixel width."""
    d = math.radians(theta_pix / ARCMIN_PER_DEGREE)
    lx = 2 * np.pi * np.fft.rfftfreq(nside, d)
    ly = 2 * np.pi * np.fft.fftfreq(nside, d)
    lx, ly = np.meshgrid(lx, ly)
    return np.sqrt(lx ** 2 + ly ** 2), d


def interp(x, xp, fp, left=None, right=None):
    """jnp.interp semantics, including left/right = "extrapolate" (linear
    continuation of the end segments)."""
    x, xp, fp = (np.asarray(v, dtype=np.float64) for v in (x, xp, fp))
    i = np.clip(np.searchsorted(xp, x, side="right"), 1, len(xp) - 1)
    df = fp[i] - fp[i - 1]
    dx = xp[i] - xp[i - 1]
    delta = x - xp[i - 1]
    f = np.where(dx == 0, fp[i], fp[i - 1] + (delta / np.where(dx == 0, 1, dx)) * df)
    if not (isinstance(left, str) and left == "extrapolate"):
        f = np.where(x < xp[0], fp[0] if left is None else left, f)
    if not (isinstance(right, str) and right == "extrapolate"):
        f = np.where(x > xp[-1], fp[-1] if right is None else right, f)
    return f


def dl2cl(dl_xx, lmax, lmax_prime, is_phi=False):
    """simulate.py:21-45"""
    ell = np.arange(2, lmax).astype(np.float64)
    ellp = np.arange(2, lmax_prime).astype(np.float64)
    dl = np.asarray(dl_xx, dtype=np.float64)
    cl = dl[2:] * 2 * np.pi / ellp ** 4 if is_phi else dl[2:] * 2 * np.pi / (ellp * (ellp + 1))
    if np.all(cl > 0):
        return np.exp(interp(np.log(ell), np.log(ellp), np.log(cl), "extrapolate", "extrapolate"))
    return interp(ell, ellp, cl, 0.0, 0.0)


CAMB_COLS = {"TT": 0, "EE": 1, "BB": 2, "TE": 3}


def extract_all_cls(unlensed_scalar, tensor, total, lens_potential, lmax, lmax_prime):
    """simulate.py:442-451 on CAMB-layout arrays ([lmax_prime, 4] D_l, muK^2)."""
    cls = {}
    for st, col in CAMB_COLS.items():
        cls[f"scalar_{st}"] = dl2cl(unlensed_scalar[:, col], lmax, lmax_prime)
        cls[f"tensor_{st}"] = dl2cl(tensor[:, col], lmax, lmax_prime)
    for st in ("TT", "TE", "EE", "BB"):
        cls[f"total_{st}"] = dl2cl(total[:, CAMB_COLS[st]], lmax, lmax_prime)
    cls["phi"] = dl2cl(lens_potential, lmax, lmax_prime, is_phi=True)
    return cls


def beam_cls(beam_fwhm, ell):
    """simulate.py:48-49"""
    return np.exp(-np.asarray(ell, dtype=np.float64) ** 2
                  * np.deg2rad(beam_fwhm / ARCMIN_PER_DEGREE) ** 2 / BEAM_TRANSFER_SCALAR)


def noise_cls(lmax_prime, uk_arcmin_t, beam_fwhm=0, l_knee=100, alpha_knee=3):
    """simulate.py:52-65"""
    ellp = np.arange(2, lmax_prime)
    bls = beam_cls(beam_fwhm, ellp)
    nls = 1 + (l_knee / ellp) ** alpha_knee
    with np.errstate(divide="ignore", invalid="ignore"):
        tt = np.nan_to_num(np.deg2rad(uk_arcmin_t / ARCMIN_PER_DEGREE) ** 2 * nls / bls,
                           nan=0.0, posinf=0.0, neginf=0.0)
    return tt, np.zeros(tt.shape), 2 * tt, 2 * tt


def covar_matrix_from_cls(nside, pix_width, ell_grid, ells, cls, origin_value=None, rescale=True,
                          use_linear_interpolation=False, fill_value=None):
    """simulate.py:70-97"""
    shape = (nside, nside // 2 + 1)
    cls = np.asarray(cls, dtype=np.float64)
    ells = np.asarray(ells, dtype=np.float64)
    lg = ell_grid.flatten()
    if np.all(cls > 0) and not use_linear_interpolation:
        ls = np.where(lg == 0, 1e-10, lg)
        res = np.exp(interp(np.log(ls), np.log(ells), np.log(cls), "extrapolate",
                            "extrapolate")).reshape(shape)
    else:
        fv = fill_value if fill_value is not None else 0.0
        res = interp(lg, ells, cls, fv, fv).reshape(shape)
    if origin_value is not None:
        res[0, 0] = origin_value
    if rescale:
        res = res / pix_width ** 2
    return res


def cos_ramp_up(length):
    """simulate.py:117-118"""
    return (np.cos(np.linspace(np.pi, 0, length)) + 1) / 2


def low_pass(l_cutoff, delta_l=50):
    """simulate.py:123-124"""
    return np.concatenate([np.ones(l_cutoff - delta_l + 1), 1 - cos_ramp_up(delta_l)])


def get_mask(l_cutoff, nside, pix_width, ell_grid):
    """simulate.py:126-130"""
    screen = low_pass(l_cutoff)
    ell = np.arange(2, len(screen)).astype(np.float64)
    return covar_matrix_from_cls(nside, pix_width, ell_grid, ell, screen[2:], origin_value=1,
                                 rescale=False)


def get_beam(nside, pix_width, ell_grid, lmax_prime, beam_fwhm=0):
    """simulate.py:133-138"""
    ellp = np.arange(2, lmax_prime).astype(np.float64)
    b = np.sqrt(beam_cls(beam_fwhm, ellp))
    return covar_matrix_from_cls(nside, pix_width, ell_grid, ellp, b, origin_value=0,
                                 rescale=False, use_linear_interpolation=True, fill_value=1.0)


# --------------------------------------------------------------------------
# block algebra (util.py:10-69) and the mixing matrices (simulate.py:142-195)
# --------------------------------------------------------------------------
def reciprocal(m):
    """util.py:45-47"""
    m = np.asarray(m)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(m != 0, 1.0 / np.where(m != 0, m, 1), 0.0)


def teb_matrix_mult(tt1, te1, et1, ee1, bb1, tt2, te2, et2, ee2, bb2):
    """util.py:10-20"""
    return (tt1 * tt2 + te1 * et2, tt1 * te2 + te1 * ee2, et1 * tt2 + ee1 * et2,
            et1 * te2 + ee1 * ee2, bb1 * bb2)


def block_matrix_sqrt(a, b, c, d):
    """util.py:25-35: the principal square root of [[a, b], [c, d]] per mode."""
    with np.errstate(invalid="ignore"):
        s = np.sqrt(a * d - b * c)
        t = np.sqrt(a + (d + 2 * s))
    t = reciprocal(t)
    return t * (a + s), t * b, t * c, t * (d + s)


def invert_block_matrix(a, b, c, d):
    """util.py:59-69 (Schur complement)."""
    rd = reciprocal(d)
    s = reciprocal(a - b * rd * c)
    return s, -s * b * rd, -rd * c * s, rd + rd * c * s * b * rd


def get_d_matrix(cf_tt, cf_te, cf_ee, cf_bb, cn_tt, cn_te, cn_ee, cn_bb):
    """simulate.py:142-166"""
    pre = np.deg2rad(5 / ARCMIN_PER_DEGREE) ** 2
    one = np.ones(cn_tt.shape)
    inv = invert_block_matrix(cf_tt, cf_te, cf_te, cf_ee)
    ibb = reciprocal(cf_bb)
    itt, ite, iet, iee, ibb = [np.nan_to_num(x, nan=0, posinf=0, neginf=0)
                               for x in (*inv, ibb)]
    s_tt = cf_tt + pre * one + 2 * cn_tt
    s_te = cf_te + 2 * cn_te
    s_ee = cf_ee + pre * one + 2 * cn_ee
    s_bb = cf_bb + pre * one + 2 * cn_bb
    d_tt, d_te, d_et, d_ee, d_bb = teb_matrix_mult(s_tt, s_te, s_te, s_ee, s_bb,
                                                   itt, ite, iet, iee, ibb)
    d_tt, d_te, _, d_ee = block_matrix_sqrt(d_tt, d_et, d_et, d_ee)
    return d_tt, d_te, d_ee, np.sqrt(d_bb)


def get_g_matrix(cphi_fid, nphi, a_phi_fid, a_phi=1):
    """simulate.py:180-185"""
    g0 = np.sqrt(1 + 2 * nphi * reciprocal(cphi_fid))
    g = np.sqrt(1 + 2 * nphi * reciprocal((a_phi / a_phi_fid) * cphi_fid))
    return reciprocal(g0) * g


# --------------------------------------------------------------------------
# quadratic-estimate normalisation (simulate.py:196-376)
# --------------------------------------------------------------------------
def get_indices(length):
    """simulate.py:198-213: every 0/1 tuple of `length`, grouped by the number
    of zeros and ordered by itertools.combinations of the zero positions."""
    rows = []
    for n0 in range(length + 1):
        for zeros in combinations(range(length), n0):
            p = [1] * length
            for i in zeros:
                p[i] = 0
            rows.append(p)
    return np.array(rows, dtype=np.int64).reshape(-1, length)


def _kij(N, pix_width):
    kx = 2 * np.pi * np.fft.fftfreq(N, pix_width)
    ky = 2 * np.pi * np.fft.rfftfreq(N, pix_width)
    return np.meshgrid(kx, ky, indexing="ij")


def get_specific_derivative(field, pix_width, t1, t2):
    """simulate.py:216-253 for two derivative indices (0 = x, 1 = y)."""
    KX, KY = _kij(field.shape[0], pix_width)
    if t1 == t2:
        return -(KX ** 2) * field if t1 == 0 else -(KY ** 2) * field
    return -(KX * KY) * field


def qe_leg(field, pix_width, encapsulated, isolated):
    """simulate.py:269-303: (i KX)^nx (i KY)^ny field / |K|^(number of isolated
    indices), nan -> 0; the identity when there are no indices at all."""
    nd = len(isolated)
    sq = list(encapsulated) + list(isolated)
    ny = int(sum(sq))
    nx = len(sq) - ny
    field = np.nan_to_num(field, nan=0)
    if nx == 0 and ny == 0:
        return field * (1.0 + 0.0j)
    KX, KY = _kij(field.shape[0], pix_width)
    lap = KX ** 2 + KY ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        res = (1j * KX) ** nx * (1j * KY) ** ny * field / np.sqrt(lap ** nd)
    return np.nan_to_num(res, nan=0)


def _ir(x):
    return np.fft.irfft2(x)


def scalar_quadratic_estimate(cn_tt, cf_tt, cfl_tt, m_tt, b_tt, pix_width):
    """simulate.py:306-316, 350-359"""
    tf = m_tt * b_tt
    sigma = tf ** 2 * cfl_tt + cn_tt
    with np.errstate(divide="ignore", invalid="ignore"):
        a = tf ** 2 * cf_tt ** 2 / sigma
        b = tf ** 2 / sigma
        c = tf ** 2 * cf_tt / sigma
    qe = np.zeros(cf_tt.shape)
    for i, j in get_indices(2):
        internal = (_ir(qe_leg(a, pix_width, [i, j], [])) * _ir(qe_leg(b, pix_width, [], []))
                    + _ir(qe_leg(c, pix_width, [i], [])) * _ir(qe_leg(c, pix_width, [j], [])))
        qe += np.abs(get_specific_derivative(np.fft.rfft2(internal), pix_width, i, j))
    with np.errstate(divide="ignore"):
        return np.nan_to_num(1 / qe, nan=0, posinf=0, neginf=0)


def levi_civita_3d():
    """simulate.py:263-266"""
    i, j, k = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
    return (i - j) * (j - k) * (k - i) / 2


def polar_quadratic_estimate(cf_ee, cf_bb, cfl_ee, cfl_bb, cn_ee, cn_bb, mask_ee, mask_bb,
                             beam_ee, beam_bb, pix_width):
    """simulate.py:319-347, 362-376 (terms whose Levi-Civita weight is zero are
    skipped: they add 0 times a finite map)."""
    eps = levi_civita_3d()
    tf2e = (mask_ee * beam_ee) ** 2
    tf2b = (mask_bb * beam_bb) ** 2
    se = tf2e * cfl_ee + cn_ee
    sb = tf2b * cfl_bb + cn_bb
    with np.errstate(divide="ignore", invalid="ignore"):
        e2, e1, e0 = tf2e * cf_ee ** 2 / se, tf2e * cf_ee / se, tf2e / se
        b2, b1, b0 = tf2b * cf_bb ** 2 / sb, tf2b * cf_bb / sb, tf2b / sb
    N = cf_ee.shape[0]
    qe = np.zeros(cf_ee.shape, dtype=np.complex128)
    pw = pix_width
    for i, j in get_indices(2):
        internal = np.zeros((N, N))
        for k, l, m, n, p, q in get_indices(6):
            wgt = 4 * eps[m, p, 2] * eps[n, q, 2]
            if wgt == 0:
                continue
            A, Bb = [k, l, m, n], [k, l, p, q]
            internal += wgt * (_ir(qe_leg(e2, pw, [i, j], A)) * _ir(qe_leg(b0, pw, [], Bb))
                               - 2 * _ir(qe_leg(e1, pw, [i], A)) * _ir(qe_leg(b1, pw, [j], Bb))
                               + _ir(qe_leg(e0, pw, [], A)) * _ir(qe_leg(b2, pw, [i, j], Bb)))
        qe += np.abs(get_specific_derivative(np.fft.rfft2(internal), pw, i, j))
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.nan_to_num(1 / qe, nan=0, posinf=0, neginf=0)


# --------------------------------------------------------------------------
# lensing of the draws and the data model (simulate.py:455-469, 648-705)
# --------------------------------------------------------------------------
def lens_fields(fields, phi_f, nside, theta_pix, n=10):
    """simulate.py:455-469 for FOURIER [C, N, N//2+1] fields (T; E, B; or T, E, B)
    and a FOURIER phi: T lensed directly, E/B through Q/U maps."""
    pw = lf.pix_width_from_theta(theta_pix)
    x = w.to_map(w.eb2qu_field(fields, nside, theta_pix))
    y = lf.lense_flow(x, np.fft.irfft2(phi_f), pw, n)
    return w.qu2eb_field(w.to_fourier(y), nside, theta_pix)


def load_sim_from_cls(nside, theta_pix, pol, seed, cls, lmax, uk_arcmin_t=3, nphi_fac=2,
                      a_phi=1, n=10, trial=0, with_qe=True):
    """load_sim (simulate.py:617-743) without CAMB, for trial `trial` of master
    seed `seed`: returns a dict of the data set's arrays."""
    lmax_prime = min(lmax, 5000)
    ell_grid, pw = gen_ell_grid(nside, theta_pix)
    ells = np.arange(2, lmax).astype(np.float64)
    unscaled_cphi = covar_matrix_from_cls(nside, pw, ell_grid, ells, cls["phi"], origin_value=0)
    cphi = a_phi * unscaled_cphi
    comps = ("TT", "TE", "EE", "BB")
    cfs = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cls[f"scalar_{c}"], origin_value=0)
           for c in comps}
    cft = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cls[f"tensor_{c}"], origin_value=0)
           for c in comps}
    cf = {c: cfs[c] + cft[c] for c in comps}
    cfl = {c: covar_matrix_from_cls(nside, pw, ell_grid, ells, cls[f"total_{c}"], origin_value=0)
           for c in comps}
    mask = get_mask(3000, nside, pw, ell_grid)
    beam = get_beam(nside, pw, ell_grid, lmax_prime)
    ncl = noise_cls(lmax_prime, uk_arcmin_t)
    ellp = np.arange(2, lmax_prime)
    cn = {c: covar_matrix_from_cls(nside, pw, ell_grid, ellp, v, origin_value=0)
          for c, v in zip(comps, ncl)}

    def draw(covs, kind0):
        return grf_fourier(np.stack(covs), 1, seed, (kind0 << 32) + trial)[0]

    phi = draw([cphi], KIND_PHI)[0]
    if pol == "I":
        fk, keys = (cf["TT"],), ("TT",)
    elif pol == "P":
        fk, keys = (cf["EE"], cf["BB"]), ("EE", "BB")
    else:
        fk, keys = (cf["TT"], cf["EE"], cf["BB"]), ("TT", "EE", "BB")
    k0 = KIND_E if pol == "P" else KIND_T
    field = draw(fk, k0)
    lensed = lens_fields(field, phi, nside, theta_pix, n)
    noise = draw([cn[k] for k in keys], k0 + 3)
    mb = mask * beam
    data = mb * lensed + noise
    d = get_d_matrix(cf["TT"], cf["TE"], cf["EE"], cf["BB"], cn["TT"], cn["TE"], cn["EE"],
                     cn["BB"])
    out = dict(phi=phi, field=field, lensed=lensed, data=data, noise=noise, cphi=cphi,
               unscaled_cphi=unscaled_cphi, cf=cf, cfs=cfs, cft=cft, cfl=cfl, cn=cn, mask=mask,
               beam=beam, d=d, pix_width=pw)
    if with_qe:
        if pol == "I":
            qe = scalar_quadratic_estimate(cn["TT"], cf["TT"], cfl["TT"], mask, beam, pw) / nphi_fac
        else:
            qe = polar_quadratic_estimate(cf["EE"], cf["BB"], cfl["EE"], cfl["BB"], cn["EE"],
                                          cn["BB"], mask, mask, beam, beam, pw) / nphi_fac
        out["qe"] = qe
        out["g"] = get_g_matrix(cphi, qe, a_phi, a_phi=a_phi)
    return out
