"""Hamiltonian Monte Carlo for the lensing potential in the mixed
parametrisation (sampling_ar.py:60-140 of the reference), on the GPU.

Each leapfrog step is one mixed phi gradient (``mixing.mixed_grad_phi_logpdf``:
the unmixing LenseFlow, the gradient ODE, gradf and the mixing Jacobian, all
in libcmbl.so); with a leading batch axis on the fields and phi, independent
chains advance together and every LenseFlow / gradient ODE launch covers all
of them.  The integrator, Hamiltonian and accept rule are the reference's:

    x <- x - eps M^-1 (p - eps/2 grad);  grad <- grad(x);  p <- p - eps/2 (grad + grad_prev)
    H(x, p) = mixed_logpdf(x) - dot(p, M^-1 p) / 2,  accept if always or log u < H1 - H0

with 30 steps of 0.01 by default and the mass matrix G^-2 (Cphi^-1 + Nphi^-1).
Momenta are drawn like simulate.py:106-113: sqrt(M/2)(a + i b), a, b standard
normal, taken to a map by an orthonormal irfft2 and back by rfft2.

The Gibbs steps of sampling_ar.py:14-62 complete the sampler: the f-step is a
constrained realisation (device draws, one batched LenseFlow, a Wiener
filter), the phi-step the HMC step above.
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .constants import DEFAULT_NSTEPS
from .fields import FlatS0, dot, fourier, map as to_map
from .matrix_operators import pinv
from .mixing import mixed_grad_phi_logpdf, mixed_logpdf


def get_mass_matrix(cphi, nphi, mixing_g):
    """sampling_ar.py:64-66"""
    return pinv(mixing_g) * pinv(mixing_g) * (pinv(cphi) + pinv(nphi))


def draw_momentum(mass_matrix, like: FlatS0, generator: torch.Generator) -> FlatS0:
    """A momentum field p ~ N(0, M) (simulate.py:106-113) shaped like `like`
    (its batch axis gives one draw per chain)."""
    phi_m = like.scalar_matrix
    shape = phi_m.shape
    N = shape[-2]
    a = torch.randn(shape, generator=generator, device=phi_m.device, dtype=torch.float64)
    b = torch.randn(shape, generator=generator, device=phi_m.device, dtype=torch.float64)
    F = torch.sqrt(mass_matrix.scalar_matrix.to(torch.float64) / 2) * torch.complex(a, b)
    # irfft2 with norm="ortho" is N x the unnormalised inverse
    m = to_map(like.replace(scalar_matrix=F.to(phi_m.dtype)))
    m = m.replace(scalar_matrix=m.scalar_matrix * N)
    return fourier(m)


def symplectic_integrate(x0, p0, mixed_field, data, noise_covariance, phi_covariance,
                         field_covariance, mask, beam, mixing_d, mixing_g, mass_matrix,
                         num_steps: int = 30, step_size: float = 0.01, n: int = DEFAULT_NSTEPS):
    """sampling_ar.py:105-140: returns (delta_h per chain, x, p)."""
    minv = pinv(mass_matrix)

    def grad(x):
        return mixed_grad_phi_logpdf(mixed_field, x, data, noise_covariance, phi_covariance,
                                     field_covariance, mask, beam, mixing_d, mixing_g, n=n)

    def hamiltonian(x, p):
        lp = mixed_logpdf(mixed_field, x, data, noise_covariance, phi_covariance,
                          field_covariance, mask, beam, mixing_g, mixing_d, n=n)
        return lp - dot(p, minv * p) / 2

    g = grad(x0)
    x, p = x0, p0
    for _ in range(num_steps):
        g_prev = g
        x = x - step_size * (minv * (p - 0.5 * step_size * g))
        g = grad(x)
        p = p - 0.5 * step_size * (g + g_prev)
    delta_h = hamiltonian(x, p) - hamiltonian(x0, p0)
    return delta_h, x, p


def hmc_step(x, always_accept: bool, mass_matrix, mixed_field, data, noise_covariance,
             phi_covariance, field_covariance, mask, beam, mixing_d, mixing_g,
             generator: torch.Generator, num_steps: int = 30, step_size: float = 0.01,
             n: int = DEFAULT_NSTEPS):
    """sampling_ar.py:69-100: one HMC proposal per chain; returns (x, delta_h,
    accept) with x updated where accepted."""
    p = draw_momentum(mass_matrix, x, generator)
    delta_h, x_test, _ = symplectic_integrate(x, p, mixed_field, data, noise_covariance,
                                              phi_covariance, field_covariance, mask, beam,
                                              mixing_d, mixing_g, mass_matrix, num_steps,
                                              step_size, n)
    delta_h = torch.as_tensor(delta_h, device=x.scalar_matrix.device).reshape(-1)
    u = torch.rand(delta_h.shape, generator=generator, device=delta_h.device, dtype=torch.float64)
    accept = torch.logical_or(torch.tensor(bool(always_accept), device=u.device),
                              torch.log(u) < delta_h.real)
    xm, tm = x.scalar_matrix, x_test.scalar_matrix
    if xm.dim() == 3:
        new = torch.where(accept[:, None, None], tm, xm)
    else:
        new = tm if bool(accept[0]) else xm
    return x.replace(scalar_matrix=new), delta_h, accept


# --------------------------------------------------------------------------
# Gibbs steps (sampling_ar.py:14-62)
# --------------------------------------------------------------------------
def _diag_planes(op, what: str) -> torch.Tensor:
    from .matrix_operators import DiagonalEB, DiagonalScalar
    if not isinstance(op, (DiagonalScalar, DiagonalEB)):
        raise TypeError(f"{what}: the f-step draws from DiagonalScalar or DiagonalEB "
                        f"operators, got {type(op).__name__}")
    return op.per_mode().to(torch.float64)


def gibbs_sample_f(field_start, data_field, phi, args: dict, seed: int, trial0: int = 0,
                   n: int = 10, maxiter: int = 500, tol: float = 1e-1):
    """sampling_ar.py:14-46: a constrained realisation of f given phi and the
    data.  Draws f' ~ N(0, Cf) and n' ~ N(0, Cn) on the device, forms their
    data d' = M B L(phi) f' + n' (one batched LenseFlow with 10 RK4 steps, the
    noise drawn into the data in place), and returns f' + WF(d - d') with the
    Wiener filter started from field_start.  field_start (and phi) may carry a
    leading chain axis: chain c then draws from the streams of trial trial0 + c
    (kinds 1..3 fields, 4..6 noise, like simulate.simulate).  args holds
    field_covariance, noise_covariance, mask and beam (DiagonalScalar for T,
    DiagonalEB for EB fields).  Returns a field like field_start."""
    from .fields import FlatS0, FlatS2
    from .lense_flow import _phi_matrix
    from .simulate import KIND_E, KIND_T, grf_fourier, lens_fields
    from .wiener_filter import wiener_filter
    Cf, Cn = args["field_covariance"], args["noise_covariance"]
    M, B = args["mask"], args["beam"]
    if isinstance(field_start, FlatS0):
        names, k0 = ["scalar_matrix"], KIND_T
    elif isinstance(field_start, FlatS2):
        names, k0 = ["polar_matrix_1", "polar_matrix_2"], KIND_E
    else:
        raise TypeError("gibbs_sample_f takes FlatS0 (T) or FlatS2 (EB) fields")
    m0 = getattr(field_start, names[0])
    lead = tuple(m0.shape[:-2])
    ntr = int(np.prod(lead)) if lead else 1
    th = field_start.theta_pix
    N = m0.shape[-2]
    rdt = torch.float32 if m0.dtype in (torch.complex64, torch.float32) else torch.float64
    new = grf_fourier(_diag_planes(Cf, "field_covariance"), ntr, seed, (k0 << 32) + trial0, th,
                      rdt)
    phi_m = _phi_matrix(phi)
    lensed = lens_fields(new, phi_m.reshape(ntr, *phi_m.shape[-2:]) if phi_m.dim() > 2 else phi_m,
                         N, th, n)
    mb = _diag_planes(M, "mask") * _diag_planes(B, "beam")
    new_data = grf_fourier(_diag_planes(Cn, "noise_covariance"), ntr, seed,
                           ((k0 + 3) << 32) + trial0, th, rdt, mul=mb, add=lensed)

    def as_field(X):
        return field_start.replace(**{nm: X[:, c].reshape(*lead, *X.shape[-2:]) if lead
                                      else X[0, c] for c, nm in enumerate(names)})
    data_diff = data_field - as_field(new_data)
    delta = wiener_filter(field_start, phi, data_diff, Cf, Cn, M, B, n=n, maxiter=maxiter,
                          tol=tol)
    return as_field(new) + delta


def gibbs_sample_phi(mixed_phi, mixed_temp, data_field, generator: torch.Generator, args: dict,
                     iteration: int, num_burn_in_always_accept: int, num_steps: int = 30,
                     step_size: float = 0.01, n: int = DEFAULT_NSTEPS):
    """sampling_ar.py:49-62: one HMC step of the mixed phi (always accepted
    during burn-in) with the mass matrix G^-2 (Cphi^-1 + QE^-1)."""
    always_accept = iteration < num_burn_in_always_accept
    mass = get_mass_matrix(args["phi_covariance"], args["quadratic_estimate"], args["mixing_g"])
    return hmc_step(mixed_phi, always_accept, mass, mixed_temp, data_field,
                    args["noise_covariance"], args["phi_covariance"], args["field_covariance"],
                    args["mask"], args["beam"], args["mixing_d"], args["mixing_g"], generator,
                    num_steps=num_steps, step_size=step_size, n=n)
