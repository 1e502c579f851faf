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
"""
from __future__ import annotations

from typing import Optional

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
