"""FP64 numpy restatement of CMBLensing.py's Gibbs f-step (sampling_ar.py:14-46),
composed of the oracle's draws, lensing and Wiener filter.

TEST INFRASTRUCTURE ONLY -- imported by tests/ (and never by the package).

Fields are FOURIER arrays [C, N, N//2+1] (C = 1: T, 2: E, B); Cf, Cn, M, B are
per-mode arrays of that shape; phi a FOURIER [N, N//2+1] array.  The draws use
the library's streams (oracle.simulate): chain t draws its field from kinds
1 (T) or 2, 3 (E, B) and its noise from kinds 4 or 5, 6, trial t.
"""
from __future__ import annotations

from . import simulate as sim
from . import wiener as w


def gibbs_sample_f(f0, d, phi, Cf, Cn, M, B, theta_pix, seed, trial=0, n=10, maxiter=500,
                   tol=1e-1):
    """f' + WF(d - (M B L(phi) f' + n')) with the Wiener filter started at f0."""
    C = f0.shape[0]
    k0 = sim.KIND_T if C == 1 else sim.KIND_E
    new = sim.grf_fourier(Cf, 1, seed, (k0 << 32) + trial)[0]
    noise = sim.grf_fourier(Cn, 1, seed, ((k0 + 3) << 32) + trial)[0]
    N = f0.shape[-2]
    new_data = M * B * sim.lens_fields(new, phi, N, theta_pix, n) + noise
    delta = w.wiener_filter(f0, phi, d - new_data, Cf, Cn, M, B, theta_pix, maxiter=maxiter,
                            tol=tol, n=n)
    return new + delta
