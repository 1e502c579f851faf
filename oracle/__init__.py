"""FP64 CPU oracle for the LenseFlow / lensed Wiener-filter path.

THIS PACKAGE IS TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker (or as the timed CPU baseline).  The product path in
``cmblensing.py_amd`` never imports it and has no CPU fallback.

It restates, in plain numpy (pocketfft, float64 / complex128), the algorithm of
CMBLensing.py (the JAX package ``cmb_lensing``).  Every function cites the
reference file:line it follows.  Arrays are batched over leading axes:
MAP arrays are ``[..., N, N]`` real, FOURIER arrays ``[..., N, N//2+1]`` complex
(numpy ``rfft2`` layout; axis -2 carries KX, axis -1 carries KY).

How each part is pinned (tests/test_oracle_*.py):

* FFT conventions ........ numpy.fft (pocketfft) + brute-force DFT at N=8/16.
* spectral derivatives ... closed forms on single plane waves, including the
                           negated KY Nyquist entry (util.py:106).
* RK4 .................... order test (error falls ~16x per halving of h),
                           phi=0 gives the identity, constant f is unchanged.
* forward LenseFlow ...... Julia CMBLensing.jl fixtures t_lensed_field_I,
                           {e,b}_lensed_field_P (reference tests/ground_truth_data)
                           and the continuum theorem Lf(x)=f(x+grad phi(x))
                           evaluated by a brute-force non-uniform DFT.
* inverse LenseFlow ...... round trip L^-1 L f = f to O(h^4); Julia fixture
                           round trip (test_lensing.py:97).
* adjoint LenseFlow ...... dot-product test <Lx,y> vs <x,L^T y> converging as
                           O(h^4); Julia t_adjoint_lensed_field_I fixture.
* gradient ODE (VJP) ..... central finite differences of a scalar built on the
                           forward LenseFlow; Julia grad_phi_* fixtures.
* gradf_logpdf / logpdf .. Julia gradf_* and logpdf_*.txt fixtures; FD of logpdf
                           along random directions (<gradf, v> = d logpdf).
* Wiener filter (PCG) .... dense direct solve at N=8; Julia f_wf_* fixtures.
"""
