// Launchers of the Wiener-filter kernels in wf.hip (internal).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace cmbl {

template <typename T>
hipError_t launch_wf_data(const cplx<T>* X, const cplx<T>* d, const T* mb, const T* bmcn,
                          cplx<T>* out, int N, int C, long nfields, hipStream_t s);
template <typename T>
hipError_t launch_fill_zero(T* x, long n, hipStream_t s);
template <typename Tb>
hipError_t launch_dot(const cplx<double>* a, const cplx<Tb>* b, int N, long len, int nrhs,
                      double* part, double* out, hipStream_t s);
template <typename Tb>
hipError_t launch_cg_xrz(cplx<double>* x, cplx<double>* r, cplx<double>* z,
                         const cplx<double>* p, const cplx<Tb>* Ap, const double* pinv,
                         const double* alpha, int N, int C, long len, int nrhs, double* part,
                         double* rz, hipStream_t s);
template <typename T>
hipError_t launch_cg_p(cplx<double>* p, const cplx<double>* z, const double* beta, cplx<T>* pT,
                       long len, int nrhs, hipStream_t s);
template <typename T>
hipError_t launch_teb_split(const cplx<T>* f, cplx<T>* XT, cplx<T>* XP, int N, long nfields,
                            hipStream_t s);
template <typename T>
hipError_t launch_wf_data_teb(const cplx<T>* XT, const cplx<T>* XP, const cplx<T>* d, const T* mb,
                              const T* bmcn, cplx<T>* outT, cplx<T>* outP, int N, long nfields,
                              hipStream_t s);
template <typename T>
hipError_t launch_wf_final_teb(const cplx<T>* XLT, const cplx<T>* XLP, const cplx<T>* F,
                               const T* cfinv, cplx<T>* out, int N, long nfields, hipStream_t s);
// In place on packed row spectra P [nfields][C][N][N/2] (a row R2C of the
// maps): column FFT, per mode C = 2: eb2qu(diag .* qu2eb(.)), C = 1:
// diag .* (.), inverse column FFT, times 1/N^2; a row C2R then gives
// irfft2(op(rfft2(x))).  diag [C][N][N/2+1].
template <typename T>
hipError_t launch_col_op(int N, cplx<T>* P, const T* diag, long nfields, int C,
                         const cplx<T>* tw, hipStream_t s);
// The same pass with the data term of the Fourier-mask posterior per mode:
// eb2qu(bmcn .* (d - mb .* qu2eb(.))) (d natural [nfields][C][N][N/2+1] or
// NULL).
template <typename T>
hipError_t launch_col_data(int N, cplx<T>* P, const cplx<T>* d, const T* mb, const T* bmcn,
                           long nfields, int C, const cplx<T>* tw, hipStream_t s);
// Natural EB (or T) half planes X -> eb2qu per mode -> inverse column FFT ->
// packed P, times 1/N^2 (the column half of irfft2(eb2qu(X))).
template <typename T>
hipError_t launch_col_inv_eb2qu(int N, const cplx<T>* X, cplx<T>* P, long nfields, int C,
                                const cplx<T>* tw, hipStream_t s);
// Packed row spectra P -> column FFT -> per mode qu2eb(.) - cfinv .* Herm(F)
// -> natural out (the column half of rfft2 and the last line of gradf_logpdf).
template <typename T>
hipError_t launch_col_fwd_final(int N, const cplx<T>* P, const cplx<T>* F, const T* cfinv,
                                cplx<T>* out, long nfields, int C, const cplx<T>* tw,
                                hipStream_t s);
hipError_t launch_cg_alpha(const double* res, const double* pAp, const unsigned char* act,
                           double* alpha, int nrhs, hipStream_t s);
hipError_t launch_cg_beta(double* res, const double* rz, const unsigned char* act, double* beta,
                          int nrhs, hipStream_t s);
hipError_t launch_cg_rel(const double* rr, const double* bb, double* rel, int nrhs, hipStream_t s);
int red_parts();

}  // namespace cmbl
