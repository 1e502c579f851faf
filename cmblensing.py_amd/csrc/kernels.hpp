// Host-side launchers of the HIP kernels in kernels.hip (internal to the
// library; the public C ABI is include/cmbl.h).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace cmbl {

// rfft2 of nmaps real N x N maps -> natural half planes X[nmaps][N][N/2+1].
// P is packed scratch [nmaps][N][N/2].
template <typename T>
hipError_t launch_rfft2(int N, const T* x, cplx<T>* P, cplx<T>* X, long nmaps,
                        const cplx<T>* tw, hipStream_t s);
// irfft2 (numpy semantics, 1/N^2) of natural half planes -> real maps.
template <typename T>
hipError_t launch_irfft2(int N, const cplx<T>* X, cplx<T>* P, T* x, long nmaps,
                         const cplx<T>* tw, hipStream_t s);
// Five derivative spectra of F (double precision, natural layout), for
// `count` phis: F [count][N][N/2+1] -> D [count][5][N][N/2+1].
hipError_t launch_phi_mult(int N, const cplx<double>* F, cplx<double>* D, double dk,
                           hipStream_t s, int count = 1);
// Deflection-operator table and per-block minimum of det(I + t Hess phi), for
// `count` phis: D [count][5][N][N] -> ptab [count][2n+1][2][N][N], minpart
// [count][*nblocks].
template <typename T>
hipError_t launch_ptab(int N, const double* D, T* ptab, int nsteps, double* minpart,
                       int* nblocks, hipStream_t s, int count = 1);
// LenseFlow stage passes.
template <typename T>
hipError_t launch_lf_row_r2c(int N, const T* y, cplx<T>* S0, long nmaps, const cplx<T>* tw,
                             hipStream_t s);
// Row C2R of packed row spectra P [nmaps][N][N/2] -> x [nmaps][N][N] (no
// scaling), with a pixel epilogue, w [N][N] for every map:
// ep 0: x; 1: w .* x; 2: d - w .* x (d NULL: zeros).
template <typename T>
hipError_t launch_row_c2r(int N, const cplx<T>* P, T* x, long nmaps, const cplx<T>* tw,
                          const T* d, const T* w, int ep, hipStream_t s);
// x-derivative column pass (src may alias dst).
template <typename T>
hipError_t launch_lf_col(int N, const cplx<T>* src, cplx<T>* dst, long nmaps, const cplx<T>* tw,
                         T dk, hipStream_t s);
template <typename T>
struct GradArgs {
  const T* zd;            // delta f stage inputs [B*C][N][N]
  const T* const* gtabs;  // per batch element: M^-1 table [(2n+1)][3][N][N]
  T* sacc;                // [B][5][N][N]: m_x, m_y, t w_xx, t (w_xy + w_yx), t w_yy
  T cs;                   // h * b_s (b = 1/6, 1/3, 1/3, 1/6)
  T tval;                 // stage time t_s
};

// Row pass of one stage.  adjoint: divergence rate, zstore (may be NULL)
// receives the next stage input.  Standard rate with ga != NULL: also
// accumulates the delta-phi sources of the gradient ODE.
// outp != NULL (standard rate, the last stage of the serving executor): the
// new state goes to the per-request maps outp[b] instead of y.  nt (FP32,
// low-storage stages 2 and 4): the accumulator and the new state are written
// and read as streaming accesses (batches larger than the last-level cache).
template <typename T>
hipError_t launch_lf_row(int N, bool adjoint, cplx<T>* S0, cplx<T>* S1, T* y, T* acc,
                         const T* const* ptabs, const cplx<T>* tw, int C, long nmaps, int stage,
                         T h, int tidx, int do_r2c, int init, T dk, hipStream_t s,
                         T* zstore = nullptr, const GradArgs<T>* ga = nullptr,
                         T* const* outp = nullptr, bool nt = false);
// Row pass of one adjoint-rate stage with the low-storage RK4 bookkeeping:
// S0 = U (x-derivative after the column pass), Z = the stage input z_s as a
// real map [nmaps][N][N] (stage 1 reads z_1 = y).  tidx: time index of this
// stage (p_y), tidx_next: of the next stage (p_x of u).  init: U of y only.
// nt: as for launch_lf_row.
template <typename T>
hipError_t launch_lf_row_adj_ls(int N, cplx<T>* S0, T* Z, T* y, T* acc, const T* const* ptabs,
                                const cplx<T>* tw, int C, long nmaps, int stage, T h, int tidx,
                                int tidx_next, int do_r2c, int init, T dk, hipStream_t s,
                                bool nt = false);
// The whole standard-rate LenseFlow of small FP32 maps (N <= 64) in one
// launch, one workgroup per component with its stage state in LDS.  src /
// outp: request pointers of the serving executor (else y in place).
template <typename T>
hipError_t launch_lf_resident(int N, const T* const* src, T* const* outp, T* y, int C,
                              const T* const* ptabs, const cplx<T>* tw, long nmaps, int nsteps,
                              int fwd_time, T h, T dk, hipStream_t s);
// First launch of the serving executor: row R2C of the requests' maps
// (src[b] + component N^2, NULL = zeros) into S0, also stored to ystage.
template <typename T>
hipError_t launch_lf_row_r2c_gather(int N, const T* const* src, int ncomp, T* ystage, cplx<T>* S0,
                                    long nmaps, const cplx<T>* tw, hipStream_t s);
template <typename T>
hipError_t launch_gather(const T* const* src, T* dst, long len, int count, hipStream_t s);
template <typename T>
hipError_t launch_scatter(const T* src, T* const* dst, long len, int count, hipStream_t s);
template <typename T>
hipError_t launch_gtab(int N, const double* D, T* gtab, int nsteps, hipStream_t s);
template <typename T>
hipError_t launch_dphi_combine(int N, const cplx<T>* F, cplx<T>* out, long nb, double dk,
                               hipStream_t s);

}  // namespace cmbl
