// Fourier-domain kernels of the lensed Wiener filter (gradients.py:12-26,
// wiener_filter.py:28-105) on natural-layout half planes X[m][kx][ky],
// ky in [0, N/2], m = batch*C + component (C = 1: T, C = 2: E,B or Q,U).
//
// EB <-> QU (util.py:140-161): psi = arctan2(lx, ly) with lx the half-axis
// frequency (its Nyquist column negated for rows kx <= N/2) and ly the full
// axis; cos 2psi = (ly^2 - lx^2) / l^2, sin 2psi = 2 lx ly / l^2 (psi = 0 at
// l = 0).  The common 2 pi / (N d) factor cancels, so integer frequencies give
// the same angles.
//   eb2qu:  Q = -E c + B s,  U = -E s - B c
//   qu2eb:  E = -Q c - U s,  B =  Q s - U c
#include <hip/hip_runtime.h>

#include "fft_lds.hpp"
#include "wf.hpp"

namespace cmbl {

// cos 2psi, sin 2psi at mode (i, j) = (kx, ky) of the half plane.  The
// frequencies are integers, so l^2 and the numerators are exact; FP64 divides,
// FP32 multiplies by the hardware reciprocal of l^2 (1 ulp).
template <typename T>
__device__ __forceinline__ void rot2psi_ij(int N, int i, int j, T& c, T& s) {
  const int ly = (i < N / 2) ? i : i - N;
  const int lx = (j == N / 2 && i <= N / 2) ? -j : j;
  const int l2 = lx * lx + ly * ly;
  if (l2 == 0) {
    c = T(1);
    s = T(0);
    return;
  }
  const int nc = ly * ly - lx * lx, ns = 2 * lx * ly;
  if constexpr (sizeof(T) == 4) {
    const float r = __builtin_amdgcn_rcpf((float)l2);
    c = (float)nc * r;
    s = (float)ns * r;
  } else {
    c = (double)nc / (double)l2;
    s = (double)ns / (double)l2;
  }
}

template <typename T>
__device__ __forceinline__ void rot2psi(int N, long idx, T& c, T& s) {
  const int H1 = N / 2 + 1;
  rot2psi_ij<T>(N, (int)(idx / H1), (int)(idx % H1), c, s);
}

template <typename T>
__device__ __forceinline__ void eb2qu_mode(cplx<T> e, cplx<T> b, T c, T s, cplx<T>& q,
                                           cplx<T>& u) {
  q = mk(-e.x * c + b.x * s, -e.y * c + b.y * s);
  u = mk(-e.x * s - b.x * c, -e.y * s - b.y * c);
}

template <typename T>
__device__ __forceinline__ void qu2eb_mode(cplx<T> q, cplx<T> u, T c, T s, cplx<T>& e,
                                           cplx<T>& b) {
  e = mk(-q.x * c - u.x * s, -q.y * c - u.y * s);
  b = mk(q.x * s - u.x * c, q.y * s - u.y * c);
}

// Data term of gradf_logpdf (gradients.py:18-22) from the rfft2 of the lensed
// QU maps X:  g = bmcn .* (d - mb .* qu2eb(X)) in EB, returned as eb2qu(g).
// mb = M B and bmcn = B M Cn^-1 are per-mode diagonals [C][N][N/2+1]; d may be
// NULL (zero data, the CG operator), X may be NULL (zero field).  out may
// alias X.
template <typename T>
__global__ void k_wf_data(const cplx<T>* X, const cplx<T>* __restrict__ d,
                          const T* __restrict__ mb, const T* __restrict__ bmcn, cplx<T>* out,
                          int N, int C, long nfields) {
  const long P = (long)N * (N / 2 + 1);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nfields * P;
       g += (long)gridDim.x * blockDim.x) {
    const long f = g / P, idx = g % P;
    const long base = f * C * P + idx;
    const cplx<T> zero = mk(T(0), T(0));
    if (C == 1) {
      cplx<T> r = X ? scale(X[base], -mb[idx]) : zero;
      if (d) r = r + d[base];
      out[base] = scale(r, bmcn[idx]);
      continue;
    }
    T c, s;
    rot2psi<T>(N, idx, c, s);
    cplx<T> e = zero, b = zero;
    if (X) qu2eb_mode(X[base], X[base + P], c, s, e, b);
    cplx<T> re = scale(e, -mb[idx]), rb = scale(b, -mb[P + idx]);
    if (d) {
      re = re + d[base];
      rb = rb + d[base + P];
    }
    re = scale(re, bmcn[idx]);
    rb = scale(rb, bmcn[P + idx]);
    cplx<T> q, u;
    eb2qu_mode(re, rb, c, s, q, u);
    out[base] = q;
    out[base + P] = u;
  }
}

// Last line of gradf_logpdf (gradients.py:25):
//   out = qu2eb(XL) - cf_inv .* qu2eb(fourier(map(eb2qu(F))))
// with XL the rfft2 of the adjoint-lensed maps and F the EB (or T) input
// field.  fourier(map(.)) of a half plane only replaces the two
// self-conjugate columns (ky = 0 and ky = N/2) by their Hermitian parts
// (Y[kx] + conj(Y[-kx]))/2, and the EB<->QU rotation is the same at kx and -kx
// on those columns, so the prior term is cf_inv .* Herm(F) computed per mode
// (herm_col; OpFinal below and k_wf_final_teb): identical to the reference's
// round trip, without leaking FFT rounding of large modes into small ones
// (which cf_inv, spanning ~1e10, would amplify).
template <typename T>
__device__ __forceinline__ cplx<T> herm_col(const cplx<T>* F, long base_f, int N, int i, int j) {
  const int H1 = N / 2 + 1;
  const cplx<T> a = F[base_f + (long)i * H1 + j];
  if (j != 0 && j != N / 2) return a;
  const cplx<T> b = F[base_f + (long)((N - i) & (N - 1)) * H1 + j];
  return mk(T(0.5) * (a.x + b.x), T(0.5) * (a.y - b.y));
}

// ---------------------------------------------------------------------------
// T + EB fields (C = 3, FlatS02) with BlockTEB operators (matrix_operators.py:
// 92-148): per mode [[TT, TE, 0], [ET, EE, 0], [0, 0, BB]], stored as five
// real planes [5][N][N/2+1] in that order.  Inside gradf the T maps and the
// Q,U pairs are lensed as two groups (C = 1 and C = 2), so these kernels split
// a [nfields][3] field into a T group [nfields] and a QU group [nfields][2]
// and join the groups back.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void teb_apply(const T* __restrict__ op, long idx, long P, cplx<T> t,
                                          cplx<T> e, cplx<T> b, cplx<T>& ot, cplx<T>& oe,
                                          cplx<T>& ob) {
  const T tt = op[idx], te = op[P + idx], et = op[2 * P + idx], ee = op[3 * P + idx];
  ot = scale(t, tt) + scale(e, te);
  oe = scale(t, et) + scale(e, ee);
  ob = scale(b, op[4 * P + idx]);
}

// f [nfields][3] (T, E, B) -> XT [nfields] (T) and XP [nfields][2] (eb2qu)
template <typename T>
__global__ void k_teb_split(const cplx<T>* __restrict__ f, cplx<T>* __restrict__ XT,
                            cplx<T>* __restrict__ XP, int N, long nfields) {
  const long P = (long)N * (N / 2 + 1);
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nfields * P;
       g += (long)gridDim.x * blockDim.x) {
    const long fi = g / P, idx = g % P;
    const long base = fi * 3 * P + idx;
    XT[fi * P + idx] = f[base];
    T c, s;
    rot2psi<T>(N, idx, c, s);
    cplx<T> q, u;
    eb2qu_mode(f[base + P], f[base + 2 * P], c, s, q, u);
    XP[fi * 2 * P + idx] = q;
    XP[fi * 2 * P + P + idx] = u;
  }
}

// Data term for T + EB: from the rfft2 of the lensed T maps XT and Q,U maps XP
// (NULL: f = 0), g = bmcn (d - mb (T, qu2eb(Q, U))), with d [nfields][3] TEB
// (NULL: d = 0); written as T -> outT and eb2qu(E, B) -> outP.  outT / outP may
// alias XT / XP.
template <typename T>
__global__ void k_wf_data_teb(const cplx<T>* XT, const cplx<T>* XP, const cplx<T>* __restrict__ d,
                              const T* __restrict__ mb, const T* __restrict__ bmcn,
                              cplx<T>* outT, cplx<T>* outP, int N, long nfields) {
  const long P = (long)N * (N / 2 + 1);
  const cplx<T> zero = mk(T(0), T(0));
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nfields * P;
       g += (long)gridDim.x * blockDim.x) {
    const long fi = g / P, idx = g % P;
    T c, s;
    rot2psi<T>(N, idx, c, s);
    cplx<T> t = zero, e = zero, b = zero;
    if (XT) {
      t = XT[fi * P + idx];
      qu2eb_mode(XP[fi * 2 * P + idx], XP[fi * 2 * P + P + idx], c, s, e, b);
    }
    cplx<T> rt, re, rb;
    teb_apply(mb, idx, P, t, e, b, rt, re, rb);
    rt = mk(-rt.x, -rt.y);
    re = mk(-re.x, -re.y);
    rb = mk(-rb.x, -rb.y);
    if (d) {
      const long base = fi * 3 * P + idx;
      rt = rt + d[base];
      re = re + d[base + P];
      rb = rb + d[base + 2 * P];
    }
    cplx<T> gt, ge, gb;
    teb_apply(bmcn, idx, P, rt, re, rb, gt, ge, gb);
    outT[fi * P + idx] = gt;
    cplx<T> q, u;
    eb2qu_mode(ge, gb, c, s, q, u);
    outP[fi * 2 * P + idx] = q;
    outP[fi * 2 * P + P + idx] = u;
  }
}

// Last line of gradf for T + EB: out [nfields][3] = (XLT, qu2eb(XLP)) minus
// cfinv (Herm F_T, Herm F_E, Herm F_B) (the prior term as in OpFinal).
// F and cfinv may be NULL.
template <typename T>
__global__ void k_wf_final_teb(const cplx<T>* XLT, const cplx<T>* XLP,
                               const cplx<T>* __restrict__ F, const T* __restrict__ cfinv,
                               cplx<T>* out, int N, long nfields) {
  const long P = (long)N * (N / 2 + 1);
  const int H1 = N / 2 + 1;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nfields * P;
       g += (long)gridDim.x * blockDim.x) {
    const long fi = g / P, idx = g % P;
    const int i = (int)(idx / H1), j = (int)(idx % H1);
    T c, s;
    rot2psi<T>(N, idx, c, s);
    cplx<T> t = XLT[fi * P + idx], e, b;
    qu2eb_mode(XLP[fi * 2 * P + idx], XLP[fi * 2 * P + P + idx], c, s, e, b);
    if (F && cfinv) {
      const long bf = fi * 3 * P;
      cplx<T> pt, pe, pb;
      teb_apply(cfinv, idx, P, herm_col(F, bf, N, i, j), herm_col(F, bf + P, N, i, j),
                herm_col(F, bf + 2 * P, N, i, j), pt, pe, pb);
      t = t - pt;
      e = e - pe;
      b = b - pb;
    }
    const long base = fi * 3 * P + idx;
    out[base] = t;
    out[base + P] = e;
    out[base + 2 * P] = b;
  }
}

template <typename T>
__global__ void k_fill_zero(T* x, long n) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < n; g += (long)gridDim.x * blockDim.x)
    x[g] = T(0);
}

// ---------------------------------------------------------------------------
// FP64 CG vector kernels.  Vectors are complex double half planes
// [nrhs][C][N][N/2+1]; the inner product is fields.py:136-142 /
// util.py:100-101: Re sum conj(a) b w / N^2, w = 1 on the ky = 0 and ky = N/2
// columns, 2 elsewhere.  Reductions are two-pass in a fixed order, so results
// do not depend on scheduling.
// ---------------------------------------------------------------------------
constexpr int kRedBlock = 256;
constexpr int kRedParts = 64;  // partial sums per right-hand side

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[kRedBlock];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kRedBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double wgt(long idx, int N) {
  const int j = (int)(idx % (N / 2 + 1));
  return (j == 0 || j == N / 2) ? 1.0 : 2.0;
}

// partial[rhs][part] of Re sum conj(a) b w; b is complex<Tb> (double or float)
template <typename Tb>
__global__ __launch_bounds__(kRedBlock) void k_dot_partial(const cplx<double>* __restrict__ a,
                                                           const cplx<Tb>* __restrict__ b, int N,
                                                           long len, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const long P = (long)N * (N / 2 + 1);
  const cplx<double>* ar = a + rhs * len;
  const cplx<Tb>* br = b + rhs * len;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kRedBlock + threadIdx.x; i < len; i += (long)kRedParts * kRedBlock) {
    const cplx<double> x = ar[i];
    const cplx<Tb> y = br[i];
    acc += (x.x * (double)y.x + x.y * (double)y.y) * wgt(i % P, N);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[rhs * kRedParts + blockIdx.x] = s;
}

__global__ void k_dot_final(const double* __restrict__ part, double scale_, double* __restrict__ out,
                            int nrhs) {
  const int rhs = blockIdx.x * blockDim.x + threadIdx.x;
  if (rhs >= nrhs) return;
  double s = 0.0;
  for (int k = 0; k < kRedParts; ++k) s += part[rhs * kRedParts + k];
  out[rhs] = s * scale_;
}

// x += alpha p ; r -= alpha Ap ; z = pinv .* r ; partial of Re sum conj(r) z w
template <typename Tb>
__global__ __launch_bounds__(kRedBlock) void k_cg_xrz(cplx<double>* __restrict__ x,
                                                      cplx<double>* __restrict__ r,
                                                      cplx<double>* __restrict__ z,
                                                      const cplx<double>* __restrict__ p,
                                                      const cplx<Tb>* __restrict__ Ap,
                                                      const double* __restrict__ pinv,
                                                      const double* __restrict__ alpha, int N,
                                                      int C, long len, double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const long P = (long)N * (N / 2 + 1);
  const double al = alpha[rhs];
  const long off = rhs * len;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kRedBlock + threadIdx.x; i < len; i += (long)kRedParts * kRedBlock) {
    const long o = off + i;
    cplx<double> xi = x[o], ri = r[o];
    const cplx<double> pi = p[o];
    const cplx<Tb> api = Ap[o];
    xi = mk(xi.x + al * pi.x, xi.y + al * pi.y);
    ri = mk(ri.x - al * (double)api.x, ri.y - al * (double)api.y);
    const double w = pinv[i % (C * P)];
    const cplx<double> zi = mk(w * ri.x, w * ri.y);
    x[o] = xi;
    r[o] = ri;
    z[o] = zi;
    acc += (ri.x * zi.x + ri.y * zi.y) * wgt(i % P, N);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[rhs * kRedParts + blockIdx.x] = s;
}

// The same update for T + EB (C = 3) with a BlockTEB pinv [5][N][N/2+1]:
// z = pinv r mixes the T and E components of each mode.
template <typename Tb>
__global__ __launch_bounds__(kRedBlock) void k_cg_xrz_teb(cplx<double>* __restrict__ x,
                                                          cplx<double>* __restrict__ r,
                                                          cplx<double>* __restrict__ z,
                                                          const cplx<double>* __restrict__ p,
                                                          const cplx<Tb>* __restrict__ Ap,
                                                          const double* __restrict__ pinv,
                                                          const double* __restrict__ alpha, int N,
                                                          double* __restrict__ part) {
  const int rhs = blockIdx.y;
  const long P = (long)N * (N / 2 + 1);
  const double al = alpha[rhs];
  const long off = rhs * 3 * P;
  double acc = 0.0;
  for (long i = blockIdx.x * (long)kRedBlock + threadIdx.x; i < P; i += (long)kRedParts * kRedBlock) {
    cplx<double> rc[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long o = off + c * P + i;
      const cplx<double> xi = x[o], pi = p[o];
      const cplx<Tb> api = Ap[o];
      const cplx<double> ri = r[o];
      x[o] = mk(xi.x + al * pi.x, xi.y + al * pi.y);
      rc[c] = mk(ri.x - al * (double)api.x, ri.y - al * (double)api.y);
      r[o] = rc[c];
    }
    cplx<double> zc[3];
    teb_apply(pinv, i, P, rc[0], rc[1], rc[2], zc[0], zc[1], zc[2]);
    const double w = wgt(i, N);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      z[off + c * P + i] = zc[c];
      acc += (rc[c].x * zc[c].x + rc[c].y * zc[c].y) * w;
    }
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) part[rhs * kRedParts + blockIdx.x] = s;
}

// p = z + beta p, and p rounded to the operator precision (pT)
template <typename T>
__global__ void k_cg_p(cplx<double>* __restrict__ p, const cplx<double>* __restrict__ z,
                       const double* __restrict__ beta, cplx<T>* __restrict__ pT, long len,
                       int nrhs) {
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < len * nrhs;
       g += (long)gridDim.x * blockDim.x) {
    const double be = beta[g / len];
    const cplx<double> zi = z[g], pi = p[g];
    const cplx<double> np = mk(zi.x + be * pi.x, zi.y + be * pi.y);
    p[g] = np;
    if (pT) pT[g] = mk((T)np.x, (T)np.y);
  }
}

// PCG scalars of one iteration (wiener_filter.py:84, 94), per right-hand
// side, with an optional active mask (inactive: alpha = beta = 0 and res kept):
//   alpha = res / (pAp != 0 ? pAp : 1)
__global__ void k_cg_alpha(const double* __restrict__ res, const double* __restrict__ pAp,
                           const unsigned char* __restrict__ act, double* __restrict__ alpha,
                           int nrhs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrhs) return;
  const double d = pAp[k];
  alpha[k] = (!act || act[k]) ? res[k] / (d != 0.0 ? d : 1.0) : 0.0;
}

//   beta = rz / (res != 0 ? res : 1) ; res = rz   (active right-hand sides)
__global__ void k_cg_beta(double* __restrict__ res, const double* __restrict__ rz,
                          const unsigned char* __restrict__ act, double* __restrict__ beta,
                          int nrhs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrhs) return;
  const bool on = !act || act[k];
  const double r0 = res[k];
  beta[k] = on ? rz[k] / (r0 != 0.0 ? r0 : 1.0) : 0.0;
  if (on) res[k] = rz[k];
}

//   rel = sqrt(rr / bb)
__global__ void k_cg_rel(const double* __restrict__ rr, const double* __restrict__ bb,
                         double* __restrict__ rel, int nrhs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrhs) return;
  rel[k] = sqrt(rr[k] / bb[k]);
}

// ---------------------------------------------------------------------------
// Column passes with the per-mode operators at their edges.  The posterior's
// chains map -> rfft2 -> per-mode operator -> irfft2 -> map and field ->
// per-mode operator -> irfft2 (and back) are run as row R2C / C2R passes
// (k_row_r2c, k_row_c2r) around ONE column pass that applies the operator
// between its column FFT and the layout change, instead of separate rfft2,
// rotation and irfft2 passes over the half planes.
//
// Packed row spectra P [nfields][C][N][N/2]: the row R2C of real maps, with
// column 0 holding DC + i Nyquist (k_row_r2c).  Natural half planes X
// [nfields][C][N][N/2+1] (numpy's rfft2 layout).  The column FFT of the packed
// column 0 gives Z = Y0 + i Yn; Y0[kx] = (Z[kx] + conj Z[-kx]) / 2 and
// Yn[kx] = (Z[kx] - conj Z[-kx]) / 2i are the ky = 0 and ky = N/2 modes, and the
// inverse way a column enters as Herm(Y0) + i Herm(Yn), Herm(Y)[kx] =
// (Y[kx] + conj Y[-kx]) / 2: the part the row C2R keeps (k_col_fwd_unpack,
// k_col_inv_pack).  Workgroups of tile 0 hold column 0; its work is spread
// over their column lanes and handed to column 0's own lanes through LDS
// (col0); the branch on the tile is uniform over a workgroup.
//
// An operator acts in place on the C values of one mode (field f, kx, ky),
// a[c] = component c; the caller keeps component cu only, so an operator may
// skip work that feeds the other components alone:
//   OpDiagRot  C = 2: eb2qu(diag .* qu2eb(a)); C = 1: diag .* a
//              (the noise / beam operators of the pixel-mask posterior,
//               matrix_operators.py:54-90 with util.py:140-161)
//   OpData     C = 2: eb2qu(bmcn .* (d - mb .* qu2eb(a))); C = 1: likewise
//              without the rotations (k_wf_data; gradients.py:16-22)
//   OpEb2Qu    eb2qu(a) (util.py:140-150)
//   OpFinal    C = 2: qu2eb(a) - cfinv .* Herm(F); C = 1: a - cfinv .* Herm(F)
//              (gradients.py:25; the prior term through herm_col)
// ---------------------------------------------------------------------------

template <typename T, int C>
struct OpDiagRot {
  const T* diag;
  int N;
  __device__ __forceinline__ void operator()(long, int kx, int ky, cplx<T> (&a)[C], int) const {
    const long PN = (long)N * (N / 2 + 1);
    const long idx = (long)kx * (N / 2 + 1) + ky;
    if constexpr (C == 1) {
      a[0] = scale(a[0], diag[idx]);
    } else {
      T cs, sn;
      rot2psi_ij<T>(N, kx, ky, cs, sn);
      cplx<T> e, b;
      qu2eb_mode(a[0], a[1], cs, sn, e, b);
      e = scale(e, diag[idx]);
      b = scale(b, diag[PN + idx]);
      eb2qu_mode(e, b, cs, sn, a[0], a[1]);
    }
  }
};

template <typename T, int C>
struct OpData {
  const cplx<T>* d;  // [nfields][C][N][N/2+1] or NULL
  const T* mb;
  const T* bmcn;
  int N;
  __device__ __forceinline__ void operator()(long f, int kx, int ky, cplx<T> (&a)[C], int) const {
    const long PN = (long)N * (N / 2 + 1);
    const long idx = (long)kx * (N / 2 + 1) + ky;
    const long base = f * C * PN + idx;
    if constexpr (C == 1) {
      cplx<T> r = scale(a[0], -mb[idx]);
      if (d) r = r + d[base];
      a[0] = scale(r, bmcn[idx]);
    } else {
      T cs, sn;
      rot2psi_ij<T>(N, kx, ky, cs, sn);
      cplx<T> e, b;
      qu2eb_mode(a[0], a[1], cs, sn, e, b);
      cplx<T> re = scale(e, -mb[idx]), rb = scale(b, -mb[PN + idx]);
      if (d) {
        re = re + d[base];
        rb = rb + d[base + PN];
      }
      re = scale(re, bmcn[idx]);
      rb = scale(rb, bmcn[PN + idx]);
      eb2qu_mode(re, rb, cs, sn, a[0], a[1]);
    }
  }
};

template <typename T, int C>
struct OpEb2Qu {
  int N;
  __device__ __forceinline__ void operator()(long, int kx, int ky, cplx<T> (&a)[C], int) const {
    if constexpr (C == 2) {
      T cs, sn;
      rot2psi_ij<T>(N, kx, ky, cs, sn);
      const cplx<T> e = a[0], b = a[1];
      eb2qu_mode(e, b, cs, sn, a[0], a[1]);
    }
  }
};

template <typename T, int C>
struct OpFinal {
  const cplx<T>* F;  // [nfields][C][N][N/2+1] or NULL (no prior term)
  const T* cfinv;
  int N;
  __device__ __forceinline__ void operator()(long f, int kx, int ky, cplx<T> (&a)[C], int cu) const {
    const long PN = (long)N * (N / 2 + 1);
    const long idx = (long)kx * (N / 2 + 1) + ky;
    if constexpr (C == 2) {
      T cs, sn;
      rot2psi_ij<T>(N, kx, ky, cs, sn);
      const cplx<T> q = a[0], u = a[1];
      qu2eb_mode(q, u, cs, sn, a[0], a[1]);
    }
    if (F && cfinv) {
      // the prior per component: only the component the caller keeps (cu)
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (c == cu)
          a[c] = a[c] - scale(herm_col(F, (f * C + c) * PN, N, kx, ky), cfinv[c * PN + idx]);
    }
  }
};

template <typename T, int C>
__device__ __forceinline__ cplx<T> pick(const cplx<T> (&a)[C], int c) {
  if constexpr (C == 1)
    return a[0];
  else
    return c == 0 ? a[0] : a[1];
}

// One column pass of an operator edge.  A workgroup holds KC columns of all C
// components of one field; thread (c, t, g) -- g fastest, so that consecutive
// lanes read consecutive columns -- transforms line (component c, column
// tile KC + g) with points t + q N/8, q < 8.  The per-mode operator mixes the
// components, which meet through LDS: after the forward transform every line
// is written to its LDS buffer and each thread reads the C values of its
// modes there.  Column 0 (packed DC + i Nyquist) reads Z[kx] and Z[-kx] and
// evaluates the operator at (kx, 0), (kx, N/2) -- and, for the repacked
// output, at (-kx, 0), (-kx, N/2) -- so no second exchange of Z is needed.
//   EDGE 0: packed -> op -> packed, in place (times scale_)
//   EDGE 1: natural -> op -> packed (times scale_)
//   EDGE 2: packed -> op -> natural (forward; scale_ unused)
// XP: tiles narrower than a 128-byte line are taken in pairs; blocks b and
// b + 8 (dispatched to the same XCD, which has its own L2) hold the two tiles
// of a pair, so each line is fetched from HBM once.  The grid is a multiple
// of 16 blocks.
template <typename T, int N, int KC, int C, int EDGE, typename Op, bool XP = false>
__global__ __launch_bounds__(C * KC * N / 8) void k_col_edge(const cplx<T>* in, cplx<T>* out,
                                                             Op op,
                                                             const cplx<T>* __restrict__ twg,
                                                             T scale_) {
  constexpr int G = N / 8, H = N / 2, TILES = H / KC;
  constexpr long PN = (long)N * (H + 1), PP = (long)N * H;
  __shared__ cplx<T> lbuf[C][KC][line_stride<N>()];
  __shared__ cplx<T> col0[EDGE == 2 ? 1 : C][EDGE == 2 ? 1 : lds_len<N, false>()];
  const int g = threadIdx.x % KC, t = (threadIdx.x / KC) % G, c = threadIdx.x / (KC * G);
  Twiddles<N, T> tw;
  tw.load(twg, t);
  long f;
  int tile;
  if constexpr (XP) {
    const long b = blockIdx.x, pair = (b / 16) * 8 + b % 8;
    f = pair / (TILES / 2);
    tile = (int)(pair % (TILES / 2)) * 2 + (int)((b % 16) / 8);
  } else {
    f = blockIdx.x / TILES;
    tile = blockIdx.x % TILES;
  }
  const int k = tile * KC + g;
  auto L = [&](int cc, int kx) -> cplx<T>& { return lbuf[cc][g][lds_idx<N, false>(kx)]; };
  cplx<T>* const bufs[1] = {lbuf[c][g]};
  const T h = T(0.5);
  cplx<T> v[1][8];
  if constexpr (EDGE != 1) {
    const cplx<T>* src = in + (f * C + c) * PP + k;
#pragma unroll
    for (int q = 0; q < 8; ++q) v[0][q] = src[(long)(t + q * G) * H];
    fft_lines<N, false, 1, false, false>(v, bufs, tw, t);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) L(c, t + q * G) = v[0][q];
    __syncthreads();
    if (k != 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        cplx<T> a[C];
#pragma unroll
        for (int cc = 0; cc < C; ++cc) a[cc] = L(cc, t + q * G);
        op(f, t + q * G, k, a, c);
        v[0][q] = pick<T, C>(a, c);
      }
    }
  } else {
    const cplx<T>* src = in + (f * C + c) * PN + k;
    if (k != 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[0][q] = src[(long)(t + q * G) * (H + 1)];
    }
    __syncthreads();
    if (k != 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) L(c, t + q * G) = v[0][q];
    }
    __syncthreads();
    if (k != 0) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        cplx<T> a[C];
#pragma unroll
        for (int cc = 0; cc < C; ++cc) a[cc] = L(cc, t + q * G);
        op(f, t + q * G, k, a, c);
        v[0][q] = pick<T, C>(a, c);
      }
    }
  }
  // Column 0, spread over the workgroup's KC column lanes of tile 0: lane g
  // takes kx = t + q N/8 for q = g, g + KC, ... (so no lane chains more than
  // 8 / KC of its operator loads); its results go through LDS (col0) to the
  // column's own lanes, or straight out.
  if (tile == 0) {
    const cplx<T>* nat = in + f * C * PN;  // EDGE 1: natural input
    for (int q = g; q < 8; q += KC) {
      const int kx = t + q * G, mx = (N - kx) & (N - 1);
      cplx<T> a0[C], an[C], b0[C], bn[C];
#pragma unroll
      for (int cc = 0; cc < C; ++cc) {
        if constexpr (EDGE == 1) {
          a0[cc] = nat[cc * PN + (long)kx * (H + 1)];
          b0[cc] = nat[cc * PN + (long)mx * (H + 1)];
          an[cc] = nat[cc * PN + (long)kx * (H + 1) + H];
          bn[cc] = nat[cc * PN + (long)mx * (H + 1) + H];
        } else {
          const cplx<T> z = lbuf[cc][0][lds_idx<N, false>(kx)];
          const cplx<T> zm = lbuf[cc][0][lds_idx<N, false>(mx)];
          const cplx<T> d = z - conj(zm), dm = zm - conj(z);
          a0[cc] = scale(z + conj(zm), h);
          an[cc] = mk(h * d.y, -h * d.x);
          b0[cc] = scale(zm + conj(z), h);
          bn[cc] = mk(h * dm.y, -h * dm.x);
        }
      }
      op(f, kx, 0, a0, c);
      op(f, kx, H, an, c);
      if constexpr (EDGE == 2) {
        cplx<T>* o = out + (f * C + c) * PN + (long)kx * (H + 1);
        o[0] = pick<T, C>(a0, c);
        o[H] = pick<T, C>(an, c);
      } else {
        op(f, mx, 0, b0, c);
        op(f, mx, H, bn, c);
        const cplx<T> y0 = scale(pick<T, C>(a0, c) + conj(pick<T, C>(b0, c)), h);
        const cplx<T> yn = scale(pick<T, C>(an, c) + conj(pick<T, C>(bn, c)), h);
        col0[c][lds_idx<N, false>(kx)] = mk(y0.x - yn.y, y0.y + yn.x);
      }
    }
  }
  if constexpr (EDGE == 2) {
    if (k != 0) {
      cplx<T>* dst = out + (f * C + c) * PN + k;
#pragma unroll
      for (int q = 0; q < 8; ++q) dst[(long)(t + q * G) * (H + 1)] = v[0][q];
    }
  } else {
    if (tile == 0) {  // uniform over the workgroup
      __syncthreads();
      if (k == 0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[0][q] = col0[c][lds_idx<N, false>(t + q * G)];
      }
    }
    fft_lines<N, true, 1, false, false>(v, bufs, tw, t);
    cplx<T>* dst = out + (f * C + c) * PP + k;
#pragma unroll
    for (int q = 0; q < 8; ++q) dst[(long)(t + q * G) * H] = scale(v[0][q], scale_);
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

template <typename T>
hipError_t launch_wf_data(const cplx<T>* X, const cplx<T>* d, const T* mb, const T* bmcn,
                          cplx<T>* out, int N, int C, long nfields, hipStream_t s) {
  const long n = nfields * (long)N * (N / 2 + 1);
  k_wf_data<T><<<grid_for(n), 256, 0, s>>>(X, d, mb, bmcn, out, N, C, nfields);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_zero(T* x, long n, hipStream_t s) {
  k_fill_zero<T><<<grid_for(n), 256, 0, s>>>(x, n);
  return hipGetLastError();
}


template <typename Tb>
hipError_t launch_dot(const cplx<double>* a, const cplx<Tb>* b, int N, long len, int nrhs,
                      double* part, double* out, hipStream_t s) {
  const double sc = 1.0 / ((double)N * (double)N);
  k_dot_partial<Tb><<<dim3(kRedParts, nrhs), kRedBlock, 0, s>>>(a, b, N, len, part);
  k_dot_final<<<(nrhs + 63) / 64, 64, 0, s>>>(part, sc, out, nrhs);
  return hipGetLastError();
}

template <typename Tb>
hipError_t launch_cg_xrz(cplx<double>* x, cplx<double>* r, cplx<double>* z,
                         const cplx<double>* p, const cplx<Tb>* Ap, const double* pinv,
                         const double* alpha, int N, int C, long len, int nrhs, double* part,
                         double* rz, hipStream_t s) {
  const double sc = 1.0 / ((double)N * (double)N);
  if (C == 3)
    k_cg_xrz_teb<Tb><<<dim3(kRedParts, nrhs), kRedBlock, 0, s>>>(x, r, z, p, Ap, pinv, alpha, N,
                                                                part);
  else
    k_cg_xrz<Tb><<<dim3(kRedParts, nrhs), kRedBlock, 0, s>>>(x, r, z, p, Ap, pinv, alpha, N, C,
                                                            len, part);
  k_dot_final<<<(nrhs + 63) / 64, 64, 0, s>>>(part, sc, rz, nrhs);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cg_p(cplx<double>* p, const cplx<double>* z, const double* beta, cplx<T>* pT,
                       long len, int nrhs, hipStream_t s) {
  k_cg_p<T><<<grid_for(len * nrhs), 256, 0, s>>>(p, z, beta, pT, len, nrhs);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_teb_split(const cplx<T>* f, cplx<T>* XT, cplx<T>* XP, int N, long nfields,
                            hipStream_t s) {
  k_teb_split<T><<<grid_for(nfields * (long)N * (N / 2 + 1)), 256, 0, s>>>(f, XT, XP, N, nfields);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_wf_data_teb(const cplx<T>* XT, const cplx<T>* XP, const cplx<T>* d, const T* mb,
                              const T* bmcn, cplx<T>* outT, cplx<T>* outP, int N, long nfields,
                              hipStream_t s) {
  k_wf_data_teb<T><<<grid_for(nfields * (long)N * (N / 2 + 1)), 256, 0, s>>>(
      XT, XP, d, mb, bmcn, outT, outP, N, nfields);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_wf_final_teb(const cplx<T>* XLT, const cplx<T>* XLP, const cplx<T>* F,
                               const T* cfinv, cplx<T>* out, int N, long nfields, hipStream_t s) {
  k_wf_final_teb<T><<<grid_for(nfields * (long)N * (N / 2 + 1)), 256, 0, s>>>(
      XLT, XLP, F, cfinv, out, N, nfields);
  return hipGetLastError();
}

// Columns per workgroup of the operator column passes: 128-byte row segments
// (16 float / 8 double complex), halved for Q/U pairs at N = 512 so that a
// workgroup (C KC N / 8 threads, C KC line buffers) stays within 1,024
// threads and half the LDS.
template <typename T, int N, int C>
constexpr int col_op_kc() {
  constexpr int base = sizeof(T) == 4 ? 16 : 8;
  constexpr int kc = (C == 2 && N >= 512) ? base / 2 : base;
  return kc < N / 2 ? kc : N / 2;
}

#define CMBL_COL_DISPATCH(N, ...)                              \
  switch (N) {                                                 \
    case 16: { constexpr int NN_ = 16; __VA_ARGS__; } break;   \
    case 32: { constexpr int NN_ = 32; __VA_ARGS__; } break;   \
    case 64: { constexpr int NN_ = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int NN_ = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int NN_ = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int NN_ = 512; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;                      \
  }

template <typename T, int C, int EDGE, template <typename, int> class OpT>
static hipError_t col_edge(int N, const cplx<T>* in, cplx<T>* out, const OpT<T, C>& op,
                           long nfields, const cplx<T>* tw, hipStream_t s) {
  CMBL_COL_DISPATCH(N, {
    constexpr int KC = col_op_kc<T, NN_, C>(), TILES = NN_ / 2 / KC;
    constexpr bool PAIRABLE = KC * sizeof(cplx<T>) < 128 && TILES % 2 == 0;
    const T sc = T(1) / (T(NN_) * T(NN_));
    const long nb = nfields * TILES;
    if (PAIRABLE && nb % 16 == 0)
      k_col_edge<T, NN_, KC, C, EDGE, OpT<T, C>, PAIRABLE>
          <<<(unsigned)nb, C * KC * NN_ / 8, 0, s>>>(in, out, op, tw, sc);
    else
      k_col_edge<T, NN_, KC, C, EDGE, OpT<T, C>>
          <<<(unsigned)nb, C * KC * NN_ / 8, 0, s>>>(in, out, op, tw, sc);
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_col_op(int N, cplx<T>* P, const T* diag, long nfields, int C,
                         const cplx<T>* tw, hipStream_t s) {
  if (C == 2) return col_edge<T, 2, 0, OpDiagRot>(N, P, P, {diag, N}, nfields, tw, s);
  if (C == 1) return col_edge<T, 1, 0, OpDiagRot>(N, P, P, {diag, N}, nfields, tw, s);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_col_data(int N, cplx<T>* P, const cplx<T>* d, const T* mb, const T* bmcn,
                           long nfields, int C, const cplx<T>* tw, hipStream_t s) {
  if (C == 2) return col_edge<T, 2, 0, OpData>(N, P, P, {d, mb, bmcn, N}, nfields, tw, s);
  if (C == 1) return col_edge<T, 1, 0, OpData>(N, P, P, {d, mb, bmcn, N}, nfields, tw, s);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_col_inv_eb2qu(int N, const cplx<T>* X, cplx<T>* P, long nfields, int C,
                                const cplx<T>* tw, hipStream_t s) {
  if (C == 2) return col_edge<T, 2, 1, OpEb2Qu>(N, X, P, {N}, nfields, tw, s);
  if (C == 1) return col_edge<T, 1, 1, OpEb2Qu>(N, X, P, {N}, nfields, tw, s);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_col_fwd_final(int N, const cplx<T>* P, const cplx<T>* F, const T* cfinv,
                                cplx<T>* out, long nfields, int C, const cplx<T>* tw,
                                hipStream_t s) {
  if (C == 2) return col_edge<T, 2, 2, OpFinal>(N, P, out, {F, cfinv, N}, nfields, tw, s);
  if (C == 1) return col_edge<T, 1, 2, OpFinal>(N, P, out, {F, cfinv, N}, nfields, tw, s);
  return hipErrorInvalidValue;
}

#undef CMBL_COL_DISPATCH

hipError_t launch_cg_alpha(const double* res, const double* pAp, const unsigned char* act,
                           double* alpha, int nrhs, hipStream_t s) {
  k_cg_alpha<<<(nrhs + 63) / 64, 64, 0, s>>>(res, pAp, act, alpha, nrhs);
  return hipGetLastError();
}

hipError_t launch_cg_beta(double* res, const double* rz, const unsigned char* act, double* beta,
                          int nrhs, hipStream_t s) {
  k_cg_beta<<<(nrhs + 63) / 64, 64, 0, s>>>(res, rz, act, beta, nrhs);
  return hipGetLastError();
}

hipError_t launch_cg_rel(const double* rr, const double* bb, double* rel, int nrhs,
                         hipStream_t s) {
  k_cg_rel<<<(nrhs + 63) / 64, 64, 0, s>>>(rr, bb, rel, nrhs);
  return hipGetLastError();
}

int red_parts() { return kRedParts; }

#define CMBL_WF_INST(T)                                                                         \
  template hipError_t launch_col_op<T>(int, cplx<T>*, const T*, long, int, const cplx<T>*,       \
                                       hipStream_t);                                            \
  template hipError_t launch_col_data<T>(int, cplx<T>*, const cplx<T>*, const T*, const T*, long, \
                                         int, const cplx<T>*, hipStream_t);                     \
  template hipError_t launch_col_inv_eb2qu<T>(int, const cplx<T>*, cplx<T>*, long, int,          \
                                              const cplx<T>*, hipStream_t);                     \
  template hipError_t launch_col_fwd_final<T>(int, const cplx<T>*, const cplx<T>*, const T*,     \
                                              cplx<T>*, long, int, const cplx<T>*, hipStream_t); \
  template hipError_t launch_wf_data<T>(const cplx<T>*, const cplx<T>*, const T*, const T*,     \
                                        cplx<T>*, int, int, long, hipStream_t);                 \
  template hipError_t launch_fill_zero<T>(T*, long, hipStream_t);                               \
  template hipError_t launch_dot<T>(const cplx<double>*, const cplx<T>*, int, long, int,        \
                                    double*, double*, hipStream_t);                             \
  template hipError_t launch_cg_xrz<T>(cplx<double>*, cplx<double>*, cplx<double>*,             \
                                       const cplx<double>*, const cplx<T>*, const double*,      \
                                       const double*, int, int, long, int, double*, double*,    \
                                       hipStream_t);                                            \
  template hipError_t launch_cg_p<T>(cplx<double>*, const cplx<double>*, const double*,         \
                                     cplx<T>*, long, int, hipStream_t);                         \
  template hipError_t launch_teb_split<T>(const cplx<T>*, cplx<T>*, cplx<T>*, int, long,        \
                                          hipStream_t);                                         \
  template hipError_t launch_wf_data_teb<T>(const cplx<T>*, const cplx<T>*, const cplx<T>*,     \
                                            const T*, const T*, cplx<T>*, cplx<T>*, int, long,  \
                                            hipStream_t);                                       \
  template hipError_t launch_wf_final_teb<T>(const cplx<T>*, const cplx<T>*, const cplx<T>*,    \
                                             const T*, cplx<T>*, int, long, hipStream_t);
CMBL_WF_INST(float)
CMBL_WF_INST(double)

}  // namespace cmbl
