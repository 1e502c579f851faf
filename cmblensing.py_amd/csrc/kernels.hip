// HIP kernels for the LenseFlow hot path on gfx950.
//
// Array layouts (row-major, contiguous):
//   map      x[m][i][j]          m = batch*C + component, i = row (KX axis),
//                                 j = column (KY axis, the rfft axis)
//   natural  X[m][kx][ky]        ky in [0, N/2], numpy rfft2 layout
//   packed   P[m][kx][ky]        ky in [0, N/2): slot ky=0 holds the DC column
//                                 in .x and the ky=N/2 (Nyquist) column in .y.
//                                 Used between the row and column passes: after
//                                 the row R2C both columns are real per row, so
//                                 one complex column carries both and the half
//                                 plane becomes a power-of-two N x N/2.
//
// Conventions are numpy's: forward unnormalised, inverse 1/N^2, the C2R along
// rows ignores the imaginary parts of the DC and Nyquist columns, KX from
// fftfreq (kx = N/2 maps to -N/2), KY from rfftfreq with the Nyquist negated.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "fft_lds.hpp"
#include "kernels.hpp"

// What this translation unit instantiates, so that the build can compile the
// FP32 and FP64 kernels concurrently: 0 everything, 1 the float kernels and
// the phi preparation, 2 the double kernels.
#ifndef CMBL_KERNELS_PART
#define CMBL_KERNELS_PART 0
#endif

namespace cmbl {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// Accesses of the RK4 accumulator and state that are next read two or more
// launches later: with NT (a call whose stage buffers take at least half the
// last-level cache) they are streaming (non-temporal) accesses, which leave
// the cache to the stage spectra and the shared p(t) rows.
template <bool NT, typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <int N>
__device__ __forceinline__ int kx_signed(int kx) {
  return kx < N / 2 ? kx : kx - N;
}

// i * s * F
template <typename T>
__device__ __forceinline__ cplx<T> imul(cplx<T> F, T s) {
  return mk(-s * F.y, s * F.x);
}

// The row kernels give each line N/8 <= 64 threads in groups aligned to the
// wavefront, so their LDS exchanges synchronise at wave scope (lds_sync).
#ifndef CMBL_WAVE_SYNC
#define CMBL_WAVE_SYNC 1
#endif
constexpr bool kRowWave = CMBL_WAVE_SYNC != 0;

// The stage-2/4 standard-rate row pass (RK = 2) runs four line transforms and
// needs more registers than 4 waves per SIMD allow when its twiddles stay in
// registers; at FP32 it reads them from an LDS table instead (2 KB at
// N = 256, 4 KB at N = 512: 4 blocks of 256 threads still fit a CU's 160 KB
// of LDS).
#ifndef CMBL_TW_LDS
#define CMBL_TW_LDS 1
#endif
template <typename T, int N, int RK>
constexpr bool row_tw_lds() {
  return CMBL_TW_LDS != 0 && RK == 2 && sizeof(T) == 4;
}
// The column pass of the LenseFlow stages (two transforms per column) holds
// 13 twiddles in registers; at FP32, N = 256 it reads them from an LDS table
// instead (2 KB beside 37 KB of line buffers), which takes it from 79 to 56
// VGPRs and from 3 to 4 blocks per CU.
#ifndef CMBL_COL_TW_LDS
#define CMBL_COL_TW_LDS 1
#endif
template <typename T, int N>
constexpr bool col_tw_lds() {
  return CMBL_COL_TW_LDS != 0 && sizeof(T) == 4 && N == 256;
}

// Row pairing used by the fused stage kernels.  Pair index p covers two
// map-rows: for C == 2 the Q and U rows of the same pixel row (they share the
// deflection operator rows), for C == 1 the pixel rows 2r and 2r+1.  The
// pairs of a workgroup are consecutive rows of one batch element (numbering
// the batch element fastest, so that a workgroup takes one pixel row of
// consecutive elements, measured slower: 58.2k against 60.7k maps/s at
// 256^2 QU batch 128 ordered by phi -- the rows of one element are
// contiguous in HBM).
struct RowPair {
  int b;       // batch element
  int ia, ib;  // pixel rows
  long ra, rb; // flat row indices into [B*C*N]
};
template <int N>
__device__ __forceinline__ RowPair row_pair(int p, int C) {
  RowPair r;
  if (C == 2) {
    r.b = p / N;
    int i = p - r.b * N;
    r.ia = r.ib = i;
    r.ra = (long)(2 * r.b) * N + i;
    r.rb = r.ra + N;
  } else {
    r.b = p / (N / 2);
    int i = 2 * (p - r.b * (N / 2));
    r.ia = i;
    r.ib = i + 1;
    r.ra = (long)r.b * N + i;
    r.rb = r.ra + 1;
  }
  return r;
}

// Place the values a = A[k], b = B[k] of two packed Hermitian half rows into
// the full complex line Z = A + i B held in LDS.  The packed slot k = 0 carries
// the DC (.x) and Nyquist (.y) values; their imaginary parts are dropped,
// which is the C2R convention.
template <int N, typename T>
__device__ __forceinline__ void put_herm(cplx<T>* buf, int k, cplx<T> a, cplx<T> b) {
  if (k == 0) {
    buf[lds_idx<N>(0)] = mk(a.x, b.x);
    buf[lds_idx<N>(N / 2)] = mk(a.y, b.y);
  } else {
    buf[lds_idx<N>(k)] = mk(a.x - b.y, a.y + b.x);      // a + i b
    buf[lds_idx<N>(N - k)] = mk(a.x + b.y, b.x - a.y);  // conj(a) + i conj(b)
  }
}

// Z = A + i B of two packed half rows into LDS, then the 8 register points of
// thread t.
template <int N, bool WAVE = false, typename T>
__device__ __forceinline__ void herm_pair_to_line(const cplx<T>* A, const cplx<T>* B, cplx<T>* buf,
                                                  cplx<T> (&v)[8], int t) {
  constexpr int G = N / 8;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = t + q * G;
    put_herm<N>(buf, k, A[k], B[k]);
  }
  lds_sync<WAVE>();
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = buf[lds_idx<N>(t + q * G)];
}

// After a forward FFT of z = a + i b (v[m] = Z[t + m*G]) split into the two
// packed half spectra and store them to A and B.
template <int N, bool WAVE = false, typename T>
__device__ __forceinline__ void split_store(const cplx<T> (&v)[8], cplx<T>* buf, cplx<T>* A,
                                            cplx<T>* B, int t, bool active) {
  constexpr int G = N / 8;
  lds_sync<WAVE>();
#pragma unroll
  for (int m = 0; m < 8; ++m) buf[lds_idx<N>(t + m * G)] = v[m];
  lds_sync<WAVE>();
  if (!active) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = t + q * G;
    cplx<T> zk = buf[lds_idx<N>(k)];
    cplx<T> a, b;
    if (k == 0) {
      cplx<T> zn = buf[lds_idx<N>(N / 2)];
      a = mk(zk.x, zn.x);
      b = mk(zk.y, zn.y);
    } else {
      cplx<T> zc = conj(buf[lds_idx<N>(N - k)]);
      const T h = T(0.5);
      a = scale(zk + zc, h);
      cplx<T> d = zk - zc;
      b = mk(h * d.y, -h * d.x);  // -i/2 * d
    }
    A[k] = a;
    B[k] = b;
  }
}

// ---------------------------------------------------------------------------
// generic rfft2 / irfft2 passes
// ---------------------------------------------------------------------------

// Row R2C of row pairs (2p, 2p+1): x[nrows][N] -> packed P[nrows][N/2].
template <typename T, int N>
__global__ __launch_bounds__(256) void k_row_r2c(const T* __restrict__ x, cplx<T>* __restrict__ P,
                                                 const cplx<T>* __restrict__ twg, long npairs) {
  constexpr int G = N / 8, PAIRS = 256 / G, H = N / 2;
  __shared__ cplx<T> lbuf[PAIRS][lds_len<N>()];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const long pp = active ? p : npairs - 1;
  const T* xa = x + (2 * pp) * N;
  const T* xb = xa + N;
  cplx<T> v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = mk(xa[t + q * G], xb[t + q * G]);
  lds_sync<kRowWave>();
  cplx<T>* const bufs[1] = {lbuf[grp]};
  fft_lines<N, false, 1, kRowWave>(v, bufs, tw, t);
  split_store<N, kRowWave>(v[0], lbuf[grp], P + (2 * pp) * H, P + (2 * pp + 1) * H, t, active);
}

// Row R2C of a LenseFlow input taken from per-request maps (the serving
// executor's first launch): row pair p covers rows 2p, 2p+1 of map
// m = 2p / N, i.e. of request m / ncomp, component m % ncomp, found at
// src[request] + component N^2 (NULL: zeros, a padding slot).  The rows are
// also stored to the staging state ystage [nmaps][N][N] (the RK4 state y).
template <typename T, int N>
__global__ __launch_bounds__(256) void k_row_r2c_gather(const T* const* __restrict__ src, int ncomp,
                                                        T* __restrict__ ystage,
                                                        cplx<T>* __restrict__ P,
                                                        const cplx<T>* __restrict__ twg,
                                                        long npairs) {
  constexpr int G = N / 8, PAIRS = 256 / G, H = N / 2;
  __shared__ cplx<T> lbuf[PAIRS][lds_len<N>()];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const long pp = active ? p : npairs - 1;
  const long row0 = 2 * pp;
  const long m = row0 / N;
  const T* base = src[m / ncomp];
  const T* xa = base ? base + ((m % ncomp) * (long)N + row0 % N) * N : nullptr;
  cplx<T> v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q)
    v[0][q] = xa ? mk(xa[t + q * G], xa[N + t + q * G]) : mk(T(0), T(0));
  if (active) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      ystage[row0 * N + t + q * G] = v[0][q].x;
      ystage[(row0 + 1) * N + t + q * G] = v[0][q].y;
    }
  }
  lds_sync<kRowWave>();
  cplx<T>* const bufs[1] = {lbuf[grp]};
  fft_lines<N, false, 1, kRowWave>(v, bufs, tw, t);
  split_store<N, kRowWave>(v[0], lbuf[grp], P + row0 * H, P + (row0 + 1) * H, t, active);
}

// Row C2R of row pairs: packed P[nrows][N/2] -> x[nrows][N] (no scaling).
// EP, a pixel-space epilogue with w [N][N] shared by all maps:
//   0: x;  1: x = w .* x;  2: x = d - w .* x  (d NULL: zeros)
// (the pixel mask's product and the data residual of the pixel-mask
// posterior, applied as the rows are stored).
template <typename T, int N, int EP = 0>
__global__ __launch_bounds__(256) void k_row_c2r(const cplx<T>* __restrict__ P, T* __restrict__ x,
                                                 const cplx<T>* __restrict__ twg, long npairs,
                                                 const T* __restrict__ d = nullptr,
                                                 const T* __restrict__ w = nullptr) {
  constexpr int G = N / 8, PAIRS = 256 / G, H = N / 2;
  __shared__ cplx<T> lbuf[PAIRS][lds_len<N>()];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const long pp = active ? p : npairs - 1;
  cplx<T> v[1][8];
  lds_sync<kRowWave>();
  herm_pair_to_line<N, kRowWave>(P + (2 * pp) * H, P + (2 * pp + 1) * H, lbuf[grp], v[0], t);
  cplx<T>* const bufs[1] = {lbuf[grp]};
  fft_lines<N, true, 1, kRowWave>(v, bufs, tw, t);
  if (!active) return;
  T* xa = x + (2 * pp) * N;
  T* xb = xa + N;
  if constexpr (EP != 0) {
    const T* wa = w + ((2 * pp) % N) * N;
    const T* da = (EP == 2 && d) ? d + (2 * pp) * N : nullptr;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int i = t + m * G;
      T a = v[0][m].x * wa[i], b = v[0][m].y * wa[N + i];
      if constexpr (EP == 2) {
        a = (da ? da[i] : T(0)) - a;
        b = (da ? da[N + i] : T(0)) - b;
      }
      xa[i] = a;
      xb[i] = b;
    }
    return;
  }
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    xa[t + m * G] = v[0][m].x;
    xb[t + m * G] = v[0][m].y;
  }
}

// Column forward C2C of packed P[m][N][N/2] -> natural X[m][N][N/2+1].
template <typename T, int N, int KC>
__global__ __launch_bounds__(KC* N / 8) void k_col_fwd_unpack(const cplx<T>* __restrict__ P,
                                                              cplx<T>* __restrict__ X,
                                                              const cplx<T>* __restrict__ twg) {
  constexpr int G = N / 8, H = N / 2, TILES = H / KC;
  __shared__ cplx<T> lbuf[KC][line_stride<N>()];
  const int g = threadIdx.x % KC, t = threadIdx.x / KC;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long m = blockIdx.x / TILES;
  const int k = (blockIdx.x % TILES) * KC + g;
  const cplx<T>* Pm = P + m * N * H;
  cplx<T>* Xm = X + m * N * (H + 1);
  cplx<T> v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = Pm[(long)(t + q * G) * H + k];
  __syncthreads();
  cplx<T>* const bufs[1] = {lbuf[g]};
  fft_lines<N, false, 1, false, false>(v, bufs, tw, t);
  // the packed column needs G[-kx]: go through LDS (uniform for all groups)
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 8; ++q) lbuf[g][lds_idx<N, false>(t + q * G)] = v[0][q];
  __syncthreads();
  if (k == 0) {
    const T h = T(0.5);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kx = t + q * G;
      cplx<T> z = v[0][q];
      cplx<T> zc = conj(lbuf[g][lds_idx<N, false>((N - kx) & (N - 1))]);
      cplx<T> a = scale(z + zc, h);
      cplx<T> d = z - zc;
      cplx<T> b = mk(h * d.y, -h * d.x);
      Xm[(long)kx * (H + 1) + 0] = a;
      Xm[(long)kx * (H + 1) + H] = b;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) Xm[(long)(t + q * G) * (H + 1) + k] = v[0][q];
  }
}

// Column inverse C2C of natural X[m][N][N/2+1] -> packed P[m][N][N/2], times
// `scale`.  The DC and Nyquist columns enter as Herm(X0) + i Herm(Xn), with
// Herm(Y)[kx] = (Y[kx] + conj(Y[-kx]))/2: exactly the part the row C2R keeps.
template <typename T, int N, int KC>
__global__ __launch_bounds__(KC* N / 8) void k_col_inv_pack(const cplx<T>* __restrict__ X,
                                                            cplx<T>* __restrict__ P,
                                                            const cplx<T>* __restrict__ twg,
                                                            T scale_) {
  constexpr int G = N / 8, H = N / 2, TILES = H / KC;
  __shared__ cplx<T> lbuf[KC][line_stride<N>()];
  const int g = threadIdx.x % KC, t = threadIdx.x / KC;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long m = blockIdx.x / TILES;
  const int k = (blockIdx.x % TILES) * KC + g;
  const cplx<T>* Xm = X + m * N * (H + 1);
  cplx<T>* Pm = P + m * N * H;
  cplx<T> v[1][8];
  if (k == 0) {
    const T h = T(0.5);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int kx = t + q * G;
      const long r0 = (long)kx * (H + 1), r1 = (long)((N - kx) & (N - 1)) * (H + 1);
      cplx<T> d0 = scale(Xm[r0] + conj(Xm[r1]), h);
      cplx<T> n0 = scale(Xm[r0 + H] + conj(Xm[r1 + H]), h);
      v[0][q] = mk(d0.x - n0.y, d0.y + n0.x);
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) v[0][q] = Xm[(long)(t + q * G) * (H + 1) + k];
  }
  __syncthreads();
  cplx<T>* const bufs[1] = {lbuf[g]};
  fft_lines<N, true, 1, false, false>(v, bufs, tw, t);
#pragma unroll
  for (int q = 0; q < 8; ++q) Pm[(long)(t + q * G) * H + k] = scale(v[0][q], scale_);
}

// ---------------------------------------------------------------------------
// phi preparation (double precision)
// ---------------------------------------------------------------------------

#if CMBL_KERNELS_PART != 2
// Five derivative spectra of F (natural layout): i KX F, i KY F, -KX^2 F,
// -KX KY F, -KY^2 F, with KY's Nyquist entry negated (util.py:103-107).
__global__ void k_phi_mult(const cplx<double>* __restrict__ F, cplx<double>* __restrict__ D, int N,
                           double dk) {
  const int H1 = N / 2 + 1;
  const long n = (long)N * H1;
  F += blockIdx.y * n;          // phi number blockIdx.y of a batch
  D += blockIdx.y * 5 * n;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x) {
    const int kx = (int)(idx / H1), ky = (int)(idx % H1);
    const double KX = (kx < N / 2 ? kx : kx - N) * dk;
    const double KY = (ky == N / 2 ? -ky : ky) * dk;
    const cplx<double> f = F[idx];
    D[0 * n + idx] = mk(-KX * f.y, KX * f.x);
    D[1 * n + idx] = mk(-KY * f.y, KY * f.x);
    D[2 * n + idx] = scale(f, -(KX * KX));
    D[3 * n + idx] = scale(f, -(KX * KY));
    D[4 * n + idx] = scale(f, -(KY * KY));
  }
}
#endif

// Deflection-operator table: for t_j = j/(2n), j = 0..2n,
//   p(t_j) = (I + t_j Hess phi)^-1 grad phi        (lense_flow.py:372-381, 397-415)
// stored as T[(2n+1)][2][N*N].  Also the per-block minimum over pixels and
// t in [0,1] of det(I + t Hess phi) = 1 + t tr + t^2 det(Hess phi).
template <typename T>
__global__ __launch_bounds__(256) void k_ptab(const double* __restrict__ Dm, T* __restrict__ ptab,
                                              int N, int nsteps, double* __restrict__ minpart) {
  const long P = (long)N * N;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  // phi number blockIdx.y of a batch: its derivatives, table and partials
  Dm += blockIdx.y * 5 * P;
  ptab += blockIdx.y * (2L * nsteps + 1) * 2 * P;
  minpart += blockIdx.y * gridDim.x;
  double mdet = 1.0;
  if (idx < P) {
    const double px = Dm[idx], py = Dm[P + idx], hxx = Dm[2 * P + idx], hxy = Dm[3 * P + idx],
                 hyy = Dm[4 * P + idx];
    const int nt = 2 * nsteps;
    for (int j = 0; j <= nt; ++j) {
      const double t = (double)j / (double)nt;
      const double mxx = 1 + t * hxx, mxy = t * hxy, myy = 1 + t * hyy;
      const double det = mxx * myy - mxy * mxy;
      const double inv = 1.0 / (det + 1e-12);
      const double ixx = myy * inv, ixy = -mxy * inv, iyy = mxx * inv;
      ptab[(2L * j) * P + idx] = (T)(px * ixx + py * ixy);
      ptab[(2L * j + 1) * P + idx] = (T)(px * ixy + py * iyy);
    }
    const double tr = hxx + hyy, dh = hxx * hyy - hxy * hxy;
    mdet = fmin(1.0, 1.0 + tr + dh);
    if (dh > 0) {
      const double ts = -tr / (2 * dh);
      if (ts > 0 && ts < 1) mdet = fmin(mdet, 1.0 - tr * tr / (4 * dh));
    }
    if (!(mdet == mdet)) mdet = -1.0;  // NaN phi is rejected like a caustic
  }
  __shared__ double red[256];
  red[threadIdx.x] = mdet;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) minpart[blockIdx.x] = red[0];
}

// Inverse-magnification table for the gradient ODE: for t_j = j/(2n),
// (M^-1_xx, M^-1_xy, M^-1_yy)(t_j) (lense_flow.py:397-415), T[(2n+1)][3][N*N].
template <typename T>
__global__ __launch_bounds__(256) void k_gtab(const double* __restrict__ Dm, T* __restrict__ gtab,
                                              int N, int nsteps) {
  const long P = (long)N * N;
  const long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= P) return;
  const double hxx = Dm[2 * P + idx], hxy = Dm[3 * P + idx], hyy = Dm[4 * P + idx];
  const int nt = 2 * nsteps;
  for (int j = 0; j <= nt; ++j) {
    const double t = (double)j / (double)nt;
    const double mxx = 1 + t * hxx, mxy = t * hxy, myy = 1 + t * hyy;
    const double inv = 1.0 / (mxx * myy - mxy * mxy + 1e-12);
    gtab[(3L * j) * P + idx] = (T)(myy * inv);
    gtab[(3L * j + 1) * P + idx] = (T)(-mxy * inv);
    gtab[(3L * j + 2) * P + idx] = (T)(mxx * inv);
  }
}

// delta-phi spectrum from the five accumulated source maps' rfft2 (natural
// layout, F[b][5][N][N/2+1]):
//   dphi = i KX F1 + i KY F2 - KX^2 F3 - KX KY F4 - KY^2 F5
// with KY's Nyquist entry negated (lense_flow.py:233-251, util.py:103-107,133-137).
template <typename T>
__global__ void k_dphi_combine(const cplx<T>* __restrict__ F, cplx<T>* __restrict__ out, int N,
                               long nb, double dk) {
  const int H1 = N / 2 + 1;
  const long P = (long)N * H1;
  for (long g = blockIdx.x * (long)blockDim.x + threadIdx.x; g < nb * P;
       g += (long)gridDim.x * blockDim.x) {
    const long b = g / P, idx = g % P;
    const int kx = (int)(idx / H1), ky = (int)(idx % H1);
    const T KX = (T)((kx < N / 2 ? kx : kx - N) * dk);
    const T KY = (T)((ky == N / 2 ? -ky : ky) * dk);
    const cplx<T>* f = F + b * 5 * P + idx;
    const cplx<T> f1 = f[0], f2 = f[P], f3 = f[2 * P], f4 = f[3 * P], f5 = f[4 * P];
    cplx<T> r = imul(f1, KX) + imul(f2, KY);
    r = r - scale(f3, KX * KX) - scale(f4, KX * KY) - scale(f5, KY * KY);
    out[b * P + idx] = r;
  }
}

// ---------------------------------------------------------------------------
// LenseFlow fused stage kernels
// ---------------------------------------------------------------------------

// Column pass: the x-derivative of a packed row spectrum.  src holds the row
// R2C of a real map; column FFT, multiply by i KX / N^2, inverse column FFT,
// store to dst (may alias src).  The y-derivative needs no column transform
// (it is diagonal in the row transform) and is applied by the row pass.
// In the packed DC|Nyquist column only the DC part of d/dx survives the C2R
// (the Nyquist part's d/dy drops out) and the self-conjugate kx = N/2 bin of
// d/dx drops out, so that bin is zeroed.
//
// XP (XCD-paired half tiles): each workgroup takes KC columns, half of a
// 2 KC-column tile whose row segments are 128 bytes, and the two halves go to
// workgroups 8 apart -- dispatched to the same XCD, so a segment's cache line
// is fetched once into that XCD's L2 and both halves' partial-line stores
// merge there.  Twice the workgroups of the full-tile shape.  The grid must
// be a multiple of 16 workgroups.
// Tile b of the column pass -> map m and the column k of lane g.
template <int N, int KC, bool XP>
__device__ __forceinline__ void col_tile(long b, int g, long& m, int& k) {
  constexpr int H = N / 2, TILES = H / KC;
  if constexpr (XP) {
    const long seg = (b / 16) * 8 + b % 8;  // 2 KC-column tile
    constexpr int SEGS = H / (2 * KC);
    m = seg / SEGS;
    k = (int)(seg % SEGS) * 2 * KC + (int)((b % 16) / 8) * KC + g;
  } else {
    m = b / TILES;
    k = (int)(b % TILES) * KC + g;
  }
}

template <typename T, int N, int KC, bool XP = false>
__global__ __launch_bounds__(KC* N / 8) void k_col_lf_dx(const cplx<T>* src, cplx<T>* dst,
                                                         const cplx<T>* __restrict__ twg, T dk) {
  constexpr int G = N / 8, H = N / 2;
  constexpr bool TWL = col_tw_lds<T, N>();
  // the twiddle table first: at the bottom of LDS its entries are reached
  // with the immediate offsets of the LDS reads
  struct alignas(16) Lds {
    alignas(16) cplx<T> tw[TWL ? TwiddlesLds<N, T>::kLen : 1];
    alignas(16) cplx<T> line[KC][line_stride<N>()];
  };
  __shared__ Lds lds;
  auto& twl = lds.tw;
  auto& lbuf = lds.line;
  const int g = threadIdx.x % KC, t = threadIdx.x / KC;
  std::conditional_t<TWL, TwiddlesLds<N, T>, Twiddles<N, T>> tw;
  if constexpr (TWL) {
    fill_twiddles_lds<N>(twl, twg, threadIdx.x, KC * G);
    tw.tab = twl;  // the __syncthreads() below orders these stores
  } else {
    tw.load(twg, t);
  }
  long m;
  int k;
  col_tile<N, KC, XP>(blockIdx.x, g, m, k);
  // the map's base is uniform over the workgroup; the lane's point offsets
  // fit 32 bits (N * N/2 points per map), so the accesses take the
  // scalar-base form with one 32-bit offset register each
  const char* s0 = reinterpret_cast<const char*>(src + m * N * H);
  char* d0 = reinterpret_cast<char*>(dst + m * N * H);
  constexpr unsigned QB = (unsigned)(G * H * sizeof(cplx<T>));  // bytes per q
  const unsigned o = (unsigned)((t * H + k) * sizeof(cplx<T>));
  unsigned oq[8];  // byte offsets of the lane's points
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    oq[q] = o + q * QB;
    // opaque to the optimiser, which would otherwise fold the constant into
    // a 64-bit address instead of the 32-bit offset register
    __asm__ volatile("" : "+v"(oq[q]));
  }
  cplx<T> v[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[0][q] = *reinterpret_cast<const cplx<T>*>(s0 + oq[q]);
  __syncthreads();
  cplx<T>* const bufs[1] = {lbuf[g]};
  fft_lines<N, false, 1, false, false>(v, bufs, tw, t);
  const T sc = T(1) / (T(N) * T(N));
  const T tf = (T)t;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    // signed frequency of row t + q N/8: t + q N/8 below N/2, else that - N
    // (t < N/8); integers, exact in T
    const T kxv = (tf + T((q < 4 ? q : q - 8) * G)) * dk * sc;
    v[0][q] = (k == 0 && q == 4 && t == 0) ? mk(T(0), T(0)) : imul(v[0][q], kxv);
  }
  fft_lines<N, true, 1, false, false>(v, bufs, tw, t);
#pragma unroll
  for (int q = 0; q < 8; ++q) *reinterpret_cast<cplx<T>*>(d0 + oq[q]) = v[0][q];
}

// Classic RK4 bookkeeping of one stage (util.py:184-199), per pixel, with
// the rate k of the stage:
//   stage 1: acc = y + h/6 k ; z = y + h/2 k
//   stage 2: acc += h/3 k    ; z = y + h/2 k
//   stage 3: acc += h/3 k    ; z = y + h k
//   stage 4: y = acc + h/6 k ; z = y
// z is the next stage input.  The row kernels branch on the stage once per
// row (it is uniform over the grid) around straight-line pixel loops.

// Extra operands of the standard-rate row pass inside the gradient ODE
// (lense_flow.py:156-252): the stage inputs of delta f, the M^-1 tables, and
// the five pixel-space source maps of the delta-phi rate, accumulated with the
// RK4 weight of each stage (the rate is linear in them, so their rfft2 and the
// derivative multipliers are applied once, after the last stage).
template <typename T>
__device__ __forceinline__ void add_sources(T* S, long NN, T cs, T tval, T fx, T fy, T px, T py,
                                            T ixx, T ixy, T iyy) {
  const T mx = ixx * fx + ixy * fy;   // lense_flow.py:221-222
  const T my = ixy * fx + iyy * fy;
  const T wxx = px * mx, wxy = px * my, wyx = py * mx, wyy = py * my;  // :226-229
  S[0] += cs * mx;
  S[NN] += cs * my;
  S[2 * NN] += cs * (tval * wxx);
  S[3 * NN] += cs * (tval * wxy + tval * wyx);
  S[4 * NN] += cs * (tval * wyy);
}

// Row pass, standard rate.  For each row pair: C2R of (dz/dx + i dz/dy) per
// row, k = p(t_s) . grad z, RK4 update of y/acc, then (unless this is the last
// stage of the last step) the row R2C of the next stage input into S0.
// S0 holds the row spectra S of the stage input, S1 their x-derivative DX;
// the line is Z = DX + i DY with DY = i ky S / N (DY's packed slot is zero:
// ky = 0 at DC, and the real Nyquist value times i drops out of the C2R).
//
// RK selects the RK4 bookkeeping.  0: the classic accumulator (acc written at
// stage 1, updated at 2 and 3, consumed at 4; the gradient ODE uses it).
// 1 and 2: the same RK4 with the accumulator rebuilt from the stage inputs
// z_s (z_1 = y): since z_2 = y + h/2 k_1, z_3 = y + h/2 k_2, z_4 = y + h k_3,
//   acc = (2 y + z_2)/3 + h/3 k_2   (stage 2; = (z_2 + 2 z_3)/3)
//   y'  = acc + (z_4 - y)/3 + h/6 k_4   (stage 4)
// and stages 1 and 3 touch no accumulator: 7 map transfers per component and
// step instead of 10.  RK = 2 (stages 2 and 4) recovers z_s from the row
// spectra S it reads anyway (one more inverse row FFT of S_a + i S_b).
//
// OUTP: the last stage of the serving executor; the new state y goes to the
// per-request outputs outp[b] + (component N + row) N (NULL: a padding slot)
// instead of the staging state.
template <typename T, int N, int C, bool GRAD, int BS, int RK = 0, bool OUTP = false,
          bool NT = false>
__global__ __launch_bounds__(BS) void k_row_lf_std(cplx<T>* __restrict__ S0,
                                                    const cplx<T>* __restrict__ S1,
                                                    T* __restrict__ y, T* __restrict__ acc,
                                                    const T* const* __restrict__ ptabs,
                                                    const cplx<T>* __restrict__ twg, long npairs,
                                                    int stage, T h, int tidx, int do_r2c, T cy,
                                                    GradArgs<T> ga, T* const* outp) {
  constexpr int G = N / 8, PAIRS = BS / G, H = N / 2;
  static_assert(BS % G == 0 && BS >= G, "a block holds whole lines");
  constexpr long NN = (long)N * N;
  constexpr bool TWL = row_tw_lds<T, N, RK>();
  // the twiddle table first: at the bottom of LDS its entries are reached
  // with the immediate offsets of the LDS reads
  struct alignas(16) Lds {
    alignas(16) cplx<T> tw[TWL ? TwiddlesLds<N, T>::kLen : 1];
    alignas(16) cplx<T> line[PAIRS][2][lds_len<N>()];
  };
  __shared__ Lds lds;
  auto& twl = lds.tw;
  auto& lbuf = lds.line;
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  std::conditional_t<TWL, TwiddlesLds<N, T>, Twiddles<N, T>> tw;
  if constexpr (TWL) {
    fill_twiddles_lds<N>(twl, twg, threadIdx.x, BS);
    __syncthreads();
    tw.tab = twl;
  } else {
    tw.load(twg, t);
  }
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const RowPair rp = row_pair<N>((int)(active ? p : npairs - 1), C);
  cplx<T>* bufa = lbuf[grp][0];
  cplx<T>* bufb = lbuf[grp][1];
  // the row spectra first; the per-pixel operands are loaded after the
  // inverse transforms, so they are not live across the FFT (register
  // pressure sets the occupancy of this kernel)
  cplx<T> dxa[4], sa[4], dxb[4], sb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = t + q * G;
    dxa[q] = S1[rp.ra * H + k];
    sa[q] = S0[rp.ra * H + k];
    dxb[q] = S1[rp.rb * H + k];
    sb[q] = S0[rp.rb * H + k];
  }
  lds_sync<kRowWave>();
  // RK == 2: the stage input of both rows, N (z_a + i z_b), from C2R of their
  // row spectra
  cplx<T> zr[1][RK == 2 ? 8 : 1];
  if constexpr (RK == 2) {
#pragma unroll
    for (int q = 0; q < 4; ++q) put_herm<N>(bufa, t + q * G, sa[q], sb[q]);
    lds_sync<kRowWave>();
#pragma unroll
    for (int q = 0; q < 8; ++q) zr[0][q] = bufa[lds_idx<N>(t + q * G)];
    cplx<T>* const bz[1] = {bufa};
    fft_lines<N, true, 1, kRowWave>(zr, bz, tw, t);
    lds_sync<kRowWave>();
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = t + q * G;
    const T c = (T)k * cy;
    put_herm<N>(bufa, k, dxa[q], k == 0 ? mk(T(0), T(0)) : imul(sa[q], c));
    put_herm<N>(bufb, k, dxb[q], k == 0 ? mk(T(0), T(0)) : imul(sb[q], c));
  }
  lds_sync<kRowWave>();
  // per-pixel work, one row at a time: the operands of row b are loaded after
  // row a is done, so at most one row's operands are live
  const T* pt = ptabs[rp.b] + (long)tidx * 2 * NN;
  T px[8], py[8];
  auto load_p = [&](long irow) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const long jj = irow * N + t + m * G;
      px[m] = pt[jj];
      py[m] = pt[NN + jj];
    }
  };
  cplx<T> z[1][8];
  // the new state of pixel j of row r (row rr of the flat state)
  auto store_y = [&](int r, long rr, int j, T v) {
    if constexpr (OUTP) {
      T* o = outp[rp.b];
      const long row = C == 2 ? (long)r * N + rp.ia : (r ? rp.ib : rp.ia);
      if (o) o[row * N + j] = v;
    } else {
      st_stream<NT>(&y[rr * N + j], v);
    }
  };
  // rate, RK4 update and next stage input of row r from its gradient g
  auto row_update = [&](int r, const cplx<T>(&g)[8]) {
    const long rr = r ? rp.rb : rp.ra;
    if (r == 0 || C == 1) load_p(r ? rp.ib : rp.ia);
    T yv[8], av[8];
    if constexpr (RK == 0) {
      // the classic bookkeeping above, one branch on the stage per row
      T zz[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) zz[m] = px[m] * g[m].x + py[m] * g[m].y;  // the rate k
      if (stage == 4) {
#pragma unroll
        for (int m = 0; m < 8; ++m) zz[m] = acc[rr * N + t + m * G] + (h / T(6)) * zz[m];
        if (active) {
#pragma unroll
          for (int m = 0; m < 8; ++m) store_y(r, rr, t + m * G, zz[m]);
        }
      } else {
        const T ha = stage == 1 ? h / T(6) : h / T(3);
        const T hz = stage == 3 ? h : h / T(2);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const int j = t + m * G;
          yv[m] = y[rr * N + j];
          av[m] = (stage == 1 ? yv[m] : acc[rr * N + j]) + ha * zz[m];
          zz[m] = yv[m] + hz * zz[m];
        }
        if (active) {
#pragma unroll
          for (int m = 0; m < 8; ++m) acc[rr * N + t + m * G] = av[m];
        }
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (r == 0)
          z[0][m].x = zz[m];
        else
          z[0][m].y = zz[m];
      }
    } else {
      // the stage is uniform over the grid: one branch per row, around
      // straight-line pixel loops (a branch per pixel keeps every operand of
      // the row live across all of them and costs occupancy)
      T zz[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int j = t + m * G;
        yv[m] = y[rr * N + j];
        zz[m] = px[m] * g[m].x + py[m] * g[m].y;  // the rate k
      }
      if constexpr (RK == 1) {  // stages 1 and 3
        const T hk = stage == 1 ? h / T(2) : h;
#pragma unroll
        for (int m = 0; m < 8; ++m) zz[m] = yv[m] + hk * zz[m];
      } else {  // stages 2 and 4
        const T inv_n = T(1) / T(N);  // exact: N is a power of two
        // thirds as products with the rounded constants (a division per pixel
        // costs ~10 VALU instructions)
        constexpr T c13 = T(1) / T(3), c23 = T(2) / T(3);
        if (stage == 2) {
          const T h3 = h / T(3);
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const T zin = (r ? zr[0][m].y : zr[0][m].x) * inv_n;
            av[m] = c23 * yv[m] + (c13 * zin + h3 * zz[m]);
            zz[m] = yv[m] + (h / T(2)) * zz[m];
          }
          if (active) {
#pragma unroll
            for (int m = 0; m < 8; ++m) st_stream<NT>(&acc[rr * N + t + m * G], av[m]);
          }
        } else {
          const T h6 = h / T(6);
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const T zin = (r ? zr[0][m].y : zr[0][m].x) * inv_n;
            zz[m] = ld_stream<NT>(&acc[rr * N + t + m * G]) + (c13 * (zin - yv[m]) + h6 * zz[m]);
          }
          if (active) {
#pragma unroll
            for (int m = 0; m < 8; ++m) store_y(r, rr, t + m * G, zz[m]);
          }
        }
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (r == 0)
          z[0][m].x = zz[m];
        else
          z[0][m].y = zz[m];
      }
    }
  };
  if constexpr (!GRAD) {
    // one line at a time: row a's inverse transform and update, then row b's
    // (only one line's gradient is live at a time)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      cplx<T>* const buf1[1] = {r ? bufb : bufa};
      cplx<T> w[1][8];
#pragma unroll
      for (int q = 0; q < 8; ++q) w[0][q] = buf1[0][lds_idx<N>(t + q * G)];
      fft_lines<N, true, 1, kRowWave>(w, buf1, tw, t);
      row_update(r, w[0]);
    }
  } else {
    cplx<T> v[2][8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      v[0][q] = bufa[lds_idx<N>(t + q * G)];
      v[1][q] = bufb[lds_idx<N>(t + q * G)];
    }
    cplx<T>* const bufs[2] = {bufa, bufb};
    fft_lines<N, true, 2, kRowWave>(v, bufs, tw, t);
    // delta-phi sources at this stage: sum_c df_c grad f_c (lense_flow.py:201-213)
    const T* gt = ga.gtabs[rp.b] + (long)tidx * 3 * NN;
    T* Sb = ga.sacc + (long)rp.b * 5 * NN;
    if constexpr (C == 2) {
      load_p(rp.ia);
      if (active) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const int j = t + m * G;
          const long ja = (long)rp.ia * N + j;
          const T zda = ga.zd[rp.ra * N + j], zdb = ga.zd[rp.rb * N + j];
          const T fx = zda * v[0][m].x + zdb * v[1][m].x;
          const T fy = zda * v[0][m].y + zdb * v[1][m].y;
          add_sources(Sb + ja, NN, ga.cs, ga.tval, fx, fy, px[m], py[m], gt[ja], gt[NN + ja],
                      gt[2 * NN + ja]);
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        const long ir = r ? rp.ib : rp.ia;
        const long rr = r ? rp.rb : rp.ra;
        load_p(ir);
        if (active) {
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const int j = t + m * G;
            const long jr = ir * N + j;
            const T zd = ga.zd[rr * N + j];
            add_sources(Sb + jr, NN, ga.cs, ga.tval, zd * v[r][m].x, zd * v[r][m].y, px[m],
                        py[m], gt[jr], gt[NN + jr], gt[2 * NN + jr]);
          }
        }
      }
    }
    row_update(0, v[0]);
    row_update(1, v[1]);
  }
  if (!do_r2c) return;  // uniform over the grid
  cplx<T>* const bufz[1] = {bufa};
  fft_lines<N, false, 1, kRowWave>(z, bufz, tw, t);
  split_store<N, kRowWave>(z[0], bufa, S0 + rp.ra * H, S0 + rp.rb * H, t, active);
}

// Row pass, adjoint rate.  init != 0: no rate, z = y (first stage input).
// Otherwise, with S0 = d/dx of U (column pass) and S1 = V, the divergence
// spectrum is W = S0 + i ky V / N per row (V's packed slot drops out as in
// grad_line); C2R of (W_a + i W_b) gives the two rates, then RK4 update.  Then,
// unless this is the very last stage, u = p_x(t_next) z and v = p_y(t_next) z
// are row-R2C'd per row (as u + i v) into S0 (U) and S1 (V).
template <typename T, int N, int C, int BS>
__global__ __launch_bounds__(BS) void k_row_lf_adj(cplx<T>* __restrict__ S0,
                                                    cplx<T>* __restrict__ S1,
                                                    T* __restrict__ y, T* __restrict__ acc,
                                                    const T* const* __restrict__ ptabs,
                                                    const cplx<T>* __restrict__ twg,
                                                    long npairs, int stage, T h, int tidx_next,
                                                    int do_r2c, int init, T cy,
                                                    T* __restrict__ zstore) {
  constexpr int G = N / 8, PAIRS = BS / G, H = N / 2;
  static_assert(BS % G == 0 && BS >= G, "a block holds whole lines");
  constexpr long NN = (long)N * N;
  __shared__ cplx<T> lbuf[PAIRS][2][lds_len<N>()];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const RowPair rp = row_pair<N>((int)(active ? p : npairs - 1), C);
  cplx<T>* bufa = lbuf[grp][0];
  cplx<T>* bufb = lbuf[grp][1];
  // the row spectra first; the per-pixel operands are loaded after the
  // inverse transform, so they are not live across it (register pressure
  // sets the occupancy of this kernel)
  cplx<T> wa[4], wb[4], va[4], vb[4];
  if (!init) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = t + q * G;
      wa[q] = S0[rp.ra * H + k];
      wb[q] = S0[rp.rb * H + k];
      va[q] = S1[rp.ra * H + k];
      vb[q] = S1[rp.rb * H + k];
    }
  }
  cplx<T> v[1][8];
  lds_sync<kRowWave>();
  if (!init) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = t + q * G;
      cplx<T> a = wa[q], b = wb[q];
      if (k != 0) {
        a = a + imul(va[q], (T)k * cy);
        b = b + imul(vb[q], (T)k * cy);
      }
      put_herm<N>(bufa, k, a, b);
    }
    lds_sync<kRowWave>();
#pragma unroll
    for (int q = 0; q < 8; ++q) v[0][q] = bufa[lds_idx<N>(t + q * G)];
    cplx<T>* const bufs[1] = {bufa};
    fft_lines<N, true, 1, kRowWave>(v, bufs, tw, t);
  }
  // the next stage's p = M(t_next)^-1 grad phi, loaded before the update so
  // that its latency overlaps the update's
  const T* pt = ptabs[rp.b] + (long)(do_r2c ? tidx_next : 0) * 2 * NN;
  T pxa[8], pya[8], pxb[8], pyb[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const long ja = (long)rp.ia * N + t + m * G, jb = (long)rp.ib * N + t + m * G;
    pxa[m] = do_r2c ? pt[ja] : T(0);
    pya[m] = do_r2c ? pt[NN + ja] : T(0);
    if constexpr (C == 2) {  // Q and U rows share the pixel row
      pxb[m] = pxa[m];
      pyb[m] = pya[m];
    } else {
      pxb[m] = do_r2c ? pt[jb] : T(0);
      pyb[m] = do_r2c ? pt[NN + jb] : T(0);
    }
  }
  // RK4 update, the classic bookkeeping (one branch on the stage)
  T za[8], zb[8];
  if (init) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      za[m] = y[rp.ra * N + t + m * G];
      zb[m] = y[rp.rb * N + t + m * G];
    }
  } else if (stage == 4) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      za[m] = acc[rp.ra * N + t + m * G] + (h / T(6)) * v[0][m].x;
      zb[m] = acc[rp.rb * N + t + m * G] + (h / T(6)) * v[0][m].y;
    }
    if (active) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        y[rp.ra * N + t + m * G] = za[m];
        y[rp.rb * N + t + m * G] = zb[m];
      }
    }
  } else {
    const T ha = stage == 1 ? h / T(6) : h / T(3);
    const T hz = stage == 3 ? h : h / T(2);
    T aa[8], ab[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int j = t + m * G;
      const T ya = y[rp.ra * N + j], yb = y[rp.rb * N + j];
      aa[m] = (stage == 1 ? ya : acc[rp.ra * N + j]) + ha * v[0][m].x;
      ab[m] = (stage == 1 ? yb : acc[rp.rb * N + j]) + ha * v[0][m].y;
      za[m] = ya + hz * v[0][m].x;
      zb[m] = yb + hz * v[0][m].y;
    }
    if (active) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        acc[rp.ra * N + t + m * G] = aa[m];
        acc[rp.rb * N + t + m * G] = ab[m];
      }
    }
  }
  if (!do_r2c) return;  // uniform over the grid
  if (zstore && active) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      zstore[rp.ra * N + t + m * G] = za[m];
      zstore[rp.rb * N + t + m * G] = zb[m];
    }
  }
  cplx<T> uv[2][8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    uv[0][m] = mk(pxa[m] * za[m], pya[m] * za[m]);
    uv[1][m] = mk(pxb[m] * zb[m], pyb[m] * zb[m]);
  }
  cplx<T>* const bufs2[2] = {bufa, bufb};
  fft_lines<N, false, 2, kRowWave>(uv, bufs2, tw, t);
  split_store<N, kRowWave>(uv[0], bufa, S0 + rp.ra * H, S1 + rp.ra * H, t, active);
  split_store<N, kRowWave>(uv[1], bufb, S0 + rp.rb * H, S1 + rp.rb * H, t, active);
}

// Row pass, adjoint rate, with the low-storage RK4 bookkeeping of
// k_row_lf_std (RK = 1, 2).  Between launches S0 holds U, the row spectra of
// u = p_x(t_s) z_s (after the column pass: their x-derivative, in place), and
// Z the stage input z_s itself as a real map (stage 1: z_1 = y, read from y).
// Storing z_s where k_row_lf_adj stores the spectra of v = p_y z_s moves the
// same bytes, and makes z_s available to rebuild the accumulator.  Per row
// pair: v = p_y(t_s) z_s of both rows as the line v_a + i v_b, forward FFT,
// multiply by i ky / N (its y-derivative: i ky keeps each row's Hermitian
// symmetry, so both rows go through one line; DC and the Nyquist bin drop out
// as in the C2R), add the Hermitian line of the two x-derivative half rows,
// one inverse FFT: the rates div(p z) of both rows as (Re, Im).  Then
//   stage 1: z_2 = y + h/2 k                                  -> Z
//   stage 2: acc = (2 y + z_2)/3 + h/3 k, z_3 = y + h/2 k     -> acc, Z
//   stage 3: z_4 = y + h k                                    -> Z
//   stage 4: y' = acc + (z_4 - y)/3 + h/6 k                   -> y (= z_1 next)
// and, unless this is the very last stage, u = p_x(t_next) z_next is row-R2C'd
// (two rows per line) into S0.  init != 0: no rate, U of z_1 = y only.
// Per launch (5.25 C + 2) N^2 s against the classic kernel's (6.5 C + 2).
template <typename T, int N, int C, int BS, bool NT = false>
__global__ __launch_bounds__(BS) void k_row_lf_adj_ls(cplx<T>* __restrict__ S0,
                                                       T* __restrict__ Z, T* __restrict__ y,
                                                       T* __restrict__ acc,
                                                       const T* const* __restrict__ ptabs,
                                                       const cplx<T>* __restrict__ twg,
                                                       long npairs, int stage, T h, int tidx,
                                                       int tidx_next, int do_r2c, int init,
                                                       T cy) {
  constexpr int G = N / 8, PAIRS = BS / G, H = N / 2;
  static_assert(BS % G == 0 && BS >= G, "a block holds whole lines");
  constexpr long NN = (long)N * N;
  __shared__ cplx<T> lbuf[PAIRS][2][lds_len<N>()];
  const int grp = threadIdx.x / G, t = threadIdx.x % G;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long p = (long)blockIdx.x * PAIRS + grp;
  const bool active = p < npairs;
  const RowPair rp = row_pair<N>((int)(active ? p : npairs - 1), C);
  cplx<T>* bufa = lbuf[grp][0];
  cplx<T>* bufb = lbuf[grp][1];
  cplx<T>* const bl[1] = {bufa};
  const T* pt = ptabs[rp.b];
  // the stage input of both rows
  const T* zsrc = (init || stage == 1) ? y : Z;
  T za[8], zb[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    za[m] = zsrc[rp.ra * N + t + m * G];
    zb[m] = zsrc[rp.rb * N + t + m * G];
  }
  cplx<T> w[1][8];
  if (!init) {
    // the x-derivative half rows as the Hermitian line W_a + i W_b (bufb)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = t + q * G;
      put_herm<N>(bufb, k, S0[rp.ra * H + k], S0[rp.rb * H + k]);
    }
    const T* py = pt + (long)tidx * 2 * NN + NN;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int j = t + m * G;
      const T pa = py[(long)rp.ia * N + j];
      const T pb = C == 2 ? pa : py[(long)rp.ib * N + j];
      w[0][m] = mk(pa * za[m], pb * zb[m]);
    }
    fft_lines<N, false, 1, kRowWave>(w, bl, tw, t);
    lds_sync<kRowWave>();
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int k = t + m * G;  // bin k of the full line; signed ky, 0 at DC and N/2
      const T c = m < 4 ? (T)k * cy : (m == 4 && t == 0 ? T(0) : (T)(k - N) * cy);
      w[0][m] = bufb[lds_idx<N>(k)] + imul(w[0][m], c);
    }
    fft_lines<N, true, 1, kRowWave>(w, bl, tw, t);
    // RK4 update, one branch on the (grid-uniform) stage around straight-line
    // pixel loops; thirds as products with the rounded constants
    constexpr T c13 = T(1) / T(3), c23 = T(2) / T(3);
    if (stage == 1) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        za[m] = za[m] + (h / T(2)) * w[0][m].x;
        zb[m] = zb[m] + (h / T(2)) * w[0][m].y;
      }
    } else if (stage == 3) {
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int j = t + m * G;
        za[m] = y[rp.ra * N + j] + h * w[0][m].x;
        zb[m] = y[rp.rb * N + j] + h * w[0][m].y;
      }
    } else if (stage == 2) {
      T aa[8], ab[8];
      const T h3 = h / T(3);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int j = t + m * G;
        const T ya = y[rp.ra * N + j], yb = y[rp.rb * N + j];
        aa[m] = c23 * ya + (c13 * za[m] + h3 * w[0][m].x);
        ab[m] = c23 * yb + (c13 * zb[m] + h3 * w[0][m].y);
        za[m] = ya + (h / T(2)) * w[0][m].x;
        zb[m] = yb + (h / T(2)) * w[0][m].y;
      }
      if (active) {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          st_stream<NT>(&acc[rp.ra * N + t + m * G], aa[m]);
          st_stream<NT>(&acc[rp.rb * N + t + m * G], ab[m]);
        }
      }
    } else {
      const T h6 = h / T(6);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int j = t + m * G;
        const T ya = y[rp.ra * N + j], yb = y[rp.rb * N + j];
        za[m] = ld_stream<NT>(&acc[rp.ra * N + j]) + (c13 * (za[m] - ya) + h6 * w[0][m].x);
        zb[m] = ld_stream<NT>(&acc[rp.rb * N + j]) + (c13 * (zb[m] - yb) + h6 * w[0][m].y);
      }
    }
    if (active) {
      if (stage == 4) {  // the new state: next read at the next step's first stage
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          st_stream<NT>(&y[rp.ra * N + t + m * G], za[m]);
          st_stream<NT>(&y[rp.rb * N + t + m * G], zb[m]);
        }
      } else {  // the stage input: read by the next row launch
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          Z[rp.ra * N + t + m * G] = za[m];
          Z[rp.rb * N + t + m * G] = zb[m];
        }
      }
    }
  }
  if (!do_r2c) return;  // uniform over the grid
  const T* px = pt + (long)tidx_next * 2 * NN;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int j = t + m * G;
    const T pa = px[(long)rp.ia * N + j];
    const T pb = C == 2 ? pa : px[(long)rp.ib * N + j];
    w[0][m] = mk(pa * za[m], pb * zb[m]);
  }
  fft_lines<N, false, 1, kRowWave>(w, bl, tw, t);
  split_store<N, kRowWave>(w[0], bufa, S0 + rp.ra * H, S0 + rp.rb * H, t, active);
}

// ---------------------------------------------------------------------------
// Resident LenseFlow of small maps (N <= 64, FP32)
// ---------------------------------------------------------------------------
// One workgroup runs the whole LenseFlow of one component (standard rate):
// its stage state -- the packed row spectra S, their x-derivative DX, the
// state y and the RK4 accumulator, 64 KB at N = 64 -- stays in LDS for all 4n
// stages, and only the p(t) rows are read from memory (from L2 when several
// components use one phi).  Per stage a column phase (every packed column:
// FFT, i KX / N^2, inverse FFT, into DX) and a row phase (per row pair: the
// lines DX + i DY, inverse FFTs, rate p(t) . grad z, classic RK4 update, the
// next stage input R2C'd into S), separated by workgroup barriers: one launch
// instead of 4n + 1 pairs of launches, for requests whose launches would be
// latency-bound.  The line transforms and the per-pixel arithmetic are those
// of the two-pass kernels; the bookkeeping is the classic accumulator (the
// state lives on chip, so rebuilding it would save no memory traffic).
// src / outp: the serving executor's request pointers (src[b] + c N^2, NULL:
// zeros; outp[b] + c N^2, NULL: not written); without them y [nmaps][N][N]
// is lensed in place.  One workgroup per component (map m = b C + c).
template <typename T, int N, int BS>
__global__ __launch_bounds__(BS) void k_lf_resident(const T* const* __restrict__ src,
                                                     T* const* __restrict__ outp,
                                                     T* __restrict__ y, int C,
                                                     const T* const* __restrict__ ptabs,
                                                     const cplx<T>* __restrict__ twg,
                                                     int nsteps, int fwd_time, T h, T dk) {
  constexpr int G = N / 8, H = N / 2, L = BS / G;  // line groups
  static_assert(BS % 64 == 0 && L >= H, "whole wavefronts, a group per pair and per column");
  constexpr int NN = N * N;
  __shared__ cplx<T> S[N * H];
  __shared__ cplx<T> DX[N * H];
  __shared__ T Y[NN];
  __shared__ T ACC[NN];
  __shared__ cplx<T> lbuf[L][2][lds_len<N>()];
  const int j = threadIdx.x / G, t = threadIdx.x % G;
  // groups beyond the H pairs / columns repeat one (their results are not
  // stored), so every lane of a wavefront runs the same line transforms
  const int jj = j % H;
  const bool owner = j < H;
  Twiddles<N, T> tw;
  tw.load(twg, t);
  const long m = blockIdx.x;
  const long b = m / C;
  const int c = (int)(m % C);
  const T* in = src ? (src[b] ? src[b] + (long)c * NN : nullptr) : y + m * NN;
  T* out = outp ? (outp[b] ? outp[b] + (long)c * NN : nullptr) : y + m * NN;
  const T* pt = ptabs[b];
  cplx<T>* bufa = lbuf[j][0];
  cplx<T>* bufb = lbuf[j][1];
  cplx<T>* const bla[1] = {bufa};
  const int ra = 2 * jj, rb = ra + 1;  // the pair's rows
  // state in, and the row spectra of the first stage input
  for (int i = threadIdx.x; i < NN; i += BS) Y[i] = in ? in[i] : T(0);
  __syncthreads();
  cplx<T> z[1][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) z[0][q] = mk(Y[ra * N + t + q * G], Y[rb * N + t + q * G]);
  fft_lines<N, false, 1, kRowWave>(z, bla, tw, t);
  split_store<N, kRowWave>(z[0], bufa, S + ra * H, S + rb * H, t, owner);
  const T sc = T(1) / (T(N) * T(N));
  const T cy = dk / T(N);
  static const int stage_half[5] = {0, 0, 1, 1, 2};
  for (int i = 0; i < nsteps; ++i) {
    for (int st = 1; st <= 4; ++st) {
      const bool last = (i == nsteps - 1) && (st == 4);
      const int jt = 2 * i + stage_half[st];
      const int tidx = fwd_time ? jt : 2 * nsteps - jt;
      // this stage's p(t) rows, loaded before the column phase so that their
      // memory latency is spent under it
      T pxr[2][8], pyr[2][8];
      {
        const T* px = pt + (long)tidx * 2 * NN;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int jp = (r ? rb : ra) * N + t + q * G;
            pxr[r][q] = px[jp];
            pyr[r][q] = px[NN + jp];
          }
      }
      __syncthreads();  // S of this stage complete
      // column phase: column jj of S -> d/dx into DX
      {
        cplx<T> v[1][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[0][q] = S[(t + q * G) * H + jj];
        fft_lines<N, false, 1, kRowWave>(v, bla, tw, t);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int kx = t + q * G;
          const T kxv = (T)kx_signed<N>(kx) * dk * sc;
          v[0][q] = (jj == 0 && kx == N / 2) ? mk(T(0), T(0)) : imul(v[0][q], kxv);
        }
        fft_lines<N, true, 1, kRowWave>(v, bla, tw, t);
        if (owner) {
#pragma unroll
          for (int q = 0; q < 8; ++q) DX[(t + q * G) * H + jj] = v[0][q];
        }
      }
      __syncthreads();  // DX complete
      // row phase: rows ra, rb
      cplx<T> dxa[4], sa[4], dxb[4], sb[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = t + q * G;
        dxa[q] = DX[ra * H + k];
        sa[q] = S[ra * H + k];
        dxb[q] = DX[rb * H + k];
        sb[q] = S[rb * H + k];
      }
      lds_sync<kRowWave>();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = t + q * G;
        const T cc = (T)k * cy;
        put_herm<N>(bufa, k, dxa[q], k == 0 ? mk(T(0), T(0)) : imul(sa[q], cc));
        put_herm<N>(bufb, k, dxb[q], k == 0 ? mk(T(0), T(0)) : imul(sb[q], cc));
      }
      lds_sync<kRowWave>();
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        cplx<T>* const b1[1] = {r ? bufb : bufa};
        cplx<T> w[1][8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[0][q] = b1[0][lds_idx<N>(t + q * G)];
        fft_lines<N, true, 1, kRowWave>(w, b1, tw, t);
        const int row = r ? rb : ra;
        // the classic RK4 bookkeeping (util.py:184-199)
        T zz[8];
        if (st == 4) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int jp = row * N + t + q * G;
            zz[q] = ACC[jp] + (h / T(6)) * (pxr[r][q] * w[0][q].x + pyr[r][q] * w[0][q].y);
          }
          if (owner) {
#pragma unroll
            for (int q = 0; q < 8; ++q) Y[row * N + t + q * G] = zz[q];
          }
        } else {
          const T ha = st == 1 ? h / T(6) : h / T(3);
          const T hz = st == 3 ? h : h / T(2);
          T av[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int jp = row * N + t + q * G;
            const T k = pxr[r][q] * w[0][q].x + pyr[r][q] * w[0][q].y;
            const T yv = Y[jp];
            av[q] = (st == 1 ? yv : ACC[jp]) + ha * k;
            zz[q] = yv + hz * k;
          }
          if (owner) {
#pragma unroll
            for (int q = 0; q < 8; ++q) ACC[row * N + t + q * G] = av[q];
          }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          if (r == 0)
            z[0][q].x = zz[q];
          else
            z[0][q].y = zz[q];
        }
      }
      if (last) break;
      fft_lines<N, false, 1, kRowWave>(z, bla, tw, t);
      split_store<N, kRowWave>(z[0], bufa, S + ra * H, S + rb * H, t, owner);
    }
  }
  __syncthreads();  // Y complete
  if (out) {
    for (int i = threadIdx.x; i < NN; i += BS) out[i] = Y[i];
  }
}

// Gather / scatter of whole maps through device pointer arrays (serving
// executor), `len` elements per slot b = blockIdx.y.  A NULL source fills the
// slot with zeros (padding); a NULL destination skips it.
template <typename T>
__global__ void k_gather(const T* const* __restrict__ src, T* __restrict__ dst, long len) {
  const int b = blockIdx.y;
  const T* s = src[b];
  T* d = dst + (long)b * len;
  const long i0 = blockIdx.x * (long)blockDim.x + threadIdx.x, di = (long)gridDim.x * blockDim.x;
  if (s) {
    for (long i = i0; i < len; i += di) d[i] = s[i];
  } else {
    for (long i = i0; i < len; i += di) d[i] = T(0);
  }
}

template <typename T>
__global__ void k_scatter(const T* __restrict__ src, T* const* __restrict__ dst, long len) {
  const int b = blockIdx.y;
  T* d = dst[b];
  if (!d) return;
  const T* s = src + (long)b * len;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x)
    d[i] = s[i];
}

// ---------------------------------------------------------------------------
// host-side launchers (dispatch on N)
// ---------------------------------------------------------------------------
// Columns per column-pass workgroup: 16 (float) / 8 (double) complex values,
// 128-byte row segments.
#ifndef CMBL_COL_KC_F32
#define CMBL_COL_KC_F32 16
#endif
#ifndef CMBL_COL_KC_F64
#define CMBL_COL_KC_F64 8
#endif
template <typename T>
constexpr int col_kc(int N) {
  return (sizeof(T) == 4 ? CMBL_COL_KC_F32 : CMBL_COL_KC_F64) < N / 2
             ? (sizeof(T) == 4 ? CMBL_COL_KC_F32 : CMBL_COL_KC_F64)
             : N / 2;
}

#define CMBL_DISPATCH_N(N, ...)                                \
  switch (N) {                                                 \
    case 16: { constexpr int NN_ = 16; __VA_ARGS__; } break;   \
    case 32: { constexpr int NN_ = 32; __VA_ARGS__; } break;   \
    case 64: { constexpr int NN_ = 64; __VA_ARGS__; } break;   \
    case 128: { constexpr int NN_ = 128; __VA_ARGS__; } break; \
    case 256: { constexpr int NN_ = 256; __VA_ARGS__; } break; \
    case 512: { constexpr int NN_ = 512; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;                      \
  }

static inline unsigned blocks_for(long n, long per) { return (unsigned)((n + per - 1) / per); }

// Launch shapes.  A grid of fewer than CMBL_SMALL_GRID default-shape blocks
// (a few requests, chains or right-hand sides) uses 4x smaller blocks, so the
// launch spreads over more CUs: 64-thread row blocks and a quarter of the
// columns per column block.
#ifndef CMBL_SMALL_GRID
#define CMBL_SMALL_GRID 256
#endif
// The column pass takes its quarter-width blocks only below this many
// default-shape blocks: narrow blocks read row segments of 32 bytes, which
// lose to the full-width shape as soon as a quarter of the CUs have a block
// (12 maps of 512^2: 16.6 us with 4-column blocks against ~11 us with 16).
#ifndef CMBL_SMALL_COL_GRID
#define CMBL_SMALL_COL_GRID 64
#endif


template <typename T>
hipError_t launch_rfft2(int N, const T* x, cplx<T>* P, cplx<T>* X, long nmaps,
                        const cplx<T>* tw, hipStream_t s) {
  CMBL_DISPATCH_N(N, {
    constexpr int KC = col_kc<T>(NN_);
    const long npairs = nmaps * NN_ / 2;
    k_row_r2c<T, NN_><<<blocks_for(npairs, 256 / (NN_ / 8)), 256, 0, s>>>(x, P, tw, npairs);
    k_col_fwd_unpack<T, NN_, KC><<<(unsigned)(nmaps * (NN_ / 2 / KC)), KC * NN_ / 8, 0, s>>>(P, X, tw);
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_irfft2(int N, const cplx<T>* X, cplx<T>* P, T* x, long nmaps,
                         const cplx<T>* tw, hipStream_t s) {
  CMBL_DISPATCH_N(N, {
    constexpr int KC = col_kc<T>(NN_);
    const long npairs = nmaps * NN_ / 2;
    const T sc = T(1) / (T(NN_) * T(NN_));
    k_col_inv_pack<T, NN_, KC><<<(unsigned)(nmaps * (NN_ / 2 / KC)), KC * NN_ / 8, 0, s>>>(X, P, tw, sc);
    k_row_c2r<T, NN_><<<blocks_for(npairs, 256 / (NN_ / 8)), 256, 0, s>>>(P, x, tw, npairs);
  });
  return hipGetLastError();
}

#if CMBL_KERNELS_PART != 2
hipError_t launch_phi_mult(int N, const cplx<double>* F, cplx<double>* D, double dk,
                           hipStream_t s, int count) {
  if (count < 1) return hipSuccess;
  const long n = (long)N * (N / 2 + 1);
  k_phi_mult<<<dim3(blocks_for(n, 256), count), 256, 0, s>>>(F, D, N, dk);
  return hipGetLastError();
}
#endif

template <typename T>
hipError_t launch_ptab(int N, const double* D, T* ptab, int nsteps, double* minpart,
                       int* nblocks, hipStream_t s, int count) {
  const long P = (long)N * N;
  *nblocks = (int)((P + 255) / 256);
  if (count < 1) return hipSuccess;
  k_ptab<T><<<dim3(*nblocks, count), 256, 0, s>>>(D, ptab, N, nsteps, minpart);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lf_row_r2c(int N, const T* y, cplx<T>* S0, long nmaps, const cplx<T>* tw,
                             hipStream_t s) {
  CMBL_DISPATCH_N(N, {
    const long npairs = nmaps * NN_ / 2;
    k_row_r2c<T, NN_><<<blocks_for(npairs, 256 / (NN_ / 8)), 256, 0, s>>>(y, S0, tw, npairs);
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_row_c2r(int N, const cplx<T>* P, T* x, long nmaps, const cplx<T>* tw,
                          const T* d, const T* w, int ep, hipStream_t s) {
  if (ep < 0 || ep > 2 || (ep != 0 && !w)) return hipErrorInvalidValue;
  CMBL_DISPATCH_N(N, {
    const long npairs = nmaps * NN_ / 2;
    const unsigned nb = blocks_for(npairs, 256 / (NN_ / 8));
    if (ep == 0)
      k_row_c2r<T, NN_, 0><<<nb, 256, 0, s>>>(P, x, tw, npairs);
    else if (ep == 1)
      k_row_c2r<T, NN_, 1><<<nb, 256, 0, s>>>(P, x, tw, npairs, d, w);
    else
      k_row_c2r<T, NN_, 2><<<nb, 256, 0, s>>>(P, x, tw, npairs, d, w);
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lf_row_r2c_gather(int N, const T* const* src, int ncomp, T* ystage, cplx<T>* S0,
                                    long nmaps, const cplx<T>* tw, hipStream_t s) {
  CMBL_DISPATCH_N(N, {
    const long npairs = nmaps * NN_ / 2;
    k_row_r2c_gather<T, NN_><<<blocks_for(npairs, 256 / (NN_ / 8)), 256, 0, s>>>(
        src, ncomp, ystage, S0, tw, npairs);
  });
  return hipGetLastError();
}

// The FP32 column pass at N = 512 runs on XCD-paired half tiles (two
// workgroups of 8 columns per CU instead of one of 16: 8-RHS 512^2 Wiener
// filter 2.11 -> 2.06 ms per iteration); at 256 the full tiles are faster
// (27.7 against 30.7 us per launch at batch 128).  CMBL_COL_PAIR=0 turns the
// pairing off (A/B runs).
static int col_pair_mode() {
  static const int mode = [] {
    const char* e = std::getenv("CMBL_COL_PAIR");
    return e ? std::atoi(e) : 1;
  }();
  return mode;
}

template <typename T>
hipError_t launch_lf_col(int N, const cplx<T>* src, cplx<T>* dst, long nmaps, const cplx<T>* tw,
                         T dk, hipStream_t s) {
  CMBL_DISPATCH_N(N, {
    constexpr int KC = col_kc<T>(NN_);
    constexpr int KCS = KC / 4 > 2 ? KC / 4 : 2;
    const unsigned nb = (unsigned)(nmaps * (NN_ / 2 / KC));
    constexpr bool PAIRABLE = sizeof(T) == 4 && NN_ == 512 && KC == 16;
    if constexpr (PAIRABLE) {
      if (col_pair_mode() && nb >= CMBL_SMALL_COL_GRID && (2 * nb) % 16 == 0) {
        k_col_lf_dx<T, NN_, KC / 2, true><<<2 * nb, KC / 2 * NN_ / 8, 0, s>>>(src, dst, tw, dk);
        break;
      }
    }
    if (nb < CMBL_SMALL_COL_GRID)
      k_col_lf_dx<T, NN_, KCS><<<(unsigned)(nmaps * (NN_ / 2 / KCS)), KCS * NN_ / 8, 0, s>>>(
          src, dst, tw, dk);
    else
      k_col_lf_dx<T, NN_, KC><<<nb, KC * NN_ / 8, 0, s>>>(src, dst, tw, dk);
  });
  return hipGetLastError();
}

// The forward/inverse LenseFlow uses the low-storage RK4 bookkeeping of
// k_row_lf_std (RK = 1, 2); the gradient ODE keeps the classic one.
#ifndef CMBL_RK4_LOWSTORE
#define CMBL_RK4_LOWSTORE 1
#endif
constexpr bool kLowStorageRK4 = CMBL_RK4_LOWSTORE != 0;

template <typename T, int N, int BS>
void row_lf_launch(bool adjoint, int C, const GradArgs<T>* ga, cplx<T>* S0, cplx<T>* S1, T* y,
                   T* acc, const T* const* ptabs, const cplx<T>* tw, long npairs, int stage, T h,
                   int tidx, int do_r2c, int init, T cy, T* zstore, hipStream_t s,
                   T* const* outp, bool nt) {
  const unsigned nb = blocks_for(npairs, BS / (N / 8));
  const GradArgs<T> g = ga ? *ga : GradArgs<T>{};
  if (outp) {  // last stage of the executor (standard rate only)
    constexpr int RKL = kLowStorageRK4 ? 2 : 0;
    if (C == 2)
      k_row_lf_std<T, N, 2, false, BS, RKL, true><<<nb, BS, 0, s>>>(
          S0, S1, y, acc, ptabs, tw, npairs, stage, h, tidx, do_r2c, cy, g, outp);
    else
      k_row_lf_std<T, N, 1, false, BS, RKL, true><<<nb, BS, 0, s>>>(
          S0, S1, y, acc, ptabs, tw, npairs, stage, h, tidx, do_r2c, cy, g, outp);
    return;
  }
  if (adjoint) {
    if (C == 2)
      k_row_lf_adj<T, N, 2, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage, h,
                                                  tidx, do_r2c, init, cy, zstore);
    else
      k_row_lf_adj<T, N, 1, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage, h,
                                                  tidx, do_r2c, init, cy, zstore);
  } else if (ga) {
    if (C == 2)
      k_row_lf_std<T, N, 2, true, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage,
                                                        h, tidx, do_r2c, cy, g, nullptr);
    else
      k_row_lf_std<T, N, 1, true, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage,
                                                        h, tidx, do_r2c, cy, g, nullptr);
  } else if (!kLowStorageRK4) {
    if (C == 2)
      k_row_lf_std<T, N, 2, false, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage,
                                                         h, tidx, do_r2c, cy, g, nullptr);
    else
      k_row_lf_std<T, N, 1, false, BS><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs, stage,
                                                         h, tidx, do_r2c, cy, g, nullptr);
  } else if (stage == 2 || stage == 4) {
    auto go = [&](auto ntc) {
      constexpr bool NT = decltype(ntc)::value;
      if (C == 2)
        k_row_lf_std<T, N, 2, false, BS, 2, false, NT><<<nb, BS, 0, s>>>(
            S0, S1, y, acc, ptabs, tw, npairs, stage, h, tidx, do_r2c, cy, g, nullptr);
      else
        k_row_lf_std<T, N, 1, false, BS, 2, false, NT><<<nb, BS, 0, s>>>(
            S0, S1, y, acc, ptabs, tw, npairs, stage, h, tidx, do_r2c, cy, g, nullptr);
    };
    if (sizeof(T) == 4 && nt)  // streaming accesses: FP32 only (no FP64 variants)
      go(std::integral_constant<bool, sizeof(T) == 4>{});
    else
      go(std::false_type{});
  } else {
    if (C == 2)
      k_row_lf_std<T, N, 2, false, BS, 1><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs,
                                                            stage, h, tidx, do_r2c, cy, g, nullptr);
    else
      k_row_lf_std<T, N, 1, false, BS, 1><<<nb, BS, 0, s>>>(S0, S1, y, acc, ptabs, tw, npairs,
                                                            stage, h, tidx, do_r2c, cy, g, nullptr);
  }
}

template <typename T>
hipError_t launch_lf_row(int N, bool adjoint, cplx<T>* S0, cplx<T>* S1, T* y, T* acc,
                         const T* const* ptabs, const cplx<T>* tw, int C, long nmaps, int stage,
                         T h, int tidx, int do_r2c, int init, T dk, hipStream_t s, T* zstore,
                         const GradArgs<T>* ga, T* const* outp, bool nt) {
  CMBL_DISPATCH_N(N, {
    const long npairs = nmaps * NN_ / 2;
    const T cy = dk / T(NN_);
    constexpr int BS_SMALL = NN_ / 8 > 64 ? NN_ / 8 : 64;
    if (blocks_for(npairs, 256 / (NN_ / 8)) < CMBL_SMALL_GRID)
      row_lf_launch<T, NN_, BS_SMALL>(adjoint, C, ga, S0, S1, y, acc, ptabs, tw, npairs, stage, h,
                                      tidx, do_r2c, init, cy, zstore, s, outp, nt);
    else
      row_lf_launch<T, NN_, 256>(adjoint, C, ga, S0, S1, y, acc, ptabs, tw, npairs, stage, h,
                                 tidx, do_r2c, init, cy, zstore, s, outp, nt);
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lf_row_adj_ls(int N, cplx<T>* S0, T* Z, T* y, T* acc, const T* const* ptabs,
                                const cplx<T>* tw, int C, long nmaps, int stage, T h, int tidx,
                                int tidx_next, int do_r2c, int init, T dk, hipStream_t s,
                                bool nt) {
  CMBL_DISPATCH_N(N, {
    const long npairs = nmaps * NN_ / 2;
    const T cy = dk / T(NN_);
    constexpr int BS_SMALL = NN_ / 8 > 64 ? NN_ / 8 : 64;
    auto go = [&](auto bs, auto ntc) {
      constexpr int BS = decltype(bs)::value;
      constexpr bool NT = decltype(ntc)::value;
      const unsigned nb = blocks_for(npairs, BS / (NN_ / 8));
      if (C == 2)
        k_row_lf_adj_ls<T, NN_, 2, BS, NT><<<nb, BS, 0, s>>>(
            S0, Z, y, acc, ptabs, tw, npairs, stage, h, tidx, tidx_next, do_r2c, init, cy);
      else
        k_row_lf_adj_ls<T, NN_, 1, BS, NT><<<nb, BS, 0, s>>>(
            S0, Z, y, acc, ptabs, tw, npairs, stage, h, tidx, tidx_next, do_r2c, init, cy);
    };
    if (blocks_for(npairs, 256 / (NN_ / 8)) < CMBL_SMALL_GRID)
      go(std::integral_constant<int, BS_SMALL>{}, std::false_type{});
    else if (sizeof(T) == 4 && nt)  // streaming accesses: FP32 only (no FP64 variants)
      go(std::integral_constant<int, 256>{}, std::integral_constant<bool, sizeof(T) == 4>{});
    else
      go(std::integral_constant<int, 256>{}, std::false_type{});
  });
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lf_resident(int N, const T* const* src, T* const* outp, T* y, int C,
                              const T* const* ptabs, const cplx<T>* tw, long nmaps, int nsteps,
                              int fwd_time, T h, T dk, hipStream_t s) {
  if constexpr (sizeof(T) != 4) {
    return hipErrorInvalidValue;
  } else {
    if (nmaps < 1) return hipSuccess;
    switch (N) {
      case 16:
        k_lf_resident<T, 16, 64><<<(unsigned)nmaps, 64, 0, s>>>(src, outp, y, C, ptabs, tw,
                                                                nsteps, fwd_time, h, dk);
        break;
      case 32:
        k_lf_resident<T, 32, 64><<<(unsigned)nmaps, 64, 0, s>>>(src, outp, y, C, ptabs, tw,
                                                                nsteps, fwd_time, h, dk);
        break;
      case 64:
        k_lf_resident<T, 64, 256><<<(unsigned)nmaps, 256, 0, s>>>(src, outp, y, C, ptabs, tw,
                                                                  nsteps, fwd_time, h, dk);
        break;
      default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
}

template <typename T>
hipError_t launch_gtab(int N, const double* D, T* gtab, int nsteps, hipStream_t s) {
  const long P = (long)N * N;
  k_gtab<T><<<blocks_for(P, 256), 256, 0, s>>>(D, gtab, N, nsteps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_dphi_combine(int N, const cplx<T>* F, cplx<T>* out, long nb, double dk,
                               hipStream_t s) {
  const long n = nb * (long)N * (N / 2 + 1);
  k_dphi_combine<T><<<(unsigned)std::min<long>((n + 255) / 256, 65536), 256, 0, s>>>(F, out, N, nb,
                                                                                     dk);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather(const T* const* src, T* dst, long len, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const unsigned bx = (unsigned)std::min<long>((len + 255) / 256, 256);
  k_gather<T><<<dim3(bx, count), 256, 0, s>>>(src, dst, len);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scatter(const T* src, T* const* dst, long len, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const unsigned bx = (unsigned)std::min<long>((len + 255) / 256, 256);
  k_scatter<T><<<dim3(bx, count), 256, 0, s>>>(src, dst, len);
  return hipGetLastError();
}

#define CMBL_INST(T)                                                                            \
  template hipError_t launch_rfft2<T>(int, const T*, cplx<T>*, cplx<T>*, long, const cplx<T>*,  \
                                      hipStream_t);                                             \
  template hipError_t launch_irfft2<T>(int, const cplx<T>*, cplx<T>*, T*, long, const cplx<T>*, \
                                       hipStream_t);                                            \
  template hipError_t launch_ptab<T>(int, const double*, T*, int, double*, int*, hipStream_t, int);  \
  template hipError_t launch_lf_row_r2c<T>(int, const T*, cplx<T>*, long, const cplx<T>*,       \
                                           hipStream_t);                                        \
  template hipError_t launch_lf_col<T>(int, const cplx<T>*, cplx<T>*, long, const cplx<T>*, T, \
                                       hipStream_t);                                            \
  template hipError_t launch_lf_row<T>(int, bool, cplx<T>*, cplx<T>*, T*, T*, const T* const*,  \
                                       const cplx<T>*, int, long, int, T, int, int, int, T,     \
                                       hipStream_t, T*, const GradArgs<T>*, T* const*, bool);   \
  template hipError_t launch_row_c2r<T>(int, const cplx<T>*, T*, long, const cplx<T>*, const T*, \
                                        const T*, int, hipStream_t);                            \
  template hipError_t launch_lf_row_r2c_gather<T>(int, const T* const*, int, T*, cplx<T>*,      \
                                                  long, const cplx<T>*, hipStream_t);           \
  template hipError_t launch_lf_row_adj_ls<T>(int, cplx<T>*, T*, T*, T*, const T* const*,       \
                                              const cplx<T>*, int, long, int, T, int, int, int, \
                                              int, T, hipStream_t, bool);                       \
  template hipError_t launch_lf_resident<T>(int, const T* const*, T* const*, T*, int,          \
                                            const T* const*, const cplx<T>*, long, int, int, T, \
                                            T, hipStream_t);                                    \
  template hipError_t launch_gtab<T>(int, const double*, T*, int, hipStream_t);                 \
  template hipError_t launch_gather<T>(const T* const*, T*, long, int, hipStream_t);            \
  template hipError_t launch_scatter<T>(const T*, T* const*, long, int, hipStream_t);           \
  template hipError_t launch_dphi_combine<T>(int, const cplx<T>*, cplx<T>*, long, double,       \
                                             hipStream_t);
#if CMBL_KERNELS_PART != 2
CMBL_INST(float)
#endif
#if CMBL_KERNELS_PART != 1
CMBL_INST(double)
#endif

}  // namespace cmbl
