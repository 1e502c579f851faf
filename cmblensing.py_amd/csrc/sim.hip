// Gaussian random fields drawn on the device, in the FOURIER basis, with the
// data model fused (simulate.py:100-113, 699-702 of the reference).
//
// The reference draws F = sqrt(C/2) (a + i b) on the rfft half plane, takes
// the map irfft2(F, norm="ortho") and later rfft2 of that map.  That round
// trip is N times the Hermitian projection of F: modes off the self-conjugate
// columns (ky = 0 and ky = N/2) are kept, and on those columns
//   G[kx] = N (F[kx] + conj F[(N - kx) % N]) / 2      (real where kx is its own partner).
// One thread computes one mode and regenerates its partner's draw from the
// counter-based generator, so the projection costs no extra memory traffic.
//
// Normal draws: Philox4x64-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11; the multipliers and Weyl key
// increments are the paper's) with key {seed, 0} and counter
// {mode, stream, 0, 0}, mode = kx (N/2+1) + ky; the first two 64-bit outputs
// give u in (0, 1] and v in [0, 1) from their top 53 bits, and Box-Muller
// gives a = sqrt(-2 ln u) cos(2 pi v), b = sqrt(-2 ln u) sin(2 pi v).
// Map m = t C + c of a batch (trial t, component c) uses stream
// stream0 + (c << 32) + t and covariance plane c.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"
#include "sim.hpp"

namespace cmbl {

__device__ __forceinline__ void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
  constexpr uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  constexpr uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    const uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

// sqrt(cv / 2) (a + i b) for one mode
__device__ __forceinline__ void draw_mode(uint64_t mode, uint64_t sid, uint64_t seed, double cv,
                                          double& re, double& im) {
  uint64_t c[4] = {mode, sid, 0, 0};
  philox4x64_10(c, seed, 0);
  const double u = (double)((c[0] >> 11) + 1) * 0x1.0p-53;
  const double v = (double)(c[1] >> 11) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u)) * sqrt(cv / 2.0);
  double sn, cs;
  sincospi(2.0 * v, &sn, &cs);
  re = r * cs;
  im = r * sn;
}

// out[m][kx][ky] = (add ? mul[c] add[m] : 0) + N herm(F_m)[kx][ky]
template <typename T>
__global__ void __launch_bounds__(256) k_grf(const double* __restrict__ cov,
                                             const double* __restrict__ mul,
                                             const cplx<T>* __restrict__ add,
                                             cplx<T>* __restrict__ out, int N, int C, long ntr,
                                             uint64_t seed, uint64_t stream0) {
  const int H1 = N / 2 + 1;
  const long P = (long)N * H1;
  const long total = ntr * C * P;
  for (long g = (long)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (long)gridDim.x * blockDim.x) {
    const long m = g / P;
    const long idx = g - m * P;
    const int c = (int)(m % C);
    const long t = m / C;
    const int i = (int)(idx / H1), j = (int)(idx - (long)i * H1);
    const uint64_t sid = stream0 + ((uint64_t)c << 32) + (uint64_t)t;
    const double* cv = cov + (long)c * P;
    double re, im;
    draw_mode((uint64_t)idx, sid, seed, cv[idx], re, im);
    if (j == 0 || j == N / 2) {
      const int ip = (N - i) % N;
      if (ip == i) {
        im = 0.0;
      } else {
        const long pidx = (long)ip * H1 + j;
        double pre, pim;
        draw_mode((uint64_t)pidx, sid, seed, cv[pidx], pre, pim);
        re = 0.5 * (re + pre);
        im = 0.5 * (im - pim);
      }
    }
    re *= N;
    im *= N;
    if (add) {
      const double w = mul ? mul[(long)c * P + idx] : 1.0;
      const cplx<T> a = add[g];
      re = w * (double)a.x + re;
      im = w * (double)a.y + im;
    }
    out[g] = mk((T)re, (T)im);
  }
}

static inline unsigned grf_grid(long n) {
  long b = (n + 255) / 256;
  return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

template <typename T>
hipError_t launch_grf(const double* cov, const double* mul, const cplx<T>* add, cplx<T>* out,
                      int N, int C, long ntrials, uint64_t seed, uint64_t stream0,
                      hipStream_t s) {
  const long n = ntrials * C * (long)N * (N / 2 + 1);
  if (n == 0) return hipSuccess;
  k_grf<T><<<grf_grid(n), 256, 0, s>>>(cov, mul, add, out, N, C, ntrials, seed, stream0);
  return hipGetLastError();
}

template hipError_t launch_grf<float>(const double*, const double*, const cplx<float>*,
                                      cplx<float>*, int, int, long, uint64_t, uint64_t,
                                      hipStream_t);
template hipError_t launch_grf<double>(const double*, const double*, const cplx<double>*,
                                       cplx<double>*, int, int, long, uint64_t, uint64_t,
                                       hipStream_t);

}  // namespace cmbl
