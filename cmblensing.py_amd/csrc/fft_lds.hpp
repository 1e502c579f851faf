// Device-side building blocks: complex arithmetic and a Stockham radix-8 line
// FFT staged through LDS, for gfx950 (CDNA4, 64-wide wavefronts).
//
// A "line" of N complex points (N = 2^m, 16 <= N <= 512) is transformed by a
// group of N/8 threads.  On entry and on exit thread t of the group holds the
// 8 points   v[q] = line[t + q*N/8],  q = 0..7   in registers, so a caller can
// load/store a line with fully coalesced accesses (consecutive threads touch
// consecutive points), and chain a forward transform, a pointwise multiply and
// an inverse transform without a trip through memory.  Between radix passes
// the line goes through an LDS buffer of lds_len<N>() complex values (padded
// or swizzled layout, below).
//
// Pass plan: floor(m/3) radix-8 passes, then one radix-2 or radix-4 pass when
// m is not a multiple of 3 (each thread then runs 4 or 2 small butterflies).
// Twiddles come from a table tw[k] = exp(-2 pi i k / N) computed on the host
// in double precision and rounded once to T; each thread keeps the ones it
// needs in registers (Twiddles<N, T>), or reads them from an LDS copy of the
// table at each transform where registers set a kernel's occupancy
// (TwiddlesLds<N, T>).
#pragma once
#include <hip/hip_runtime.h>

namespace cmbl {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> conj(cplx<T> a) {
  return mk(a.x, -a.y);
}
template <typename T>
__device__ __forceinline__ cplx<T> scale(cplx<T> a, T s) {
  return mk(a.x * s, a.y * s);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV, typename T>
__device__ __forceinline__ cplx<T> rot90(cplx<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// LDS layout of a line.  A line is read contiguously (point t + q*N/8 by
// thread t) and written with the strides of the Stockham passes (8t + q,
// then (t/8)*64 + t%8 + 8q, ...), in 8-byte (float2) accesses served by
// 4-byte banks.  The padded layout (one slot every 8 points) keeps the
// strided writes on distinct banks but puts point 16 of every contiguous
// run of 16 lanes on the bank of point 0.  The swizzled layout maps point i
// to i ^ ((i >> 3) & 31), a permutation of each aligned block of 32 under
// which both the contiguous and the strided accesses of 16 (or 32)
// consecutive lanes hit distinct banks, only the mirrored N - k accesses of
// the Hermitian pack/unpack keeping a 2-way conflict, and needs no padding.
// On the 256^2 row kernels it removes most of the LDS bank-conflict cycles
// (rocprofv3 SQ_LDS_BANK_CONFLICT).  It is used at N = 256 only: shorter
// lines, of which several share a wavefront, are served better by the padded
// layout, and at N = 512 its per-access index registers cost the row kernels
// a wave per SIMD.
#ifndef CMBL_LDS_SWIZZLE
#define CMBL_LDS_SWIZZLE 1
#endif
template <int N>
__host__ __device__ constexpr bool lds_swizzled() {
  return CMBL_LDS_SWIZZLE != 0 && N == 256;
}
template <int N, bool SW = lds_swizzled<N>()>
__host__ __device__ constexpr int lds_len() {
  return SW ? N : N + N / 8;
}
template <int N, bool SW = lds_swizzled<N>()>
__device__ __forceinline__ int lds_idx(int i) {
  if constexpr (SW)
    return i ^ ((i >> 3) & 31);
  else
    return i + (i >> 3);
}

// Index of point i0 + q * step, q < 8, for a step that is a multiple of 8,
// or 1 with i0 a multiple of 8 (i0 + q then stays in i0's group of 8): in
// the padded layout the index is then affine in q, lds_idx(i0) +
// q * lds_idx(step).  Written that way the compiler forms one address and
// takes the others as immediate offsets of the LDS instructions.
template <int N, bool SW>
__device__ __forceinline__ int lds_idx_q(int i0, int q, int step) {
  if constexpr (!SW)
    return lds_idx<N, false>(i0) + q * (step + step / 8);
  else
    return lds_idx<N, SW>(i0 + q * step);
}

// Index of point t + q N/8 of a line (thread t's points in the natural
// order): affine in q when N/8 is a multiple of 8.
template <int N, bool SW>
__device__ __forceinline__ int line_idx(int t, int q) {
  constexpr int G = N / 8;
  if constexpr (G % 8 == 0)
    return lds_idx_q<N, SW>(t, q, G);
  else
    return lds_idx<N, SW>(t + q * G);
}

// Stride between the LDS buffers of neighbouring lines in the kernels whose
// consecutive lanes work on different lines (the column passes, which use
// the padded layout: lane g of KC consecutive lanes works on line g).  An
// odd stride, one complex more than a multiple of 32, puts those lines on
// distinct banks; a stride of 2 more doubled the measured bank-conflict
// cycles of the 256^2 column pass.
#ifndef CMBL_LINE_PAD
#define CMBL_LINE_PAD 1
#endif
template <int N>
__host__ __device__ constexpr int line_stride() {
  return ((lds_len<N, false>() + 31) / 32) * 32 + CMBL_LINE_PAD;
}

template <int N>
__host__ __device__ constexpr int ilog2() {
  return N <= 1 ? 0 : 1 + ilog2<N / 2>();
}

template <bool INV, typename T>
__device__ __forceinline__ void dft2(cplx<T>& a, cplx<T>& b) {
  cplx<T> t = a - b;
  a = a + b;
  b = t;
}

// In-place 4-point DFT, natural order in and out.
template <bool INV, typename T>
__device__ __forceinline__ void dft4(cplx<T>& y0, cplx<T>& y1, cplx<T>& y2, cplx<T>& y3) {
  cplx<T> c0 = y0 + y2, c1 = y0 - y2, c2 = y1 + y3, c3 = rot90<INV>(y1 - y3);
  y0 = c0 + c2;
  y2 = c0 - c2;
  y1 = c1 + c3;
  y3 = c1 - c3;
}

// In-place 8-point DFT, natural order in and out.
template <bool INV, typename T>
__device__ __forceinline__ void dft8(cplx<T>* v) {
  const T r = T(0.70710678118654752440084436210485);
  cplx<T> a0 = v[0] + v[4], a1 = v[1] + v[5], a2 = v[2] + v[6], a3 = v[3] + v[7];
  cplx<T> b0 = v[0] - v[4], b1 = v[1] - v[5], b2 = v[2] - v[6], b3 = v[3] - v[7];
  // b_k *= W8^k, W8 = exp(-+2 pi i / 8)
  b1 = INV ? mk(r * (b1.x - b1.y), r * (b1.x + b1.y)) : mk(r * (b1.x + b1.y), r * (b1.y - b1.x));
  b2 = rot90<INV>(b2);
  b3 = INV ? mk(-r * (b3.x + b3.y), r * (b3.x - b3.y)) : mk(r * (b3.y - b3.x), -r * (b3.x + b3.y));
  dft4<INV>(a0, a1, a2, a3);
  dft4<INV>(b0, b1, b2, b3);
  v[0] = a0;
  v[2] = a1;
  v[4] = a2;
  v[6] = a3;
  v[1] = b0;
  v[3] = b1;
  v[5] = b2;
  v[7] = b3;
}

// Pass plan of a length-N line.
template <int N>
struct Passes {
  static constexpr int M = ilog2<N>();
  static constexpr int P8 = M / 3;                       // radix-8 passes
  static constexpr int REM = M % 3;                      // trailing radix-2^REM pass
  static constexpr int R = 1 << REM;                     // its radix (1: none)
  static constexpr int NB = REM ? 8 / R : 0;             // its butterflies per thread
  static constexpr int G = N / 8;                        // threads per line
  static constexpr int NTW = 7 * (P8 - 1) + NB * (R - 1);  // twiddles per thread
};

// The twiddles one thread needs across all passes of a forward transform
// (an inverse transform uses their conjugates), read once from the table
// tw[k] = exp(-2 pi i k / N) in global memory and kept in registers, so
// every transform the thread runs reuses them.
template <int N, typename T>
struct Twiddles {
  static constexpr bool kReload = false;
  static constexpr bool kPacked = false;
  cplx<T> w[Passes<N>::NTW > 0 ? Passes<N>::NTW : 1];
  // twiddle number wi of this thread (idx: its index in the table)
  __device__ __forceinline__ cplx<T> at(int wi, int) const { return w[wi]; }
  __device__ __forceinline__ void load(const cplx<T>* __restrict__ tw, int t) {
    using P = Passes<N>;
    int i = 0;
    int Ns = 8;
#pragma unroll
    for (int p = 1; p < P::P8; ++p) {
      const int k = t & (Ns - 1);
#pragma unroll
      for (int q = 1; q < 8; ++q) w[i++] = tw[(q * k) * (N / (Ns * 8))];
      Ns *= 8;
    }
    if constexpr (P::REM != 0) {
#pragma unroll
      for (int b = 0; b < P::NB; ++b) {
        const int k = (t + b * P::G) & (Ns - 1);
#pragma unroll
        for (int q = 1; q < P::R; ++q) w[i++] = tw[(q * k) * (N / (Ns * P::R))];
      }
    }
  }
};

// The same twiddles read at each use from a table in LDS, for kernels whose
// occupancy is set by register pressure and which run several transforms:
// the registers are then taken only while a transform runs.  The table holds
// each thread group's twiddles side by side -- radix-8 pass p (1 <= p < P8),
// group ktw < 8^p: W^(q ktw N / 8^(p+1)), q = 1..7, at pass_off(p) + 7 ktw +
// q - 1; the trailing radix-R pass, group ktw < N/R: W^(q ktw), q = 1..R-1,
// at kRemOff + (R-1) ktw + q - 1 -- so a pass reads them from one address
// with immediate offsets (fill with fill_twiddles_lds).  The table lies in
// its own __shared__ array, so the compiler could prove its values unchanged
// from one transform to the next and keep them in registers; fft_lines
// therefore passes the thread index through an empty asm statement first
// (kReload), which makes every transform's table addresses opaque to it.
template <int N, typename T>
struct TwiddlesLds {
  using P = Passes<N>;
  static constexpr bool kReload = true;
  static constexpr bool kPacked = true;
  static constexpr __host__ __device__ int pass_off(int p) {
    int o = 0, ns = 8;
    for (int i = 1; i < p; ++i) {
      o += 7 * ns;
      ns *= 8;
    }
    return o;
  }
  static constexpr int kRemOff = pass_off(P::P8);
  static constexpr int kLen = kRemOff + (P::REM ? (P::R - 1) * (N / P::R) : 0);
  const cplx<T>* tab;
  __device__ __forceinline__ cplx<T> pass_tw(int p, int ktw, int q) const {
    return tab[pass_off(p) + 7 * ktw + (q - 1)];
  }
  __device__ __forceinline__ cplx<T> rem_tw(int ktw, int q) const {
    return tab[kRemOff + (P::R - 1) * ktw + (q - 1)];
  }
};

// Fill the grouped twiddle table of TwiddlesLds<N, T> (kLen entries) from the
// table tw[k] = exp(-2 pi i k / N) in global memory; threads tid of nthr.
template <int N, typename T>
__device__ __forceinline__ void fill_twiddles_lds(cplx<T>* tab, const cplx<T>* __restrict__ twg,
                                                  int tid, int nthr) {
  using TL = TwiddlesLds<N, T>;
  constexpr int R = Passes<N>::R;
  for (int e = tid; e < TL::kLen; e += nthr) {
    int idx;
    if (e < TL::kRemOff) {
      int o = 0, ns = 8;
      while (e >= o + 7 * ns) {
        o += 7 * ns;
        ns *= 8;
      }
      const int ktw = (e - o) / 7, q = (e - o) % 7 + 1;
      idx = (q * ktw) * (N / (ns * 8));
    } else {
      const int ktw = (e - TL::kRemOff) / (R - 1), q = (e - TL::kRemOff) % (R - 1) + 1;
      idx = q * ktw;
    }
    tab[e] = twg[idx];
  }
}

// Synchronisation of an LDS exchange.  Block scope: __syncthreads().  Wave
// scope (WAVE): for kernels in which the threads of every line lie inside one
// wavefront (N/8 threads per line, groups aligned to the 64 lanes).  LDS
// operations of a wavefront execute in issue order, so only the compiler has
// to be kept from moving LDS accesses across the exchange -- the fences and
// wave barrier do that (the idiom of rocPRIM's wave_barrier()), and the
// other wavefronts of the block run on unhindered.  Every lane of the
// wavefront must reach it.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

template <bool INV, typename T>
__device__ __forceinline__ cplx<T> cmul_tw(cplx<T> a, cplx<T> w) {
  return INV ? cmulc(a, w) : cmul(a, w);
}

// NL independent lines, each handled by the same group of N/8 threads
// (thread index t within the group).  bufs[l] points to that line's LDS
// buffer (lds_len<N, SW>() entries in layout SW), tw to the thread's
// twiddles (Twiddles or TwiddlesLds).  Every thread of the workgroup (WAVE: of the wavefront) must
// call this the same number of times: the passes synchronise with
// lds_sync<WAVE>().
template <int N, bool INV, int NL, bool WAVE = false, bool SW = lds_swizzled<N>(), typename T,
          typename TW>
__device__ __forceinline__ void fft_lines(cplx<T> (&v)[NL][8], cplx<T>* const (&bufs)[NL],
                                          const TW& tw, int t) {
  using P = Passes<N>;
  constexpr int P8 = P::P8, REM = P::REM, G = P::G;
  static_assert(N >= 16 && N <= 512 && (N & (N - 1)) == 0, "N must be 2^m, 16..512");
  int tt = t;  // thread index for the twiddle addresses
  if constexpr (TW::kReload) __asm__ volatile("" : "+v"(tt));

  int Ns = 1;
  int wi = 0;
#pragma unroll
  for (int p = 0; p < P8; ++p) {
    // radix-8 butterfly j = t on points t + q*N/8
    const int k = t & (Ns - 1);
    if (p > 0) {
      const int ktw = tt & (Ns - 1);
#pragma unroll
      for (int q = 1; q < 8; ++q) {
        cplx<T> w;
        if constexpr (TW::kPacked)
          w = tw.pass_tw(p, ktw, q);
        else
          w = tw.at(wi, (q * ktw) * (N / (Ns * 8)));
        ++wi;
#pragma unroll
        for (int l = 0; l < NL; ++l) v[l][q] = cmul_tw<INV>(v[l][q], w);
      }
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) dft8<INV>(v[l]);
    const bool last = (p == P8 - 1) && (REM == 0);
    if (!last) {
      // Ns = 8^p; with Ns == 1 the base 8 t is a multiple of 8
      const int base = (t / Ns) * Ns * 8 + k;
      lds_sync<WAVE>();
#pragma unroll
      for (int l = 0; l < NL; ++l)
#pragma unroll
        for (int q = 0; q < 8; ++q) bufs[l][lds_idx_q<N, SW>(base, q, Ns)] = v[l][q];
      lds_sync<WAVE>();
      Ns *= 8;
      if (p < P8 - 1) {
#pragma unroll
        for (int l = 0; l < NL; ++l)
#pragma unroll
          for (int q = 0; q < 8; ++q) v[l][q] = bufs[l][line_idx<N, SW>(t, q)];
      }
    }
  }
  if constexpr (REM != 0) {
    constexpr int R = P::R;     // 2 or 4
    constexpr int NB = P::NB;   // butterflies per thread
    // Ns == N / R here.  Butterfly b handles j = t + b*G on points
    // j + q*N/R = t + (b + q*NB)*G -> register index b + q*NB.
#pragma unroll
    for (int l = 0; l < NL; ++l)
#pragma unroll
      for (int m = 0; m < 8; ++m) v[l][m] = bufs[l][line_idx<N, SW>(t, m)];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ktw = (tt + b * G) & (Ns - 1);  // Ns * R == N
#pragma unroll
      for (int q = 1; q < R; ++q) {
        cplx<T> w;
        if constexpr (TW::kPacked)
          w = tw.rem_tw(ktw, q);
        else
          w = tw.at(wi, q * ktw);
        ++wi;
#pragma unroll
        for (int l = 0; l < NL; ++l) v[l][b + q * NB] = cmul_tw<INV>(v[l][b + q * NB], w);
      }
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        if constexpr (R == 2) {
          dft2<INV>(v[l][b], v[l][b + NB]);
        } else {
          dft4<INV>(v[l][b], v[l][b + NB], v[l][b + 2 * NB], v[l][b + 3 * NB]);
        }
      }
    }
    // outputs of butterfly j land on j + q*N/R: the same register slots.
  }
}

// Cooperative load of the twiddle table into LDS (whole workgroup).
template <int N, typename T>
__device__ __forceinline__ void load_twiddles(cplx<T>* tw_lds, const cplx<T>* tw_glob) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) tw_lds[i] = tw_glob[i];
}

}  // namespace cmbl
