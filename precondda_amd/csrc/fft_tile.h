// LDS-resident mixed-radix Stockham FFT over a tile of independent columns.
//
// A tile holds `ncol` columns of length L in LDS as data[pos * ncol + col], so the
// global rows that feed it (contiguous runs of the fastest axes) land in contiguous
// LDS words and consecutive lanes of a wave touch consecutive banks. Every radix pass
// reads all of a thread's butterflies into registers, barriers, then writes them back
// in autosort order: in place in LDS, with one barrier pair per pass.
//
// Butterfly j of a pass with radix R after Ns already-combined points:
//   in:  data[j + r*L/R],  r = 0..R-1
//   twiddle r by w_L^(r*k*L/(Ns*R)) with k = j mod Ns
//   out: data[(j/Ns)*Ns*R + k + r*Ns]
// (Stockham autosort; the result is in natural order after the last pass.)
// The inverse pass uses IDFT(x) = conj(DFT(conj(x))) with the same forward table.
#pragma once

#include "pdda_common.h"

namespace pdda {

constexpr int FFT_MAX_PASSES = 16;
constexpr int FFT_NT = 256;  // threads per workgroup of every FFT kernel

// Exact unsigned division by a divisor fixed for a launch: q = (x * m) >> 32 with
// m = ceil(2^32 / d), exact whenever x * d < 2^32 (all tile and pass indices here are
// below 2^13 and divisors below 2^13).
struct FastDiv {
  unsigned long long m;
  __host__ __device__ static FastDiv make(unsigned d) {
    FastDiv f;
    f.m = ((1ull << 32) + d - 1) / d;
    return f;
  }
  __device__ inline unsigned div(unsigned x) const { return (unsigned)((x * m) >> 32); }
};

struct FftPlan {
  int L;
  int npass;
  int radix[FFT_MAX_PASSES];
  FastDiv ns_div[FFT_MAX_PASSES];  // division by Ns (product of the radices before pass p)
  int gen;  // 1 if a radix 7, 11 or 13 pass is present
};

// Elements of one tile per thread (tile capacity = FFT_NT * EPT) and the resulting
// tile capacity per precision: 48 KiB of LDS for the tile in both precisions.
template <typename T> struct TileCfg;
template <> struct TileCfg<float> { static constexpr int EPT = 24; };
template <> struct TileCfg<double> { static constexpr int EPT = 12; };
template <typename T> constexpr int tile_cap() { return FFT_NT * TileCfg<T>::EPT; }

// ---------------------------------------------------------------------------------
// Forward (exp(-2 pi i ...)) DFTs of small radices on registers.
// ---------------------------------------------------------------------------------
template <typename T, int R> struct Dft {
  // Generic radix: direct sum with roots taken from the length-L table.
  static __device__ inline void run(C<T>* v, const C<T>* tw, int L) {
    C<T> o[R], w[R];
    const int step = L / R;
    w[0] = mk<T>(T(1), T(0));
#pragma unroll
    for (int q = 1; q < R; ++q) w[q] = tw[q * step];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      C<T> acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], w[(r * m) % R]));
      o[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = o[m];
  }
};

template <typename T> struct Dft<T, 2> {
  static __device__ inline void run(C<T>* v, const C<T>*, int) {
    C<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T> struct Dft<T, 3> {
  static __device__ inline void run(C<T>* v, const C<T>*, int) {
    const T h = T(0.86602540378443864676);  // sin(2 pi / 3)
    C<T> s = cadd(v[1], v[2]), d = csub(v[1], v[2]);
    C<T> x0 = v[0];
    v[0] = cadd(x0, s);
    C<T> t = mk<T>(x0.x - T(0.5) * s.x, x0.y - T(0.5) * s.y);
    C<T> u = mk<T>(h * d.y, -h * d.x);  // -i * h * d
    v[1] = cadd(t, u);
    v[2] = csub(t, u);
  }
};

template <typename T> __device__ inline void dft4(C<T>& a0, C<T>& a1, C<T>& a2, C<T>& a3) {
  C<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <typename T> struct Dft<T, 4> {
  static __device__ inline void run(C<T>* v, const C<T>*, int) { dft4<T>(v[0], v[1], v[2], v[3]); }
};

template <typename T> struct Dft<T, 5> {
  static __device__ inline void run(C<T>* v, const C<T>*, int) {
    const T c1 = T(0.30901699437494742410), c2 = T(-0.80901699437494742410);
    const T s1 = T(0.95105651629515357212), s2 = T(0.58778525229247312917);
    C<T> a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    C<T> a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    C<T> x0 = v[0];
    v[0] = cadd(x0, cadd(a1, a2));
    C<T> t1 = mk<T>(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
    C<T> t2 = mk<T>(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
    C<T> u1 = mk<T>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    C<T> u2 = mk<T>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    // X1 = t1 - i u1, X4 = t1 + i u1, X2 = t2 - i u2, X3 = t2 + i u2
    v[1] = mk<T>(t1.x + u1.y, t1.y - u1.x);
    v[4] = mk<T>(t1.x - u1.y, t1.y + u1.x);
    v[2] = mk<T>(t2.x + u2.y, t2.y - u2.x);
    v[3] = mk<T>(t2.x - u2.y, t2.y + u2.x);
  }
};

template <typename T> struct Dft<T, 8> {
  static __device__ inline void run(C<T>* v, const C<T>*, int) {
    const T c = T(0.70710678118654752440);
    C<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T>(e0, e1, e2, e3);
    dft4<T>(o0, o1, o2, o3);
    // o_k *= w8^k, w8 = exp(-i pi/4)
    o1 = mk<T>(c * (o1.x + o1.y), c * (o1.y - o1.x));
    o2 = cmul_mi(o2);
    o3 = mk<T>(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
  }
};

// ---------------------------------------------------------------------------------
// One Stockham pass of radix R over the whole tile.
// ---------------------------------------------------------------------------------
template <typename T, int R, int EPT, bool INV>
__device__ inline void stockham_pass(C<T>* data, int L, int ncol, FastDiv dcol, int Ns, FastDiv dns,
                                     const C<T>* tw) {
  constexpr int MAXB = (EPT + R - 1) / R;
  const int LR = L / R;
  const int nb = LR * ncol;
  const int tstride = L / (Ns * R);
  C<T> v[MAXB][R];
  int jj[MAXB], cc[MAXB];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * FFT_NT;
    jj[q] = (int)dcol.div((unsigned)b);
    cc[q] = b - jj[q] * ncol;
    if (b < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = data[(jj[q] + r * LR) * ncol + cc[q]];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * FFT_NT;
    if (b < nb) {
      const int j = jj[q], col = cc[q];
      const int jd = (int)dns.div((unsigned)j);
      const int k = j - jd * Ns;
      if (INV) v[q][0] = cconj(v[q][0]);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const C<T> w = tw[r * k * tstride];
        v[q][r] = INV ? cmul(cconj(v[q][r]), w) : cmul(v[q][r], w);
      }
      Dft<T, R>::run(v[q], tw, L);
      const int base = jd * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) data[(base + r * Ns) * ncol + col] = INV ? cconj(v[q][r]) : v[q][r];
    }
  }
  __syncthreads();
}

// Full transform of every column of the tile. `data` must be complete in LDS
// (caller barriers after loading); returns after a barrier. GEN enables the direct
// radix-7/11/13 butterflies, which roughly double the register footprint, so kernels
// are instantiated both ways and the host picks GEN only for lengths that need them.
template <typename T, bool INV, bool GEN>
__device__ inline void tile_fft(C<T>* data, int ncol, const FftPlan& pl, const C<T>* tw) {
  constexpr int EPT = TileCfg<T>::EPT;
  const FastDiv dcol = FastDiv::make((unsigned)ncol);
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    const FastDiv dns = pl.ns_div[p];
    switch (R) {
      case 2: stockham_pass<T, 2, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
      case 3: stockham_pass<T, 3, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
      case 4: stockham_pass<T, 4, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
      case 5: stockham_pass<T, 5, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
      case 8: stockham_pass<T, 8, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
      default:
        if constexpr (GEN) {
          switch (R) {
            case 7: stockham_pass<T, 7, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
            case 11: stockham_pass<T, 11, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
            case 13: stockham_pass<T, 13, EPT, INV>(data, pl.L, ncol, dcol, Ns, dns, tw); break;
            default: break;  // unreachable: plans are validated on the host
          }
        }
        break;
    }
    Ns *= R;
  }
}

// Copies the length-L root table into LDS (caller barriers before use).
// Product of two non-negative indices whose product stays below 2^31 and whose factors
// are below 2^24: a full-rate 24-bit multiply (the 32-bit multiply issues at a quarter
// of the VALU rate). Tile and row offsets in the FFT passes are such products.
__device__ inline int imul24(int a, int b) { return __umul24((unsigned)a, (unsigned)b); }

template <typename T>
__device__ inline void load_twiddles(C<T>* dst, const C<T>* __restrict__ src, int L) {
  for (int i = threadIdx.x; i < L; i += blockDim.x) dst[i] = src[i];
}

}  // namespace pdda
