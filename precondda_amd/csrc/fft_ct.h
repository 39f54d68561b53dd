// Compile-time-planned Stockham FFT over a tile of columns, for the FFT lengths of the
// solve workloads (16 ... 512). Radices, strides and twiddle strides are template
// constants, so the per-butterfly index work is shifts and multiplies by constants;
// the first pass may read its inputs straight from global memory and the last pass may
// write straight to global memory, which leaves one LDS exchange for two-pass lengths.
//
// Tile columns: C = c3 * 2^lg (c3 in {1, 3}); thread t serves column t mod C and the
// butterflies j = t / C + q * (NT / C) of every pass.
#pragma once

#include "fft_tile.h"

namespace pdda {

// cos(j pi / 48), j = 0..24: roots of unity of order 96 (covers radices 6, 12, 16, 32).
__host__ __device__ constexpr double kCos96[25] = {
    1.0, 0.99785892323860350674, 0.99144486137381041115, 0.98078528040323044913, 0.96592582628906828675,
    0.94693012949510566426, 0.92387953251128675613, 0.89687274153268830389, 0.86602540378443864676,
    0.83146961230254523708, 0.79335334029123516458, 0.75183980747897739641, 0.70710678118654752440,
    0.65934581510006886843, 0.60876142900872063942, 0.55557023301960222474, 0.5,
    0.44228869021900128200, 0.38268343236508977173, 0.32143946530316158070, 0.25881904510252076235,
    0.19509032201612826785, 0.13052619222005159155, 0.06540312923014306682, 0.0};

__host__ __device__ constexpr double cos96(int a) {
  a = ((a % 96) + 96) % 96;
  return a <= 24 ? kCos96[a] : a <= 48 ? -kCos96[48 - a] : a <= 72 ? -kCos96[a - 48] : kCos96[96 - a];
}
// exp(-2 pi i m / N) for N dividing 96
template <int N> __host__ __device__ constexpr double root_re(int m) {
  static_assert(96 % N == 0, "exact roots of unity of order dividing 96 only");
  return cos96(m * (96 / N));
}
template <int N> __host__ __device__ constexpr double root_im(int m) {
  static_assert(96 % N == 0, "exact roots of unity of order dividing 96 only");
  return -cos96(24 - m * (96 / N));
}

template <typename T, int N> struct DftC;  // forward DFT of N registers

template <typename T> struct DftC<T, 2> { static __device__ inline void run(C<T>* v) { Dft<T, 2>::run(v, nullptr, 0); } };
template <typename T> struct DftC<T, 3> { static __device__ inline void run(C<T>* v) { Dft<T, 3>::run(v, nullptr, 0); } };
template <typename T> struct DftC<T, 4> { static __device__ inline void run(C<T>* v) { Dft<T, 4>::run(v, nullptr, 0); } };
template <typename T> struct DftC<T, 5> { static __device__ inline void run(C<T>* v) { Dft<T, 5>::run(v, nullptr, 0); } };
template <typename T> struct DftC<T, 8> { static __device__ inline void run(C<T>* v) { Dft<T, 8>::run(v, nullptr, 0); } };

// N = N1 * N2 by one Cooley-Tukey step: n = n2 + N2 n1, k = k1 + N1 k2.
template <typename T, int N1, int N2> __device__ inline void dft_compose(C<T>* v) {
  constexpr int N = N1 * N2;
  C<T> a[N2][N1];
#pragma unroll
  for (int n2 = 0; n2 < N2; ++n2) {
    C<T> t[N1];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) t[n1] = v[n2 + N2 * n1];
    DftC<T, N1>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
      const int m = (n2 * k1) % N;
      a[n2][k1] = (m == 0) ? t[k1] : cmul(t[k1], mk<T>(T(root_re<N>(m)), T(root_im<N>(m))));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < N1; ++k1) {
    C<T> t[N2];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) t[n2] = a[n2][k1];
    DftC<T, N2>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) v[k1 + N1 * k2] = t[k2];
  }
}

template <typename T> struct DftC<T, 6> { static __device__ inline void run(C<T>* v) { dft_compose<T, 3, 2>(v); } };
template <typename T> struct DftC<T, 12> { static __device__ inline void run(C<T>* v) { dft_compose<T, 4, 3>(v); } };
template <typename T> struct DftC<T, 16> { static __device__ inline void run(C<T>* v) { dft_compose<T, 4, 4>(v); } };
template <typename T> struct DftC<T, 32> { static __device__ inline void run(C<T>* v) { dft_compose<T, 8, 4>(v); } };

// Compile-time radix plans of the fast path.
template <int... R> struct RList {};
template <int L> struct LPlan { static constexpr bool ok = false; };
#define PDDA_LPLAN(Lv, ...)                      \
  template <> struct LPlan<Lv> {                 \
    static constexpr bool ok = true;             \
    using type = RList<__VA_ARGS__>;             \
  };
PDDA_LPLAN(16, 4, 4)
PDDA_LPLAN(24, 8, 3)
PDDA_LPLAN(32, 8, 4)
PDDA_LPLAN(48, 8, 6)
PDDA_LPLAN(64, 8, 8)
PDDA_LPLAN(96, 16, 6)
PDDA_LPLAN(128, 16, 8)
PDDA_LPLAN(192, 16, 12)
PDDA_LPLAN(256, 16, 16)
PDDA_LPLAN(384, 8, 8, 6)
PDDA_LPLAN(512, 8, 8, 8)
#undef PDDA_LPLAN

// Column map: C = c3 * 2^lg columns; thread t -> (col, jb), stride jstep = NT / C.
struct ColMap {
  int lg, three, C, jstep;
  // nt: threads of the workgroup (FFT_NT, or a wider tile's 512)
  __device__ static ColMap make(int C_, int nt = FFT_NT) {
    ColMap m;
    m.C = C_;
    m.three = (C_ % 3 == 0) ? 1 : 0;
    int p = m.three ? C_ / 3 : C_;
    m.lg = 0;
    while ((1 << m.lg) < p) ++m.lg;
    m.jstep = nt / C_;
    return m;
  }
  __device__ inline int div(int x) const { return three ? (x >> lg) / 3 : (x >> lg); }
};

// Elements of a tile per thread on the fast path (register budget per pass).
template <typename T, int L> constexpr int ct_ept() {
  return sizeof(T) == 8 ? (L >= 384 ? 16 : 12) : (L >= 384 ? 32 : 24);
}

// One Stockham pass: element loads through `ld(pos, col)`, stores through
// `st(pos, col, v)`. MID_SYNC barriers between the loads and the stores (in-place LDS).
template <typename T, int L, int R, int NS, bool INV, int EPT, bool MID_SYNC, class LD, class ST>
__device__ inline void ct_pass(const ColMap& cm, const C<T>* tw, LD ld, ST st) {
  constexpr int LR = L / R;
  constexpr int MAXB = (EPT + R - 1) / R;
  constexpr int TS = L / (NS * R);
  const int t = threadIdx.x;
  const int jb = cm.div(t);
  const int col = t - jb * cm.C;
  const bool active = jb < cm.jstep;
  C<T> v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int j = jb + q * cm.jstep;
    if (active && j < LR) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = ld(j + r * LR, col);
    }
  }
  if (MID_SYNC) __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int j = jb + q * cm.jstep;
    if (active && j < LR) {
      const int k = j % NS;
      if (INV) v[q][0] = cconj(v[q][0]);
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const C<T> w = tw[r * k * TS];
          v[q][r] = INV ? cmul(cconj(v[q][r]), w) : cmul(v[q][r], w);
        }
      } else if (INV) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = cconj(v[q][r]);
      }
      DftC<T, R>::run(v[q]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) st(base + r * NS, col, INV ? cconj(v[q][r]) : v[q][r]);
    }
  }
}

// Runs the passes of RList<R...> with Ns = NS. First pass: loads via ld0 (barrier
// between its loads and stores only if FIRST_SYNC, i.e. ld0 reads the LDS tile);
// middle passes: LDS in place; last pass: stores via stN. Barriers separate passes.
template <typename T, int L, bool INV, int EPT, bool FIRST_SYNC, bool FIRST, int NS, class LD0, class STN, int R0,
          int... RR>
__device__ inline void ct_run(const ColMap& cm, const C<T>* tw, C<T>* lds, LD0 ld0, STN stN) {
  auto ldl = [&](int pos, int col) { return lds[imul24(pos, cm.C) + col]; };
  auto stl = [&](int pos, int col, C<T> v) { lds[imul24(pos, cm.C) + col] = v; };
  constexpr bool last = sizeof...(RR) == 0;
  if constexpr (FIRST) {
    if constexpr (last) ct_pass<T, L, R0, NS, INV, EPT, FIRST_SYNC>(cm, tw, ld0, stN);
    else ct_pass<T, L, R0, NS, INV, EPT, FIRST_SYNC>(cm, tw, ld0, stl);
  } else {
    if constexpr (last) ct_pass<T, L, R0, NS, INV, EPT, true>(cm, tw, ldl, stN);
    else ct_pass<T, L, R0, NS, INV, EPT, true>(cm, tw, ldl, stl);
  }
  if constexpr (!last) {
    __syncthreads();
    ct_run<T, L, INV, EPT, FIRST_SYNC, false, NS * R0, LD0, STN, RR...>(cm, tw, lds, ld0, stN);
  }
}

template <typename T, int L, bool INV, int EPT, bool FIRST_SYNC, class LD0, class STN, int... R>
__device__ inline void ct_fft_impl(RList<R...>, const ColMap& cm, const C<T>* tw, C<T>* lds, LD0 ld0, STN stN) {
  ct_run<T, L, INV, EPT, FIRST_SYNC, true, 1, LD0, STN, R...>(cm, tw, lds, ld0, stN);
}

// Full transform of every column: inputs from ld0(pos, col), outputs to stN(pos, col, v).
// `lds` must hold L * C elements (intermediate passes). FIRST_SYNC must be set when
// ld0 reads the LDS tile itself. No barrier at entry or exit; the twiddle table must
// be complete in LDS by the second pass (the first pass uses no twiddles).
template <typename T, int L, bool INV, bool FIRST_SYNC, class LD0, class STN>
__device__ inline void ct_fft(const ColMap& cm, const C<T>* tw, C<T>* lds, LD0 ld0, STN stN) {
  ct_fft_impl<T, L, INV, ct_ept<T, L>(), FIRST_SYNC>(typename LPlan<L>::type{}, cm, tw, lds, ld0, stN);
}

}  // namespace pdda
