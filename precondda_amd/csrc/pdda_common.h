// Shared device/host helpers for the gfx950 DDA kernels: complex arithmetic on
// float2/double2, wave64 reductions, error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace pdda {

// ---------------------------------------------------------------------------------
// Complex helpers. Vectors are interleaved (re, im) pairs, so a complex element is a
// float2 / double2 and loads/stores are 8- or 16-byte vector accesses.
// ---------------------------------------------------------------------------------
template <typename T> struct cvec;
template <> struct cvec<float> { using type = float2; };
template <> struct cvec<double> { using type = double2; };
template <typename T> using C = typename cvec<T>::type;

template <typename T> __host__ __device__ inline C<T> mk(T re, T im) { C<T> r; r.x = re; r.y = im; return r; }
template <typename V> __host__ __device__ inline V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __host__ __device__ inline V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename V> __host__ __device__ inline V cmul(V a, V b) {
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
// conj(a) * b
template <typename V> __host__ __device__ inline V cmulc(V a, V b) {
  V r; r.x = a.x * b.x + a.y * b.y; r.y = a.x * b.y - a.y * b.x; return r;
}
template <typename V> __host__ __device__ inline V cconj(V a) { a.y = -a.y; return a; }
template <typename V, typename S> __host__ __device__ inline V cscale(V a, S s) { a.x *= s; a.y *= s; return a; }
// multiply by -i (forward-direction quarter turn)
template <typename V> __host__ __device__ inline V cmul_mi(V a) { V r; r.x = a.y; r.y = -a.x; return r; }

// complex<double> scalar used for every Krylov scalar and reduction result
struct cd { double x, y; };
__host__ __device__ inline cd cd_add(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd cd_sub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd cd_mul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline cd cd_div(cd a, cd b) {
  // Smith's algorithm keeps the quotient accurate when |b| is far from 1.
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, den = b.x + b.y * r;
    return cd{(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  double r = b.x / b.y, den = b.x * r + b.y;
  return cd{(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}

// ---------------------------------------------------------------------------------
// Wave64 / workgroup reductions of complex<double> partial sums.
// ---------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Reduces ND complex values across the workgroup; result valid in thread 0.
// scratch must hold 2*ND*(blockDim.x/64) doubles.
template <int ND>
__device__ inline void block_sum(double (&re)[ND], double (&im)[ND], double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    re[d] = wave_sum(re[d]);
    im[d] = wave_sum(im[d]);
  }
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      scratch[(2 * d) * nw + wid] = re[d];
      scratch[(2 * d + 1) * nw + wid] = im[d];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      double a = 0, b = 0;
      for (int w = 0; w < nw; ++w) { a += scratch[(2 * d) * nw + w]; b += scratch[(2 * d + 1) * nw + w]; }
      re[d] = a; im[d] = b;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------
// Host-side error state (thread-local message behind pdda_last_error()).
// ---------------------------------------------------------------------------------
void set_error(const std::string& msg);
const char* last_error();

}  // namespace pdda

#define PDDA_OK 0
#define PDDA_ERR_ARG (-1)
#define PDDA_ERR_SIZE (-2)
#define PDDA_ERR_HIP (-3)
#define PDDA_ERR_OOM (-4)
#define PDDA_ERR_STATE (-5)

#define PDDA_HIP_CHECK(expr)                                                                  \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      pdda::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                     \
      return _e == hipErrorOutOfMemory ? PDDA_ERR_OOM : PDDA_ERR_HIP;                         \
    }                                                                                         \
  } while (0)

#define PDDA_LAUNCH_CHECK() PDDA_HIP_CHECK(hipGetLastError())
