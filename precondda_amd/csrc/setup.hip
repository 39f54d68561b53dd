// Target and lattice setup on the device: occupancy marking of sphere / ellipsoid /
// rectangle targets, bounding box of the occupied sites, crop-and-pad to even
// dimensions, and plane-wave incident fields (any number of directions).
#include "../../include/pdda.h"
#include "pdda_common.h"

#include <algorithm>

using namespace pdda;

namespace {

// Shape records as laid out by the host (doubles):
//   [0] type (0 sphere, 1 ellipsoid, 2 rectangle)
//   [1..3] centre, [4..6] radius | semi-axes | dimensions, [7..8] value (re, im)
//   [9..14] window lo/hi per axis (site indices, hi exclusive)
constexpr int SHAPE_REC = 16;

#pragma clang fp contract(off)
__device__ inline bool inside(const double* s, double px, double py, double pz) {
  const int type = (int)s[0];
  const double eps = 1e-7;
  if (type == 0) {
    const double dx = px - s[1], dy = py - s[2], dz = pz - s[3];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    return (dist < s[4]) || (fabs(dist - s[4]) < eps);
  }
  if (type == 1) {
    const double dx = (px - s[1]) / s[4], dy = (py - s[2]) / s[5], dz = (pz - s[3]) / s[6];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    return dist <= 1.0 + eps;
  }
  const double hx = s[4] / 2, hy = s[5] / 2, hz = s[6] / 2;
  const double dx = fabs(px - s[1]), dy = fabs(py - s[2]), dz = fabs(pz - s[3]);
  const bool in = (dx <= hx) && (dy <= hy) && (dz <= hz);
  const bool edge = (fabs(dx - hx) < eps && dy <= hy && dz <= hz) || (fabs(dy - hy) < eps && dx <= hx && dz <= hz) ||
                    (fabs(dz - hz) < eps && dx <= hx && dy <= hy);
  return in || edge;
}
#pragma clang fp contract(on)

// Up to MARK_CHUNK shape records passed by value (no device buffer, no host sync).
constexpr int MARK_CHUNK = 16;
struct ShapeChunk {
  double rec[MARK_CHUNK * SHAPE_REC];
};

// first: start from zero; otherwise later shapes overwrite what earlier chunks wrote.
__global__ void k_mark(ShapeChunk chunk, int nshapes, int first, int nx, int ny, int nz, double d, double offset,
                       double2* __restrict__ out) {
  const long long n = (long long)nx * ny * nz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i / ((long long)ny * nz));
    const int iy = (int)((i / nz) % ny);
    const int iz = (int)(i % nz);
    const double px = (ix + offset) * d, py = (iy + offset) * d, pz = (iz + offset) * d;
    double2 v = first ? make_double2(0.0, 0.0) : out[i];
    for (int q = 0; q < nshapes; ++q) {
      const double* s = chunk.rec + q * SHAPE_REC;
      if (ix < (int)s[9] || ix >= (int)s[10] || iy < (int)s[11] || iy >= (int)s[12] || iz < (int)s[13] ||
          iz >= (int)s[14])
        continue;
      if (inside(s, px, py, pz) && (s[7] != 0.0 || s[8] != 0.0)) v = make_double2(s[7], s[8]);
    }
    out[i] = v;
  }
}

__global__ void k_bbox_init(int* bbox, int nx, int ny, int nz) {
  if (threadIdx.x == 0) {
    bbox[0] = nx; bbox[1] = ny; bbox[2] = nz;
    bbox[3] = -1; bbox[4] = -1; bbox[5] = -1;
  }
}

// bbox[0..2] = min index per axis, bbox[3..5] = max index per axis of nonzero sites.
// Reduced within each wave first: one lane per wave issues the six atomics.
__global__ void k_bbox(const double2* __restrict__ a, int nx, int ny, int nz, int* bbox) {
  const long long n = (long long)nx * ny * nz;
  int lo[3] = {nx, ny, nz}, hi[3] = {-1, -1, -1};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double2 v = a[i];
    if (v.x != 0.0 || v.y != 0.0) {
      const int c[3] = {(int)(i / ((long long)ny * nz)), (int)((i / nz) % ny), (int)(i % nz)};
      for (int q = 0; q < 3; ++q) {
        lo[q] = min(lo[q], c[q]);
        hi[q] = max(hi[q], c[q]);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int q = 0; q < 3; ++q) {
      lo[q] = min(lo[q], __shfl_xor(lo[q], off, 64));
      hi[q] = max(hi[q], __shfl_xor(hi[q], off, 64));
    }
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 3; ++q) {
      atomicMin(bbox + q, lo[q]);
      atomicMax(bbox + 3 + q, hi[q]);
    }
}

__global__ void k_crop(const double2* __restrict__ in, int nx, int ny, int nz, int x0, int y0, int z0, int cx, int cy,
                       int cz, int ox, int oy, int oz, double2* __restrict__ out) {
  const long long n = (long long)ox * oy * oz;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int ix = (int)(i / ((long long)oy * oz));
    const int iy = (int)((i / oz) % oy);
    const int iz = (int)(i % oz);
    double2 v = make_double2(0.0, 0.0);
    if (ix < cx && iy < cy && iz < cz) v = in[((long long)(x0 + ix) * ny + (y0 + iy)) * nz + (z0 + iz)];
    out[i] = v;
  }
}

#pragma clang fp contract(off)
// E[r][i,j,l,c] = e_r[c] * exp(i k_r . (i,j,l) d); dirs holds (kx,ky,kz,ex,ey,ez) per rhs,
// k already multiplied in. Phase accumulated in FP64 from integer indices.
// Up to WAVE_CHUNK directions passed by value (no device buffer, no host sync).
constexpr int WAVE_CHUNK = 64;
struct WaveChunk {
  double v[WAVE_CHUNK * 6];
};

// E holds the chunk's first RHS at offset 0.
__global__ void k_plane_wave(WaveChunk dirs, int nrhs, int nx, int ny, int nz, double d, double2* __restrict__ E) {
  const long long n = (long long)nx * ny * nz;
  const long long tot = n * nrhs;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(t / n);
    const long long i = t - (long long)r * n;
    const int ix = (int)(i / ((long long)ny * nz));
    const int iy = (int)((i / nz) % ny);
    const int iz = (int)(i % nz);
    const double* q = dirs.v + 6 * r;
    const double ph = q[0] * (ix * d) + q[1] * (iy * d) + q[2] * (iz * d);
    double sn, cs;
    sincos(ph, &sn, &cs);
    double2* e = E + 3 * t;
    for (int c = 0; c < 3; ++c) e[c] = make_double2(q[3 + c] * cs, q[3 + c] * sn);
  }
}
#pragma clang fp contract(on)

}  // namespace

extern "C" {

int pdda_mark_shapes(const double* shapes_host, int nshapes, int nx, int ny, int nz, double d, int voxel_center,
                     void* out_c128, void* stream) {
  if (!shapes_host || nshapes < 0 || nx <= 0 || ny <= 0 || nz <= 0 || !out_c128) {
    set_error("pdda_mark_shapes: bad arguments");
    return PDDA_ERR_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  // records go in kernel arguments, MARK_CHUNK at a time, in shape order
  for (int q0 = 0; q0 == 0 || q0 < nshapes; q0 += MARK_CHUNK) {
    ShapeChunk ch{};
    const int nq = std::min(MARK_CHUNK, nshapes - q0);
    for (int i = 0; i < nq * SHAPE_REC; ++i) ch.rec[i] = shapes_host[(size_t)q0 * SHAPE_REC + i];
    hipLaunchKernelGGL(k_mark, dim3(2048), dim3(256), 0, s, ch, std::max(nq, 0), q0 == 0 ? 1 : 0, nx, ny, nz, d,
                       voxel_center ? 0.5 : 0.0, (double2*)out_c128);
    PDDA_LAUNCH_CHECK();
  }
  return PDDA_OK;
}

int pdda_trim_bbox(const void* a_c128, int nx, int ny, int nz, int* bbox_host, void* stream) {
  if (!a_c128 || !bbox_host) { set_error("pdda_trim_bbox: bad arguments"); return PDDA_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  int* db = nullptr;
  // stream-ordered allocation: no device-wide synchronisation (other streams keep running)
  PDDA_HIP_CHECK(hipMallocAsync((void**)&db, 6 * sizeof(int), s));
  hipLaunchKernelGGL(k_bbox_init, dim3(1), dim3(64), 0, s, db, nx, ny, nz);
  hipLaunchKernelGGL(k_bbox, dim3(1024), dim3(256), 0, s, (const double2*)a_c128, nx, ny, nz, db);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(bbox_host, db, 6 * sizeof(int), hipMemcpyDeviceToHost, s);
  const hipError_t ef = hipFreeAsync(db, s);
  if (e == hipSuccess) e = ef;
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { set_error(std::string("pdda_trim_bbox: ") + hipGetErrorString(e)); return PDDA_ERR_HIP; }
  return PDDA_OK;
}

int pdda_crop_pad(const void* in_c128, int nx, int ny, int nz, int x0, int y0, int z0, int cx, int cy, int cz,
                  int ox, int oy, int oz, void* out_c128, void* stream) {
  if (!in_c128 || !out_c128 || cx > ox || cy > oy || cz > oz || x0 < 0 || y0 < 0 || z0 < 0 || x0 + cx > nx ||
      y0 + cy > ny || z0 + cz > nz) {
    set_error("pdda_crop_pad: bad arguments");
    return PDDA_ERR_ARG;
  }
  hipLaunchKernelGGL(k_crop, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const double2*)in_c128, nx, ny, nz, x0,
                     y0, z0, cx, cy, cz, ox, oy, oz, (double2*)out_c128);
  PDDA_LAUNCH_CHECK();
  return PDDA_OK;
}

int pdda_plane_wave(const double* dirs_host, int nrhs, int nx, int ny, int nz, double d, void* out_c128,
                    void* stream) {
  if (!dirs_host || nrhs <= 0 || !out_c128) { set_error("pdda_plane_wave: bad arguments"); return PDDA_ERR_ARG; }
  hipStream_t s = (hipStream_t)stream;
  const long long n3 = 3LL * nx * ny * nz;
  hipError_t e = hipSuccess;
  for (int r0 = 0; r0 < nrhs && e == hipSuccess; r0 += WAVE_CHUNK) {
    WaveChunk ch{};
    const int nr = std::min(WAVE_CHUNK, nrhs - r0);
    for (int i = 0; i < nr * 6; ++i) ch.v[i] = dirs_host[(size_t)r0 * 6 + i];
    hipLaunchKernelGGL(k_plane_wave, dim3(2048), dim3(256), 0, s, ch, nr, nx, ny, nz, d,
                       (double2*)out_c128 + (size_t)r0 * n3);
    e = hipGetLastError();
  }
  if (e != hipSuccess) { set_error(std::string("pdda_plane_wave: ") + hipGetErrorString(e)); return PDDA_ERR_HIP; }
  return PDDA_OK;
}

}  // extern "C"
