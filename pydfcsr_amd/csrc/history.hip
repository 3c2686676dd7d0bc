// a8-a10: re-gridding of density-function slices into the device history.
//
// Replaces DF_tracker.DF_interp (deposit.py:296-309) as used by
// append_interpolant for the new slice (deposit.py:328-337) and for a full
// re-grid of every logged slice (deposit.py:379-390).  Each target point
// (xi, zi) of the common interp grid takes scipy RegularGridInterpolator's
// linear value on the slice's own linspace grids: interval search, norm
// distances y0, y1 and the corner sum
//   v00(1-y0)(1-y1) + v01(1-y0)y1 + v10 y0(1-y1) + v11 y0 y1,
// with fill 0 outside the grid (mean(vx_x) of the slice for vx_x).
//
// History layout: corner-packed records (record (x, z) holds field f at
// (x, z), (x, z1), (x1, z), (x1, z1) in elements f*4 + 0..3, x1 and z1
// clamped to the last plane).  float64: whole records [x][z][slot][20];
// float32 chunked: 16-byte chunks (one field's four corners each) stored
// [x][slot][f][z], where threads of consecutive z write consecutive chunks;
// float32 time-innermost: whole records [x][z][slot][20].  Any of the three
// rings may be absent.
#include "common.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxJobs = 4096;

// per-job mean of the raw vx_x field (np.mean(vx_x), deposit.py:332, 384)
__global__ __launch_bounds__(kThreads) void vxx_mean_kernel(const double *__restrict__ src,
                                                            int64_t slot_stride,
                                                            const DfcsrRegridJob *__restrict__ jobs,
                                                            double *__restrict__ means) {
  __shared__ double red[kThreads / 64];
  const DfcsrRegridJob J = jobs[blockIdx.x];
  const int64_t N = (int64_t)J.nbx * J.nbz;
  const double *f = src + (int64_t)J.src_slot * slot_stride + (int64_t)DFCSR_F_VX_X * N;
  double s = 0.0;
  for (int64_t e = threadIdx.x; e < N; e += kThreads) s += f[e];
  s = block_sum(s, red);
  if (threadIdx.x == 0) means[blockIdx.x] = s / (double)N;
}

// RegularGridInterpolator value of the five fields of job J at (xq, zq).
__device__ inline void rgi5(const DfcsrRegridJob &J, const double *__restrict__ base,
                            double fill_vxx, double xq, double zq, double v[5]) {
#pragma clang fp contract(off)
  const int64_t Nsrc = (int64_t)J.nbx * J.nbz;
  bool oob = !(xq >= J.x0 && xq <= J.x1 && zq >= J.z0 && zq <= J.z1);
  if (oob || J.nbx < 2 || J.nbz < 2) {
    for (int f = 0; f < 5; ++f) v[f] = (f == DFCSR_F_VX_X) ? fill_vxx : 0.0;
    return;
  }
  int i = rgi_interval(xq, J.x0, J.x1, J.nbx);
  int j = rgi_interval(zq, J.z0, J.z1, J.nbz);
  double gi = linspace_at(J.x0, J.x1, J.nbx, i), gi1 = linspace_at(J.x0, J.x1, J.nbx, i + 1);
  double gj = linspace_at(J.z0, J.z1, J.nbz, j), gj1 = linspace_at(J.z0, J.z1, J.nbz, j + 1);
  double y0 = (xq - gi) / (gi1 - gi);
  double y1 = (zq - gj) / (gj1 - gj);
  int64_t o = (int64_t)i * J.nbz + j;
  for (int f = 0; f < 5; ++f) {
    const double *F = base + (int64_t)f * Nsrc;
    double r = F[o] * (1 - y0) * (1 - y1);
    r = r + F[o + 1] * (1 - y0) * y1;
    r = r + F[o + J.nbz] * y0 * (1 - y1);
    r = r + F[o + J.nbz + 1] * y0 * y1;
    v[f] = r;
  }
}

// One thread per target record: evaluates the four corner points of the
// record and writes the corner-packed record into each ring present, and
// the record's own grid node (its first corner) into each cell-major ring
// (at the last z also into the row's extra cell, its copy).
__global__ __launch_bounds__(kThreads) void regrid_kernel(
    const double *__restrict__ src, int64_t slot_stride, const DfcsrRegridJob *__restrict__ jobs,
    const double *__restrict__ means, double *__restrict__ h64, float *__restrict__ h32,
    float *__restrict__ h32t, double *__restrict__ c64, float *__restrict__ c32, int Xi, int Zi,
    int cap, double xi0, double xi1, double zi0, double zi1) {
  const DfcsrRegridJob J = jobs[blockIdx.y];
  const int64_t cells = (int64_t)Xi * Zi;
  const double *base = src + (int64_t)J.src_slot * slot_stride;
  const double fill_vxx = means[blockIdx.y];
  for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < cells;
       c += (int64_t)gridDim.x * kThreads) {
    const int ix = (int)(c / Zi), iz = (int)(c % Zi);
    const int ix1 = ix + 1 < Xi ? ix + 1 : ix, iz1 = iz + 1 < Zi ? iz + 1 : iz;
    const double x0 = linspace_at(xi0, xi1, Xi, ix), x1 = linspace_at(xi0, xi1, Xi, ix1);
    const double z0 = linspace_at(zi0, zi1, Zi, iz), z1 = linspace_at(zi0, zi1, Zi, iz1);
    double v[4][5];
    rgi5(J, base, fill_vxx, x0, z0, v[0]);
    rgi5(J, base, fill_vxx, x0, z1, v[1]);
    rgi5(J, base, fill_vxx, x1, z0, v[2]);
    rgi5(J, base, fill_vxx, x1, z1, v[3]);
    if (h64) {
      double2 *d2 = reinterpret_cast<double2 *>(h64 + (c * cap + J.dst_slot) * DFCSR_HIST_REC64);
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        d2[2 * f] = make_double2(v[0][f], v[1][f]);
        d2[2 * f + 1] = make_double2(v[2][f], v[3][f]);
      }
    }
    if (h32t) {
      float4 *t4 = reinterpret_cast<float4 *>(h32t + (c * cap + J.dst_slot) * DFCSR_HIST_REC32);
#pragma unroll
      for (int f = 0; f < 5; ++f)
        t4[f] = make_float4((float)v[0][f], (float)v[1][f], (float)v[2][f], (float)v[3][f]);
    }
    if (h32) {
      const int64_t blk = (int64_t)ix * cap + J.dst_slot;
      float4 *s4 = reinterpret_cast<float4 *>(h32) + blk * DFCSR_HIST_CHUNKS32 * Zi + iz;
#pragma unroll
      for (int f = 0; f < 5; ++f)
        s4[f * (int64_t)Zi] = make_float4((float)v[0][f], (float)v[1][f], (float)v[2][f],
                                          (float)v[3][f]);
    }
    const int64_t cell =
        (((int64_t)J.dst_slot * Xi + ix) * DFCSR_CELL_ROW(Zi) + iz) * DFCSR_CELL;
    const bool pad = iz == Zi - 1;   // also the row's copy of its last cell
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      if (c64) {
        c64[cell + f] = v[0][f];
        if (pad) c64[cell + DFCSR_CELL + f] = v[0][f];
      }
      if (c32) {
        c32[cell + f] = (float)v[0][f];
        if (pad) c32[cell + DFCSR_CELL + f] = (float)v[0][f];
      }
    }
  }
}

// dcells32 plane of logical slice i (first <= i < nt): cells32 of slice
// i + 1 minus that of slice i, 0 for the newest.  n: floats per plane, read
// as float4 when kVec (n % 4 == 0, so every plane starts 16-byte aligned).
template <bool kVec>
__global__ __launch_bounds__(kThreads) void tdiff_kernel(const float *__restrict__ c,
                                                         float *__restrict__ d, int64_t n,
                                                         int cap, int head, int nt, int first) {
  for (int i = first + (int)blockIdx.y; i < nt; i += (int)gridDim.y) {
    const int s0 = (head + i) % cap, s1 = (head + i + 1) % cap;
    const bool last = i == nt - 1;
    if (kVec) {
      const float4 *a = reinterpret_cast<const float4 *>(c + (int64_t)s0 * n);
      const float4 *b = reinterpret_cast<const float4 *>(c + (int64_t)s1 * n);
      float4 *o = reinterpret_cast<float4 *>(d + (int64_t)s0 * n);
      for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n / 4;
           e += (int64_t)gridDim.x * kThreads) {
        if (last) {
          o[e] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
          const float4 u = a[e], v = b[e];
          o[e] = make_float4(v.x - u.x, v.y - u.y, v.z - u.z, v.w - u.w);
        }
      }
    } else {
      const float *a = c + (int64_t)s0 * n, *b = c + (int64_t)s1 * n;
      float *o = d + (int64_t)s0 * n;
      for (int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x; e < n;
           e += (int64_t)gridDim.x * kThreads)
        o[e] = last ? 0.f : b[e] - a[e];
    }
  }
}

// Whether a cell can contribute to a wake integrand: density, density_x or
// density_z nonzero, or vx or vx_x not finite (beyond the float32 range for
// a float64 cell, so that a span holds for float32 copies too).
template <typename T>
__device__ inline bool cell_contributes(const T *c) {
  const T big = (T)3.4028234663852886e38;   // FLT_MAX
  const T vx = c[DFCSR_F_VX], vxx = c[DFCSR_F_VX_X];
  return c[DFCSR_F_DENSITY] != (T)0 || c[DFCSR_F_DENSITY_X] != (T)0 ||
         c[DFCSR_F_DENSITY_Z] != (T)0 || !(vx <= big && vx >= -big) ||
         !(vxx <= big && vxx >= -big);
}

// x rows per workgroup of zspan_kernel: a step refreshes two slices, so 4
// rows give 2 x 125 workgroups on a 500-row grid (16 gave 64: 49 us per step)
constexpr int kSpanRows = 4;

// zspan entries (x, slot of slice i) for x in one block of kSpanRows rows:
// each wave finds the z span of whole rows (its own cells) of slices i and
// i + 1 into LDS, then each entry joins rows x, x + 1 of both slices.
template <typename T>
__global__ __launch_bounds__(kThreads) void zspan_kernel(const T *__restrict__ cells,
                                                         uint32_t *__restrict__ zspan, int Xi,
                                                         int Zi, int cap, int head, int nt,
                                                         int first) {
  __shared__ int lo_s[2][kSpanRows + 1], hi_s[2][kSpanRows + 1];
  const int xb = blockIdx.x * kSpanRows;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = kThreads / 64;
  const int nrow = min(kSpanRows + 1, Xi - xb);   // rows xb .. xb + nrow - 1 (x + 1 clamped)
  const int64_t plane = (int64_t)Xi * DFCSR_CELL_ROW(Zi) * DFCSR_CELL;
  for (int i = first + (int)blockIdx.y; i < nt; i += (int)gridDim.y) {
    const int np = i + 1 < nt ? 2 : 1;
    for (int w = wave; w < np * nrow; w += nwave) {
      const int p = w / nrow, r = w - p * nrow;
      const T *row = cells + (int64_t)((head + i + p) % cap) * plane +
                     (int64_t)(xb + r) * DFCSR_CELL_ROW(Zi) * DFCSR_CELL;
      int lo = 0x7fffffff, hi = -1;
      for (int z = lane; z < Zi; z += 64) {
        if (cell_contributes(row + (int64_t)z * DFCSR_CELL)) {
          lo = min(lo, z);
          hi = max(hi, z);
        }
      }
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        lo = min(lo, __shfl_xor(lo, d, 64));
        hi = max(hi, __shfl_xor(hi, d, 64));
      }
      if (lane == 0) {
        lo_s[p][r] = lo;
        hi_s[p][r] = hi;
      }
    }
    __syncthreads();
    if ((int)threadIdx.x < min(kSpanRows, Xi - xb)) {
      const int r = threadIdx.x, r1 = min(r + 1, nrow - 1);
      int lo = min(lo_s[0][r], lo_s[0][r1]), hi = max(hi_s[0][r], hi_s[0][r1]);
      if (np == 2) {
        lo = min(lo, min(lo_s[1][r], lo_s[1][r1]));
        hi = max(hi, max(hi_s[1][r], hi_s[1][r1]));
      }
      // no contributing cell: zlo = 0xffff > zhi = 0
      const uint32_t e = hi < 0 ? 0xffffu : ((uint32_t)lo | ((uint32_t)hi << 16));
      zspan[(int64_t)(xb + r) * cap + (head + i) % cap] = e;
    }
    __syncthreads();   // before the next slice overwrites the row spans
  }
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int dfcsr_history_zspan(const double *cells, const float *cells32, uint32_t *zspan,
                                   int Xi, int Zi, int cap, int head, int nt, int first,
                                   void *stream) {
  if ((!cells && !cells32) || !zspan) return set_error(DFCSR_EINVAL, "history_zspan: NULL ring");
  if (Xi < 1 || Zi < 1 || Zi > 0xffff)
    return set_error(DFCSR_EINVAL, "history_zspan: grid %d x %d (Zi <= 65535)", Xi, Zi);
  if (cap < 1 || nt < 0 || nt > cap || head < 0 || head >= cap)
    return set_error(DFCSR_EINVAL, "history_zspan: cap %d head %d nt %d", cap, head, nt);
  if (first < 0) first = 0;
  if (first >= nt) return DFCSR_OK;
  const int gx = (Xi + kSpanRows - 1) / kSpanRows;
  int gy = nt - first;
  if (gy > 1024) gy = 1024;
  hipStream_t s = as_stream(stream);
  if (cells)
    hipLaunchKernelGGL(zspan_kernel<double>, dim3(gx, gy), dim3(kThreads), 0, s, cells, zspan,
                       Xi, Zi, cap, head, nt, first);
  else
    hipLaunchKernelGGL(zspan_kernel<float>, dim3(gx, gy), dim3(kThreads), 0, s, cells32, zspan,
                       Xi, Zi, cap, head, nt, first);
  DFCSR_LAUNCH_CHECK("history_zspan");
  return DFCSR_OK;
}

extern "C" int dfcsr_cells32_tdiff(const float *cells32, float *dcells32, int Xi, int Zi, int cap,
                                   int head, int nt, int first, void *stream) {
  if (!cells32 || !dcells32) return set_error(DFCSR_EINVAL, "cells32_tdiff: NULL ring");
  if (Xi < 1 || Zi < 1) return set_error(DFCSR_EINVAL, "cells32_tdiff: grid %d x %d", Xi, Zi);
  if (cap < 1 || nt < 0 || nt > cap || head < 0 || head >= cap)
    return set_error(DFCSR_EINVAL, "cells32_tdiff: cap %d head %d nt %d", cap, head, nt);
  if (first < 0) first = 0;
  if (first >= nt) return DFCSR_OK;
  const int64_t n = (int64_t)Xi * DFCSR_CELL_ROW(Zi) * DFCSR_CELL;   // the row copies too
  const bool vec = n % 4 == 0;
  const int64_t per = vec ? n / 4 : n;
  int gx = (int)((per + kThreads - 1) / kThreads);
  if (gx > 512) gx = 512;
  int gy = nt - first;
  if (gy > 1024) gy = 1024;
  hipStream_t s = as_stream(stream);
  if (vec)
    hipLaunchKernelGGL(tdiff_kernel<true>, dim3(gx, gy), dim3(kThreads), 0, s, cells32, dcells32,
                       n, cap, head, nt, first);
  else
    hipLaunchKernelGGL(tdiff_kernel<false>, dim3(gx, gy), dim3(kThreads), 0, s, cells32, dcells32,
                       n, cap, head, nt, first);
  DFCSR_LAUNCH_CHECK("cells32_tdiff");
  return DFCSR_OK;
}

extern "C" int64_t dfcsr_regrid_workspace_bytes(int njobs) {
  if (njobs < 1) njobs = 1;
  return (int64_t)njobs * (int64_t)(sizeof(DfcsrRegridJob) + sizeof(double)) + 256;
}

extern "C" int dfcsr_regrid(const double *src, int64_t src_slot_stride, const DfcsrRegridJob *jobs,
                            int njobs, double *hist64, float *hist32, float *hist32t,
                            double *cells, float *cells32, int Xi, int Zi, int cap, double xi0,
                            double xi1, double zi0, double zi1, void *ws, int64_t ws_bytes,
                            void *stream) {
  if (!src || !jobs || !ws) return set_error(DFCSR_EINVAL, "regrid: NULL pointer");
  if (!hist64 && !hist32 && !hist32t && !cells && !cells32)
    return set_error(DFCSR_EINVAL, "regrid: no history ring");
  if (njobs < 1) return DFCSR_OK;
  if (njobs > kMaxJobs) return set_error(DFCSR_ERANGE, "regrid: %d jobs > %d", njobs, kMaxJobs);
  if (Xi < 2 || Zi < 2) return set_error(DFCSR_EINVAL, "regrid: interp grid %d x %d", Xi, Zi);
  if (cap < 1) return set_error(DFCSR_EINVAL, "regrid: ring capacity %d", cap);
  if (ws_bytes < dfcsr_regrid_workspace_bytes(njobs))
    return set_error(DFCSR_ENOWS, "regrid: workspace too small");
  for (int k = 0; k < njobs; ++k) {
    const DfcsrRegridJob &J = jobs[k];
    if (J.nbx < 1 || J.nbz < 1 || (int64_t)5 * J.nbx * J.nbz > src_slot_stride || J.src_slot < 0 ||
        J.dst_slot < 0 || J.dst_slot >= cap)
      return set_error(DFCSR_EINVAL, "regrid: bad job %d (nbx %d nbz %d src %d dst %d)", k, J.nbx,
                       J.nbz, J.src_slot, J.dst_slot);
  }
  hipStream_t s = as_stream(stream);
  // Job descriptors travel in the workspace: written by a host->device copy
  // ordered on the stream.
  DfcsrRegridJob *djobs = static_cast<DfcsrRegridJob *>(ws);
  double *means = reinterpret_cast<double *>(
      static_cast<char *>(ws) + (((int64_t)njobs * sizeof(DfcsrRegridJob) + 63) / 64) * 64);
  DFCSR_CHECK_HIP(hipMemcpyAsync(djobs, jobs, (size_t)njobs * sizeof(DfcsrRegridJob),
                                 hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(vxx_mean_kernel, dim3(njobs), dim3(kThreads), 0, s, src, src_slot_stride,
                     (const DfcsrRegridJob *)djobs, means);
  const int64_t npts = (int64_t)Xi * Zi;
  int gx = (int)((npts + kThreads - 1) / kThreads);
  if (gx > 2048) gx = 2048;
  hipLaunchKernelGGL(regrid_kernel, dim3(gx, njobs), dim3(kThreads), 0, s, src, src_slot_stride,
                     (const DfcsrRegridJob *)djobs, (const double *)means, hist64, hist32, hist32t,
                     cells, cells32, Xi, Zi, cap, xi0, xi1, zi0, zi1);
  DFCSR_LAUNCH_CHECK("regrid");
  return DFCSR_OK;
}
