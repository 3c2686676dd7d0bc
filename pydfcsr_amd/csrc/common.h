// Shared device/host helpers for the gfx950 CSR step library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "dfcsr.h"

#define DFCSR_WAVE 64

namespace dfcsr {

// Records the message for dfcsr_last_error() and returns code.
int set_error(int code, const char *fmt, ...);

#define DFCSR_CHECK_HIP(expr)                                                    \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess)                                                        \
      return ::dfcsr::set_error(DFCSR_EHIP, "%s failed: %s", #expr,              \
                                hipGetErrorString(_e));                          \
  } while (0)

#define DFCSR_LAUNCH_CHECK(name)                                                 \
  do {                                                                           \
    hipError_t _e = hipGetLastError();                                           \
    if (_e != hipSuccess)                                                        \
      return ::dfcsr::set_error(DFCSR_EHIP, "launch of %s failed: %s", name,     \
                                hipGetErrorString(_e));                          \
  } while (0)

// np.linspace(start, stop, num)[i]: y = i*step + start with step =
// (stop - start)/(num - 1), two roundings (no fused multiply-add), and the
// last point pinned to stop.  num == 1 gives start.
__host__ __device__ inline double linspace_at(double start, double stop, int num, int i) {
#pragma clang fp contract(off)
  if (num <= 1) return start;
  if (i == num - 1) return stop;
  double step = (stop - start) / (double)(num - 1);
  double y = (double)i * step;
  return y + start;
}

// Index rule of pyDFCSR's interpolate3D / interpolate1D (interp3D.py:30-52):
// i0 = int(u) (truncation), i1 = i0 at the last plane else i0 + 1, valid iff
// i0 >= 0 and i1 < size.  Equivalent to -1 < u < size; NaN is invalid.
struct TruncIdx {
  int i0, i1;
  double frac;
  bool ok;
};

__host__ __device__ inline TruncIdx trunc_index(double u, int size) {
#pragma clang fp contract(off)
  TruncIdx r;
  r.ok = (u > -1.0) && (u < (double)size);
  double f = r.ok ? trunc(u) : 0.0;
  r.i0 = (int)f;
  r.i1 = (r.i0 == size - 1) ? r.i0 : r.i0 + 1;
  r.frac = r.ok ? (u - f) : 0.0;
  return r;
}

// Interval search of scipy's RegularGridInterpolator on a linspace axis:
// the i with g[i] <= v < g[i+1], clipped to [0, num-2] (v == g[num-1] gives
// num-2), using the exact linspace values.  Caller handles out-of-bounds.
__host__ __device__ inline int rgi_interval(double v, double g0, double g1, int num) {
#pragma clang fp contract(off)
  double step = (g1 - g0) / (double)(num - 1);
  double u = (v - g0) / step;
  int i = (u >= 0.0) ? (int)u : 0;
  if (i > num - 2) i = num - 2;
  if (i < 0) i = 0;
  while (i > 0 && linspace_at(g0, g1, num, i) > v) --i;
  while (i < num - 2 && linspace_at(g0, g1, num, i + 1) <= v) ++i;
  return i;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, DFCSR_WAVE);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, DFCSR_WAVE));
  return v;
}

// Deterministic block sum (fixed tree), result valid in all threads.
// smem: at least blockDim.x/64 doubles.
__device__ inline double block_sum(double v, double *smem) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double r = 0.0;
  for (int i = 0; i < nw; ++i) r += smem[i];
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double *smem) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  double r = smem[0];
  for (int i = 1; i < nw; ++i) r = fmax(r, smem[i]);
  __syncthreads();
  return r;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace dfcsr
