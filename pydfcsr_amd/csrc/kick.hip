// a16: mesh -> particle kick.
//
// Replaces Beam.apply_wakes (beams.py:108-131): the wake grids are scaled
// (step_size * W * 1e6 / E0, numpy's left-to-right order) and interpolated
// with scipy RegularGridInterpolator's linear rule on the linspace axes
// (xrange, zrange) at (x_transform, z), fill 0 outside; pz and (if
// transverse_on) px are incremented in place.
#include "common.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;

__device__ inline double scaled(const double *W, int64_t o, double step, double E0) {
#pragma clang fp contract(off)
  double v = step * W[o];
  v = v * 1e6;
  return v / E0;
}

__global__ __launch_bounds__(kThreads) void kick_kernel(
    double *__restrict__ px, double *__restrict__ pz, const double *__restrict__ x,
    const double *__restrict__ z, int64_t n, const double *__restrict__ dE,
    const double *__restrict__ xk, int nx, int nz, double xr0, double xr1, double zr0,
    double zr1, double p0, double p1, double step, double E0, int transverse_on) {
#pragma clang fp contract(off)
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kThreads) {
    const double zv = z[p];
    double pv = p0 * zv;
    pv = pv + p1;
    const double xt = x[p] - pv;   // x - np.polyval(slope, z)
    double dEv = 0.0, dxv = 0.0;
    const bool in = xt >= xr0 && xt <= xr1 && zv >= zr0 && zv <= zr1 && nx >= 2 && nz >= 2;
    if (in) {
      int i = rgi_interval(xt, xr0, xr1, nx);
      int j = rgi_interval(zv, zr0, zr1, nz);
      double gi = linspace_at(xr0, xr1, nx, i), gi1 = linspace_at(xr0, xr1, nx, i + 1);
      double gj = linspace_at(zr0, zr1, nz, j), gj1 = linspace_at(zr0, zr1, nz, j + 1);
      double y0 = (xt - gi) / (gi1 - gi);
      double y1 = (zv - gj) / (gj1 - gj);
      int64_t o = (int64_t)i * nz + j;
      double r = scaled(dE, o, step, E0) * (1 - y0) * (1 - y1);
      r = r + scaled(dE, o + 1, step, E0) * (1 - y0) * y1;
      r = r + scaled(dE, o + nz, step, E0) * y0 * (1 - y1);
      r = r + scaled(dE, o + nz + 1, step, E0) * y0 * y1;
      dEv = r;
      if (transverse_on) {
        double q = scaled(xk, o, step, E0) * (1 - y0) * (1 - y1);
        q = q + scaled(xk, o + 1, step, E0) * (1 - y0) * y1;
        q = q + scaled(xk, o + nz, step, E0) * y0 * (1 - y1);
        q = q + scaled(xk, o + nz + 1, step, E0) * y0 * y1;
        dxv = q;
      }
    }
    pz[p] = pz[p] + dEv;
    if (transverse_on) px[p] = px[p] + dxv;
  }
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int dfcsr_apply_wakes(double *px, double *pz, const double *x, const double *z,
                                 int64_t n, const double *dE_dct, const double *x_kick, int xbins,
                                 int zbins, double xr0, double xr1, double zr0, double zr1,
                                 double slope0, double slope1, double step_size,
                                 double init_energy, int transverse_on, void *stream) {
  if (!pz || !x || !z || !dE_dct || (transverse_on && (!px || !x_kick)))
    return set_error(DFCSR_EINVAL, "apply_wakes: NULL pointer");
  if (xbins < 1 || zbins < 1) return set_error(DFCSR_EINVAL, "apply_wakes: bad mesh");
  if (n <= 0) return DFCSR_OK;
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(kick_kernel, dim3((int)b), dim3(kThreads), 0, as_stream(stream), px, pz, x, z,
                     n, dE_dct, x_kick, xbins, zbins, xr0, xr1, zr0, zr1, slope0, slope1,
                     step_size, init_energy, transverse_on);
  DFCSR_LAUNCH_CHECK("apply_wakes");
  return DFCSR_OK;
}
