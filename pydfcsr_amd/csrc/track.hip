// Particle tracking through one lattice-element piece, replacing
// bmadx.track_element as CSR2D drives it (CSR.py:146-199 builds the element
// of each step piece, CSR.py:227-294 tracks it).  bmadx (Bmad-X) is absent;
// these are Bmad's maps in its canonical coordinates (x, px = Px/P0, y, py,
// z = -beta c (t - t_ref), pz = dP/P0) with P0 = p0c:
//   drift      exact field-free drift;
//   quadrupole energy-scaled linear map per step, the second-order path
//              length terms in z and the low-energy z correction;
//   sbend      exact motion in the uniform field (the particle's circle of
//              radius P_h/g meets the exit face at angle g L), hard-edge
//              pole faces rotated by e1 / e2 (linear edge kicks, vertical
//              focusing softened by the fringe integral FINT x 2 HGAP);
//   sextupole  drift - thin kick - drift per step.
// One thread per particle, coordinates updated in place ([6] rows of n).
// Built without FMA contraction: the operation order is the oracle's.
#include "common.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;

struct Coords {
  double x, px, y, py, z, pz;
};

__device__ inline double sqrt_one(double v) { return v / (sqrt(1.0 + v) + 1.0); }

// ds (beta/beta0 - 1), its pz series where mc2 (beta0 pz)^2 < 3e-7 E_tot.
__device__ inline double low_energy_dz(double pz, double p0c, double mc2, double ds) {
  const double P = 1.0 + pz;
  const double beta = P * p0c / sqrt((P * p0c) * (P * p0c) + mc2 * mc2);
  const double beta0 = p0c / sqrt(p0c * p0c + mc2 * mc2);
  const double e_tot = sqrt(p0c * p0c + mc2 * mc2);
  const double b2 = beta0 * beta0;
  const double r = mc2 / e_tot;
  if (mc2 * ((beta0 * pz) * (beta0 * pz)) < 3e-7 * e_tot) {
    const double poly = 1.0 - 3.0 * (pz * b2) / 2.0 + pz * pz * b2 * (2.0 * b2 - r * r / 2.0);
    return ds * pz * poly * (r * r);
  }
  return ds * (beta - beta0) / beta0;
}

__device__ inline void drift(Coords &c, double L, double p0c, double mc2) {
  const double P = 1.0 + c.pz;
  const double Px = c.px / P, Py = c.py / P;
  const double pxy2 = Px * Px + Py * Py;
  const double Pl = sqrt(1.0 - pxy2);
  c.x = c.x + L * Px / Pl;
  c.y = c.y + L * Py / Pl;
  const double a = (mc2 * mc2 * (2.0 * c.pz + c.pz * c.pz)) / ((p0c * P) * (p0c * P) + mc2 * mc2);
  c.z = c.z + L * (sqrt_one(a) + sqrt_one(-pxy2) / Pl);
}

// One quadrupole plane, u'' = -k u (k > 0 focusing), u' = pu / P.
struct Plane {
  double m11, m12, m21, m22, d0, d1, d2;
};

__device__ inline Plane quad_plane(double k, double L, double P) {
  const double sk = sqrt(fabs(k));
  const double th = sk * L;
  double cx, sx;
  if (k > 0) {
    cx = cos(th);
    sx = sin(th) / sk;
  } else if (k < 0) {
    cx = cosh(th);
    sx = sinh(th) / sk;
  } else {
    cx = 1.0;
    sx = L;
  }
  Plane p;
  p.m11 = cx;
  p.m12 = sx / P;
  p.m21 = -k * sx * P;
  p.m22 = cx;
  p.d0 = k * (cx * sx - L) / 4.0;
  p.d1 = k * (sx * sx) / (2.0 * P);
  p.d2 = -(L + cx * sx) / (4.0 * (P * P));
  return p;
}

__device__ inline void offset_set(Coords &c, double xo, double yo, double tilt) {
  const double x = c.x - xo, y = c.y - yo;
  const double ct = cos(tilt), st = sin(tilt);
  const double px = c.px, py = c.py;
  c.x = x * ct + y * st;
  c.px = px * ct + py * st;
  c.y = -x * st + y * ct;
  c.py = -px * st + py * ct;
}

__device__ inline void offset_unset(Coords &c, double xo, double yo, double tilt) {
  const double ct = cos(tilt), st = sin(tilt);
  const double x = c.x, px = c.px, y = c.y, py = c.py;
  c.x = x * ct - y * st + xo;
  c.px = px * ct - py * st;
  c.y = x * st + y * ct + yo;
  c.py = px * st + py * ct;
}

__device__ inline void quadrupole(Coords &c, const DfcsrElement &e, double p0c, double mc2) {
  offset_set(c, e.x_offset, e.y_offset, e.tilt);
  const double ds = e.L / e.num_steps;
  for (int i = 0; i < e.num_steps; ++i) {
    const double P = 1.0 + c.pz;
    const double k = e.k1 / P;
    const Plane tx = quad_plane(k, ds, P), ty = quad_plane(-k, ds, P);
    c.z = c.z + tx.d0 * (c.x * c.x) + tx.d1 * c.x * c.px + tx.d2 * (c.px * c.px) +
          ty.d0 * (c.y * c.y) + ty.d1 * c.y * c.py + ty.d2 * (c.py * c.py);
    const double x = tx.m11 * c.x + tx.m12 * c.px, px = tx.m21 * c.x + tx.m22 * c.px;
    const double y = ty.m11 * c.y + ty.m12 * c.py, py = ty.m21 * c.y + ty.m22 * c.py;
    c.x = x;
    c.px = px;
    c.y = y;
    c.py = py;
    c.z = c.z + low_energy_dz(c.pz, p0c, mc2, ds);
  }
  offset_unset(c, e.x_offset, e.y_offset, e.tilt);
}

// Hard-edge pole face rotated by e; fint * hgap > 0 softens the vertical
// edge focusing by psi = fint (2 hgap) g (1 + sin^2 e) / cos e.
__device__ inline void bend_edge(Coords &c, double g, double e, double fint, double hgap) {
  const double psi = fint * (2.0 * hgap) * g * (1.0 + sin(e) * sin(e)) / cos(e);
  if (e == 0.0 && psi == 0.0) return;
  c.px = c.px + c.x * g * tan(e);
  c.py = c.py - c.y * g * tan(e - psi);
}

__device__ inline void sbend(Coords &c, const DfcsrElement &e, double p0c, double mc2) {
  const double g = e.g, L = e.L;
  bend_edge(c, g, e.e1, e.fint, e.hgap);
  if (g == 0.0) {
    drift(c, L, p0c, mc2);
  } else {
    const double pz = c.pz, P = 1.0 + pz;
    const double th = g * L;
    const double ct = cos(th), st = sin(th);
    const double pxy2 = c.px * c.px + c.py * c.py;
    const double ps_1 = sqrt_one(2.0 * pz + pz * pz - pxy2);   // p_s - 1
    const double Ph = sqrt(P * P - c.py * c.py);
    const double px1 = c.px * ct + (ps_1 - g * c.x) * st;
    const double ps1_1 = sqrt_one(2.0 * pz + pz * pz - c.py * c.py - px1 * px1);
    const double x1 = c.x * ct + (ps1_1 - ct * ps_1 + c.px * st) / g;
    const double dphi = asin(c.px / Ph) - asin(px1 / Ph);   // phi - g L
    const double beta = P * p0c / sqrt((P * p0c) * (P * p0c) + mc2 * mc2);
    const double beta0 = p0c / sqrt(p0c * p0c + mc2 * mc2);
    c.y = c.y + c.py * (L + dphi / g);
    c.z = c.z + L * beta / beta0 - P * (L + dphi / g);
    c.x = x1;
    c.px = px1;
    c.pz = P - 1.0;
  }
  bend_edge(c, g, e.e2, e.fintx, e.hgapx);
}

__device__ inline void sextupole(Coords &c, const DfcsrElement &e, double p0c, double mc2) {
  offset_set(c, e.x_offset, e.y_offset, e.tilt);
  const double ds = e.L / e.num_steps;
  for (int i = 0; i < e.num_steps; ++i) {
    drift(c, ds / 2.0, p0c, mc2);
    const double x = c.x, y = c.y;
    c.px = c.px - e.k2 * ds * (x * x - y * y) / 2.0;
    c.py = c.py + e.k2 * ds * x * y;
    drift(c, ds / 2.0, p0c, mc2);
  }
  offset_unset(c, e.x_offset, e.y_offset, e.tilt);
}

__global__ __launch_bounds__(kThreads) void track_kernel(double *__restrict__ x,
                                                         double *__restrict__ px,
                                                         double *__restrict__ y,
                                                         double *__restrict__ py,
                                                         double *__restrict__ z,
                                                         double *__restrict__ pz, int64_t n,
                                                         DfcsrElement e, double p0c, double mc2) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    Coords c = {x[i], px[i], y[i], py[i], z[i], pz[i]};
    switch (e.type) {
      case DFCSR_ELE_DRIFT: drift(c, e.L, p0c, mc2); break;
      case DFCSR_ELE_SBEND: sbend(c, e, p0c, mc2); break;
      case DFCSR_ELE_QUAD: quadrupole(c, e, p0c, mc2); break;
      default: sextupole(c, e, p0c, mc2); break;
    }
    x[i] = c.x;
    px[i] = c.px;
    y[i] = c.y;
    py[i] = c.py;
    z[i] = c.z;
    pz[i] = c.pz;
  }
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int dfcsr_track_element(double *x, double *px, double *y, double *py, double *z,
                                   double *pz, int64_t n, const DfcsrElement *ele, double p0c,
                                   double mc2, void *stream) {
  if (!x || !px || !y || !py || !z || !pz || !ele)
    return set_error(DFCSR_EINVAL, "track_element: NULL pointer");
  if (ele->type < DFCSR_ELE_DRIFT || ele->type > DFCSR_ELE_SEXTUPOLE)
    return set_error(DFCSR_EINVAL, "track_element: unknown element type %d", ele->type);
  if ((ele->type == DFCSR_ELE_QUAD || ele->type == DFCSR_ELE_SEXTUPOLE) &&
      (ele->num_steps < 1 || ele->num_steps > 1000))
    return set_error(DFCSR_EINVAL, "track_element: num_steps %d", ele->num_steps);
  if (!(p0c > 0.0) || !(mc2 > 0.0))
    return set_error(DFCSR_EINVAL, "track_element: p0c %g, mc2 %g", p0c, mc2);
  if (n <= 0) return DFCSR_OK;
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b > 8192) b = 8192;
  hipLaunchKernelGGL(track_kernel, dim3((unsigned)b), dim3(kThreads), 0, as_stream(stream), x, px,
                     y, py, z, pz, n, *ele, p0c, mc2);
  DFCSR_LAUNCH_CHECK("track_element");
  return DFCSR_OK;
}
