// Device functions shared by the wake kernels and the parity entry points:
// lattice lookups (interpolate1D, interp1D.py:13-36), the trilinear history
// lookup (interpolate3D, interp3D.py:18-66) and one integrand point of
// get_CSR_integrand (CSR.py:605-782).
//
// The floating-point contraction mode is the including file's: the parity
// entry points compile this without fused multiply-add and with true
// divisions and square roots (kRecip = false) so they reproduce the
// reference's rounding; the wake kernels multiply by precomputed reciprocals
// and take |r - r'|, 1/|r - r'| and 1/scale from the hardware estimates
// refined by Newton steps (kRecip = true, a few ulp).
#pragma once

#include "common.h"

namespace dfcsr {

struct Lat6 {
  double X0, Y0, nx, ny, tx, ty;
};

// interpolate1D of all six lattice columns at s.
template <bool kRecip>
__device__ inline Lat6 lattice_at(const DfcsrLattice &L, double s, double inv_ds) {
  double u = kRecip ? (s - L.min_s) * inv_ds : (s - L.min_s) / L.ds;
  TruncIdx t = trunc_index(u, L.ns);
  Lat6 r;
  if (!t.ok) {
    r.X0 = r.Y0 = r.nx = r.ny = r.tx = r.ty = 0.0;
    return r;
  }
  const double *a = L.table + (int64_t)t.i0 * 6;
  const double *b = L.table + (int64_t)t.i1 * 6;
  const double w = 1 - t.frac, f = t.frac;
  r.X0 = a[0] * w + b[0] * f;
  r.Y0 = a[1] * w + b[1] * f;
  r.nx = a[2] * w + b[2] * f;
  r.ny = a[3] * w + b[3] * f;
  r.tx = a[4] * w + b[4] * f;
  r.ty = a[5] * w + b[5] * f;
  return r;
}

// rho(s') by the element masks of CSR.py:651-656 (later matches win; 0 past
// the lattice end).
__device__ inline double rho_at(const DfcsrLattice &L, double sp) {
  double r = 0.0;
  for (int k = 0; k < L.nelem; ++k) {
    bool in = (k == 0) ? (sp < L.distance[0]) : (sp < L.distance[k] && sp >= L.distance[k - 1]);
    if (in) r = L.rho[k];
  }
  return r;
}

// History axes with the per-call reciprocals.
struct HistAxes {
  double min_t, min_x, min_z, dt, dx, dz, it, ix, iz;
};

__host__ __device__ inline HistAxes make_axes(const DfcsrHistory &H) {
  HistAxes a;
  a.min_t = H.min_t;
  a.min_x = H.min_x;
  a.min_z = H.min_z;
  a.dt = H.dt;
  a.dx = H.dx;
  a.dz = H.dz;
  a.it = 1.0 / H.dt;
  a.ix = 1.0 / H.dx;
  a.iz = 1.0 / H.dz;
  return a;
}

// Ring slot of logical slice ti.
__device__ inline int hist_slot(const DfcsrHistory &H, int ti) {
  int slot = H.head + ti;
  return slot >= H.cap ? slot - H.cap : slot;
}

// The newest slices whose span entries a near-region workgroup stages in
// LDS for each of its x rows.
constexpr int kSpanPlanes = 16;

// Whether a valid lookup at x row xi, first time plane ti and z indices Z
// may gather a contributing cell: whether its z corners reach the span of
// (xi, slot of ti), which covers rows xi, xi + 1 and planes ti, ti + 1.
// zrow: that row's entries for slices zt0 .. zt0 + kSpanPlanes - 1 (LDS);
// older slices read the table itself (with zt0 = nt, every slice does: the
// region-0 form, which keeps no LDS copy).  When false, every gathered density,
// density_x and density_z is 0 and vx, vx_x are finite, so the integrand is
// exactly 0.
__device__ inline bool zspan_holds(const DfcsrHistory &H, const uint32_t *zrow, int zt0, int xi,
                                   int ti, const TruncIdx &Z) {
  const uint32_t e = ti >= zt0 ? zrow[ti - zt0] : H.zspan[(int64_t)xi * H.cap + hist_slot(H, ti)];
  return Z.i1 >= (int)(e & 0xffffu) && Z.i0 <= (int)(e >> 16);
}

// FP64 record (x, z, slot of ti): [x][z][slot][20] layout.
__device__ inline const double *hist_rec(const DfcsrHistory &H, int ti, int xi, int zi) {
  return H.data + (((int64_t)xi * H.Zi + zi) * H.cap + hist_slot(H, ti)) * DFCSR_HIST_REC64;
}

// Byte offsets of an x row within a plane of a cell-major ring: row i0, and
// from row i0 to row i1.  A plane spans at most 4 GiB (checked on the
// host), so lookups address cells with 32-bit offsets from the plane.
struct CellRow {
  uint32_t rb, dxb;
};

// The corners (x0,z), (x0,z+1), (x1,z), (x1,z+1) of a float64 cell plane
// from byte oa (x0, z) and ob (x1, z): each x row's two cells as 80
// contiguous bytes.
__device__ inline void load_pairs5(const char *plane, uint32_t oa, uint32_t ob, double v[4][5]) {
  const double *a = reinterpret_cast<const double *>(plane + oa);
  const double *b = reinterpret_cast<const double *>(plane + ob);
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    v[0][f] = a[f];
    v[1][f] = a[5 + f];
    v[2][f] = b[f];
    v[3][f] = b[5 + f];
  }
}

// interpolate3D's lerp a (1 - w) + b w (interp3D.py:53-61).  kFused (the
// wake kernels): b w added to the rounded a (1 - w) by one fused
// multiply-add; else the reference's two rounded products and a sum.
__device__ inline double lerp_exact(double a, double b, double w) {
#pragma clang fp contract(off)
  return a * (1 - w) + b * w;
}

template <bool kFused>
__device__ inline double lerp(double a, double b, double w) {
  return kFused ? fma(b, w, a * (1 - w)) : lerp_exact(a, b, w);
}

// The trilinear value from the corners (x0,z0), (x0,z1), (x1,z0), (x1,z1) of
// planes t0 (a*) and t1 (b*): along t, then x, then z, interpolate3D's order
// (the exact lookups; the wake kernels use trilerp_w below).  Every layout's
// lookup calls one form, so equal corner values give equal results.
template <bool kFused>
__device__ inline double trilerp(double a00, double a01, double a10, double a11, double b00,
                                 double b01, double b10, double b11, double td, double xd,
                                 double zd) {
  const double c00 = lerp<kFused>(a00, b00, td), c01 = lerp<kFused>(a01, b01, td);
  const double c10 = lerp<kFused>(a10, b10, td), c11 = lerp<kFused>(a11, b11, td);
  return lerp<kFused>(lerp<kFused>(c00, c10, xd), lerp<kFused>(c01, c11, xd), zd);
}

// The wake kernels' trilinear value: the eight corner weights, products of
// the t, x and z weights computed once per lookup, and one weighted sum per
// field (8 operations against 14 for t, x and z lerps).  Every corner is
// weighted, as in interpolate3D, so a non-finite corner gives a non-finite
// value there too; at the last plane, x row or z the two planes, rows or z
// cells are the same values, as its clamped indices give them.
struct TriWeights {
  double w[8];   // (t0|t1) x (x0,z0), (x0,z1), (x1,z0), (x1,z1)
};

__device__ inline TriWeights tri_weights(double td, double xd, double zd) {
  const double t0 = 1.0 - td, x0 = 1.0 - xd, z0 = 1.0 - zd;
  const double p00 = t0 * x0, p01 = t0 * xd, p10 = td * x0, p11 = td * xd;
  return {{p00 * z0, p00 * zd, p01 * z0, p01 * zd, p10 * z0, p10 * zd, p11 * z0, p11 * zd}};
}

// a: plane t0, b: plane t1; corners (x0,z0), (x0,z1), (x1,z0), (x1,z1)
__device__ inline double trilerp_w(const TriWeights &w, double a00, double a01, double a10,
                                   double a11, double b00, double b01, double b10, double b11) {
  double acc = w.w[0] * a00;
  acc = fma(w.w[1], a01, acc);
  acc = fma(w.w[2], a10, acc);
  acc = fma(w.w[3], a11, acc);
  acc = fma(w.w[4], b00, acc);
  acc = fma(w.w[5], b01, acc);
  acc = fma(w.w[6], b10, acc);
  return fma(w.w[7], b11, acc);
}

// Trilinear lookup of all five fields at (t, x, z) with interpolate3D's
// truncation and validity rules.  Returns false (and zeros) when invalid.
// kRecip: the wake kernels' form (axes scaled by reciprocals, fused lerps);
// else the reference's divisions and unfused lerps.
// kCells: read the cell-major FP64 ring H.cells instead of the corner-packed
// one; the values, and so the result, are the same.
// Xpre: the x index of xq when the caller has it already (the same value);
// rr: that row's offsets in a cell plane (computed here when absent).
// zrow (source points of the near regions, with a span table): that row's
// staged span entries (zspan_holds); a lookup they show to gather no
// contributing cell returns false (zeros) -- its integrand would be exactly 0.
template <bool kRecip, bool kCells = false>
__device__ inline bool hist_lookup5(const DfcsrHistory &H, const HistAxes &A, double tq,
                                    double xq, double zq, double out[5],
                                    const TruncIdx *Xpre = nullptr,
                                    const CellRow *rr = nullptr,
                                    const uint32_t *zrow = nullptr, int zt0 = 0) {
  double ut = kRecip ? (tq - A.min_t) * A.it : (tq - A.min_t) / A.dt;
  double uz = kRecip ? (zq - A.min_z) * A.iz : (zq - A.min_z) / A.dz;
  TruncIdx X;
  if (Xpre) {
    X = *Xpre;
  } else {
    const double ux = kRecip ? (xq - A.min_x) * A.ix : (xq - A.min_x) / A.dx;
    X = trunc_index(ux, H.Xi);
  }
  TruncIdx T = trunc_index(ut, H.nt), Z = trunc_index(uz, H.Zi);
  if (!(T.ok && X.ok && Z.ok) || (zrow && !zspan_holds(H, zrow, zt0, X.i0, T.i0, Z))) {
    for (int f = 0; f < 5; ++f) out[f] = 0.0;
    return false;
  }
  const double td = T.frac, xd = X.frac, zd = Z.frac;
  if (kCells) {
    // cell (slot, x, z) at byte ((slot Xi + x) (Zi + 1) + z) * 40 of the
    // cell-major ring; each x row's cells z0 and z0 + 1 are read as one
    // 80-byte pair (five 16-byte loads, against three per 40-byte cell).  At
    // the grid's last z the pair is that cell and the row's copy of it, the
    // clamped corners of interpolate3D, so the result is the corner-packed
    // ring's exactly.
    constexpr uint32_t cb = DFCSR_CELL * sizeof(double);
    const uint32_t rowB = (uint32_t)DFCSR_CELL_ROW(H.Zi) * cb;
    const CellRow r = rr ? *rr : CellRow{(uint32_t)X.i0 * rowB, (uint32_t)(X.i1 - X.i0) * rowB};
    const int64_t planeB = (int64_t)H.Xi * rowB;
    const char *ring = reinterpret_cast<const char *>(H.cells);
    const uint32_t o0 = r.rb + (uint32_t)Z.i0 * cb;
    // both planes read on every lookup (at the last plane the second is the
    // first again, a cache hit): no branch between the two forms
    const char *p0 = ring + hist_slot(H, T.i0) * planeB;
    const char *p1 = T.i1 != T.i0 ? ring + hist_slot(H, T.i1) * planeB : p0;
    double v0[4][5], v1[4][5];
    load_pairs5(p0, o0, o0 + r.dxb, v0);
    load_pairs5(p1, o0, o0 + r.dxb, v1);
    if constexpr (kRecip) {
      const TriWeights w = tri_weights(td, xd, zd);
#pragma unroll
      for (int f = 0; f < 5; ++f)
        out[f] = trilerp_w(w, v0[0][f], v0[1][f], v0[2][f], v0[3][f], v1[0][f], v1[1][f],
                           v1[2][f], v1[3][f]);
    } else {
#pragma unroll
      for (int f = 0; f < 5; ++f)
        out[f] = trilerp<false>(v0[0][f], v0[1][f], v0[2][f], v0[3][f], v1[0][f], v1[1][f],
                                v1[2][f], v1[3][f], td, xd, zd);
    }
    return true;
  }
  // corner-packed records: p0 at (t0, x0, z0), p1 at (t1, x0, z0); element
  // f*4 + {0,1,2,3} = (x0,z0), (x0,z1), (x1,z0), (x1,z1).  A node's records
  // of consecutive slots are consecutive.  At the last plane (i1 == i0) the
  // second read is the first record again, a cache hit: selecting its
  // address, not copying its values, lets both reads be issued before
  // either is waited for.
  const int s0 = hist_slot(H, T.i0);
  const double *node = H.data + ((int64_t)X.i0 * H.Zi + Z.i0) * H.cap * DFCSR_HIST_REC64;
  const double *r0 = node + (int64_t)s0 * DFCSR_HIST_REC64;
  const double *r1 = T.i1 == T.i0 ? r0 : (s0 + 1 == H.cap ? node : r0 + DFCSR_HIST_REC64);
  const double2 *p0 = reinterpret_cast<const double2 *>(r0);
  const double2 *p1 = reinterpret_cast<const double2 *>(r1);
  double2 a[DFCSR_HIST_REC64 / 2], b[DFCSR_HIST_REC64 / 2];
#pragma unroll
  for (int k = 0; k < DFCSR_HIST_REC64 / 2; ++k) a[k] = p0[k];
#pragma unroll
  for (int k = 0; k < DFCSR_HIST_REC64 / 2; ++k) b[k] = p1[k];
  if constexpr (kRecip) {
    const TriWeights w = tri_weights(td, xd, zd);
#pragma unroll
    for (int f = 0; f < 5; ++f)
      out[f] = trilerp_w(w, a[2 * f].x, a[2 * f].y, a[2 * f + 1].x, a[2 * f + 1].y, b[2 * f].x,
                         b[2 * f].y, b[2 * f + 1].x, b[2 * f + 1].y);
  } else {
#pragma unroll
    for (int f = 0; f < 5; ++f)
      out[f] = trilerp<false>(a[2 * f].x, a[2 * f].y, a[2 * f + 1].x, a[2 * f + 1].y,
                              b[2 * f].x, b[2 * f].y, b[2 * f + 1].x, b[2 * f + 1].y, td, xd, zd);
  }
  return true;
}

// Observer-side quantities of one mesh point (CSR.py:608-642).
struct Observer {
  double s, x, t;
  Lat6 L;
  double vx;   // vx at (t, x, s - t)
};

template <bool kRecip>
__device__ inline Observer make_observer(const DfcsrHistory &H, const HistAxes &A,
                                         const DfcsrLattice &L, double s, double x, double t) {
  Observer o;
  o.s = s;
  o.x = x;
  o.t = t;
  o.L = lattice_at<kRecip>(L, s, 1.0 / L.ds);
  double v[5];
  hist_lookup5<kRecip>(H, A, t, x, s - t, v);
  o.vx = v[DFCSR_F_VX];
  return o;
}

// Per-column (fixed s') quantities.
struct Column {
  double sp, rho;
  Lat6 L;
  double ax, ay;   // (X0_s - X0_sp) + x n_s  (r - r' before the -x' n' term)
};

template <bool kRecip>
__device__ inline Column make_column(const DfcsrLattice &L, const Observer &o, double sp,
                                     double inv_ds) {
  Column c;
  c.sp = sp;
  c.L = lattice_at<kRecip>(L, sp, inv_ds);
  c.rho = rho_at(L, sp);
  c.ax = o.L.X0 - c.L.X0 + o.x * o.L.nx;
  c.ay = o.L.Y0 - c.L.Y0 + o.x * o.L.ny;
  return c;
}

// 1/sqrt(a) for a normal a > 0: the hardware estimate and two Newton steps
// y += y (1 - a y^2) / 2, each squaring the relative error (a few ulp).
__device__ inline double rsqrt_newton(double a) {
  double y = __builtin_amdgcn_rsq(a);
  double e = fma(-(a * y), y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-(a * y), y, 1.0);
  return fma(0.5 * y, e, y);
}

// 1/a for a normal a: the hardware estimate and two Newton steps r += r (1 - a r).
__device__ inline double recip_newton(double a) {
  double r = __builtin_amdgcn_rcp(a);
  double e = fma(-a, r, 1.0);
  r = fma(r, e, r);
  e = fma(-a, r, 1.0);
  return fma(r, e, r);
}

// One integrand point (I_z, I_x) at source (x', s'), CSR.py:645-775 in the
// reference's operation order.  Returns false (I = 0) when the retarded
// lookup falls outside the history, where every gathered field is zero and
// the reference's formula gives exactly 0, and (with zrow) when the span
// table shows that the gathered density and its gradients are all 0, which
// gives exactly 0 too -- except at R == 0, where the reference's divisions
// give 0/0: that point is evaluated (with zero fields) as it is.
template <bool kRecip, bool kCells = false>
__device__ inline bool csr_point(const DfcsrHistory &H, const HistAxes &A, const Observer &o,
                                 const Column &c, double xp, double &Iz, double &Ix,
                                 const TruncIdx *Xpre = nullptr,
                                 const CellRow *rr = nullptr,
                                 const uint32_t *zrow = nullptr, int zt0 = 0) {
  const double rx = c.ax - xp * c.L.nx;
  const double ry = c.ay - xp * c.L.ny;
  double R, iR = 0.0;
  if (kRecip) {
    const double R2 = rx * rx + ry * ry;
    iR = R2 != 0.0 ? rsqrt_newton(R2) : 0.0;
    R = R2 * iR;
  } else {
    R = sqrt(rx * rx + ry * ry);
  }
  const double t_ret = o.t - R;
  double f[5];
  if (!hist_lookup5<kRecip, kCells>(H, A, t_ret, xp, c.sp - t_ret, f, Xpre, rr, zrow, zt0) &&
      R != 0.0) {
    Iz = 0.0;
    Ix = 0.0;
    return false;
  }
  const double dens = f[DFCSR_F_DENSITY], dens_x = f[DFCSR_F_DENSITY_X],
               dens_z = f[DFCSR_F_DENSITY_Z], vx_ret = f[DFCSR_F_VX], div_v = f[DFCSR_F_VX_X];
  const double scale = 1 + xp * c.rho;
  const double dzs = kRecip ? dens_z * recip_newton(scale) : dens_z / scale;
  const double vel_x = o.L.tx + o.vx * o.L.nx;
  const double vel_y = o.L.ty + o.vx * o.L.ny;
  const double vr_x = c.L.tx + vx_ret * c.L.nx;
  const double vr_y = c.L.ty + vx_ret * c.L.ny;
  const double nab_x = dens_x * c.L.nx + dzs * c.L.tx;
  const double nab_y = dens_x * c.L.ny + dzs * c.L.ty;
  const double part1 = vel_x * vr_x + vel_y * vr_y;
  const double num1 = scale * ((vel_x - part1 * vr_x) * nab_x + (vel_y - part1 * vr_y) * nab_y);
  const double num2 = -scale * part1 * dens * div_v;
  const double nmx = o.L.nx - c.L.nx, nmy = o.L.ny - c.L.ny;
  const double p1 = rx * nmx + ry * nmy;
  const double p2 = o.L.nx * c.L.tx + o.L.ny * c.L.ty;
  const double pd = -(vr_x * nab_x + vr_y * nab_y) - dens * div_v;
  if (kRecip && R != 0.0) {
    const double iR2 = iR * iR;
    Iz = (num1 + num2) * iR;
    Ix = scale * p1 * (iR2 * iR) * dens + scale * p1 * iR2 * pd + -scale * p2 * iR * pd;
  } else {
    Iz = num1 / R + (num2) / R;
    const double W1 = scale * p1 / (R * R * R) * dens;
    const double W2 = scale * p1 / (R * R) * pd;
    const double W3 = -scale * p2 / R * pd;
    Ix = W1 + W2 + W3;
  }
  return true;
}

}  // namespace dfcsr
