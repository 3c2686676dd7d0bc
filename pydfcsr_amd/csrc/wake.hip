// a13-a15: the retarded CSR integral on the wake mesh.
//
// Replaces CSR2D.get_CSR_mesh / calculate_2D_CSR / get_CSR_wake /
// get_CSR_integrand (CSR.py:361-782).  A mesh point's integral over its
// regions (3 without chirp, 4 with) is split into phases: one phase is one
// region and one chunk of its x' rows, over all s' columns.
//   observer   : lattice lookups at s, vx at the observer and the region
//                table, once per mesh point (wake_observer_kernel).
//   workgroup  = one (mesh point, phase); lanes of a wavefront take
//                consecutive s' columns, so they share x' rows (the x-range
//                test is wave-uniform) and their retarded lookups are
//                neighbouring records of the time-innermost history (region
//                0) or neighbouring cells of the cell-major one (regions 1..,
//                when the history has it).
//   per column : lattice lookups at s', rho(s'), the trapezoid weight in s'.
//   per point  : |r - r'|, t_ret, the trilinear gather of five fields, the
//                integrand; points whose gather is outside the history add 0
//                and skip the arithmetic.
//   reduction  : sum_j,k wx_j wz_k I_jk per lane, a fixed-order block tree
//                per phase, then the phases of a mesh point in phase order:
//                deterministic, the same every launch.
// Launch order: both launches run the mesh points of one z line phase by
// phase.  Region 0 (the far region) runs first and each XCD takes a
// contiguous range (its L2 holds what a z line shares: the same x' rows and
// retarded planes, nearly the same z').  The near regions' z line shares
// ~130 MB, more than an L2 but within the 256 MB last-level cache: all XCDs
// sweep the same z line together, each taking groups of 16 neighbouring
// workgroups.
#include <mutex>
#include <vector>

#include "csr_point.h"

namespace dfcsr {

int check_history(const DfcsrHistory *H, const char *who);
int check_lattice(const DfcsrLattice *L, const char *who);

namespace {

constexpr int kThreads = 256;
#ifndef DFCSR_WAKE_CHUNK_FAR
#define DFCSR_WAKE_CHUNK_FAR 64
#endif
#ifndef DFCSR_WAKE_CHUNK_NEAR
#define DFCSR_WAKE_CHUNK_NEAR 64
#endif
#ifndef DFCSR_WAKE_MINBLOCKS
#define DFCSR_WAKE_MINBLOCKS 2
#endif

// at most this many x' rows per phase, in region 0 / in the later regions
constexpr int kChunkFar = DFCSR_WAKE_CHUNK_FAR;
constexpr int kChunkNear = DFCSR_WAKE_CHUNK_NEAR;
constexpr int kMaxRows = 64;   // a phase's row table holds at most this many rows
static_assert(kChunkFar <= kMaxRows && kChunkNear <= kMaxRows, "phase rows exceed the row table");

// A phase's x' rows, computed once per workgroup: x', its trapezoid weight
// and its history index (the same expressions the points would evaluate).
struct RowTable {
  double xp[kMaxRows];
  double w[kMaxRows];
  double frac[kMaxRows];
  int i0[kMaxRows];
  int i1[kMaxRows];
  int ok[kMaxRows];
  // byte offsets of row i0, and from row i0 to row i1, within a plane of the
  // cell-major ring (a plane spans at most 4 GiB, so 32 bits hold both)
  CellRow rr[kMaxRows];
  // the mixed form's float32 copies of x', the weight and the x fraction
  float xpf[kMaxRows];
  float wf[kMaxRows];
  float fracf[kMaxRows];
};
constexpr int kNumXcd = 8;
constexpr int kNearGroup = 16;   // consecutive near-phase workgroups per XCD

struct LinAxis {
  double a, b, step;
  int n;
  __device__ void init(double a_, double b_, int n_) {
#pragma clang fp contract(off)
    a = a_;
    b = b_;
    n = n_;
    step = n_ > 1 ? (b_ - a_) / (double)(n_ - 1) : 0.0;
  }
  __device__ double at(int i) const {
#pragma clang fp contract(off)
    if (n <= 1) return a;
    if (i == n - 1) return b;
    double y = (double)i * step;
    return y + a;
  }
  // trapezoid weight of node i (np.trapz over this axis)
  __device__ double w(int i) const {
#pragma clang fp contract(off)
    if (n < 2) return 0.0;
    double g = at(i);
    double dl = i > 0 ? g - at(i - 1) : 0.0;
    double dr = i < n - 1 ? at(i + 1) - g : 0.0;
    return 0.5 * (dl + dr);
  }
};

struct Regions {
  int n;
  LinAxis xa[4], sa[4];
};

__device__ inline double npmax0(double v) { return (v != v) ? v : fmax(0.0, v); }

// get_CSR_wake region bounds (CSR.py:461-553).
__device__ void make_regions(const DfcsrWakeParams &P, double s, double x, double t,
                             Regions &rg) {
#pragma clang fp contract(off)
  const double sz = P.sigma_z, sx = P.sigma_x, tan_theta = P.slope0;
  const double x0 = (s - t) * P.slope0;
  const int Nx = P.Nx, Nz = P.Nz;
  if (fabs(tan_theta) <= 1) {
    double s2 = s - 500 * sz, s3 = s - 20 * sz, s4 = s + 5 * sz;
    double s1 = npmax0(s2 - P.n_formation_length * P.formation_length);
    rg.n = 3;
    rg.xa[0].init(x0 - 20 * sx, x0 + 20 * sx, 2 * Nx);
    rg.sa[0].init(s1, s2, Nz);
    rg.xa[1].init(x0 - 10 * sx, x0 + 10 * sx, Nx);
    rg.sa[1].init(s2, s3, Nz);
    rg.xa[2].init(x0 - 10 * sx, x0 + 10 * sx, Nx);
    rg.sa[2].init(s3, s4, Nz);
    return;
  }
  double d, x1_l, x1_r, x2_l, x2_r;
  if (tan_theta > 0) {
    double tan_alpha = -2 * tan_theta / (1 - tan_theta * tan_theta);
    d = (10 * sx + P.mean_x - x) / tan_alpha;
    x1_l = x + 0.1 * sx;
    x1_r = x + 10 * sx;
    x2_l = x - 3 * sx;
    x2_r = x1_l;
  } else {
    double tan_alpha = 2 * tan_theta / (1 - tan_theta * tan_theta);
    d = -(P.mean_x - x - 10 * sx) / tan_alpha;
    x1_l = x - 10 * sx;
    x1_r = x - 1 * sx;
    x2_l = x1_r;
    x2_r = x + 3 * sx;
  }
  double s4 = s + 3 * sz;
  double s3 = npmax0(s - d);
  double s2 = s3 - 200 * sz;
  double s1 = npmax0(s2 - P.n_formation_length * P.formation_length);
  rg.n = 4;
  rg.xa[0].init(x0 - 20 * sx, x0 + 20 * sx, 2 * Nx);   // xp4 over sp1
  rg.sa[0].init(s1, s2, Nz);
  rg.xa[1].init(x0 - 5 * sx, x0 + 5 * sx, Nx);         // xp3 over sp2
  rg.sa[1].init(s2, s3, Nz);
  rg.xa[2].init(x1_l, x1_r, Nx);                       // xp1 over sp3
  rg.sa[2].init(s3, s4, Nz);
  rg.xa[3].init(x2_l, x2_r, Nx);                       // xp2 over sp3
  rg.sa[3].init(s3, s4, Nz);
}

// Mesh point i of get_CSR_mesh (CSR.py:380-389): zmesh = zrange[iz],
// xmesh = xrange[ix] + polyval(slope, zmesh).
__device__ inline void mesh_point(const DfcsrWakeParams &P, int i, double &xm, double &zm) {
#pragma clang fp contract(off)
  if (P.observers) {
    zm = P.observers[2 * i];
    xm = P.observers[2 * i + 1];
    return;
  }
  int ix = i / P.mesh_zbins, iz = i % P.mesh_zbins;
  zm = linspace_at(P.zr0, P.zr1, P.mesh_zbins, iz);
  double xt = linspace_at(P.xr0, P.xr1, P.mesh_xbins, ix);
  double pv = P.slope0 * zm;
  pv = pv + P.slope1;
  xm = xt + pv;
}

// Groups of G consecutive logical workgroups on one XCD, consecutive groups
// on consecutive XCDs: the XCDs together sweep the logical order while each
// L2 sees G neighbours.  Workgroups past the last full round keep their index.
__device__ inline int xcd_group(int b, int nb, int G) {
  const int round = kNumXcd * G;
  if (b >= (nb / round) * round) return b;
  const int q = b / kNumXcd, x = b % kNumXcd;
  return ((q / G) * kNumXcd + x) * G + q % G;
}

__device__ inline int xcd_remap(int b, int nb) {
  int q = nb / kNumXcd, r = nb % kNumXcd;
  int xcd = b % kNumXcd, pos = b / kNumXcd;
  return xcd * q + (xcd < r ? xcd : r) + pos;
}

// ---- mixed precision --------------------------------------------------------
//
// Same integral, regrouped so that FP32 only sees well-conditioned terms.
// With v = tau + vx n (observer), v' = tau' + vx' n' and grad = dx n' + g tau'
// (g = dz/scale), the reference's numerators expand exactly into per-column
// coefficients (A = v.tau', B = v.n', e = tau'.n', a = tau'.tau', b = n'.n'):
//   v.grad - (v.v') v'.grad = dx (D0 + vx'(D1 + vx' D2)) + g (G0 + vx'(G1 + vx' G2))
//     D0 = B - A e, D1 = -(A b + B e), D2 = -B b,
//     G0 = A (1 - a), G1 = -(A e + B a), G2 = -B e,
//   v.v' = A + vx' B,   v'.grad = dx (e + vx' b) + g (a + vx' e).
// Those coefficients, |r - r'|, t_ret, z' = s' - t_ret, the history indices
// and q = (r-r').(n-n')/R - n.tau' (whose two terms cancel to leading order)
// are float64; the gathered fields, the coefficients' products and the
// integrand are float32.  1/R comes from the FP32 hardware reciprocal square
// root with one FP64 Newton step (a nonzero |r - r'| between positions of
// metres is at least ~1e-16 m, so R^2 is never an FP32 denormal); 1/scale is
// the FP32 hardware reciprocal.  The
// retarded indices come from one FMA each: (t_ret - min_t)/dt = ut0 - R/dt
// and (z' - min_z)/dz = uz0 + R/dz with ut0 per wake and uz0 per column.
// Each point's weighted FP32 integrand is added to float64 row sums of its
// column, which are weighted and accumulated in float64; the block and phase
// reductions are float64.
struct ColumnMixed {
  double sp, ax, ay, npx, npy, P0, P1, p2;
  double uz0;   // ((s' - t) - min_z) / dz: z' index of the column before + R / dz
  float rho, D0, D1, D2, G0, G1, G2, e, a, b, A, B;
};

__device__ inline ColumnMixed make_column_mixed(const DfcsrLattice &L, const Observer &o,
                                                const HistAxes &HA, double sp, double inv_ds) {
  const Column c = make_column<true>(L, o, sp, inv_ds);
  ColumnMixed m;
  m.sp = sp;
  m.uz0 = ((sp - o.t) - HA.min_z) * HA.iz;
  m.ax = c.ax;
  m.ay = c.ay;
  m.npx = c.L.nx;
  m.npy = c.L.ny;
  const double vx_ = o.L.tx + o.vx * o.L.nx, vy_ = o.L.ty + o.vx * o.L.ny;
  const double A = vx_ * c.L.tx + vy_ * c.L.ty, B = vx_ * c.L.nx + vy_ * c.L.ny;
  const double e = c.L.tx * c.L.nx + c.L.ty * c.L.ny;
  const double a = c.L.tx * c.L.tx + c.L.ty * c.L.ty;
  const double b = c.L.nx * c.L.nx + c.L.ny * c.L.ny;
  const double nmx = o.L.nx - c.L.nx, nmy = o.L.ny - c.L.ny;
  m.P0 = c.ax * nmx + c.ay * nmy;
  m.P1 = c.L.nx * nmx + c.L.ny * nmy;
  m.p2 = o.L.nx * c.L.tx + o.L.ny * c.L.ty;
  m.rho = (float)c.rho;
  m.D0 = (float)(B - A * e);
  m.D1 = (float)(-(A * b + B * e));
  m.D2 = (float)(-B * b);
  m.G0 = (float)(A * (1.0 - a));
  m.G1 = (float)(-(A * e + B * a));
  m.G2 = (float)(-B * e);
  m.e = (float)e;
  m.a = (float)a;
  m.b = (float)b;
  m.A = (float)A;
  m.B = (float)B;
  return m;
}

// Chunk 0 (field 0's four corners) of the FP32 record (x, slot of ti, z)
// in the chunked ring; field f's chunk is f * Zi float4 further on.
__device__ inline const float4 *hist_rec32(const DfcsrHistory &H, int ti, int xi, int zi) {
  const int64_t blk = ((int64_t)xi * H.cap + hist_slot(H, ti)) * DFCSR_HIST_CHUNKS32;
  return reinterpret_cast<const float4 *>(H.data32) + blk * H.Zi + zi;
}

// The same record in the time-innermost ring: five consecutive chunks.
__device__ inline const float4 *hist_rec32t(const DfcsrHistory &H, int ti, int xi, int zi) {
  const int64_t rec = ((int64_t)xi * H.Zi + zi) * H.cap + hist_slot(H, ti);
  return reinterpret_cast<const float4 *>(H.data32t) + rec * DFCSR_HIST_CHUNKS32;
}

// vx at (tq, xq, zq) from an FP32 ring (the time-innermost one when present),
// with interpolate3D's truncation and validity rules and its lerp order.
__device__ inline double lookup_vx32(const DfcsrHistory &H, const HistAxes &A, double tq,
                                     double xq, double zq) {
  const TruncIdx T = trunc_index((tq - A.min_t) * A.it, H.nt);
  const TruncIdx X = trunc_index((xq - A.min_x) * A.ix, H.Xi);
  const TruncIdx Z = trunc_index((zq - A.min_z) * A.iz, H.Zi);
  if (!(T.ok && X.ok && Z.ok)) return 0.0;
  float4 a, b;
  if (H.data32t) {
    a = hist_rec32t(H, T.i0, X.i0, Z.i0)[DFCSR_F_VX];
    b = hist_rec32t(H, T.i1, X.i0, Z.i0)[DFCSR_F_VX];
  } else {
    a = hist_rec32(H, T.i0, X.i0, Z.i0)[(int64_t)DFCSR_F_VX * H.Zi];
    b = hist_rec32(H, T.i1, X.i0, Z.i0)[(int64_t)DFCSR_F_VX * H.Zi];
  }
  const double td = T.frac, xd = X.frac, zd = Z.frac;
  const double c00 = (double)a.x * (1 - td) + (double)b.x * td;
  const double c01 = (double)a.y * (1 - td) + (double)b.y * td;
  const double c10 = (double)a.z * (1 - td) + (double)b.z * td;
  const double c11 = (double)a.w * (1 - td) + (double)b.w * td;
  const double c0 = c00 * (1 - xd) + c10 * xd;
  const double c1 = c01 * (1 - xd) + c11 * xd;
  return c0 * (1 - zd) + c1 * zd;
}

// The corners (x0,z0), (x0,z1), (x1,z0), (x1,z1) of a float32 cell plane
// from the cell at byte oa of base (x0, z0) and the one at byte ob (x1,
// z0): each x row's cells z0 and z1 = z0 + 1 read as 40 contiguous bytes.
// 32-bit offsets from one base let the loads add them to its scalar address.
__device__ inline void load_pairs32(const char *base, uint32_t oa, uint32_t ob, float v[4][5]) {
  const float *a = reinterpret_cast<const float *>(base + (uint64_t)oa);
  const float *b = reinterpret_cast<const float *>(base + (uint64_t)ob);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    v[0][k] = a[k];
    v[1][k] = a[5 + k];
    v[2][k] = b[k];
    v[3][k] = b[5 + k];
  }
}

// Ways a lookup finds the history: the corner-packed rings, the cell-major
// rings with a 64-bit plane address per lookup, or (mixed precision,
// float32 cell-major rings of at most 4 GiB) 32-bit byte offsets from the
// rings' starts, which the loads add to the rings' scalar addresses.
enum CellMode { kRings = 0, kCellsPlane = 1, kCellsRing32 = 2 };

// One source point in mixed precision; false when outside the history or
// (zrow: the x row's staged span entries) when the span table shows an
// integrand of exactly 0 (zspan_holds).
// kTI: read the time-innermost FP32 ring, else the chunked one.  kCells
// (a CellMode): read the float32 cell-major ring H.cells32 and its time
// differences H.dcells32 instead.
template <bool kTI, int kCells = kRings>
__device__ inline bool csr_point_mixed(const DfcsrHistory &H, const HistAxes &A, double ut0,
                                       const ColumnMixed &c, double xp, float xpf, int xi0,
                                       int xi1, CellRow rr, float xd, float &Iz, float &Ix,
                                       const uint32_t *zrow = nullptr, int zt0 = 0) {
  const double rx = c.ax - xp * c.npx;
  const double ry = c.ay - xp * c.npy;
  const double R2 = rx * rx + ry * ry;
  if (R2 == 0.0) {   // the reference's 0/0 (or +-inf): not a finite contribution
    Iz = Ix = NAN;
    return true;
  }
  double y = (double)__builtin_amdgcn_rsqf((float)R2);
  y = y * (1.5 - 0.5 * R2 * y * y);               // one Newton step: ~1e-14 relative
  const double R = R2 * y;
  const double ut = ut0 - R * A.it;               // (t - R - min_t) / dt
  const double uz = c.uz0 + R * A.iz;             // (s' - (t - R) - min_z) / dz
  const TruncIdx T = trunc_index(ut, H.nt), Z = trunc_index(uz, H.Zi);
  if (!(T.ok && Z.ok) || (zrow && !zspan_holds(H, zrow, zt0, xi0, T.i0, Z))) return false;
  const float td = (float)T.frac;
  const float zd = (float)Z.frac;
  float f[5];
  if (kCells != kRings) {
    // float32 cell (slot, x, z) at byte ((slot Xi + x) (Zi + 1) + z) * 20 of
    // the cell-major rings.  The time lerp v0 + td (v1 - v0) takes v1 - v0
    // from H.dcells32 at the same offset: the float32 difference of the next
    // slice's cell and this one's, 0 at the newest slice, where
    // interpolate3D clamps i1 to i0 -- so one plane of each ring serves every
    // lookup, with no branch.  Each x row's two z cells are one 40-byte pair;
    // at the grid's last z the pair is that cell and the row's copy of it,
    // interpolate3D's clamped corners, so the result is the corner-packed
    // rings' exactly.
    constexpr uint32_t cb = DFCSR_CELL * sizeof(float);
    const uint32_t planeB =
        (uint32_t)H.Xi * (uint32_t)DFCSR_CELL_ROW(H.Zi) * cb;   // < 4 GiB (checked)
    const char *ring = reinterpret_cast<const char *>(H.cells32);
    const char *dring = reinterpret_cast<const char *>(H.dcells32);
    const uint32_t o = rr.rb + (uint32_t)Z.i0 * cb;
    const int s0 = hist_slot(H, T.i0);
    float v[4][5], d[4][5];
    if (kCells == kCellsRing32) {
      const uint32_t q = (uint32_t)s0 * planeB + o;
      load_pairs32(ring, q, q + rr.dxb, v);
      load_pairs32(dring, q, q + rr.dxb, d);
    } else {
      const uint64_t pl = (uint64_t)(uint32_t)s0 * planeB;
      load_pairs32(ring + pl, o, o + rr.dxb, v);
      load_pairs32(dring + pl, o, o + rr.dxb, d);
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const float c00 = fmaf(td, d[0][k], v[0][k]);
      const float c01 = fmaf(td, d[1][k], v[1][k]);
      const float c10 = fmaf(td, d[2][k], v[2][k]);
      const float c11 = fmaf(td, d[3][k], v[3][k]);
      const float c0 = fmaf(xd, c10 - c00, c00);
      const float c1 = fmaf(xd, c11 - c01, c01);
      f[k] = fmaf(zd, c1 - c0, c0);
    }
  } else {
    // one corner-packed record per time plane, one 16-byte chunk per field
    const float4 *q0 = kTI ? hist_rec32t(H, T.i0, xi0, Z.i0) : hist_rec32(H, T.i0, xi0, Z.i0);
    const float4 *q1 = kTI ? hist_rec32t(H, T.i1, xi0, Z.i0) : hist_rec32(H, T.i1, xi0, Z.i0);
    const int64_t cs = kTI ? 1 : H.Zi;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const float4 a = q0[k * cs], b = q1[k * cs];   // (x0,z0), (x0,z1), (x1,z0), (x1,z1)
      const float c00 = fmaf(td, b.x - a.x, a.x);
      const float c01 = fmaf(td, b.y - a.y, a.y);
      const float c10 = fmaf(td, b.z - a.z, a.z);
      const float c11 = fmaf(td, b.w - a.w, a.w);
      const float c0 = fmaf(xd, c10 - c00, c00);
      const float c1 = fmaf(xd, c11 - c01, c01);
      f[k] = fmaf(zd, c1 - c0, c0);
    }
  }
  const float dens = f[DFCSR_F_DENSITY], dx = f[DFCSR_F_DENSITY_X], dz = f[DFCSR_F_DENSITY_Z];
  const float vr = f[DFCSR_F_VX], dv = f[DFCSR_F_VX_X];
  const float scale = fmaf(xpf, c.rho, 1.0f);
  const float g = dz * __builtin_amdgcn_rcpf(scale);
  const float Cd = fmaf(vr, fmaf(vr, c.D2, c.D1), c.D0);
  const float Cg = fmaf(vr, fmaf(vr, c.G2, c.G1), c.G0);
  const float part1 = fmaf(vr, c.B, c.A);
  const float vgrad = fmaf(dx, fmaf(vr, c.b, c.e), g * fmaf(vr, c.e, c.a));
  const float ddv = dens * dv;
  const float pd = -vgrad - ddv;
  const double p1 = c.P0 - xp * c.P1;
  const double p1R = p1 * y;
  const float q = (float)(p1R - c.p2);
  const float ir = (float)y;
  const float sir = scale * ir;
  Iz = sir * (fmaf(dx, Cd, g * Cg) - part1 * ddv);
  Ix = sir * fmaf((float)p1R * ir, dens, pd * q);
  return true;
}

// ---- phase decomposition ----------------------------------------------------
//
// Region 0 has 2 Nx rows, the others Nx (make_regions); each region's rows
// are split into ceil(rows / kChunk*) chunks of equal length.  Phase order of
// a mesh point: the chunks of region 0, then those of regions 1, 2 (, 3).
//
// Region 0 of a mesh (observers from the mesh formula) is integrated
// observer-major instead: every observer of a z line has the same region-0
// x' rows (x0 depends on s only) and, off the chirp branch, the same s'
// columns, so a wavefront's lanes take neighbouring observers of one z line
// and evaluate the same source point together -- their retarded lookups fall
// on the same or neighbouring records.  A workgroup covers one z line, a
// block of W observers along x (W = 64, or the mesh's x bins rounded up to a
// power of two) and 4 S column blocks (S = 64 / W lane groups per wave); its
// row table holds the region's rows inside the history, 64 at a time.  The
// partial sums of region 0 are then per (observer, column block): nA = nCB.
struct PhasePlan {
  int nreg;                 // 3 without chirp, 4 with
  int rows0, ch0, per0;     // region 0 by row chunks: rows, chunks, rows per chunk
  int rows1, ch1, per1;     // each later region
  int nA, nB;               // partial sums of region 0 / phases of the later regions
  bool obs_major;           // region 0 observer-major (mesh observers)
  int W, S, nob, q;         // observer-major: lanes per observer block, column blocks
                            // per wave, observer blocks per z line, workgroups per block
};

// Waves of an observer-major region-0 launch wanted per simulation: about
// 2048 workgroups of 4 waves, 8 per CU, at most 32 per (z line, observer
// block).  Measured: the large run (128 x 128, 8 per block) is fastest there
// and slows with 16 or 32 per block (FP64 12.43 -> 12.51 / 12.65 ms); the
// chicane (64 x 64) gains from 16 to 32 per block (FP64 6.29 -> 5.97 ms).
constexpr int kFarWaveTarget = 8192;

__host__ __device__ inline PhasePlan make_plan(const DfcsrWakeParams &P) {
  PhasePlan pl;
  pl.nreg = (fabs(P.slope0) <= 1) ? 3 : 4;   // the branch test of make_regions
  pl.rows0 = 2 * P.Nx;
  pl.ch0 = (pl.rows0 + kChunkFar - 1) / kChunkFar;
  pl.per0 = (pl.rows0 + pl.ch0 - 1) / pl.ch0;
  pl.rows1 = P.Nx;
  pl.ch1 = (pl.rows1 + kChunkNear - 1) / kChunkNear;
  pl.per1 = (pl.rows1 + pl.ch1 - 1) / pl.ch1;
  pl.nB = (pl.nreg - 1) * pl.ch1;
  pl.obs_major = P.observers == nullptr;
  if (pl.obs_major) {
    const int xb = P.mesh_xbins, zb = P.mesh_zbins;
    int W = 1;
    while (W < xb && W < 64) W <<= 1;
    pl.W = W;
    pl.S = 64 / W;
    pl.nob = (xb + W - 1) / W;
    const int64_t lines = (int64_t)zb * pl.nob;
    int q = (int)((kFarWaveTarget / 4 + lines - 1) / lines);
    const int qmax = (P.Nz + 4 * pl.S - 1) / (4 * pl.S);   // no more blocks than columns
    q = q < 1 ? 1 : q;
    q = q > 32 ? 32 : q;
    q = q > qmax ? (qmax < 1 ? 1 : qmax) : q;
    pl.q = q;
    pl.nA = 4 * pl.S * q;
  } else {
    pl.W = pl.S = pl.nob = pl.q = 0;
    pl.nA = pl.ch0;
  }
  return pl;
}

// Workgroups of the region-0 launch of one simulation.
__host__ __device__ inline int64_t far_blocks(const PhasePlan &pl, const DfcsrWakeParams &P) {
  return pl.obs_major ? (int64_t)P.mesh_zbins * pl.nob * pl.q
                      : (int64_t)P.mesh_xbins * P.mesh_zbins * pl.nA;
}

// Region and row range [j0, j1) of phase ph (row-chunk phases: region 0 when
// it is not observer-major, and the later regions).
__host__ __device__ inline void phase_rows(const PhasePlan &pl, int ph, int &r, int &j0, int &j1) {
  if (ph < pl.nA) {
    r = 0;
    j0 = ph * pl.per0;
    j1 = min(j0 + pl.per0, pl.rows0);
  } else {
    const int q = ph - pl.nA;
    r = 1 + q / pl.ch1;
    j0 = (q % pl.ch1) * pl.per1;
    j1 = min(j0 + pl.per1, pl.rows1);
  }
}

// One simulation's wake: descriptors, outputs and its place in a launch.
struct WakeTask {
  DfcsrHistory H;
  DfcsrLattice L;
  DfcsrWakeParams P;
  double *dE, *xk;        // outputs [M]
  double *xmesh, *zmesh;  // optional outputs [M] (NULL: not written)
  int64_t part;           // first of its M * (nA + nB) * 2 partial sums
  int64_t obs;            // first of its M * kObsStride observer values
  int M, nA, nB;
  int blockA0, blockB0, obs0;   // first workgroup / observer in a batched launch
};

// Observer data of one mesh point, computed once by wake_observer_kernel
// and read by every phase of that point (kObsStride doubles):
//   [0] s, [1] x, [2..7] lattice X0, Y0, n_x, n_y, tau_x, tau_y at s,
//   [8] vx at the observer, [9] 1 if s lies outside the lattice table (its
//   lookups are 0 there), [16 + 4 r] x' axis of region r (a, b, step, n),
//   [32 + 4 r] s' axis of region r.
constexpr int kObsStride = 48;

__device__ inline void store_axis(double *o, const LinAxis &a) {
  o[0] = a.a;
  o[1] = a.b;
  o[2] = a.step;
  o[3] = (double)a.n;
}

__device__ inline LinAxis load_axis(const double *o) {
  LinAxis a;
  a.a = o[0];
  a.b = o[1];
  a.step = o[2];
  a.n = (int)o[3];
  return a;
}

__device__ inline void observer_point(const WakeTask &T, int i, double *__restrict__ od) {
  double xm, zm;
  mesh_point(T.P, i, xm, zm);
  const double t = T.P.t;
  const double s = t + zm, x = xm;
  const HistAxes A = make_axes(T.H);
  Observer o;
  if (T.H.data) {
    o = make_observer<true>(T.H, A, T.L, s, x, t);
  } else {   // FP32 rings only (mixed precision)
    o.s = s;
    o.x = x;
    o.t = t;
    o.L = lattice_at<true>(T.L, s, 1.0 / T.L.ds);
    o.vx = lookup_vx32(T.H, A, t, x, s - t);
  }
  Regions rg = {};
  make_regions(T.P, s, x, t, rg);
  od[0] = s;
  od[1] = x;
  od[2] = o.L.X0;
  od[3] = o.L.Y0;
  od[4] = o.L.nx;
  od[5] = o.L.ny;
  od[6] = o.L.tx;
  od[7] = o.L.ty;
  od[8] = o.vx;
  od[9] = trunc_index((s - T.L.min_s) * (1.0 / T.L.ds), T.L.ns).ok ? 0.0 : 1.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    store_axis(od + 16 + 4 * r, rg.xa[r]);
    store_axis(od + 32 + 4 * r, rg.sa[r]);
  }
}

// Integrates phase ph of the mesh point whose observer data is od; every
// thread returns the block sums of wz_k wx_j (I_z, I_x) over the phase.
// kTI (mixed only): read the time-innermost FP32 ring.  kCells (a
// CellMode): read the cell-major ring (H.cells for FP64; H.cells32 and its
// time differences H.dcells32 for mixed).
template <bool kMixed, bool kTI, int kCells = kRings, bool kSkip = false>
__device__ void wake_phase(const WakeTask &T, const double *__restrict__ od, int ph, double *red,
                           RowTable &rt, double &out_z, double &out_x) {
  const DfcsrHistory &H = T.H;
  const DfcsrLattice &L = T.L;
  const double t = T.P.t;
  const HistAxes A = make_axes(H);
  const double inv_ds = 1.0 / L.ds;
  int r, j0, j1;
  phase_rows(make_plan(T.P), ph, r, j0, j1);
  const LinAxis xa = load_axis(od + 16 + 4 * r);
  const LinAxis sa = load_axis(od + 32 + 4 * r);
  // (all branches below are uniform over the workgroup)
  if (xa.n < 2 || sa.n < 2) {   // trapz over a single sample is 0, NaN or not
    out_z = out_x = 0.0;
    return;
  }
  if (od[9] != 0.0) {
    // Observer and a source column both outside the lattice table: both
    // positions and normals look up as 0, so R == 0 on every row of that
    // column and the reference's integrand is 0/0 there -- the region's
    // integral is NaN.  s' is monotone, so its end nodes decide.
    const bool in0 = trunc_index((sa.at(0) - L.min_s) * inv_ds, L.ns).ok;
    const bool in1 = trunc_index((sa.at(sa.n - 1) - L.min_s) * inv_ds, L.ns).ok;
    if (!in0 || !in1) {
      out_z = out_x = NAN;
      return;
    }
  }
  // hull of the rows whose x' lies inside the history (x' is monotone in j);
  // a phase without one adds exactly 0 and skips its column setup
  int v0 = j1, v1 = j0;
  for (int j = j0; j < j1; ++j) {
    const double ux = (xa.at(j) - A.min_x) * A.ix;
    if (ux > -1.0 && ux < (double)H.Xi) {
      v0 = min(v0, j);
      v1 = j + 1;
    }
  }
  if (v0 >= v1) {   // uniform over the workgroup
    out_z = out_x = 0.0;
    return;
  }
  Observer o;
  o.s = od[0];
  o.x = od[1];
  o.t = t;
  o.L.X0 = od[2];
  o.L.Y0 = od[3];
  o.L.nx = od[4];
  o.L.ny = od[5];
  o.L.tx = od[6];
  o.L.ty = od[7];
  o.vx = od[8];
  // kSkip (the near regions, whose lookups fall on the newest slices): the
  // span entries of the phase's rows for the newest kSpanPlanes slices
  __shared__ uint32_t zs[kSkip ? kMaxRows : 1][kSpanPlanes];
  const bool skip = kSkip && H.zspan != nullptr;   // uniform
  const int zt0 = max(0, H.nt - kSpanPlanes);
  if (skip) {
    for (int e = threadIdx.x; e < (v1 - v0) * kSpanPlanes; e += blockDim.x) {
      const int r = e / kSpanPlanes, j = e - r * kSpanPlanes;
      const int xi = trunc_index((xa.at(v0 + r) - A.min_x) * A.ix, H.Xi).i0;
      zs[r][j] = zt0 + j < H.nt ? H.zspan[(int64_t)xi * H.cap + hist_slot(H, zt0 + j)] : 0xffffu;
    }
  }
  for (int r = threadIdx.x; r < v1 - v0; r += blockDim.x) {
    const double xp = xa.at(v0 + r);
    const TruncIdx X = trunc_index((xp - A.min_x) * A.ix, H.Xi);
    rt.xp[r] = xp;
    rt.w[r] = xa.w(v0 + r);
    rt.frac[r] = X.frac;
    rt.i0[r] = X.i0;
    rt.i1[r] = X.i1;
    rt.ok[r] = X.ok;
    rt.xpf[r] = (float)xp;
    rt.wf[r] = (float)rt.w[r];
    rt.fracf[r] = (float)X.frac;
    if (kCells != kRings) {
      const uint32_t rowB = (uint32_t)DFCSR_CELL_ROW(H.Zi) * (uint32_t)(DFCSR_CELL *
                                                        (kMixed ? sizeof(float) : sizeof(double)));
      rt.rr[r].rb = (uint32_t)X.i0 * rowB;
      rt.rr[r].dxb = (uint32_t)(X.i1 - X.i0) * rowB;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int ngroup = (sa.n + 63) >> 6;
  double acc_z = 0.0, acc_x = 0.0;
  unsigned evaluated = 0;   // points this lane evaluated (for the optional counter)
  for (int g = wave; g < ngroup; g += nwave) {
    // a group of m < 64 columns (the last one) spreads its rows over 64 / m
    // lane sets, so the wave still does a full wave's work
    const int m = min(64, sa.n - 64 * g);
    const int nq = 64 / m;
    const int q = lane / m;
    if (q >= nq) continue;
    const int k = 64 * g + (lane - q * m);
    const double sp = sa.at(k);
    const double wz = sa.w(k);
    if (kMixed) {
      const ColumnMixed col = make_column_mixed(L, o, A, sp, inv_ds);
      const double ut0 = (t - A.min_t) * A.it;
      double cz = 0.0, cx = 0.0;
      for (int r = q; r < v1 - v0; r += nq) {
        if (!rt.ok[r]) continue;
        float Iz, Ix;
        if (!csr_point_mixed<kTI, kCells>(H, A, ut0, col, rt.xp[r], rt.xpf[r], rt.i0[r],
                                           rt.i1[r], rt.rr[r], rt.fracf[r], Iz, Ix,
                                           skip ? zs[r] : nullptr, zt0))
          continue;
        ++evaluated;
        const float wx = rt.wf[r];
        cz += (double)(wx * Iz);
        cx += (double)(wx * Ix);
      }
      acc_z += wz * cz;
      acc_x += wz * cx;
    } else {
      const Column col = make_column<true>(L, o, sp, inv_ds);
      double cz = 0.0, cx = 0.0;
      for (int r = q; r < v1 - v0; r += nq) {
        if (!rt.ok[r]) continue;
        const TruncIdx X = {rt.i0[r], rt.i1[r], rt.frac[r], true};
        double Iz, Ix;
        if (!csr_point<true, kCells != kRings>(H, A, o, col, rt.xp[r], Iz, Ix, &X, &rt.rr[r],
                                                skip ? zs[r] : nullptr, zt0))
          continue;
        ++evaluated;
        const double wx = rt.w[r];
        cz += wx * Iz;
        cx += wx * Ix;
      }
      acc_z += wz * cz;
      acc_x += wz * cx;
    }
  }
  if (T.P.points_evaluated) {   // uniform: a diagnostic call
    unsigned long long c = evaluated;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0 && c) atomicAdd(T.P.points_evaluated, c);
  }
  out_z = block_sum(acc_z, red);
  out_x = block_sum(acc_x, red);
}

// Workgroup l of a task's far (kNear = false) or near launch -> (mesh
// point, phase); writes the phase's partial sums.  kTI (mixed only): the
// phases read the time-innermost FP32 ring, else the chunked one; kCells (a
// CellMode): the cell-major ring.
template <bool kMixed, bool kNear, bool kTI, int kCells = kRings>
__device__ inline void run_phase_block(const WakeTask &T, int l, double *__restrict__ partials,
                                       double *red) {
  const int xb = T.P.mesh_xbins, zb = T.P.mesh_zbins;
  int i, ph;
  // z line major in both launches: the mesh points of one z line, phase by
  // phase
  const int np = kNear ? T.nB : T.nA;
  const int per_line = np * xb;
  const int zl = l / per_line, rem = l - zl * per_line;
  const int q = rem / xb;
  i = (rem - q * xb) * zb + zl;
  ph = kNear ? T.nA + q : q;
  __shared__ RowTable rt;
  double sz, sx;
  wake_phase<kMixed, kTI, kCells, kNear>(T, partials + T.obs + (int64_t)i * kObsStride, ph, red,
                                         rt, sz, sx);
  if (threadIdx.x == 0) {
    double *p = partials + T.part + ((int64_t)i * (T.nA + T.nB) + ph) * 2;
    p[0] = sz;
    p[1] = sx;
  }
}

// Workgroup l of a task's observer-major region-0 launch: z line, block of
// observers, group of column blocks (see PhasePlan).  Lane (sub, ol) of wave
// w integrates, for observer ol of the block, column block (4 g + w) S + sub
// over every row of region 0 inside the history, and writes that partial sum
// (NaN where the reference's integrand is 0/0 over the region, as
// wake_phase).  Summation: per column over the rows in order, weighted sums
// over the block's columns in order.  kSkip: lookups the history's span
// table shows to gather zero density and gradients are skipped (an integrand
// of exactly 0, as in the near regions); every entry is read from the table
// itself (zrow = the table, zt0 = nt: no LDS copy of the newest slices).
template <bool kMixed, bool kTI, bool kSkip = false>
__device__ void far_obs_block(const WakeTask &T, int l, double *__restrict__ partials,
                              RowTable &rt) {
  const DfcsrHistory &H = T.H;
  const DfcsrLattice &L = T.L;
  const PhasePlan pl = make_plan(T.P);
  const int xb = T.P.mesh_xbins, zb = T.P.mesh_zbins;
  const int per_line = pl.nob * pl.q;
  const int zl = l / per_line, rem = l - zl * per_line;
  const int ob = rem / pl.q, g = rem - ob * pl.q;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane / pl.W, ol = lane - sub * pl.W;
  const int ix = ob * pl.W + ol;
  const bool valid = ix < xb;
  const int i = (valid ? ix : ob * pl.W) * zb + zl;   // padding lanes follow the block's first
  const int cb = (4 * g + wave) * pl.S + sub;
  const int Nz = T.P.Nz;
  const int k0 = (int)((int64_t)cb * Nz / pl.nA), k1 = (int)((int64_t)(cb + 1) * Nz / pl.nA);
  const double *__restrict__ od = partials + T.obs + (int64_t)i * kObsStride;
  const double t = T.P.t;
  const HistAxes A = make_axes(H);
  const double inv_ds = 1.0 / L.ds;
  // the z line's x' rows (the same for every observer of the line) and this
  // observer's s' columns (the same too, off the chirp branch)
  const LinAxis xa = load_axis(partials + T.obs + (int64_t)(ob * pl.W * zb + zl) * kObsStride + 16);
  const LinAxis sa = load_axis(od + 32);
  double *out = partials + T.part + ((int64_t)i * (pl.nA + pl.nB) + cb) * 2;
  if (xa.n < 2 || sa.n < 2) {   // trapz over a single sample (uniform)
    if (valid) out[0] = out[1] = 0.0;
    return;
  }
  bool nan_region = false;
  if (od[9] != 0.0) {   // as wake_phase: observer and an end column outside the table
    const bool in0 = trunc_index((sa.at(0) - L.min_s) * inv_ds, L.ns).ok;
    const bool in1 = trunc_index((sa.at(sa.n - 1) - L.min_s) * inv_ds, L.ns).ok;
    nan_region = !in0 || !in1;
  }
  // hull of the rows inside the history (x' is monotone in j)
  __shared__ int hull[2];
  if (threadIdx.x == 0) {
    hull[0] = xa.n;
    hull[1] = -1;
  }
  __syncthreads();
  {
    int lo = xa.n, hi = -1;
    for (int j = threadIdx.x; j < xa.n; j += blockDim.x) {
      const double ux = (xa.at(j) - A.min_x) * A.ix;
      if (ux > -1.0 && ux < (double)H.Xi) {
        lo = min(lo, j);
        hi = max(hi, j);
      }
    }
    if (lo <= hi) {
      atomicMin(&hull[0], lo);
      atomicMax(&hull[1], hi);
    }
  }
  __syncthreads();
  const int v0 = hull[0], v1 = hull[1] + 1;
  Observer o;
  o.s = od[0];
  o.x = od[1];
  o.t = t;
  o.L.X0 = od[2];
  o.L.Y0 = od[3];
  o.L.nx = od[4];
  o.L.ny = od[5];
  o.L.tx = od[6];
  o.L.ty = od[7];
  o.vx = od[8];
  const double ut0 = (t - A.min_t) * A.it;
  double acc_z = 0.0, acc_x = 0.0;
  unsigned evaluated = 0;
  for (int b0 = v0; b0 < v1; b0 += kMaxRows) {   // uniform over the workgroup
    const int nr = min(kMaxRows, v1 - b0);
    if (threadIdx.x < nr) {
      const int r = threadIdx.x;
      const double xp = xa.at(b0 + r);
      const TruncIdx X = trunc_index((xp - A.min_x) * A.ix, H.Xi);
      rt.xp[r] = xp;
      rt.w[r] = xa.w(b0 + r);
      rt.frac[r] = X.frac;
      rt.i0[r] = X.i0;
      rt.i1[r] = X.i1;
      rt.ok[r] = X.ok;
      rt.xpf[r] = (float)xp;
      rt.wf[r] = (float)rt.w[r];
      rt.fracf[r] = (float)X.frac;
    }
    __syncthreads();
    for (int k = k0; k < k1; ++k) {
      const double sp = sa.at(k);
      const double wz = sa.w(k);
      double cz = 0.0, cx = 0.0;
      if (kMixed) {
        const ColumnMixed col = make_column_mixed(L, o, A, sp, inv_ds);
        for (int r = 0; r < nr; ++r) {
          if (!rt.ok[r]) continue;
          float Iz, Ix;
          if (!csr_point_mixed<kTI, kRings>(H, A, ut0, col, rt.xp[r], rt.xpf[r], rt.i0[r],
                                            rt.i1[r], CellRow{0u, 0u}, rt.fracf[r], Iz, Ix,
                                            kSkip ? H.zspan : nullptr, H.nt))
            continue;
          ++evaluated;
          const float wx = rt.wf[r];
          cz += (double)(wx * Iz);
          cx += (double)(wx * Ix);
        }
      } else {
        const Column col = make_column<true>(L, o, sp, inv_ds);
        for (int r = 0; r < nr; ++r) {
          if (!rt.ok[r]) continue;
          const TruncIdx X = {rt.i0[r], rt.i1[r], rt.frac[r], true};
          double Iz, Ix;
          if (!csr_point<true, false>(H, A, o, col, rt.xp[r], Iz, Ix, &X, nullptr,
                                      kSkip ? H.zspan : nullptr, H.nt))
            continue;
          ++evaluated;
          const double wx = rt.w[r];
          cz += wx * Iz;
          cx += wx * Ix;
        }
      }
      acc_z += wz * cz;
      acc_x += wz * cx;
    }
    __syncthreads();   // before the next rows overwrite the table
  }
  if (T.P.points_evaluated) {   // uniform: a diagnostic call
    unsigned long long c = valid ? evaluated : 0u;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0 && c) atomicAdd(T.P.points_evaluated, c);
  }
  if (valid) {
    out[0] = nan_region ? NAN : acc_z;
    out[1] = nan_region ? NAN : acc_x;
  }
}

// The FP32 ring a mixed launch reads: region 0 (many retarded planes per
// wavefront) the time-innermost one, the later regions (one plane,
// consecutive z') the chunked one, each when present.
__host__ __device__ inline bool mixed_reads_ti(const DfcsrHistory &H, bool near) {
  return near ? (H.data32 == nullptr) : (H.data32t != nullptr);
}

// Whether a mixed wake's near regions read the float32 cell-major ring and
// its time differences (both must be present).
__host__ __device__ inline bool mixed_cells(const DfcsrHistory &H) {
  return H.cells32 != nullptr && H.dcells32 != nullptr;
}

// Whether every byte of the float32 cell-major rings has a 32-bit offset.
__host__ __device__ inline bool cells32_ring32(const DfcsrHistory &H) {
  return (int64_t)H.cap * H.Xi * DFCSR_CELL_ROW(H.Zi) * (int64_t)(DFCSR_CELL * sizeof(float)) <=
         0xffffffffLL;
}

// Workgroups per CU the register allocation must allow: the near phases on
// the cell-major ring 3 (FP64) or 4 (mixed: 128 registers), the others 2.
template <bool kMixed, bool kNear, bool kTI, int kCells = kRings>
__global__ __launch_bounds__(kThreads, kCells ? (kMixed ? 4 : 3) : DFCSR_WAKE_MINBLOCKS) void wake_phase_kernel(
    WakeTask T, double *__restrict__ partials) {
  __shared__ double red[kThreads / 64];
  const int b = blockIdx.x, nb = gridDim.x;
  run_phase_block<kMixed, kNear, kTI, kCells>(
      T, kNear ? xcd_group(b, nb, kNearGroup) : xcd_remap(b, nb), partials, red);
}

// The observer-major region-0 launch: each XCD a contiguous range of z lines.
// Three workgroups per CU (FP64: 168 VGPRs and a few spilled, against 191 at
// two; large run FP64 far phases -10 %).  kSkip: the span-table skip
// (far_skip).
template <bool kMixed, bool kTI, bool kSkip = false>
__global__ __launch_bounds__(kThreads, 3) void wake_far_kernel(
    WakeTask T, double *__restrict__ partials) {
  __shared__ RowTable rt;
  far_obs_block<kMixed, kTI, kSkip>(T, xcd_remap(blockIdx.x, gridDim.x), partials, rt);
}

// Whether a wake's observer-major region 0 skips lookups on the span table:
// on the chirp branch only.  Measured, bit-identical either way: the chicane
// (chirp branch, 1e6 particles, 64 x 64 mesh) FP64 wake 4.60 -> 4.14 ms and
// mixed 2.52 -> 2.35 ms; off the chirp branch the 1e7-particle 128 x 128 run
// loses (FP64 wake 11.73 -> 12.13 ms): its skipped points read cached zero
// records, and the span read lengthens every other point's path.
__host__ __device__ inline bool far_skip(const DfcsrHistory &H, const PhasePlan &pl) {
  return H.zspan != nullptr && pl.nreg == 4;
}

// Last task whose first index (workgroup or observer) is <= b.
template <typename F>
__device__ inline int find_task(const WakeTask *__restrict__ tasks, int ntask, int b, F first) {
  int lo = 0, hi = ntask - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first(tasks[mid]) <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool kMixed, bool kNear>
__global__ __launch_bounds__(kThreads, kMixed && kNear ? 4 : DFCSR_WAKE_MINBLOCKS) void wake_phase_batched_kernel(
    const WakeTask *__restrict__ tasks, int ntask, double *__restrict__ partials) {
  __shared__ double red[kThreads / 64];
  const int b = kNear ? xcd_group(blockIdx.x, gridDim.x, kNearGroup)
                      : xcd_remap(blockIdx.x, gridDim.x);
  const int j = find_task(tasks, ntask, b,
                          [](const WakeTask &T) { return kNear ? T.blockB0 : T.blockA0; });
  const WakeTask &T = tasks[j];
  const int l = b - (kNear ? T.blockB0 : T.blockA0);
  if (!kNear && T.P.observers == nullptr) {   // observer-major region 0
    __shared__ RowTable rt;
    const bool skip = far_skip(T.H, make_plan(T.P));   // per task, as dfcsr_wake_mesh
    if constexpr (kMixed) {
      const bool ti = mixed_reads_ti(T.H, false);
      if (ti && skip)
        far_obs_block<true, true, true>(T, l, partials, rt);
      else if (ti)
        far_obs_block<true, true>(T, l, partials, rt);
      else if (skip)
        far_obs_block<true, false, true>(T, l, partials, rt);
      else
        far_obs_block<true, false>(T, l, partials, rt);
    } else if (skip) {
      far_obs_block<false, false, true>(T, l, partials, rt);
    } else {
      far_obs_block<false, false>(T, l, partials, rt);
    }
    return;
  }
  if constexpr (kMixed) {   // per task: its FP32 rings, cell-major or not
    if (kNear && mixed_cells(T.H) && cells32_ring32(T.H))
      run_phase_block<true, kNear, false, kCellsRing32>(T, l, partials, red);
    else if (kNear && mixed_cells(T.H))
      run_phase_block<true, kNear, false, kCellsPlane>(T, l, partials, red);
    else if (mixed_reads_ti(T.H, kNear))
      run_phase_block<true, kNear, true>(T, l, partials, red);
    else
      run_phase_block<true, kNear, false>(T, l, partials, red);
  } else if (kNear && T.H.cells) {   // per task: the cell-major ring or not
    run_phase_block<false, kNear, false, kCellsPlane>(T, l, partials, red);
  } else {
    run_phase_block<false, kNear, false>(T, l, partials, red);
  }
}

__global__ __launch_bounds__(kThreads) void wake_observer_kernel(WakeTask T,
                                                                 double *__restrict__ partials) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < T.M) observer_point(T, i, partials + T.obs + (int64_t)i * kObsStride);
}

__global__ __launch_bounds__(kThreads) void wake_observer_batched_kernel(
    const WakeTask *__restrict__ tasks, int ntask, int nobs, double *__restrict__ partials) {
  const int g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= nobs) return;
  const int j = find_task(tasks, ntask, g, [](const WakeTask &T) { return T.obs0; });
  const WakeTask &T = tasks[j];
  const int i = g - T.obs0;
  observer_point(T, i, partials + T.obs + (int64_t)i * kObsStride);
}

// Sum of the phases of mesh point i in phase order; the reference's signs
// and scaling (CSR.py:599-602).
__device__ inline void finalize_point(const WakeTask &T, int i, const double *__restrict__ partials) {
  const int nph = T.nA + T.nB;
  const double *p = partials + T.part + (int64_t)i * nph * 2;
  double sz = 0.0, sx = 0.0;
  for (int ph = 0; ph < nph; ++ph) {
    sz += p[2 * ph];
    sx += p[2 * ph + 1];
  }
  T.dE[i] = -T.P.csr_scaling * sz;
  T.xk[i] = T.P.csr_scaling * sx;
  if (T.xmesh) {
    double xm, zm;
    mesh_point(T.P, i, xm, zm);
    T.xmesh[i] = xm;
    T.zmesh[i] = zm;
  }
}

__global__ __launch_bounds__(kThreads) void wake_finalize_kernel(WakeTask T,
                                                                 const double *__restrict__ partials) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < T.M) finalize_point(T, i, partials);
}

__global__ __launch_bounds__(kThreads) void wake_finalize_batched_kernel(
    const WakeTask *__restrict__ tasks, int ntask, int nobs, const double *__restrict__ partials) {
  const int g = blockIdx.x * kThreads + threadIdx.x;
  if (g >= nobs) return;
  const int j = find_task(tasks, ntask, g, [](const WakeTask &T) { return T.obs0; });
  finalize_point(tasks[j], g - tasks[j].obs0, partials);
}

int check_params(const DfcsrWakeParams *p, const char *who) {
  if (!p) return set_error(DFCSR_EINVAL, "%s: NULL params", who);
  if (p->Nx < 1 || p->Nz < 1 || p->mesh_xbins < 1 || p->mesh_zbins < 1)
    return set_error(DFCSR_EINVAL, "%s: bad bins (Nx %d Nz %d mesh %d x %d)", who, p->Nx, p->Nz,
                     p->mesh_xbins, p->mesh_zbins);
  if ((int64_t)p->mesh_xbins * p->mesh_zbins > (int64_t)1 << 24)
    return set_error(DFCSR_ERANGE, "%s: mesh too large", who);

  return DFCSR_OK;
}

// The rings a wake of this precision reads: FP64 the float64 ring, mixed
// one of the float32 rings.
int check_rings(const DfcsrHistory &H, int precision, const char *who) {
  if (precision == DFCSR_PREC_FP64 && !H.data)
    return set_error(DFCSR_EINVAL, "%s: FP64 precision needs the FP64 history ring", who);
  // the wake addresses a cell within a plane of a cell-major ring with 32-bit
  // byte offsets
  const bool cells = precision == DFCSR_PREC_MIXED ? mixed_cells(H) : H.cells != nullptr;
  const int64_t cellB = DFCSR_CELL * (precision == DFCSR_PREC_MIXED ? 4 : 8);
  if (cells && (int64_t)H.Xi * DFCSR_CELL_ROW(H.Zi) * cellB > ((int64_t)1 << 32))
    return set_error(DFCSR_EINVAL, "%s: a cell-major plane of %d x %d cells exceeds 4 GiB", who,
                     H.Xi, DFCSR_CELL_ROW(H.Zi));
  if (precision == DFCSR_PREC_MIXED && !H.data32 && !H.data32t)
    return set_error(DFCSR_EINVAL, "%s: mixed precision needs an FP32 history ring", who);
  return DFCSR_OK;
}

// check_params without setting the error message (for the size queries).
bool params_ok(const DfcsrWakeParams &p) {
  return p.Nx >= 1 && p.Nz >= 1 && p.mesh_xbins >= 1 && p.mesh_zbins >= 1 &&
         (int64_t)p.mesh_xbins * p.mesh_zbins <= ((int64_t)1 << 24);
}

// Partial sums and observer data of one simulation, in doubles.
int64_t partial_doubles(const DfcsrWakeParams &p) {
  const PhasePlan pl = make_plan(p);
  return (int64_t)p.mesh_xbins * p.mesh_zbins * ((pl.nA + pl.nB) * 2 + kObsStride);
}

// Fills a task; rejects launches whose workgroup counts overflow int.
int make_task(const DfcsrHistory &H, const DfcsrLattice &L, const DfcsrWakeParams &P,
              double *dE, double *xk, double *xmesh, double *zmesh, WakeTask &T, const char *who) {
  const PhasePlan pl = make_plan(P);
  T.H = H;
  T.L = L;
  T.P = P;
  T.dE = dE;
  T.xk = xk;
  T.xmesh = xmesh;
  T.zmesh = zmesh;
  T.M = P.mesh_xbins * P.mesh_zbins;
  T.part = 0;
  T.obs = (int64_t)T.M * (pl.nA + pl.nB) * 2;
  T.nA = pl.nA;
  T.nB = pl.nB;
  T.blockA0 = T.blockB0 = T.obs0 = 0;
  if ((int64_t)T.M * (pl.nA + pl.nB) > ((int64_t)1 << 30) || far_blocks(pl, P) > ((int64_t)1 << 30))
    return set_error(DFCSR_ERANGE, "%s: %d mesh points x %d phases is too many workgroups", who,
                     T.M, pl.nA + pl.nB);
  return DFCSR_OK;
}

constexpr int64_t kTableAlign = 256;
constexpr int kMaxTasks = DFCSR_WAKE_MAX_TASKS;

int64_t table_bytes(int k) {
  return (((int64_t)k * (int64_t)sizeof(WakeTask) + kTableAlign - 1) / kTableAlign) * kTableAlign;
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_wake_workspace_bytes(const DfcsrWakeParams *p) {
  if (!p || !params_ok(*p)) return 256;
  return partial_doubles(*p) * (int64_t)sizeof(double) + 256;
}

extern "C" int dfcsr_wake_mesh(const DfcsrHistory *hist, const DfcsrLattice *lat,
                               const DfcsrWakeParams *p, int precision, double *xmesh,
                               double *zmesh, double *dE_dct, double *x_kick, void *ws,
                               int64_t ws_bytes, void *stream) {
  int rc;
  if ((rc = check_history(hist, "wake_mesh"))) return rc;
  if ((rc = check_lattice(lat, "wake_mesh"))) return rc;
  if ((rc = check_params(p, "wake_mesh"))) return rc;
  if (!dE_dct || !x_kick) return set_error(DFCSR_EINVAL, "wake_mesh: NULL output");
  if ((xmesh == nullptr) != (zmesh == nullptr))
    return set_error(DFCSR_EINVAL, "wake_mesh: xmesh and zmesh must both be given or both NULL");
  if (precision != DFCSR_PREC_FP64 && precision != DFCSR_PREC_MIXED)
    return set_error(DFCSR_EINVAL, "wake_mesh: unknown precision %d", precision);
  if ((rc = check_rings(*hist, precision, "wake_mesh"))) return rc;
  if (!ws || ws_bytes < dfcsr_wake_workspace_bytes(p))
    return set_error(DFCSR_ENOWS, "wake_mesh: workspace of %lld bytes, %lld needed",
                     (long long)ws_bytes, (long long)dfcsr_wake_workspace_bytes(p));
  WakeTask T;
  if ((rc = make_task(*hist, *lat, *p, dE_dct, x_kick, xmesh, zmesh, T, "wake_mesh"))) return rc;
  double *part = static_cast<double *>(ws);
  hipStream_t s = as_stream(stream);
  const PhasePlan pl = make_plan(*p);
  const dim3 ga((unsigned)far_blocks(pl, *p)), gb(T.M * T.nB), gf((T.M + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(wake_observer_kernel, gf, dim3(kThreads), 0, s, T, part);
  if (precision == DFCSR_PREC_MIXED) {
    const bool ti_far = mixed_reads_ti(*hist, false);
    const bool skip = far_skip(*hist, pl);
    if (pl.obs_major && ti_far && skip)
      hipLaunchKernelGGL((wake_far_kernel<true, true, true>), ga, dim3(kThreads), 0, s, T, part);
    else if (pl.obs_major && ti_far)
      hipLaunchKernelGGL((wake_far_kernel<true, true>), ga, dim3(kThreads), 0, s, T, part);
    else if (pl.obs_major && skip)
      hipLaunchKernelGGL((wake_far_kernel<true, false, true>), ga, dim3(kThreads), 0, s, T, part);
    else if (pl.obs_major)
      hipLaunchKernelGGL((wake_far_kernel<true, false>), ga, dim3(kThreads), 0, s, T, part);
    else if (ti_far)
      hipLaunchKernelGGL((wake_phase_kernel<true, false, true>), ga, dim3(kThreads), 0, s, T, part);
    else
      hipLaunchKernelGGL((wake_phase_kernel<true, false, false>), ga, dim3(kThreads), 0, s, T, part);
    const bool ti = mixed_reads_ti(*hist, true);
    if (mixed_cells(*hist) && cells32_ring32(*hist)) {
      hipLaunchKernelGGL((wake_phase_kernel<true, true, false, kCellsRing32>), gb, dim3(kThreads),
                         0, s, T, part);
    } else if (mixed_cells(*hist)) {
      hipLaunchKernelGGL((wake_phase_kernel<true, true, false, kCellsPlane>), gb, dim3(kThreads),
                         0, s, T, part);
    } else if (ti) {
      hipLaunchKernelGGL((wake_phase_kernel<true, true, true>), gb, dim3(kThreads), 0, s, T, part);
    } else {
      hipLaunchKernelGGL((wake_phase_kernel<true, true, false>), gb, dim3(kThreads), 0, s, T, part);
    }
  } else {
    if (pl.obs_major && far_skip(*hist, pl))
      hipLaunchKernelGGL((wake_far_kernel<false, false, true>), ga, dim3(kThreads), 0, s, T, part);
    else if (pl.obs_major)
      hipLaunchKernelGGL((wake_far_kernel<false, false>), ga, dim3(kThreads), 0, s, T, part);
    else
      hipLaunchKernelGGL((wake_phase_kernel<false, false, false>), ga, dim3(kThreads), 0, s, T, part);
    if (hist->cells) {
      hipLaunchKernelGGL((wake_phase_kernel<false, true, false, kCellsPlane>), gb, dim3(kThreads),
                         0, s, T, part);
    } else {
      hipLaunchKernelGGL((wake_phase_kernel<false, true, false>), gb, dim3(kThreads), 0, s, T, part);
    }
  }
  hipLaunchKernelGGL(wake_finalize_kernel, gf, dim3(kThreads), 0, s, T, (const double *)part);
  DFCSR_LAUNCH_CHECK("wake_mesh");
  return DFCSR_OK;
}

extern "C" int64_t dfcsr_wake_task_table_bytes(int k) { return table_bytes(k < 1 ? 1 : k); }

extern "C" int64_t dfcsr_wake_batched_workspace_bytes(int k, const DfcsrWakeParams *p) {
  if (k < 1 || !p) return table_bytes(1);
  int64_t n = 0;
  for (int j = 0; j < k; ++j)
    if (params_ok(p[j])) n += partial_doubles(p[j]);
  return table_bytes(k) + n * (int64_t)sizeof(double);
}

extern "C" int dfcsr_wake_tasks_pack(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                                     const DfcsrWakeParams *p, int precision,
                                     double *const *dE_dct, double *const *x_kick, void *tasks,
                                     int64_t tasks_bytes, DfcsrWakeBatch *batch) {
  if (k < 1 || k > kMaxTasks)
    return set_error(DFCSR_ERANGE, "wake_tasks_pack: %d tasks (1 to %d)", k, kMaxTasks);
  if (!hist || !lat || !p || !dE_dct || !x_kick || !tasks || !batch)
    return set_error(DFCSR_EINVAL, "wake_tasks_pack: NULL pointer");
  if (tasks_bytes < (int64_t)k * (int64_t)sizeof(WakeTask))
    return set_error(DFCSR_ENOWS, "wake_tasks_pack: task buffer too small");
  if (precision != DFCSR_PREC_FP64 && precision != DFCSR_PREC_MIXED)
    return set_error(DFCSR_EINVAL, "wake_tasks_pack: unknown precision %d", precision);
  WakeTask *th = static_cast<WakeTask *>(tasks);
  int64_t na = 0, nb = 0, nobs = 0, part = 0;
  for (int j = 0; j < k; ++j) {
    int rc;
    if ((rc = check_history(&hist[j], "wake_tasks_pack"))) return rc;
    if ((rc = check_lattice(&lat[j], "wake_tasks_pack"))) return rc;
    if ((rc = check_params(&p[j], "wake_tasks_pack"))) return rc;
    if (!dE_dct[j] || !x_kick[j]) return set_error(DFCSR_EINVAL, "wake_tasks_pack: NULL output");
    if ((rc = check_rings(hist[j], precision, "wake_tasks_pack"))) return rc;
    WakeTask &T = th[j];
    if ((rc = make_task(hist[j], lat[j], p[j], dE_dct[j], x_kick[j], nullptr, nullptr, T,
                        "wake_tasks_pack")))
      return rc;
    T.obs += part;
    T.part = part;
    T.blockA0 = (int)na;
    T.blockB0 = (int)nb;
    T.obs0 = (int)nobs;
    part += partial_doubles(p[j]);
    na += far_blocks(make_plan(p[j]), p[j]);
    nb += (int64_t)T.M * T.nB;
    nobs += T.M;
    if (na > ((int64_t)1 << 30) || nb > ((int64_t)1 << 30))
      return set_error(DFCSR_ERANGE, "wake_tasks_pack: too many workgroups");
  }
  batch->tasks = k;
  batch->blocks_far = (int32_t)na;
  batch->blocks_near = (int32_t)nb;
  batch->observers = (int32_t)nobs;
  batch->partial_bytes = part * (int64_t)sizeof(double);
  batch->precision = precision;
  return DFCSR_OK;
}

extern "C" int dfcsr_wake_batched_launch(const void *dev_tasks, const DfcsrWakeBatch *batch,
                                         int precision, void *partials, int64_t partials_bytes,
                                         void *stream) {
  if (!dev_tasks || !batch || batch->tasks < 1 || batch->blocks_far < 1 ||
      batch->blocks_near < 1 || batch->observers < 1)
    return set_error(DFCSR_EINVAL, "wake_batched_launch: bad arguments");
  if (!partials || partials_bytes < batch->partial_bytes)
    return set_error(DFCSR_ENOWS, "wake_batched_launch: partial-sum buffer too small");
  if (precision != batch->precision)
    return set_error(DFCSR_EINVAL, "wake_batched_launch: tasks packed for precision %d, not %d",
                     batch->precision, precision);
  hipStream_t s = as_stream(stream);
  const WakeTask *dt = static_cast<const WakeTask *>(dev_tasks);
  double *part = static_cast<double *>(partials);
  const int k = batch->tasks;
  const dim3 ga(batch->blocks_far), gb(batch->blocks_near);
  const dim3 gf((batch->observers + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(wake_observer_batched_kernel, gf, dim3(kThreads), 0, s, dt, k,
                     batch->observers, part);
  if (precision == DFCSR_PREC_MIXED) {
    hipLaunchKernelGGL((wake_phase_batched_kernel<true, false>), ga, dim3(kThreads), 0, s, dt, k, part);
    hipLaunchKernelGGL((wake_phase_batched_kernel<true, true>), gb, dim3(kThreads), 0, s, dt, k, part);
  } else if (precision == DFCSR_PREC_FP64) {
    hipLaunchKernelGGL((wake_phase_batched_kernel<false, false>), ga, dim3(kThreads), 0, s, dt, k, part);
    hipLaunchKernelGGL((wake_phase_batched_kernel<false, true>), gb, dim3(kThreads), 0, s, dt, k, part);
  } else {
    return set_error(DFCSR_EINVAL, "wake_batched_launch: unknown precision %d", precision);
  }
  hipLaunchKernelGGL(wake_finalize_batched_kernel, gf, dim3(kThreads), 0, s, dt, k,
                     batch->observers, (const double *)part);
  DFCSR_LAUNCH_CHECK("wake_batched_launch");
  return DFCSR_OK;
}

extern "C" int dfcsr_wake_mesh_batched(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                                       const DfcsrWakeParams *p, int precision,
                                       double *const *dE_dct, double *const *x_kick, void *ws,
                                       int64_t ws_bytes, void *stream) {
  if (k < 1) return DFCSR_OK;
  if (k > kMaxTasks)
    return set_error(DFCSR_ERANGE, "wake_mesh_batched: %d tasks > %d", k, kMaxTasks);
  if (!ws) return set_error(DFCSR_EINVAL, "wake_mesh_batched: NULL workspace");
  if (!p) return set_error(DFCSR_EINVAL, "wake_mesh_batched: NULL params");
  if (ws_bytes < dfcsr_wake_batched_workspace_bytes(k, p))
    return set_error(DFCSR_ENOWS, "wake_mesh_batched: workspace too small");
  std::vector<WakeTask> tasks_h((size_t)k);
  DfcsrWakeBatch batch;
  int rc = dfcsr_wake_tasks_pack(k, hist, lat, p, precision, dE_dct, x_kick, tasks_h.data(),
                                 (int64_t)(tasks_h.size() * sizeof(WakeTask)), &batch);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  // ordered on the stream after whatever still reads the workspace's table;
  // the host table must outlive the copy, so wait for it before returning
  DFCSR_CHECK_HIP(hipMemcpyAsync(ws, tasks_h.data(), (size_t)k * sizeof(WakeTask),
                                 hipMemcpyHostToDevice, s));
  DFCSR_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t tb = table_bytes(k);
  return dfcsr_wake_batched_launch(ws, &batch, precision, static_cast<char *>(ws) + tb,
                                   ws_bytes - tb, stream);
}

// ---- the entry point that allocates nothing --------------------------------
//
// dfcsr_wake_mesh_noalloc keeps its partial sums and observer data in a
// file-scope device array instead of a caller's workspace.  Each GPU the
// library runs on holds its own copy of the array until the process ends;
// host code reaches it only through the per-device address hipGetSymbolAddress
// returns (the array's name in host code is not a device address).  Every
// phase workgroup writes its partial sums and the observer launch its
// observer data before they are read, so nothing is cleared between calls;
// calls that overlap on one GPU share the array and race.
namespace dfcsr {
namespace {

static __device__ __align__(256) double scratch[DFCSR_WAKE_SCRATCH_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_dev[kMaxDevices] = {};

// Address of the current device's copy of scratch, looked up once per device.
int scratch_address(void **out, const char *who) {
  int dev = -1;
  DFCSR_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(DFCSR_ERANGE, "%s: device %d (at most %d devices)", who, dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(DFCSR_EHIP, "%s: hipGetSymbolAddress(scratch) on device %d: %s", who, dev,
                       hipGetErrorString(e));
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return DFCSR_OK;
}

}  // namespace
}  // namespace dfcsr

extern "C" int dfcsr_wake_mesh_noalloc(const DfcsrHistory *hist, const DfcsrLattice *lat,
                                       const DfcsrWakeParams *p, int precision, double *xmesh,
                                       double *zmesh, double *dE_dct, double *x_kick,
                                       void *stream) {
  int rc;
  if ((rc = check_params(p, "wake_mesh_noalloc"))) return rc;
  const int64_t need = dfcsr_wake_workspace_bytes(p);
  if (need > (int64_t)DFCSR_WAKE_SCRATCH_BYTES)
    return set_error(DFCSR_ERANGE,
                     "wake_mesh_noalloc: %d x %d mesh points need %lld bytes of scratch, more "
                     "than the %lld the library holds (use dfcsr_wake_mesh with a workspace)",
                     p->mesh_xbins, p->mesh_zbins, (long long)need,
                     (long long)DFCSR_WAKE_SCRATCH_BYTES);
  void *ws = nullptr;
  if ((rc = scratch_address(&ws, "wake_mesh_noalloc"))) return rc;
  return dfcsr_wake_mesh(hist, lat, p, precision, xmesh, zmesh, dE_dct, x_kick, ws,
                         (int64_t)DFCSR_WAKE_SCRATCH_BYTES, stream);
}
