// a13-a15: the retarded CSR integral on the wake mesh.
//
// Replaces CSR2D.get_CSR_mesh / calculate_2D_CSR / get_CSR_wake /
// get_CSR_integrand (CSR.py:361-782).  One workgroup integrates one mesh
// point over all of its regions (3 without chirp, 4 with):
//   work item  = (region, row chunk, s' column k); lanes of a wavefront take
//                consecutive columns, so they share x' rows (the x-range test
//                is wave-uniform) and their retarded lookups land on the same
//                history rows (z' = s' - t_ret barely moves along s').
//   per column : lattice lookups at s', rho(s'), the trapezoid weight in s'.
//   per point  : |r - r'|, t_ret, the trilinear gather of five fields, the
//                integrand; points whose gather is outside the history add 0
//                and skip the arithmetic.
//   reduction  : sum_j,k wx_j wz_k I_jk per lane, then a fixed-order block
//                tree: deterministic, the same every launch.
// Workgroups are mapped so each XCD integrates a contiguous range of mesh
// points (neighbouring observers read the same history lines through that
// XCD's L2).
#include "csr_point.h"

namespace dfcsr {

int check_history(const DfcsrHistory *H, const char *who);
int check_lattice(const DfcsrLattice *L, const char *who);

namespace {

constexpr int kThreads = 256;
#ifndef DFCSR_WAKE_CHUNK
#define DFCSR_WAKE_CHUNK 64
#endif
#ifndef DFCSR_WAKE_MINBLOCKS
#define DFCSR_WAKE_MINBLOCKS 2
#endif
#ifndef DFCSR_WAKE_ORDER
#define DFCSR_WAKE_ORDER 0
#endif

constexpr int kChunk = DFCSR_WAKE_CHUNK;   // x' rows per work item
constexpr int kNumXcd = 8;

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

__device__ inline int xcd_remap(int b, int nb) {
  int q = nb / kNumXcd, r = nb % kNumXcd;
  int xcd = b % kNumXcd, pos = b / kNumXcd;
  return xcd * q + (xcd < r ? xcd : r) + pos;
}

// Region table of one mesh point, shared by the workgroup through LDS
// (dynamic indexing of a private array would put it in scratch memory).
struct RegionLds {
  Regions rg;
  int off[5];
};

// Integrates mesh point i; every thread returns the block sums.
__device__ void wake_point_fp64(const DfcsrHistory &H, const DfcsrLattice &L,
                                const DfcsrWakeParams &P, int i, double *red, RegionLds &R,
                                double &out_z, double &out_x, double &xm_out, double &zm_out) {
  double xm, zm;
  mesh_point(P, i, xm, zm);
  xm_out = xm;
  zm_out = zm;
  const double t = P.t;
  const double s = t + zm, x = xm;
  const HistAxes A = make_axes(H);
  const double inv_ds = 1.0 / L.ds;
  const Observer o = make_observer<true>(H, A, L, s, x, t);
  if (threadIdx.x == 0) {
    make_regions(P, s, x, t, R.rg);
    R.off[0] = 0;
    for (int r = 0; r < R.rg.n; ++r)
      R.off[r + 1] = R.off[r] + ((R.rg.xa[r].n + kChunk - 1) / kChunk) * R.rg.sa[r].n;
  }
  __syncthreads();
  const int nreg = R.rg.n;
  const int total = R.off[nreg];
  double acc_z = 0.0, acc_x = 0.0;
  for (int it = threadIdx.x; it < total; it += blockDim.x) {
    int r = 0;
    while (r + 1 < nreg && it >= R.off[r + 1]) ++r;
    const int rem = it - R.off[r];
    const LinAxis sa = R.rg.sa[r];
    const LinAxis xa = R.rg.xa[r];
    const int c = rem / sa.n, k = rem % sa.n;
    const double sp = sa.at(k);
    const double wz = sa.w(k);
    const Column col = make_column<true>(L, o, sp, inv_ds);
    const int j0 = c * kChunk;
    const int j1 = min(j0 + kChunk, xa.n);
    double cz = 0.0, cx = 0.0;
    for (int j = j0; j < j1; ++j) {
      const double xp = xa.at(j);
      const double ux = (xp - A.min_x) * A.ix;
      if (!(ux > -1.0 && ux < (double)H.Xi)) continue;   // wave-uniform: rows are shared
      double Iz, Ix;
      if (!csr_point<true>(H, A, o, col, xp, Iz, Ix)) continue;
      const double wx = xa.w(j);
      cz += wx * Iz;
      cx += wx * Ix;
    }
    acc_z += wz * cz;
    acc_x += wz * cx;
  }
  out_z = block_sum(acc_z, red);
  out_x = block_sum(acc_x, red);
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
// integrand are float32.  1/R comes from an FP32 reciprocal square root with
// one FP64 Newton step.  Each 32-row chunk is summed in FP32 with Neumaier's
// compensated summation (Neumaier 1974, the Kahan-Babuska improvement); chunk
// sums are weighted and accumulated in float64, and the block reduction is
// float64.
struct ColumnMixed {
  double sp, ax, ay, npx, npy, P0, P1, p2;
  float rho, D0, D1, D2, G0, G1, G2, e, a, b, A, B;
};

__device__ inline ColumnMixed make_column_mixed(const DfcsrLattice &L, const Observer &o,
                                                double sp, double inv_ds) {
  const Column c = make_column<true>(L, o, sp, inv_ds);
  ColumnMixed m;
  m.sp = sp;
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

__device__ inline const float *hist_rec32(const DfcsrHistory &H, int ti, int xi, int zi) {
  int slot = H.head + ti;
  if (slot >= H.cap) slot -= H.cap;
  return H.data32 + (((int64_t)slot * H.Xi + xi) * H.Zi + zi) * DFCSR_HIST_REC32;
}

struct Neumaier {
  float s = 0.f, c = 0.f;
  __device__ void add(float x) {
    float t = s + x;
    c += (fabsf(s) >= fabsf(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
  }
  __device__ float value() const { return s + c; }
};

// One source point in mixed precision; false when outside the history.
__device__ inline bool csr_point_mixed(const DfcsrHistory &H, const HistAxes &A, double t,
                                       const ColumnMixed &c, double xp, int xi0, float xd,
                                       float &Iz, float &Ix) {
  const double rx = c.ax - xp * c.npx;
  const double ry = c.ay - xp * c.npy;
  const double R2 = rx * rx + ry * ry;
  double y = (double)rsqrtf((float)R2);
  y = y * (1.5 - 0.5 * R2 * y * y);               // one Newton step: ~1e-14 relative
  const double R = R2 * y;
  const double t_ret = t - R;
  const double ut = (t_ret - A.min_t) * A.it;
  const double uz = ((c.sp - t_ret) - A.min_z) * A.iz;
  const TruncIdx T = trunc_index(ut, H.nt), Z = trunc_index(uz, H.Zi);
  if (!(T.ok && Z.ok)) return false;
  const float td = (float)T.frac, zd = (float)Z.frac;
  // one corner-packed record (one cache line) per time plane
  const float4 *q0 = reinterpret_cast<const float4 *>(hist_rec32(H, T.i0, xi0, Z.i0));
  const float4 *q1 = reinterpret_cast<const float4 *>(hist_rec32(H, T.i1, xi0, Z.i0));
  float f[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const float4 a = q0[k], b = q1[k];   // (x0,z0), (x0,z1), (x1,z0), (x1,z1)
    const float c00 = fmaf(td, b.x - a.x, a.x);
    const float c01 = fmaf(td, b.y - a.y, a.y);
    const float c10 = fmaf(td, b.z - a.z, a.z);
    const float c11 = fmaf(td, b.w - a.w, a.w);
    const float c0 = fmaf(xd, c10 - c00, c00);
    const float c1 = fmaf(xd, c11 - c01, c01);
    f[k] = fmaf(zd, c1 - c0, c0);
  }
  const float dens = f[DFCSR_F_DENSITY], dx = f[DFCSR_F_DENSITY_X], dz = f[DFCSR_F_DENSITY_Z];
  const float vr = f[DFCSR_F_VX], dv = f[DFCSR_F_VX_X];
  const float scale = fmaf((float)xp, c.rho, 1.0f);
  const float g = dz / scale;
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

__device__ void wake_point_mixed(const DfcsrHistory &H, const DfcsrLattice &L,
                                 const DfcsrWakeParams &P, int i, double *red, RegionLds &R,
                                 double &out_z, double &out_x, double &xm_out, double &zm_out) {
  double xm, zm;
  mesh_point(P, i, xm, zm);
  xm_out = xm;
  zm_out = zm;
  const double t = P.t;
  const double s = t + zm, x = xm;
  const HistAxes A = make_axes(H);
  const double inv_ds = 1.0 / L.ds;
  const Observer o = make_observer<true>(H, A, L, s, x, t);
  if (threadIdx.x == 0) {
    make_regions(P, s, x, t, R.rg);
    R.off[0] = 0;
    for (int r = 0; r < R.rg.n; ++r)
      R.off[r + 1] = R.off[r] + ((R.rg.xa[r].n + kChunk - 1) / kChunk) * R.rg.sa[r].n;
  }
  __syncthreads();
  const int nreg = R.rg.n;
  const int total = R.off[nreg];
  double acc_z = 0.0, acc_x = 0.0;
  for (int it = threadIdx.x; it < total; it += blockDim.x) {
    int r = 0;
    while (r + 1 < nreg && it >= R.off[r + 1]) ++r;
    const int rem = it - R.off[r];
    const LinAxis sa = R.rg.sa[r];
    const LinAxis xa = R.rg.xa[r];
    const int c = rem / sa.n, k = rem % sa.n;
    const double sp = sa.at(k);
    const double wz = sa.w(k);
    const ColumnMixed col = make_column_mixed(L, o, sp, inv_ds);
    const int j0 = c * kChunk;
    const int j1 = min(j0 + kChunk, xa.n);
    Neumaier cz, cx;
    for (int j = j0; j < j1; ++j) {
      const double xp = xa.at(j);
      const TruncIdx X = trunc_index((xp - A.min_x) * A.ix, H.Xi);
      if (!X.ok) continue;   // wave-uniform: rows are shared
      float Iz, Ix;
      if (!csr_point_mixed(H, A, t, col, xp, X.i0, (float)X.frac, Iz, Ix)) continue;
      const float wx = (float)xa.w(j);
      cz.add(wx * Iz);
      cx.add(wx * Ix);
    }
    acc_z += wz * (double)cz.value();
    acc_x += wz * (double)cx.value();
  }
  out_z = block_sum(acc_z, red);
  out_x = block_sum(acc_x, red);
}

// Mesh point of block b.  Blocks of one XCD integrate a contiguous range of
// mesh points (row-major: neighbouring z at one x, whose retarded lookups
// overlap), so they share that XCD's L2.  DFCSR_WAKE_ORDER=1 instead runs
// over x first (measured slower on the chicane workload).
__device__ inline int mesh_of_block(const DfcsrWakeParams &P, int b, int nb) {
  const int l = xcd_remap(b, nb);
  if (DFCSR_WAKE_ORDER == 0 || P.observers) return l;
  const int ix = l % P.mesh_xbins, zt = l / P.mesh_xbins;
  int iz = zt;
  if (DFCSR_WAKE_ORDER == 2 && P.mesh_zbins % kNumXcd == 0) {
    // XCD k owns z-lines k, k+8, ...: the same z at all x runs together on
    // one XCD while every XCD gets observers from the whole bunch
    const int q = P.mesh_zbins / kNumXcd;
    iz = (zt % q) * kNumXcd + zt / q;
  }
  return ix * P.mesh_zbins + iz;
}

__global__ __launch_bounds__(kThreads, DFCSR_WAKE_MINBLOCKS) void wake_fp64_kernel(DfcsrHistory H, DfcsrLattice L,
                                                             DfcsrWakeParams P,
                                                             double *__restrict__ xmesh,
                                                             double *__restrict__ zmesh,
                                                             double *__restrict__ dE,
                                                             double *__restrict__ xk) {
  __shared__ double red[kThreads / 64];
  const int i = mesh_of_block(P, blockIdx.x, gridDim.x);
  __shared__ RegionLds R;
  double sz, sx, xm, zm;
  wake_point_fp64(H, L, P, i, red, R, sz, sx, xm, zm);
  if (threadIdx.x == 0) {
    dE[i] = -P.csr_scaling * sz;
    xk[i] = P.csr_scaling * sx;
    if (xmesh) xmesh[i] = xm;
    if (zmesh) zmesh[i] = zm;
  }
}

__global__ __launch_bounds__(kThreads, DFCSR_WAKE_MINBLOCKS) void wake_mixed_kernel(
    DfcsrHistory H, DfcsrLattice L, DfcsrWakeParams P, double *__restrict__ xmesh,
    double *__restrict__ zmesh, double *__restrict__ dE, double *__restrict__ xk) {
  __shared__ double red[kThreads / 64];
  const int i = mesh_of_block(P, blockIdx.x, gridDim.x);
  __shared__ RegionLds R;
  double sz, sx, xm, zm;
  wake_point_mixed(H, L, P, i, red, R, sz, sx, xm, zm);
  if (threadIdx.x == 0) {
    dE[i] = -P.csr_scaling * sz;
    xk[i] = P.csr_scaling * sx;
    if (xmesh) xmesh[i] = xm;
    if (zmesh) zmesh[i] = zm;
  }
}

struct WakeTask {
  DfcsrHistory H;
  DfcsrLattice L;
  DfcsrWakeParams P;
  double *dE, *xk;
  int block0;   // first block of this task in the flattened launch
  int M;
};

template <bool kMixed>
__global__ __launch_bounds__(kThreads, DFCSR_WAKE_MINBLOCKS) void wake_batched_kernel(
    const WakeTask *__restrict__ tasks, int ntask) {
  __shared__ double red[kThreads / 64];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  int lo = 0, hi = ntask - 1;
  while (lo < hi) {   // last task with block0 <= b
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].block0 <= b) lo = mid;
    else hi = mid - 1;
  }
  const WakeTask &T = tasks[lo];
  const int i = b - T.block0;
  __shared__ RegionLds R;
  double sz, sx, xm, zm;
  if (kMixed) wake_point_mixed(T.H, T.L, T.P, i, red, R, sz, sx, xm, zm);
  else wake_point_fp64(T.H, T.L, T.P, i, red, R, sz, sx, xm, zm);
  if (threadIdx.x == 0) {
    T.dE[i] = -T.P.csr_scaling * sz;
    T.xk[i] = T.P.csr_scaling * sx;
  }
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

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_wake_workspace_bytes(int mesh_points) {
  (void)mesh_points;
  return 256;
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
  (void)ws;
  (void)ws_bytes;
  const int M = p->mesh_xbins * p->mesh_zbins;
  if (precision == DFCSR_PREC_FP64) {
    hipLaunchKernelGGL(wake_fp64_kernel, dim3(M), dim3(kThreads), 0, as_stream(stream), *hist,
                       *lat, *p, xmesh, zmesh, dE_dct, x_kick);
  } else if (precision == DFCSR_PREC_MIXED) {
    if (!hist->data32)
      return set_error(DFCSR_EINVAL, "wake_mesh: mixed precision needs the FP32 history");
    hipLaunchKernelGGL(wake_mixed_kernel, dim3(M), dim3(kThreads), 0, as_stream(stream), *hist,
                       *lat, *p, xmesh, zmesh, dE_dct, x_kick);
  } else {
    return set_error(DFCSR_EINVAL, "wake_mesh: unknown precision %d", precision);
  }
  DFCSR_LAUNCH_CHECK("wake_mesh");
  return DFCSR_OK;
}

extern "C" int64_t dfcsr_wake_batched_workspace_bytes(int k) {
  if (k < 1) k = 1;
  return (int64_t)k * (int64_t)sizeof(WakeTask) + 256;
}

extern "C" int dfcsr_wake_tasks_pack(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                                     const DfcsrWakeParams *p, int precision,
                                     double *const *dE_dct, double *const *x_kick, void *tasks,
                                     int64_t tasks_bytes, int *blocks) {
  if (k < 1 || k > 4096) return set_error(DFCSR_ERANGE, "wake_tasks_pack: %d tasks", k);
  if (!hist || !lat || !p || !dE_dct || !x_kick || !tasks || !blocks)
    return set_error(DFCSR_EINVAL, "wake_tasks_pack: NULL pointer");
  if (tasks_bytes < (int64_t)k * (int64_t)sizeof(WakeTask))
    return set_error(DFCSR_ENOWS, "wake_tasks_pack: buffer too small");
  if (precision != DFCSR_PREC_FP64 && precision != DFCSR_PREC_MIXED)
    return set_error(DFCSR_EINVAL, "wake_tasks_pack: unknown precision %d", precision);
  WakeTask *th = static_cast<WakeTask *>(tasks);
  int64_t nb = 0;
  for (int j = 0; j < k; ++j) {
    int rc;
    if ((rc = check_history(&hist[j], "wake_tasks_pack"))) return rc;
    if ((rc = check_lattice(&lat[j], "wake_tasks_pack"))) return rc;
    if ((rc = check_params(&p[j], "wake_tasks_pack"))) return rc;
    if (!dE_dct[j] || !x_kick[j]) return set_error(DFCSR_EINVAL, "wake_tasks_pack: NULL output");
    if (precision == DFCSR_PREC_MIXED && !hist[j].data32)
      return set_error(DFCSR_EINVAL, "wake_tasks_pack: task %d lacks the FP32 history", j);
    th[j].H = hist[j];
    th[j].L = lat[j];
    th[j].P = p[j];
    th[j].dE = dE_dct[j];
    th[j].xk = x_kick[j];
    th[j].block0 = (int)nb;
    th[j].M = p[j].mesh_xbins * p[j].mesh_zbins;
    nb += th[j].M;
    if (nb > ((int64_t)1 << 30)) return set_error(DFCSR_ERANGE, "wake_tasks_pack: too many blocks");
  }
  *blocks = (int)nb;
  return DFCSR_OK;
}

extern "C" int dfcsr_wake_batched_launch(const void *dev_tasks, int k, int blocks, int precision,
                                         void *stream) {
  if (!dev_tasks || k < 1 || blocks < 1)
    return set_error(DFCSR_EINVAL, "wake_batched_launch: bad arguments");
  hipStream_t s = as_stream(stream);
  const WakeTask *dt = static_cast<const WakeTask *>(dev_tasks);
  if (precision == DFCSR_PREC_MIXED)
    hipLaunchKernelGGL(wake_batched_kernel<true>, dim3(blocks), dim3(kThreads), 0, s, dt, k);
  else if (precision == DFCSR_PREC_FP64)
    hipLaunchKernelGGL(wake_batched_kernel<false>, dim3(blocks), dim3(kThreads), 0, s, dt, k);
  else
    return set_error(DFCSR_EINVAL, "wake_batched_launch: unknown precision %d", precision);
  DFCSR_LAUNCH_CHECK("wake_batched_launch");
  return DFCSR_OK;
}

extern "C" int dfcsr_wake_mesh_batched(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                                       const DfcsrWakeParams *p, int precision,
                                       double *const *dE_dct, double *const *x_kick, void *ws,
                                       int64_t ws_bytes, void *stream) {
  if (k < 1) return DFCSR_OK;
  if (!ws) return set_error(DFCSR_EINVAL, "wake_mesh_batched: NULL workspace");
  if (ws_bytes < dfcsr_wake_batched_workspace_bytes(k))
    return set_error(DFCSR_ENOWS, "wake_mesh_batched: workspace too small");
  if (k > 256) return set_error(DFCSR_ERANGE, "wake_mesh_batched: %d tasks > 256", k);
  WakeTask tasks_h[256];
  int blocks = 0;
  int rc = dfcsr_wake_tasks_pack(k, hist, lat, p, precision, dE_dct, x_kick, tasks_h,
                                 (int64_t)sizeof(tasks_h), &blocks);
  if (rc) return rc;
  hipStream_t s = as_stream(stream);
  DFCSR_CHECK_HIP(hipMemcpyAsync(ws, tasks_h, (size_t)k * sizeof(WakeTask), hipMemcpyHostToDevice, s));
  return dfcsr_wake_batched_launch(ws, k, blocks, precision, stream);
}
