/* CPU oracle -- TEST INFRASTRUCTURE ONLY (timed CPU baseline and a second
 * restatement of the integral).  Never linked into the product.
 *
 * Float64 C restatement of pyDFCSR's retarded integral, paths relative to
 * the reference's pyDFCSR_2D/:
 *   interpolate1D            interp1D.py:13-36
 *   interpolate3D            interp3D.py:18-66
 *   get_CSR_integrand        CSR.py:605-782 (same operation order)
 *   get_CSR_wake regions     CSR.py:454-553, nested trapz CSR.py:556-597
 *   calculate_2D_CSR loop    CSR.py:397-418, split over OpenMP threads the
 *                            way calculate_2D_CSR_parallel splits mesh
 *                            points over MPI ranks (CSR.py:420-451).
 * Built with -ffp-contract=off so every a*b+c rounds twice, as in numpy.
 * The history is read through strides, so both the reference's stacked
 * (Nt, Xi, Zi) arrays and a ring of interleaved records can be passed.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  const double *base[5];   /* density, density_x, density_z, vx, vx_x */
  int64_t st, sx, sz;      /* element strides of slot, x, z */
  int32_t nt, Xi, Zi, cap, head;
  double min_t, min_x, min_z, dt, dx, dz;
} RefHist;

typedef struct {
  const double *cx, *cy, *nx, *ny, *tx, *ty;   /* lattice tables, ns each */
  int32_t ns, nelem;
  double min_s, ds;
  const double *distance, *rho;
} RefLat;

typedef struct {
  double t, sigma_x, sigma_z, slope0, mean_x, formation_length, n_formation_length, csr_scaling;
  int32_t Nx, Nz;
} RefWake;

static void lin(double a, double b, int n, double *y) {
  if (n == 1) {
    y[0] = a;
    return;
  }
  double step = (b - a) / (double)(n - 1);
  for (int i = 0; i < n; ++i) y[i] = (double)i * step + a;
  y[n - 1] = b;
}

static double interp1(const double *v, int n, double min, double d, double x) {
  double u = (x - min) / d;
  if (!(u > -1.0 && u < (double)n)) return 0.0;
  int i0 = (int)u;
  int i1 = (i0 == n - 1) ? i0 : i0 + 1;
  double xd = u - i0;
  return v[i0] * (1 - xd) + v[i1] * xd;
}

static int slot_of(const RefHist *H, int i) {
  int s = H->head + i;
  return s >= H->cap ? s - H->cap : s;
}

static void interp3(const RefHist *H, double tq, double xq, double zq, double out[5]) {
  double ut = (tq - H->min_t) / H->dt;
  double ux = (xq - H->min_x) / H->dx;
  double uz = (zq - H->min_z) / H->dz;
  if (!(ut > -1.0 && ut < (double)H->nt && ux > -1.0 && ux < (double)H->Xi && uz > -1.0 &&
        uz < (double)H->Zi)) {
    for (int f = 0; f < 5; ++f) out[f] = 0.0;
    return;
  }
  int t0 = (int)ut, x0 = (int)ux, z0 = (int)uz;
  int t1 = t0 == H->nt - 1 ? t0 : t0 + 1;
  int x1 = x0 == H->Xi - 1 ? x0 : x0 + 1;
  int z1 = z0 == H->Zi - 1 ? z0 : z0 + 1;
  double xd = ut - t0, yd = ux - x0, zd = uz - z0;
  int64_t a0 = (int64_t)slot_of(H, t0) * H->st, a1 = (int64_t)slot_of(H, t1) * H->st;
  int64_t b0 = (int64_t)x0 * H->sx, b1 = (int64_t)x1 * H->sx;
  int64_t c0 = (int64_t)z0 * H->sz, c1 = (int64_t)z1 * H->sz;
  for (int f = 0; f < 5; ++f) {
    const double *d = H->base[f];
    double c00 = d[a0 + b0 + c0] * (1 - xd) + d[a1 + b0 + c0] * xd;
    double c01 = d[a0 + b0 + c1] * (1 - xd) + d[a1 + b0 + c1] * xd;
    double c10 = d[a0 + b1 + c0] * (1 - xd) + d[a1 + b1 + c0] * xd;
    double c11 = d[a0 + b1 + c1] * (1 - xd) + d[a1 + b1 + c1] * xd;
    double e0 = c00 * (1 - yd) + c10 * yd;
    double e1 = c01 * (1 - yd) + c11 * yd;
    out[f] = e0 * (1 - zd) + e1 * zd;
  }
}

/* integrand grids Iz, Ix [nxp][nsp] of one observer: CSR.py:605-782 */
static void integrand(const RefHist *H, const RefLat *L, double s, double x, double t,
                      const double *xp, int nxp, const double *sp, int nsp, double *Iz,
                      double *Ix) {
  double v5[5];
  interp3(H, t, x, s - t, v5);
  const double vx = v5[3];
  const double X0s = interp1(L->cx, L->ns, L->min_s, L->ds, s);
  const double Y0s = interp1(L->cy, L->ns, L->min_s, L->ds, s);
  const double nsx = interp1(L->nx, L->ns, L->min_s, L->ds, s);
  const double nsy = interp1(L->ny, L->ns, L->min_s, L->ds, s);
  const double tsx = interp1(L->tx, L->ns, L->min_s, L->ds, s);
  const double tsy = interp1(L->ty, L->ns, L->min_s, L->ds, s);
  for (int k = 0; k < nsp; ++k) {
    const double spk = sp[k];
    const double X0p = interp1(L->cx, L->ns, L->min_s, L->ds, spk);
    const double Y0p = interp1(L->cy, L->ns, L->min_s, L->ds, spk);
    const double npx = interp1(L->nx, L->ns, L->min_s, L->ds, spk);
    const double npy = interp1(L->ny, L->ns, L->min_s, L->ds, spk);
    const double tpx = interp1(L->tx, L->ns, L->min_s, L->ds, spk);
    const double tpy = interp1(L->ty, L->ns, L->min_s, L->ds, spk);
    double rho = 0.0;
    for (int e = 0; e < L->nelem; ++e) {
      int in = e == 0 ? (spk < L->distance[0])
                      : (spk < L->distance[e] && spk >= L->distance[e - 1]);
      if (in) rho = L->rho[e];
    }
    for (int i = 0; i < nxp; ++i) {
      const double xpi = xp[i];
      const double rx = X0s - X0p + x * nsx - xpi * npx;
      const double ry = Y0s - Y0p + x * nsy - xpi * npy;
      const double R = sqrt(rx * rx + ry * ry);
      const double t_ret = t - R;
      double f[5];
      interp3(H, t_ret, xpi, spk - t_ret, f);
      const double dens = f[0], dx = f[1], dz = f[2], vr = f[3], dv = f[4];
      const double scale = 1 + xpi * rho;
      const double vel_x = tsx + vx * nsx, vel_y = tsy + vx * nsy;
      const double vrx = tpx + vr * npx, vry = tpy + vr * npy;
      const double nbx = dx * npx + dz / scale * tpx;
      const double nby = dx * npy + dz / scale * tpy;
      const double part1 = vel_x * vrx + vel_y * vry;
      const double num1 = scale * ((vel_x - part1 * vrx) * nbx + (vel_y - part1 * vry) * nby);
      const double num2 = -scale * part1 * dens * dv;
      const double nmx = nsx - npx, nmy = nsy - npy;
      const double p1 = rx * nmx + ry * nmy;
      const double p2 = nsx * tpx + nsy * tpy;
      const double pd = -(vrx * nbx + vry * nby) - dens * dv;
      const double W1 = scale * p1 / (R * R * R) * dens;
      const double W2 = scale * p1 / (R * R) * pd;
      const double W3 = -scale * p2 / R * pd;
      Iz[(int64_t)i * nsp + k] = num1 / R + (num2) / R;
      Ix[(int64_t)i * nsp + k] = W1 + W2 + W3;
    }
  }
}

/* trapz(trapz(I, xp, axis=0), sp): inner sum row by row, outer in order */
static double trapz2(const double *I, const double *xp, int nxp, const double *sp, int nsp,
                     double *col) {
  for (int k = 0; k < nsp; ++k) col[k] = 0.0;
  for (int i = 0; i + 1 < nxp; ++i) {
    double d = xp[i + 1] - xp[i];
    for (int k = 0; k < nsp; ++k) {
      double v = d * (I[(int64_t)(i + 1) * nsp + k] + I[(int64_t)i * nsp + k]) / 2.0;
      col[k] = (i == 0) ? v : col[k] + v;
    }
  }
  double acc = 0.0;
  for (int k = 0; k + 1 < nsp; ++k) acc += (sp[k + 1] - sp[k]) * (col[k + 1] + col[k]) / 2.0;
  return acc;
}

static double npmax0(double v) { return v != v ? v : (v > 0 ? v : 0.0); }

/* get_CSR_wake for observer (s, x): CSR.py:454-602 */
static void wake_point(const RefHist *H, const RefLat *L, const RefWake *P, double s, double x,
                       double *dE, double *xk, double *buf) {
  const double t = P->t, sz = P->sigma_z, sx = P->sigma_x, th = P->slope0;
  const double x0 = (s - t) * P->slope0;
  const int Nx = P->Nx, Nz = P->Nz;
  double xa[4][2], sa[4][2];
  int nx[4], nreg;
  const double nfl = P->n_formation_length * P->formation_length;
  if (fabs(th) <= 1) {
    double s2 = s - 500 * sz, s3 = s - 20 * sz, s4 = s + 5 * sz, s1 = npmax0(s2 - nfl);
    nreg = 3;
    xa[0][0] = x0 - 20 * sx; xa[0][1] = x0 + 20 * sx; nx[0] = 2 * Nx; sa[0][0] = s1; sa[0][1] = s2;
    xa[1][0] = x0 - 10 * sx; xa[1][1] = x0 + 10 * sx; nx[1] = Nx; sa[1][0] = s2; sa[1][1] = s3;
    xa[2][0] = x0 - 10 * sx; xa[2][1] = x0 + 10 * sx; nx[2] = Nx; sa[2][0] = s3; sa[2][1] = s4;
  } else {
    double d, x1l, x1r, x2l, x2r;
    if (th > 0) {
      double ta = -2 * th / (1 - th * th);
      d = (10 * sx + P->mean_x - x) / ta;
      x1l = x + 0.1 * sx; x1r = x + 10 * sx; x2l = x - 3 * sx; x2r = x1l;
    } else {
      double ta = 2 * th / (1 - th * th);
      d = -(P->mean_x - x - 10 * sx) / ta;
      x1l = x - 10 * sx; x1r = x - 1 * sx; x2l = x1r; x2r = x + 3 * sx;
    }
    double s4 = s + 3 * sz, s3 = npmax0(s - d), s2 = s3 - 200 * sz, s1 = npmax0(s2 - nfl);
    nreg = 4;
    xa[0][0] = x0 - 20 * sx; xa[0][1] = x0 + 20 * sx; nx[0] = 2 * Nx; sa[0][0] = s1; sa[0][1] = s2;
    xa[1][0] = x0 - 5 * sx; xa[1][1] = x0 + 5 * sx; nx[1] = Nx; sa[1][0] = s2; sa[1][1] = s3;
    xa[2][0] = x1l; xa[2][1] = x1r; nx[2] = Nx; sa[2][0] = s3; sa[2][1] = s4;
    xa[3][0] = x2l; xa[3][1] = x2r; nx[3] = Nx; sa[3][0] = s3; sa[3][1] = s4;
  }
  double *xp = buf, *sp = xp + 2 * Nx, *Iz = sp + Nz, *Ix = Iz + (int64_t)2 * Nx * Nz,
         *col = Ix + (int64_t)2 * Nx * Nz;
  double e = 0.0, k = 0.0;
  for (int r = 0; r < nreg; ++r) {
    lin(xa[r][0], xa[r][1], nx[r], xp);
    lin(sa[r][0], sa[r][1], Nz, sp);
    integrand(H, L, s, x, t, xp, nx[r], sp, Nz, Iz, Ix);
    e = e + (-P->csr_scaling * trapz2(Iz, xp, nx[r], sp, Nz, col));
    k = k + P->csr_scaling * trapz2(Ix, xp, nx[r], sp, Nz, col);
  }
  *dE = e;
  *xk = k;
}

/* calculate_2D_CSR over M observers obs[i] = (s - t, x).  Returns 0. */
int csr_ref_wake(const RefHist *H, const RefLat *L, const RefWake *P, const double *obs, int M,
                 double *dE, double *xk, int nthreads) {
  const int64_t per = (int64_t)2 * P->Nx + P->Nz + 4 * (int64_t)P->Nx * P->Nz + P->Nz + 8;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *buf = (double *)malloc(sizeof(double) * per);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int i = 0; i < M; ++i) {
      double s = P->t + obs[2 * i];
      wake_point(H, L, P, s, obs[2 * i + 1], &dE[i], &xk[i], buf);
    }
    free(buf);
  }
  return 0;
}

int csr_ref_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
