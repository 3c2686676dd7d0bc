// Parity entry points for the lookup primitives and the integrand:
// interpolate3D (interp3D.py:18-66), interpolate1D (interp1D.py:13-36) and
// get_CSR_integrand on an explicit (x', s') grid (CSR.py:605-782).  Built
// without fused multiply-add and with true divisions, so each value follows
// the reference's IEEE operation sequence.
#include "csr_point.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void interp3d_kernel(DfcsrHistory H, int field,
                                                            const double *__restrict__ q,
                                                            int64_t n, double *__restrict__ out) {
  HistAxes A = make_axes(H);
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kThreads) {
    double v[5];
    hist_lookup5<false>(H, A, q[3 * p], q[3 * p + 1], q[3 * p + 2], v);
    out[p] = v[field];
  }
}

__global__ __launch_bounds__(kThreads) void interp1d_kernel(DfcsrLattice L, int col,
                                                            const double *__restrict__ s,
                                                            int64_t n, double *__restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kThreads) {
    Lat6 r = lattice_at<false>(L, s[p], 0.0);
    const double v[6] = {r.X0, r.Y0, r.nx, r.ny, r.tx, r.ty};
    out[p] = v[col];
  }
}

__global__ __launch_bounds__(kThreads) void integrand_kernel(DfcsrHistory H, DfcsrLattice L,
                                                             double s, double x, double t,
                                                             const double *__restrict__ xp,
                                                             int nxp, const double *__restrict__ sp,
                                                             int nsp, double *__restrict__ Iz,
                                                             double *__restrict__ Ix) {
  HistAxes A = make_axes(H);
  Observer o = make_observer<false>(H, A, L, s, x, t);
  int64_t n = (int64_t)nxp * nsp;
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kThreads) {
    int i = (int)(p / nsp), k = (int)(p % nsp);
    Column c = make_column<false>(L, o, sp[k], 0.0);
    double a, b;
    csr_point<false>(H, A, o, c, xp[i], a, b);
    Iz[p] = a;
    Ix[p] = b;
  }
}

int grid_for(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

}  // namespace

int check_history(const DfcsrHistory *H, const char *who) {
  if (!H) return set_error(DFCSR_EINVAL, "%s: NULL history", who);
  if (!H->data && !H->data32 && !H->data32t)
    return set_error(DFCSR_EINVAL, "%s: history without a ring", who);
  if (H->nt < 2)
    return set_error(DFCSR_EINVAL,
                     "%s: history holds %d slice(s); the time axis needs two "
                     "(deposit.py:419 divides by len - 1)",
                     who, H->nt);
  if (H->cap < H->nt || H->head < 0 || H->head >= H->cap || H->Xi < 2 || H->Zi < 2)
    return set_error(DFCSR_EINVAL, "%s: bad history geometry (cap %d head %d nt %d Xi %d Zi %d)",
                     who, H->cap, H->head, H->nt, H->Xi, H->Zi);
  return DFCSR_OK;
}

int check_lattice(const DfcsrLattice *L, const char *who) {
  if (!L || !L->table || !L->distance || !L->rho)
    return set_error(DFCSR_EINVAL, "%s: NULL lattice", who);
  if (L->ns < 2 || L->nelem < 1 || L->nelem > DFCSR_MAX_ELEMENTS)
    return set_error(DFCSR_EINVAL, "%s: bad lattice (ns %d, nelem %d)", who, L->ns, L->nelem);
  return DFCSR_OK;
}

}  // namespace dfcsr

using namespace dfcsr;

extern "C" int dfcsr_interpolate3d(const DfcsrHistory *hist, int field, const double *q,
                                   int64_t n, double *out, void *stream) {
  if (!q || !out) return set_error(DFCSR_EINVAL, "interpolate3d: NULL pointer");
  if (field < 0 || field > 4) return set_error(DFCSR_EINVAL, "interpolate3d: field %d", field);
  if (!hist || !hist->data || hist->cap < hist->nt || hist->nt < 1 || hist->head < 0 ||
      hist->head >= hist->cap || hist->Xi < 1 || hist->Zi < 1)
    return set_error(DFCSR_EINVAL, "interpolate3d: bad history");
  if (n <= 0) return DFCSR_OK;
  hipLaunchKernelGGL(interp3d_kernel, dim3(grid_for(n)), dim3(kThreads), 0, as_stream(stream),
                     *hist, field, q, n, out);
  DFCSR_LAUNCH_CHECK("interpolate3d");
  return DFCSR_OK;
}

extern "C" int dfcsr_interpolate1d(const DfcsrLattice *lat, int col, const double *s, int64_t n,
                                   double *out, void *stream) {
  if (!lat || !lat->table || !s || !out) return set_error(DFCSR_EINVAL, "interpolate1d: NULL");
  if (col < 0 || col > 5) return set_error(DFCSR_EINVAL, "interpolate1d: column %d", col);
  if (lat->ns < 1) return set_error(DFCSR_EINVAL, "interpolate1d: empty table");
  if (n <= 0) return DFCSR_OK;
  hipLaunchKernelGGL(interp1d_kernel, dim3(grid_for(n)), dim3(kThreads), 0, as_stream(stream),
                     *lat, col, s, n, out);
  DFCSR_LAUNCH_CHECK("interpolate1d");
  return DFCSR_OK;
}

extern "C" int dfcsr_csr_integrand(const DfcsrHistory *hist, const DfcsrLattice *lat, double s,
                                   double x, double t, const double *xp, int nxp, const double *sp,
                                   int nsp, double *Iz, double *Ix, void *stream) {
  int rc = check_history(hist, "csr_integrand");
  if (rc) return rc;
  if (!hist->data) return set_error(DFCSR_EINVAL, "csr_integrand: needs the FP64 ring");
  if ((rc = check_lattice(lat, "csr_integrand"))) return rc;
  if (!xp || !sp || !Iz || !Ix) return set_error(DFCSR_EINVAL, "csr_integrand: NULL pointer");
  if (nxp < 1 || nsp < 1) return set_error(DFCSR_EINVAL, "csr_integrand: empty grid");
  int64_t n = (int64_t)nxp * nsp;
  hipLaunchKernelGGL(integrand_kernel, dim3(grid_for(n)), dim3(kThreads), 0, as_stream(stream),
                     *hist, *lat, s, x, t, xp, nxp, sp, nsp, Iz, Ix);
  DFCSR_LAUNCH_CHECK("csr_integrand");
  return DFCSR_OK;
}
