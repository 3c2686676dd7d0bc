// a4-a6: density-function post-processing of one deposited slice.
//
// Replaces deposit.py:183-235.  The grids are small (<= 512^2) and the steps
// depend on each other through global reductions (max, trapezoid sum, masked
// mean), so the whole sequence runs as ONE workgroup of 1024 threads with
// barriers between phases; the arrays stay in L2.
//
// Savitzky-Golay: scipy.signal.savgol_filter(mode='interp') is linear: the
// interior is the symmetric smoothing stencil c[] (the centre row of the
// least-squares hat matrix H = V (V^T V)^-1 V^T of the window) and the
// first/last window//2 outputs are rows of the same H applied to the first/
// last `window` samples (scipy fits and evaluates a polynomial there,
// scipy/signal/_savitzky_golay.py _fit_edge).  H is computed on the host in
// extended precision and passed by value.
#include <string.h>

#include "common.h"

namespace dfcsr {

struct SGOp {
  int W, h;
  double c[DFCSR_MAX_SG_WINDOW];
  double E[DFCSR_MAX_SG_WINDOW / 2][DFCSR_MAX_SG_WINDOW];   // left-edge rows of H
};

// Hat matrix rows of a degree-`order` least-squares fit over `W` points,
// by Gauss-Jordan on the normal equations of the centred monomial basis
// (long double; the matrices are integer valued).
static int make_sg(int W, int order, SGOp *op) {
  if (W < 1 || (W % 2) == 0 || W > DFCSR_MAX_SG_WINDOW)
    return set_error(DFCSR_EINVAL, "savgol: window %d must be odd and <= %d", W,
                     DFCSR_MAX_SG_WINDOW);
  if (order < 0 || order >= W)
    return set_error(DFCSR_EINVAL, "savgol: polyorder %d must be < window %d", order, W);
  const int m = order + 1, h = W / 2;
  long double V[DFCSR_MAX_SG_WINDOW][DFCSR_MAX_SG_WINDOW];
  for (int j = 0; j < W; ++j) {
    long double t = (long double)(j - h), p = 1.0L;
    for (int k = 0; k < m; ++k) {
      V[j][k] = p;
      p *= t;
    }
  }
  // A = V^T V (m x m), B = V^T (m x W); solve A X = B in place.
  long double A[DFCSR_MAX_SG_WINDOW][DFCSR_MAX_SG_WINDOW];
  long double B[DFCSR_MAX_SG_WINDOW][DFCSR_MAX_SG_WINDOW];
  for (int a = 0; a < m; ++a) {
    for (int b = 0; b < m; ++b) {
      long double s = 0.0L;
      for (int j = 0; j < W; ++j) s += V[j][a] * V[j][b];
      A[a][b] = s;
    }
    for (int j = 0; j < W; ++j) B[a][j] = V[j][a];
  }
  for (int col = 0; col < m; ++col) {
    int piv = col;
    for (int r = col + 1; r < m; ++r)
      if (fabsl(A[r][col]) > fabsl(A[piv][col])) piv = r;
    if (A[piv][col] == 0.0L) return set_error(DFCSR_EINVAL, "savgol: singular fit");
    if (piv != col) {
      for (int k = 0; k < m; ++k) { long double t = A[col][k]; A[col][k] = A[piv][k]; A[piv][k] = t; }
      for (int k = 0; k < W; ++k) { long double t = B[col][k]; B[col][k] = B[piv][k]; B[piv][k] = t; }
    }
    long double d = A[col][col];
    for (int k = 0; k < m; ++k) A[col][k] /= d;
    for (int k = 0; k < W; ++k) B[col][k] /= d;
    for (int r = 0; r < m; ++r) {
      if (r == col) continue;
      long double f = A[r][col];
      if (f == 0.0L) continue;
      for (int k = 0; k < m; ++k) A[r][k] -= f * A[col][k];
      for (int k = 0; k < W; ++k) B[r][k] -= f * B[col][k];
    }
  }
  memset(op, 0, sizeof(*op));
  op->W = W;
  op->h = h;
  // H[i][j] = sum_k V[i][k] X[k][j]
  for (int i = 0; i <= h; ++i)
    for (int j = 0; j < W; ++j) {
      long double s = 0.0L;
      for (int k = 0; k < m; ++k) s += V[i][k] * B[k][j];
      if (i == h) op->c[j] = (double)s;
      else op->E[i][j] = (double)s;
    }
  return DFCSR_OK;
}

namespace {

constexpr int kThreads = 1024;

// One Savitzky-Golay pass along `axis` (0: along x rows, 1: along z) of a
// [nx][nz] array.
__device__ void sg_pass(const SGOp &op, const double *__restrict__ in, double *__restrict__ out,
                        int nx, int nz, int axis) {
  const int W = op.W, h = op.h;
  const int n = axis == 0 ? nx : nz;
  const int64_t N = (int64_t)nx * nz;
  const int64_t stride_ax = axis == 0 ? nz : 1;
  for (int64_t e = threadIdx.x; e < N; e += blockDim.x) {
    int i = axis == 0 ? (int)(e / nz) : (int)(e % nz);   // position along the axis
    const double *line = in + (e - (int64_t)i * stride_ax);   // element 0 of this line
    double acc = 0.0;
    if (i < h) {
      for (int k = 0; k < W; ++k) acc += op.E[i][k] * line[k * stride_ax];
    } else if (i >= n - h) {
      int r = n - 1 - i;   // mirror row of the left edge
      for (int k = 0; k < W; ++k) acc += op.E[r][W - 1 - k] * line[(int64_t)(n - W + k) * stride_ax];
    } else {
      for (int k = 0; k < W; ++k) acc += op.c[k] * line[(int64_t)(i - h + k) * stride_ax];
    }
    out[e] = acc;
  }
}

__device__ inline bool axis_uniform(double g0, double g1, int n) {
  // np.gradient collapses to the scalar-spacing formula when every diff of
  // the coordinate array equals the first one.
  if (n < 2) return true;
  double d0 = linspace_at(g0, g1, n, 1) - linspace_at(g0, g1, n, 0);
  int ok = 1;
  for (int i = threadIdx.x; i < n - 1; i += blockDim.x)
    if (linspace_at(g0, g1, n, i + 1) - linspace_at(g0, g1, n, i) != d0) ok = 0;
  return __syncthreads_and(ok);
}

// np.gradient(f, g) along `axis` (edge_order 1), numpy/lib/_function_base_impl.py.
__device__ void gradient_pass(const double *__restrict__ f, double *__restrict__ out, int nx,
                              int nz, int axis, double g0, double g1, bool uniform) {
#pragma clang fp contract(off)
  const int n = axis == 0 ? nx : nz;
  const int64_t N = (int64_t)nx * nz;
  const int64_t st = axis == 0 ? nz : 1;
  const double dxu = linspace_at(g0, g1, n, 1) - linspace_at(g0, g1, n, 0);
  for (int64_t e = threadIdx.x; e < N; e += blockDim.x) {
    int i = axis == 0 ? (int)(e / nz) : (int)(e % nz);
    double r;
    if (n < 2) {
      r = 0.0;
    } else if (i == 0) {
      r = (f[e + st] - f[e]) / dxu;
    } else if (i == n - 1) {
      double dn = uniform ? dxu : linspace_at(g0, g1, n, n - 1) - linspace_at(g0, g1, n, n - 2);
      r = (f[e] - f[e - st]) / dn;
    } else if (uniform) {
      r = (f[e + st] - f[e - st]) / (2. * dxu);
    } else {
      double xm = linspace_at(g0, g1, n, i - 1), x0 = linspace_at(g0, g1, n, i),
             xp = linspace_at(g0, g1, n, i + 1);
      double dx1 = x0 - xm, dx2 = xp - x0;
      double a = -(dx2) / (dx1 * (dx1 + dx2));
      double b = (dx2 - dx1) / (dx1 * dx2);
      double c = dx1 / (dx2 * (dx1 + dx2));
      r = a * f[e - st] + b * f[e] + c * f[e + st];
    }
    out[e] = r;
  }
}

__global__ __launch_bounds__(kThreads) void df_kernel(
    SGOp op, double *__restrict__ dens, double *__restrict__ vx, double *__restrict__ dxo,
    double *__restrict__ dzo, double *__restrict__ vxxo, int nx, int nz, double x0, double x1,
    double z0, double z1, double vt, double *__restrict__ T0, double *__restrict__ T1,
    double *__restrict__ T2, double *__restrict__ info) {
#pragma clang fp contract(off)
  __shared__ double red[kThreads / 64];
  const int64_t N = (int64_t)nx * nz;
  const int tid = threadIdx.x;

  // deposit.py:183-184 velocity = sum(px)/count above max/vt
  double m = -INFINITY;
  for (int64_t e = tid; e < N; e += kThreads) m = fmax(m, dens[e]);
  const double dmax_raw = block_max(m, red);
  const double thr = dmax_raw / vt;
  for (int64_t e = tid; e < N; e += kThreads)
    if (dens[e] > thr) vx[e] = vx[e] / dens[e];
  __syncthreads();

  // deposit.py:195-199 savgol axis 0 then axis 1
  sg_pass(op, dens, T0, nx, nz, 0);
  sg_pass(op, vx, T1, nx, nz, 0);
  __syncthreads();
  sg_pass(op, T0, dens, nx, nz, 1);
  sg_pass(op, T1, vx, nx, nz, 1);
  __syncthreads();

  // deposit.py:201 dsum = trapz(trapz(density, x_grids, axis=0), z_grids):
  // the inner trapz adds the rows in order (numpy reduces axis 0 row by row).
  for (int j = tid; j < nz; j += kThreads) {
    double col = 0.0;
    for (int i = 0; i + 1 < nx; ++i) {
      double d = linspace_at(x0, x1, nx, i + 1) - linspace_at(x0, x1, nx, i);
      double t = d * (dens[(int64_t)i * nz + j] + dens[(int64_t)(i + 1) * nz + j]) / 2.0;
      col = (i == 0) ? t : col + t;
    }
    T2[j] = col;
  }
  __syncthreads();
  double part = 0.0;
  for (int j = tid; j + 1 < nz; j += kThreads) {
    double d = linspace_at(z0, z1, nz, j + 1) - linspace_at(z0, z1, nz, j);
    part += d * (T2[j] + T2[j + 1]) / 2.0;
  }
  const double dsum = block_sum(part, red);
  for (int64_t e = tid; e < N; e += kThreads) {
    double d = dens[e] / dsum;
    dens[e] = d;
    if (d <= thr) vx[e] = 0.0;   // raw-count threshold on the normalised density (deposit.py:204)
  }
  __syncthreads();

  // deposit.py:212-213 gradients
  const bool ux = axis_uniform(x0, x1, nx);
  const bool uz = axis_uniform(z0, z1, nz);
  gradient_pass(dens, T0, nx, nz, 0, x0, x1, ux);
  gradient_pass(dens, T1, nx, nz, 1, z0, z1, uz);
  gradient_pass(vx, T2, nx, nz, 0, x0, x1, ux);
  __syncthreads();

  // deposit.py:215-224 smoothing of the derivatives
  sg_pass(op, T0, dxo, nx, nz, 0);
  sg_pass(op, T1, dzo, nx, nz, 0);
  sg_pass(op, T2, vxxo, nx, nz, 0);
  __syncthreads();
  sg_pass(op, dxo, T0, nx, nz, 1);
  sg_pass(op, dzo, T1, nx, nz, 1);
  sg_pass(op, vxxo, T2, nx, nz, 1);
  __syncthreads();

  // deposit.py:233-235 vx_x below 8*max/vt takes the mean of vx_x above it
  m = -INFINITY;
  for (int64_t e = tid; e < N; e += kThreads) m = fmax(m, dens[e]);
  const double dmax = block_max(m, red);
  const double thr2 = dmax / vt * 8;
  double s = 0.0, c = 0.0;
  for (int64_t e = tid; e < N; e += kThreads)
    if (dens[e] > thr2) {
      s += T2[e];
      c += 1.0;
    }
  const double ssum = block_sum(s, red);
  const double cnt = block_sum(c, red);
  const double fill = cnt > 0.0 ? ssum / cnt : NAN;   // np.mean of an empty selection
  for (int64_t e = tid; e < N; e += kThreads) {
    dxo[e] = T0[e];
    dzo[e] = T1[e];
    vxxo[e] = (dens[e] < thr2) ? fill : T2[e];
  }
  if (tid == 0 && info) {
    info[0] = dmax_raw;
    info[1] = dsum;
    info[2] = dmax;
    info[3] = fill;
  }
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_df_workspace_bytes(int nbins_x, int nbins_z) {
  if (nbins_x < 1 || nbins_z < 1) return 0;
  int64_t N = (int64_t)nbins_x * nbins_z;
  return 3 * N * 8 + 64;
}

extern "C" int dfcsr_df_process(double *density, double *vx, double *density_x,
                                double *density_z, double *vx_x, int nbins_x, double x_start,
                                double x_end, int nbins_z, double z_start, double z_end,
                                double velocity_threshold, int window, int order, double *info,
                                void *ws, int64_t ws_bytes, void *stream) {
  if (!density || !vx || !density_x || !density_z || !vx_x || !ws)
    return set_error(DFCSR_EINVAL, "df_process: NULL pointer");
  if (nbins_x < 1 || nbins_z < 1) return set_error(DFCSR_EINVAL, "df_process: bad grid");
  if ((int64_t)nbins_x * nbins_z > (int64_t)1 << 22)
    return set_error(DFCSR_ERANGE, "df_process: grid %d x %d above 4M cells", nbins_x, nbins_z);
  if (window > nbins_x || window > nbins_z)
    return set_error(DFCSR_EINVAL,
                     "df_process: If mode is 'interp', window_length must be less than or "
                     "equal to the size of x (window %d, grid %d x %d)",
                     window, nbins_x, nbins_z);
  if (ws_bytes < dfcsr_df_workspace_bytes(nbins_x, nbins_z))
    return set_error(DFCSR_ENOWS, "df_process: workspace too small");
  SGOp op;
  int rc = make_sg(window, order, &op);
  if (rc) return rc;
  int64_t N = (int64_t)nbins_x * nbins_z;
  double *T0 = static_cast<double *>(ws);
  double *T1 = T0 + N, *T2 = T1 + N;
  hipLaunchKernelGGL(df_kernel, dim3(1), dim3(kThreads), 0, as_stream(stream), op, density, vx,
                     density_x, density_z, vx_x, nbins_x, nbins_z, x_start, x_end, z_start,
                     z_end, velocity_threshold, T0, T1, T2, info);
  DFCSR_LAUNCH_CHECK("df_process");
  return DFCSR_OK;
}

extern "C" int dfcsr_savgol_operator(int window, int order, double *c, double *edge) {
  // Host-side helper (no device work): interior stencil c[window] and the
  // left-edge rows edge[(window/2) * window] of the smoothing operator.
  SGOp op;
  int rc = make_sg(window, order, &op);
  if (rc) return rc;
  for (int j = 0; j < window; ++j) c[j] = op.c[j];
  for (int i = 0; i < window / 2; ++i)
    for (int j = 0; j < window; ++j) edge[i * window + j] = op.E[i][j];
  return DFCSR_OK;
}
