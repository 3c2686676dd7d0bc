// a4-a6: density-function post-processing of one deposited slice.
//
// Replaces deposit.py:183-235 as seven launches on one stream: the two
// separable smoothings (density and vx, then the three derivatives) run as
// LDS-tiled stencils, one workgroup per output tile; the gradients and the
// element-wise steps run over the whole grid; the global steps between them
// (the maxima, the trapezoid normalisation, the masked mean of vx_x) read
// partials or run as one fixed-order workgroup, so results are the same
// every launch.  The velocity division and the normalisation are applied
// where their inputs are read, in the reference's operation order.
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

constexpr int kBlock = 1024;   // single-workgroup reductions (fixed order)
constexpr int kT = 256;        // grid-wide element kernels
constexpr int kMaxGrid = 1024;

// Scalars passed between the launches (workspace tail).
enum { S_THR = 0, S_DSUM = 1, S_THR2 = 2, S_FILL = 3, S_COUNT = 8 };

__host__ __device__ inline int grid_of(int64_t N) {
  int64_t g = (N + kT - 1) / kT;
  return (int)(g < 1 ? 1 : (g > kMaxGrid ? kMaxGrid : g));
}

// Savitzky-Golay output at element e of a [nx][nz] array along `axis`
// (0: along x rows, 1: along z); in(e) gives the input at element e.
template <typename In>
__device__ inline double sg_at(const SGOp &op, const In &in, int nx, int nz, int axis, int64_t e) {
#pragma clang fp contract(off)
  const int W = op.W, h = op.h;
  const int n = axis == 0 ? nx : nz;
  const int64_t st = axis == 0 ? nz : 1;
  const int i = axis == 0 ? (int)(e / nz) : (int)(e % nz);   // position along the axis
  const int64_t e0 = e - (int64_t)i * st;                     // element 0 of this line
  double acc = 0.0;
  if (i < h) {
    for (int k = 0; k < W; ++k) acc += op.E[i][k] * in(e0 + k * st);
  } else if (i >= n - h) {
    const int r = n - 1 - i;   // mirror row of the left edge
    for (int k = 0; k < W; ++k) acc += op.E[r][W - 1 - k] * in(e0 + (int64_t)(n - W + k) * st);
  } else {
    for (int k = 0; k < W; ++k) acc += op.c[k] * in(e0 + (int64_t)(i - h + k) * st);
  }
  return acc;
}

// np.gradient spacing test: the coordinate diffs all equal the first one
// (then numpy uses the scalar-spacing formula).  Evaluated on the host.
__host__ __device__ inline bool axis_uniform(double g0, double g1, int n) {
#pragma clang fp contract(off)
  if (n < 2) return true;
  const double d0 = linspace_at(g0, g1, n, 1) - linspace_at(g0, g1, n, 0);
  for (int i = 1; i < n - 1; ++i)
    if (linspace_at(g0, g1, n, i + 1) - linspace_at(g0, g1, n, i) != d0) return false;
  return true;
}

// np.gradient(f, g) along `axis` at element e (edge_order 1),
// numpy/lib/_function_base_impl.py.
template <typename In>
__device__ inline double gradient_at(const In &f, int nx, int nz, int axis, double g0, double g1,
                                     bool uniform, int64_t e) {
#pragma clang fp contract(off)
  const int n = axis == 0 ? nx : nz;
  const int64_t st = axis == 0 ? nz : 1;
  const int i = axis == 0 ? (int)(e / nz) : (int)(e % nz);
  const double dxu = linspace_at(g0, g1, n, 1) - linspace_at(g0, g1, n, 0);
  if (n < 2) return 0.0;
  if (i == 0) return (f(e + st) - f(e)) / dxu;
  if (i == n - 1) {
    const double dn = uniform ? dxu : linspace_at(g0, g1, n, n - 1) - linspace_at(g0, g1, n, n - 2);
    return (f(e) - f(e - st)) / dn;
  }
  if (uniform) return (f(e + st) - f(e - st)) / (2. * dxu);
  const double xm = linspace_at(g0, g1, n, i - 1), x0 = linspace_at(g0, g1, n, i),
               xp = linspace_at(g0, g1, n, i + 1);
  const double dx1 = x0 - xm, dx2 = xp - x0;
  const double a = -(dx2) / (dx1 * (dx1 + dx2));
  const double b = (dx2 - dx1) / (dx1 * dx2);
  const double c = dx1 / (dx2 * (dx1 + dx2));
  return a * f(e - st) + b * f(e) + c * f(e + st);
}

// Block max of a grid-strided pass, one partial per workgroup.
__device__ inline void max_partial(double m, double *part) {
  __shared__ double red[kT / 64];
  m = block_max(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__device__ inline double max_of(const double *part, int n) {
  double m = -INFINITY;
  for (int k = 0; k < n; ++k) m = fmax(m, part[k]);
  return m;
}

// 1. max of the raw density (deposit.py:183)
__global__ __launch_bounds__(kT) void df_max_kernel(const double *__restrict__ dens, int64_t N,
                                                     double *__restrict__ part) {
  double m = -INFINITY;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < N; e += (int64_t)gridDim.x * kT)
    m = fmax(m, dens[e]);
  max_partial(m, part);
}

// Separable Savitzky-Golay smoothing of a tile: axis 0 (along x) into LDS,
// then axis 1 (along z) from LDS.  A workgroup owns kTileX x kTileZ outputs;
// its LDS rows hold the axis-0 results of its rows over every column the
// axis-1 windows of its outputs read (interior window, or the first / last
// `W` columns for the edge rows of the fit).  Each output is the same sum in
// the same order as two grid-wide passes would give.
constexpr int kTileX = 8, kTileZ = 32;
constexpr int kTileW = kTileZ + 2 * DFCSR_MAX_SG_WINDOW;
constexpr int kTileFields = 3;

// Input field of a tile pass: a grid, optionally divided by the density
// where that exceeds thr (the velocity division of deposit.py:183-184).
struct TileIn {
  const double *a, *dens;
  double thr;
  __device__ double operator()(int64_t e) const {
    if (!dens) return a[e];
    const double d = dens[e];
    return d > thr ? a[e] / d : a[e];
  }
};

// One LDS row of axis-0 results: element e of grid row i is at row[e - off].
struct LdsRow {
  const double *row;
  int64_t off;
  __device__ double operator()(int64_t e) const { return row[e - off]; }
};

template <int NF>
__device__ void sg_tile(const SGOp &op, const TileIn (&in)[NF], double *const (&out)[NF], int nx,
                        int nz) {
  __shared__ double T[kTileFields][kTileX][kTileW];
  static_assert(NF <= kTileFields, "tile fields");
  const int x0 = blockIdx.y * kTileX, z0 = blockIdx.x * kTileZ;
  const int h = op.h, W = op.W;
  const int lo = max(0, min(z0 - h, nz - W));
  const int hi = min(nz, max(z0 + kTileZ + h, W));
  const int w = hi - lo;   // <= kTileZ + 2 W
  const int rows = min(kTileX, nx - x0), cols = min(kTileZ, nz - z0);
  for (int t = threadIdx.x; t < rows * w; t += blockDim.x) {
    const int r = t / w, c = t - r * w;
    const int64_t e = (int64_t)(x0 + r) * nz + (lo + c);
#pragma unroll
    for (int f = 0; f < NF; ++f) T[f][r][c] = sg_at(op, in[f], nx, nz, 0, e);
  }
  __syncthreads();
  for (int t = threadIdx.x; t < rows * cols; t += blockDim.x) {
    const int r = t / cols, c = t - r * cols;
    const int64_t i = x0 + r;
    const int64_t e = i * nz + (z0 + c);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const LdsRow row{&T[f][r][0], i * nz + lo};
      out[f][e] = sg_at(op, row, nx, nz, 1, e);
    }
  }
}

// 2. the velocity division and both smoothing axes of density and vx
// (deposit.py:183-184, 195-199), into S0 (density) and S1 (vx)
__global__ __launch_bounds__(kT) void df_sg_kernel(SGOp op, const double *__restrict__ dens,
                                                   const double *__restrict__ vx, int nx, int nz,
                                                   double vt, const double *__restrict__ part,
                                                   int npart, double *__restrict__ S0,
                                                   double *__restrict__ S1,
                                                   double *__restrict__ scal) {
  const double thr = max_of(part, npart) / vt;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) scal[S_THR] = thr;
  const TileIn in[2] = {{dens, nullptr, 0.0}, {vx, dens, thr}};
  double *const out[2] = {S0, S1};
  sg_tile<2>(op, in, out, nx, nz);
}

// 4. dsum = trapz(trapz(density, x_grids, axis=0), z_grids) (deposit.py:201):
// the inner trapz adds the rows of each column in order (numpy reduces axis
// 0 row by row), the outer one is a fixed-order block sum.  One workgroup.
__global__ __launch_bounds__(kBlock) void df_dsum_kernel(const double *__restrict__ dens, int nx,
                                                         int nz, double x0, double x1, double z0,
                                                         double z1, double *__restrict__ colsum,
                                                         double *__restrict__ scal) {
#pragma clang fp contract(off)
  __shared__ double red[kBlock / 64];
  const int tid = threadIdx.x;
  for (int j = tid; j < nz; j += kBlock) {
    double col = 0.0;
#pragma unroll 8
    for (int i = 0; i + 1 < nx; ++i) {
      const double d = linspace_at(x0, x1, nx, i + 1) - linspace_at(x0, x1, nx, i);
      const double t = d * (dens[(int64_t)i * nz + j] + dens[(int64_t)(i + 1) * nz + j]) / 2.0;
      col = (i == 0) ? t : col + t;
    }
    colsum[j] = col;
  }
  __syncthreads();
  double part = 0.0;
  for (int j = tid; j + 1 < nz; j += kBlock) {
    const double d = linspace_at(z0, z1, nz, j + 1) - linspace_at(z0, z1, nz, j);
    part += d * (colsum[j] + colsum[j + 1]) / 2.0;
  }
  const double dsum = block_sum(part, red);
  if (tid == 0) scal[S_DSUM] = dsum;
}

// Normalised density and its thresholded vx (deposit.py:202, 204), as read
// by the gradients before they are stored.
struct Normed {
  const double *dens;
  double dsum;
  __device__ double operator()(int64_t e) const { return dens[e] / dsum; }
};
struct VxCut {
  const double *vx, *dens;
  double dsum, thr;
  __device__ double operator()(int64_t e) const { return (dens[e] / dsum <= thr) ? 0.0 : vx[e]; }
};

// 5. gradients of the normalised density and of vx (deposit.py:212-213);
// the density and vx outputs take their normalised / thresholded values
// (deposit.py:202, 204), with max partials of the normalised density
__global__ __launch_bounds__(kT) void df_grad_kernel(const double *__restrict__ S0,
                                                     const double *__restrict__ S1, int nx, int nz,
                                                     double x0, double x1, double z0, double z1,
                                                     int ux, int uz,
                                                     const double *__restrict__ scal,
                                                     double *__restrict__ G0,
                                                     double *__restrict__ G1,
                                                     double *__restrict__ G2,
                                                     double *__restrict__ dens,
                                                     double *__restrict__ vx,
                                                     double *__restrict__ part) {
  const int64_t N = (int64_t)nx * nz;
  const double dsum = scal[S_DSUM], thr = scal[S_THR];
  const Normed d{S0, dsum};
  const VxCut v{S1, S0, dsum, thr};
  double m = -INFINITY;
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < N; e += (int64_t)gridDim.x * kT) {
    G0[e] = gradient_at(d, nx, nz, 0, x0, x1, ux, e);
    G1[e] = gradient_at(d, nx, nz, 1, z0, z1, uz, e);
    G2[e] = gradient_at(v, nx, nz, 0, x0, x1, ux, e);
    const double dn = d(e);
    dens[e] = dn;
    vx[e] = v(e);
    m = fmax(m, dn);
  }
  max_partial(m, part);
}

// 6. both smoothing axes of the three derivatives (deposit.py:215-224)
__global__ __launch_bounds__(kT) void df_sg_grad_kernel(SGOp op, const double *__restrict__ G0,
                                                        const double *__restrict__ G1,
                                                        const double *__restrict__ G2, int nx,
                                                        int nz, double *__restrict__ o0,
                                                        double *__restrict__ o1,
                                                        double *__restrict__ o2) {
  const TileIn in[3] = {{G0, nullptr, 0.0}, {G1, nullptr, 0.0}, {G2, nullptr, 0.0}};
  double *const out[3] = {o0, o1, o2};
  sg_tile<3>(op, in, out, nx, nz);
}

// 7. vx_x below 8*max/vt takes the mean of vx_x above it (deposit.py:233-235):
// fixed-order sums in one workgroup.
__global__ __launch_bounds__(kBlock) void df_fill_kernel(const double *__restrict__ dens,
                                                         const double *__restrict__ G2, int64_t N,
                                                         double vt, const double *__restrict__ part,
                                                         int npart, double *__restrict__ scal,
                                                         double *__restrict__ info) {
  __shared__ double red[kBlock / 64];
  const double dmax = max_of(part, npart);
  const double thr2 = dmax / vt * 8;
  double s = 0.0, c = 0.0;
  for (int64_t e = threadIdx.x; e < N; e += kBlock)
    if (dens[e] > thr2) {
      s += G2[e];
      c += 1.0;
    }
  const double ssum = block_sum(s, red);
  const double cnt = block_sum(c, red);
  const double fill = cnt > 0.0 ? ssum / cnt : NAN;   // np.mean of an empty selection
  if (threadIdx.x == 0) {
    scal[S_THR2] = thr2;
    scal[S_FILL] = fill;
    if (info) {
      info[1] = scal[S_DSUM];
      info[2] = dmax;
      info[3] = fill;
    }
  }
}

// 8. outputs
__global__ __launch_bounds__(kT) void df_out_kernel(const double *__restrict__ dens,
                                                    const double *__restrict__ G0,
                                                    const double *__restrict__ G1,
                                                    const double *__restrict__ G2, int64_t N,
                                                    const double *__restrict__ scal,
                                                    double *__restrict__ dxo,
                                                    double *__restrict__ dzo,
                                                    double *__restrict__ vxxo) {
  const double thr2 = scal[S_THR2], fill = scal[S_FILL];
  for (int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x; e < N; e += (int64_t)gridDim.x * kT) {
    dxo[e] = G0[e];
    dzo[e] = G1[e];
    vxxo[e] = (dens[e] < thr2) ? fill : G2[e];
  }
}

__global__ void df_info_kernel(const double *__restrict__ part, int npart, double *__restrict__ info) {
  if (threadIdx.x == 0) info[0] = max_of(part, npart);
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_df_workspace_bytes(int nbins_x, int nbins_z) {
  if (nbins_x < 1 || nbins_z < 1) return 0;
  const int64_t N = (int64_t)nbins_x * nbins_z;
  // three grid temporaries, column sums, two sets of max partials, scalars
  return (3 * N + nbins_z + 2 * (int64_t)kMaxGrid + S_COUNT) * 8 + 64;
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
  const int nx = nbins_x, nz = nbins_z;
  const int64_t N = (int64_t)nx * nz;
  double *T0 = static_cast<double *>(ws);
  double *T1 = T0 + N, *T2 = T1 + N, *col = T2 + N;
  double *partA = col + nz, *partB = partA + kMaxGrid, *scal = partB + kMaxGrid;
  const int g = grid_of(N);
  const dim3 gt((nz + kTileZ - 1) / kTileZ, (nx + kTileX - 1) / kTileX);
  hipStream_t s = as_stream(stream);
  // T0/T1: density and vx smoothed; density_x/z, vx_x: the gradients; then
  // density/vx: normalised and thresholded; T0,T1,T2: the gradients smoothed.
  hipLaunchKernelGGL(df_max_kernel, dim3(g), dim3(kT), 0, s, density, N, partA);
  hipLaunchKernelGGL(df_sg_kernel, gt, dim3(kT), 0, s, op, density, vx, nx, nz,
                     velocity_threshold, partA, g, T0, T1, scal);
  hipLaunchKernelGGL(df_dsum_kernel, dim3(1), dim3(kBlock), 0, s, T0, nx, nz, x_start, x_end,
                     z_start, z_end, col, scal);
  const int ux = axis_uniform(x_start, x_end, nx), uz = axis_uniform(z_start, z_end, nz);
  hipLaunchKernelGGL(df_grad_kernel, dim3(g), dim3(kT), 0, s, T0, T1, nx, nz, x_start, x_end,
                     z_start, z_end, ux, uz, scal, density_x, density_z, vx_x, density, vx, partB);
  hipLaunchKernelGGL(df_sg_grad_kernel, gt, dim3(kT), 0, s, op, density_x, density_z, vx_x, nx,
                     nz, T0, T1, T2);
  hipLaunchKernelGGL(df_fill_kernel, dim3(1), dim3(kBlock), 0, s, density, T2, N,
                     velocity_threshold, partB, g, scal, info);
  hipLaunchKernelGGL(df_out_kernel, dim3(g), dim3(kT), 0, s, density, T0, T1, T2, N, scal,
                     density_x, density_z, vx_x);
  if (info) hipLaunchKernelGGL(df_info_kernel, dim3(1), dim3(64), 0, s, partA, g, info);
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
