// 2D Savitzky-Golay filter of pyDFCSR's SGolay_filter.sgolay2d
// (SGolay_filter.py:3-81): a least-squares polynomial fit over a square
// window, applied as one convolution of the border-padded grid.
//
// Host: the convolution kernels, rows of the pseudo-inverse of the window's
// monomial matrix (window coordinates scaled to [-1, 1] so the normal
// equations stay well conditioned, solved in long double, then rescaled).
// Device: one launch per call.  Each workgroup stages its padded tile
// (32 + 2h) x (64 + 2h) in LDS, evaluating the reference's mirrored-band
// padding on the fly at the border, and writes one or two outputs (the
// 'both' derivative shares the tile).
#include <string.h>

#include <vector>

#include "common.h"

namespace dfcsr {
namespace {

constexpr int kTileW = 64, kTileH = 32;   // outputs per tile: columns (one wave), rows
constexpr int kThreads = 256;
constexpr int kRows = kTileH * kTileW / kThreads;   // rows per thread
constexpr int kMaxTaps = DFCSR_SG2D_MAX_WINDOW * DFCSR_SG2D_MAX_WINDOW;
static_assert(kTileW == 64 && kRows * kThreads == kTileH * kTileW, "tile shape");

struct SG2DKernels {
  int w, h, nout;
  double k[2][kMaxTaps];   // k[o][a * w + b]: the convolution kernel of output o
};

// Exponent pairs (x power, y power) of the fit, in the reference's order:
// degree k = 0..order, and within it x power k - n for n = 0..k
// (SGolay_filter.py:24).
int fit_terms(int order, int ex[], int ey[]) {
  int n = 0;
  for (int k = 0; k <= order; ++k)
    for (int j = 0; j <= k; ++j) {
      ex[n] = k - j;
      ey[n] = j;
      ++n;
    }
  return n;
}

// Rows `rows` of pinv(A), A[p][t] = dx_p^ex_t * dy_p^ey_t over the window
// points p = a * w + b (dx = a - h, dy = b - h; SGolay_filter.py:27-34).
// A has full column rank for order < w, so pinv(A) = (A^T A)^-1 A^T.
int pinv_rows(int w, int order, int nrows, const int rows[], double out[][kMaxTaps]) {
  constexpr int kMaxTerms = DFCSR_SG2D_MAX_WINDOW * (DFCSR_SG2D_MAX_WINDOW + 1) / 2;
  int ex[kMaxTerms], ey[kMaxTerms];
  const int nt = fit_terms(order, ex, ey);
  const int h = w / 2, P = w * w;
  // scaled monomials: u = dx / h, v = dy / h; column t of A = column t of
  // the scaled matrix times h^(ex + ey)
  std::vector<long double> Sv((size_t)P * nt), Gv((size_t)nt * nt), Xv((size_t)nt * P);
  auto S = [&](int p, int t) -> long double & { return Sv[(size_t)p * nt + t]; };
  auto G = [&](int r, int c) -> long double & { return Gv[(size_t)r * nt + c]; };
  auto X = [&](int r, int p) -> long double & { return Xv[(size_t)r * P + p]; };
  for (int a = 0; a < w; ++a)
    for (int b = 0; b < w; ++b) {
      const long double u = (long double)(a - h) / h, v = (long double)(b - h) / h;
      for (int t = 0; t < nt; ++t) {
        long double m = 1.0L;
        for (int e = 0; e < ex[t]; ++e) m *= u;
        for (int e = 0; e < ey[t]; ++e) m *= v;
        S(a * w + b, t) = m;
      }
    }
  for (int r = 0; r < nt; ++r) {
    for (int c = 0; c < nt; ++c) {
      long double s = 0.0L;
      for (int p = 0; p < P; ++p) s += S(p, r) * S(p, c);
      G(r, c) = s;
    }
    for (int p = 0; p < P; ++p) X(r, p) = S(p, r);
  }
  // Gauss-Jordan with partial pivoting: X <- G^-1 S^T
  for (int col = 0; col < nt; ++col) {
    int piv = col;
    for (int r = col + 1; r < nt; ++r)
      if (fabsl(G(r, col)) > fabsl(G(piv, col))) piv = r;
    if (G(piv, col) == 0.0L) return set_error(DFCSR_EINVAL, "sgolay2d: singular fit");
    if (piv != col) {
      for (int k = 0; k < nt; ++k) std::swap(G(col, k), G(piv, k));
      for (int k = 0; k < P; ++k) std::swap(X(col, k), X(piv, k));
    }
    const long double d = G(col, col);
    for (int k = 0; k < nt; ++k) G(col, k) /= d;
    for (int k = 0; k < P; ++k) X(col, k) /= d;
    for (int r = 0; r < nt; ++r) {
      if (r == col) continue;
      const long double f = G(r, col);
      if (f == 0.0L) continue;
      for (int k = 0; k < nt; ++k) G(r, k) -= f * G(col, k);
      for (int k = 0; k < P; ++k) X(r, k) -= f * X(col, k);
    }
  }
  // pinv(A) = D pinv(S), D = diag(h^-(ex + ey))
  for (int i = 0; i < nrows; ++i) {
    const int t = rows[i];
    long double scale = 1.0L;
    for (int e = 0; e < ex[t] + ey[t]; ++e) scale /= h;
    for (int p = 0; p < P; ++p) out[i][p] = (double)(X(t, p) * scale);
  }
  return DFCSR_OK;
}

// The kernels the reference convolves with (SGolay_filter.py:69-81):
// smoothing pinv row 0; 'col' -(row 1), the x (first-axis) slope; 'row'
// -(row 2), the y slope; 'both' the two.
int make_kernels(int w, int order, int derivative, SG2DKernels *K) {
  if (w < 3 || (w % 2) == 0)
    return set_error(DFCSR_EINVAL, "sgolay2d: window_size %d must be odd and >= 3", w);
  if (w > DFCSR_SG2D_MAX_WINDOW)
    return set_error(DFCSR_ERANGE, "sgolay2d: window_size %d exceeds %d", w,
                     DFCSR_SG2D_MAX_WINDOW);
  if (order < 0 || order >= w)
    return set_error(DFCSR_EINVAL, "sgolay2d: order %d must be >= 0 and < window_size %d",
                     order, w);
  if (derivative < DFCSR_SG2D_NONE || derivative > DFCSR_SG2D_BOTH)
    return set_error(DFCSR_EINVAL, "sgolay2d: derivative %d is not one of 0..3", derivative);
  if (derivative != DFCSR_SG2D_NONE && order < 1)
    return set_error(DFCSR_EINVAL, "sgolay2d: a derivative needs order >= 1");
  int rows[2], nrows;
  if (derivative == DFCSR_SG2D_NONE) {
    rows[0] = 0;
    nrows = 1;
  } else if (derivative == DFCSR_SG2D_COL) {
    rows[0] = 1;
    nrows = 1;
  } else if (derivative == DFCSR_SG2D_ROW) {
    rows[0] = 2;
    nrows = 1;
  } else {
    rows[0] = 1;
    rows[1] = 2;
    nrows = 2;
  }
  memset(K, 0, sizeof(*K));
  K->w = w;
  K->h = w / 2;
  K->nout = nrows;
  int rc = pinv_rows(w, order, nrows, rows, K->k);
  if (rc) return rc;
  if (derivative != DFCSR_SG2D_NONE)
    for (int i = 0; i < nrows; ++i)
      for (int p = 0; p < w * w; ++p) K->k[i][p] = -K->k[i][p];
  return DFCSR_OK;
}

__device__ inline double at(const double *__restrict__ z, int nz, int r, int c) {
  return z[(int64_t)r * nz + c];
}

// Right band (SGolay_filter.py:49-50): column nz + j, j < h, of real row r.
__device__ inline double right_band(const double *__restrict__ z, int nz, int r, int j) {
  const double e = at(z, nz, r, nz - 1);
  return e + fabs(at(z, nz, r, nz - 2 - j) - e);
}

// Bottom band (SGolay_filter.py:43-44): row nx + i, i < h, of real column c.
__device__ inline double bottom_band(const double *__restrict__ z, int nx, int nz, int i, int c) {
  const double e = at(z, nz, nx - 1, c);
  return e + fabs(at(z, nz, nx - 2 - i, c) - e);
}

// The padded grid Z at real coordinates (r, c), -h <= r < nx + h,
// -h <= c < nz + h (SGolay_filter.py:36-66): the grid itself inside; the
// four bands mirror the first h rows/columns about the edge value, below it
// at the top/left edges and above it at the bottom/right; the top-left and
// bottom-right corners mirror about the corner value; the top-right and
// bottom-left corners mirror the right and bottom bands about their first
// row and column.
__device__ double padded(const double *__restrict__ z, int nx, int nz, int r, int c) {
  if (r >= 0 && r < nx) {
    if (c < 0) {
      const double e = at(z, nz, r, 0);
      return e - fabs(at(z, nz, r, -c) - e);
    }
    if (c >= nz) return right_band(z, nz, r, c - nz);
    return at(z, nz, r, c);
  }
  if (r < 0) {
    if (c < 0) {
      const double e = at(z, nz, 0, 0);
      return e - fabs(at(z, nz, -r, -c) - e);
    }
    if (c >= nz) {
      const double e = right_band(z, nz, 0, c - nz);
      return e - fabs(right_band(z, nz, -r, c - nz) - e);
    }
    const double e = at(z, nz, 0, c);
    return e - fabs(at(z, nz, -r, c) - e);
  }
  const int i = r - nx;
  if (c >= nz) {
    const double e = at(z, nz, nx - 1, nz - 1);
    return e + fabs(at(z, nz, nx - 2 - i, nz - 2 - (c - nz)) - e);
  }
  if (c < 0) {
    const double e = bottom_band(z, nx, nz, i, 0);
    return e - fabs(bottom_band(z, nx, nz, i, -c) - e);
  }
  return bottom_band(z, nx, nz, i, c);
}

// One 64 x 32 output tile per workgroup of 256 threads: lane = column,
// each wave 8 consecutive rows per thread.  out[i][j] = sum over a, b of
// k[a][b] * Z[i + w - 1 - a][j + w - 1 - b] (scipy.signal.fftconvolve,
// mode 'valid'), accumulated with fused multiply-adds in a fixed order.
// Per tap column a thread reads its 8 + w - 1 tile values once and applies
// them to its 8 rows.  Tiles are dealt to the XCDs in contiguous bands
// (block l runs on XCD l % 8) so neighbouring tiles share an L2 for their
// halos.
template <int W, int NOUT>
__global__ __launch_bounds__(kThreads) void sgolay2d_kernel(SG2DKernels K,
                                                           const double *__restrict__ z, int nx,
                                                           int nz, int tiles_x, int ntiles,
                                                           double *__restrict__ out0,
                                                           double *__restrict__ out1) {
  constexpr int H = W / 2, TWP = kTileW + 2 * H, THP = kTileH + 2 * H;
  __shared__ double tile[THP * TWP];
  const int per = (ntiles + 7) / 8;
  const int t = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (t >= ntiles) return;   // the whole workgroup: no barrier is pending
  const int ty = t / tiles_x, tx = t - ty * tiles_x;
  const int i0 = ty * kTileH, j0 = tx * kTileW;
  const bool inner = i0 >= H && j0 >= H && i0 + kTileH + H <= nx && j0 + kTileW + H <= nz;
  for (int e = threadIdx.x; e < THP * TWP; e += kThreads) {
    const int a = e / TWP, b = e - a * TWP;
    const int r = i0 - H + a, c = j0 - H + b;
    double v;
    if (inner)
      v = at(z, nz, r, c);
    else
      v = (r < nx + H && c < nz + H) ? padded(z, nx, nz, r, c) : 0.0;
    tile[e] = v;
  }
  __syncthreads();
  const int tj = threadIdx.x & (kTileW - 1), ti = threadIdx.x / kTileW;
  const int j = j0 + tj, ib = i0 + ti * kRows;
  if (j >= nz || ib >= nx) return;
  double acc[NOUT][kRows];
#pragma unroll
  for (int o = 0; o < NOUT; ++o)
#pragma unroll
    for (int r = 0; r < kRows; ++r) acc[o][r] = 0.0;
  // the tap-column loop stays rolled: unrolled, the compiler keeps every
  // column's values live (up to 256 registers; window 7 'both' 3x slower)
#pragma unroll 1
  for (int b = 0; b < W; ++b) {
    double v[kRows + W - 1];
    const double *col = tile + (ti * kRows) * TWP + tj + W - 1 - b;
#pragma unroll
    for (int k = 0; k < kRows + W - 1; ++k) v[k] = col[k * TWP];
#pragma unroll
    for (int a = 0; a < W; ++a)
#pragma unroll
      for (int o = 0; o < NOUT; ++o) {
        const double kk = K.k[o][a * W + b];
#pragma unroll
        for (int r = 0; r < kRows; ++r) acc[o][r] = fma(kk, v[r + W - 1 - a], acc[o][r]);
      }
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    if (ib + r >= nx) break;
    const int64_t idx = (int64_t)(ib + r) * nz + j;
    out0[idx] = acc[0][r];
    if (NOUT == 2) out1[idx] = acc[NOUT - 1][r];
  }
}

template <int W>
void launch_window(const SG2DKernels &K, const double *z, int nx, int nz, double *out0,
                   double *out1, hipStream_t s) {
  const int tiles_x = (nz + kTileW - 1) / kTileW, tiles_y = (nx + kTileH - 1) / kTileH;
  const int ntiles = tiles_x * tiles_y;
  const dim3 grid(8 * ((ntiles + 7) / 8));
  if (K.nout == 2)
    hipLaunchKernelGGL((sgolay2d_kernel<W, 2>), grid, dim3(kThreads), 0, s, K, z, nx, nz,
                       tiles_x, ntiles, out0, out1);
  else
    hipLaunchKernelGGL((sgolay2d_kernel<W, 1>), grid, dim3(kThreads), 0, s, K, z, nx, nz,
                       tiles_x, ntiles, out0, out1);
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int dfcsr_sgolay2d_kernels(int window_size, int order, int derivative, double *k0,
                                      double *k1) {
  // Host-side helper (no device work).
  SG2DKernels K;
  int rc = make_kernels(window_size, order, derivative, &K);
  if (rc) return rc;
  if (!k0 || (K.nout == 2 && !k1))
    return set_error(DFCSR_EINVAL, "sgolay2d_kernels: NULL output");
  const int P = window_size * window_size;
  for (int p = 0; p < P; ++p) k0[p] = K.k[0][p];
  if (K.nout == 2)
    for (int p = 0; p < P; ++p) k1[p] = K.k[1][p];
  return DFCSR_OK;
}

extern "C" int dfcsr_sgolay2d(const double *z, int nx, int nz, int window_size, int order,
                              int derivative, double *out, double *out2, void *stream) {
  SG2DKernels K;
  int rc = make_kernels(window_size, order, derivative, &K);
  if (rc) return rc;
  if (!z || !out || (derivative == DFCSR_SG2D_BOTH && !out2))
    return set_error(DFCSR_EINVAL, "sgolay2d: NULL array");
  if (nx < K.h + 1 || nz < K.h + 1)
    return set_error(DFCSR_EINVAL, "sgolay2d: a %d x %d grid is smaller than half the window + 1 (%d)",
                     nx, nz, K.h + 1);
  if ((int64_t)((nx + kTileH - 1) / kTileH) * ((nz + kTileW - 1) / kTileW) > (1 << 30))
    return set_error(DFCSR_ERANGE, "sgolay2d: a %d x %d grid exceeds the launch grid", nx, nz);
  double *o2 = derivative == DFCSR_SG2D_BOTH ? out2 : nullptr;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (window_size) {
    case 3: launch_window<3>(K, z, nx, nz, out, o2, s); break;
    case 5: launch_window<5>(K, z, nx, nz, out, o2, s); break;
    case 7: launch_window<7>(K, z, nx, nz, out, o2, s); break;
    case 9: launch_window<9>(K, z, nx, nz, out, o2, s); break;
    case 11: launch_window<11>(K, z, nx, nz, out, o2, s); break;
    default: launch_window<13>(K, z, nx, nz, out, o2, s); break;
  }
  DFCSR_LAUNCH_CHECK("sgolay2d");
  return DFCSR_OK;
}
