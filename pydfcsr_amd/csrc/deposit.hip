// a3: cloud-in-cell deposition with two-word fixed-point atomics.
//
// Replaces histogram_cic_2d (deposit.py:42-87) and the pair of calls in
// DF_tracker.get_DF (deposit.py:172-182).  Cell indices use the reference's
// FP64 expression cell = (q - start) * (1 / ((end - start) / nbins)),
// i = floor(cell), S = 1 - (cell - i), evaluated without fused multiply-add
// so they are bit-identical to the reference.  Each corner weight w*S1*S2 is
// scaled by 2^k and split into a signed high word (multiples of 2^32) and an
// unsigned low word (< 2^32, rounded); both are summed with 64-bit integer
// atomics.  Integer addition is associative, so the grid is the same for
// every launch order, and with k chosen from n * max|w| (n*max|w|*2^k < 2^94)
// no word can overflow while the quantum 2^-k is ~1e-22 of the largest
// weight: the sums are exact to far below float64 rounding.
#include "common.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 2048;

// ws layout: [0] scale for grid A, [1] scale for grid B (doubles),
//            [8 .. 8+kMaxBlocks) block max|w| partials,
//            then per grid (A, B): nb1*nb2 int64 high words and nb1*nb2
//            uint64 low words.
constexpr int kHdr = 8 + kMaxBlocks;

int nblocks_for(int64_t n) {
  int64_t b = (n + kThreads * 4 - 1) / (kThreads * 4);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

__global__ __launch_bounds__(kThreads) void absmax_partial(const double *__restrict__ w, int64_t n,
                                                           double *__restrict__ part) {
  __shared__ double red[kThreads / 64];
  double m = 0.0;
  int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    m = fmax(m, fabs(w[i]));
  m = block_max(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// Picks 2^k with n*max|w|*2^k < 2^94: every scaled weight s has |s| < 2^94/n,
// so n high words (|s|/2^32 each) stay below 2^62 and n low words (< 2^32
// each, n < 2^31) below 2^63.  scale[0]: unit weights, scale[1]: w.
__device__ double pick_scale(double bound) {
  if (!(bound > 0.0) || !isfinite(bound)) return ldexp(1.0, 60);
  int e;
  frexp(bound, &e);   // bound < 2^e
  int k = 94 - e;
  if (k > 900) k = 900;
  return ldexp(1.0, k);
}

__global__ void scale_final(const double *__restrict__ part, int nblk, int64_t n, int has_w,
                            double *__restrict__ scale) {
  __shared__ double red[kThreads / 64];
  double m = 0.0;
  if (has_w)
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) m = fmax(m, part[b]);
  m = block_max(m, red);
  if (threadIdx.x == 0) {
    scale[0] = pick_scale((double)n);
    scale[1] = pick_scale((double)n * m);
  }
}

__global__ __launch_bounds__(kThreads) void zero_i64(int64_t *__restrict__ a, int64_t n,
                                                     int32_t *__restrict__ c, int64_t nc) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0;
  if (c)
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += stride) c[i] = 0;
}

// acc[0..ncell) high words, acc[ncell..2 ncell) low words
__device__ inline void fx_add(int64_t *acc, int64_t cell, int64_t ncell, double v, double scale) {
  const double s = v * scale;                   // exact: scale is a power of two
  const double h = floor(s * 0x1p-32);
  const double l = s - h * 0x1p32;              // in [0, 2^32]
  const long long hi = (long long)h;
  const unsigned long long lo = (unsigned long long)rint(l);
  if (hi != 0) atomicAdd(reinterpret_cast<unsigned long long *>(acc + cell), (unsigned long long)hi);
  if (lo != 0) atomicAdd(reinterpret_cast<unsigned long long *>(acc + ncell + cell), lo);
}

// One pass over the particles.  Grid A: weights wA (unit weights when wA is
// null) at scale[sA].  Grid B (only when wB is non-null): weights wB at
// scale[1].
__global__ __launch_bounds__(kThreads) void cic_kernel(
    const double *__restrict__ q1, const double *__restrict__ q2,
    const double *__restrict__ wA, const double *__restrict__ wB, int64_t n, int nb1,
    double start1, double inv1, int nb2, double start2, double inv2,
    const double *__restrict__ scale, int sA, int64_t *__restrict__ accA,
    int64_t *__restrict__ accB, int32_t *__restrict__ cell_index,
    int32_t *__restrict__ cell_count) {
#pragma clang fp contract(off)
  const double scA = scale[sA], scB = scale[1];
  int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < n; p += stride) {
    double c1 = (q1[p] - start1) * inv1;
    double c2 = (q2[p] - start2) * inv2;
    double f1 = floor(c1), f2 = floor(c2);
    double S1 = 1.0 - (c1 - f1);
    double S2 = 1.0 - (c2 - f2);
    // Indices far off the grid only need to fail the range tests below.
    int i1 = (f1 > -4.0 && f1 < (double)nb1 + 4.0) ? (int)f1 : -8;
    int i2 = (f2 > -4.0 && f2 < (double)nb2 + 4.0) ? (int)f2 : -8;
    if (cell_index) {
      // floor() of the exact cell coordinate, saturated to int32
      double g1 = fmin(fmax(f1, -2147483648.0), 2147483647.0);
      double g2 = fmin(fmax(f2, -2147483648.0), 2147483647.0);
      cell_index[2 * p] = isnan(f1) ? INT32_MIN : (int32_t)g1;
      cell_index[2 * p + 1] = isnan(f2) ? INT32_MIN : (int32_t)g2;
    }
    bool a0 = (i1 >= 0) && (i1 < nb1), a1 = (i1 + 1 >= 0) && (i1 + 1 < nb1);
    bool b0 = (i2 >= 0) && (i2 < nb2), b1 = (i2 + 1 >= 0) && (i2 + 1 < nb2);
    if (cell_count && a0 && b0) atomicAdd(cell_count + (int64_t)i1 * nb2 + i2, 1);
    double wa = wA ? wA[p] : 1.0;
    // Corner weights in the reference's association: (w*S1)*S2.
    double t00 = wa * S1 * S2, t01 = wa * S1 * (1.0 - S2);
    double t10 = wa * (1.0 - S1) * S2, t11 = wa * (1.0 - S1) * (1.0 - S2);
    const int64_t nc = (int64_t)nb1 * nb2;
    const int64_t base = (int64_t)i1 * nb2 + i2;
    if (a0 && b0) fx_add(accA, base, nc, t00, scA);
    if (a0 && b1) fx_add(accA, base + 1, nc, t01, scA);
    if (a1 && b0) fx_add(accA, base + nb2, nc, t10, scA);
    if (a1 && b1) fx_add(accA, base + nb2 + 1, nc, t11, scA);
    if (wB) {
      double wb = wB[p];
      double u00 = wb * S1 * S2, u01 = wb * S1 * (1.0 - S2);
      double u10 = wb * (1.0 - S1) * S2, u11 = wb * (1.0 - S1) * (1.0 - S2);
      if (a0 && b0) fx_add(accB, base, nc, u00, scB);
      if (a0 && b1) fx_add(accB, base + 1, nc, u01, scB);
      if (a1 && b0) fx_add(accB, base + nb2, nc, u10, scB);
      if (a1 && b1) fx_add(accB, base + nb2 + 1, nc, u11, scB);
    }
  }
}

// total = hi * 2^32 + lo, normalised so the low part is < 2^32, then one
// rounding to float64 and an exact power-of-two rescale.
// LDS-privatised form of cic_kernel for grids whose row tiles fit in LDS.
// Workgroup (chunk c, tile y) takes particles [c*per, (c+1)*per) and the
// grid rows [y*rows, y*rows + rows): it accumulates the same two-word
// fixed-point words in LDS with 64-bit LDS atomics and adds its non-zero
// cells to the global words once.  Integer sums: the grid is identical to
// the direct form's, for every launch order.
constexpr int kTileLdsBytes = 48 * 1024;

__device__ inline void lds_fx_add(unsigned long long *hi, unsigned long long *lo, int cell,
                                  double v, double scale) {
  const double s = v * scale;
  const double h = floor(s * 0x1p-32);
  const double l = s - h * 0x1p32;
  const long long hv = (long long)h;
  const unsigned long long lv = (unsigned long long)rint(l);
  if (hv != 0) atomicAdd(hi + cell, (unsigned long long)hv);
  if (lv != 0) atomicAdd(lo + cell, lv);
}

__global__ __launch_bounds__(kThreads) void cic_tile_kernel(
    const double *__restrict__ q1, const double *__restrict__ q2,
    const double *__restrict__ wA, const double *__restrict__ wB, int64_t n, int64_t per,
    int nb1, double start1, double inv1, int nb2, double start2, double inv2, int rows,
    const double *__restrict__ scale, int sA, int64_t *__restrict__ accA,
    int64_t *__restrict__ accB, int32_t *__restrict__ cell_index,
    int32_t *__restrict__ cell_count) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned long long lds[];
  const int r0 = blockIdx.y * rows;
  const int r1 = min(r0 + rows, nb1);
  const int tcells = (r1 - r0) * nb2;
  const int ngrid = wB ? 2 : 1;
  unsigned long long *hA = lds, *lA = lds + tcells;
  unsigned long long *hB = lds + 2 * tcells, *lB = lds + 3 * tcells;
  int *cnt = reinterpret_cast<int *>(lds + 2 * ngrid * tcells);
  for (int e = threadIdx.x; e < 2 * ngrid * tcells; e += kThreads) lds[e] = 0ull;
  if (cell_count)
    for (int e = threadIdx.x; e < tcells; e += kThreads) cnt[e] = 0;
  __syncthreads();
  const double scA = scale[sA], scB = scale[1];
  const int64_t p0 = (int64_t)blockIdx.x * per;
  const int64_t p1 = min(p0 + per, n);
  for (int64_t p = p0 + threadIdx.x; p < p1; p += kThreads) {
    double c1 = (q1[p] - start1) * inv1;
    double f1 = floor(c1);
    // rows touched: f1 and f1 + 1; skip particles that miss this tile
    const bool owner_low = (f1 >= (double)r0 && f1 < (double)r1) ||
                           (blockIdx.y == 0 && !(f1 >= 0.0)) ||
                           (r1 == nb1 && f1 >= (double)nb1);
    const bool touches = f1 >= (double)(r0 - 1) && f1 < (double)r1;
    if (!touches && !(owner_low && cell_index)) continue;
    double c2 = (q2[p] - start2) * inv2;
    double f2 = floor(c2);
    double S1 = 1.0 - (c1 - f1);
    double S2 = 1.0 - (c2 - f2);
    int i1 = (f1 > -4.0 && f1 < (double)nb1 + 4.0) ? (int)f1 : -8;
    int i2 = (f2 > -4.0 && f2 < (double)nb2 + 4.0) ? (int)f2 : -8;
    if (cell_index && owner_low) {
      double g1 = fmin(fmax(f1, -2147483648.0), 2147483647.0);
      double g2 = fmin(fmax(f2, -2147483648.0), 2147483647.0);
      cell_index[2 * p] = isnan(f1) ? INT32_MIN : (int32_t)g1;
      cell_index[2 * p + 1] = isnan(f2) ? INT32_MIN : (int32_t)g2;
    }
    if (!touches) continue;
    const bool a0 = (i1 >= r0) && (i1 < r1), a1 = (i1 + 1 >= r0) && (i1 + 1 < r1);
    const bool b0 = (i2 >= 0) && (i2 < nb2), b1 = (i2 + 1 >= 0) && (i2 + 1 < nb2);
    const int lrow = i1 - r0;   // may be -1 when only row i1+1 is in the tile
    if (cell_count && a0 && b0) atomicAdd(cnt + lrow * nb2 + i2, 1);
    const double wa = wA ? wA[p] : 1.0;
    const double t00 = wa * S1 * S2, t01 = wa * S1 * (1.0 - S2);
    const double t10 = wa * (1.0 - S1) * S2, t11 = wa * (1.0 - S1) * (1.0 - S2);
    const int base = lrow * nb2 + i2;
    if (a0 && b0) lds_fx_add(hA, lA, base, t00, scA);
    if (a0 && b1) lds_fx_add(hA, lA, base + 1, t01, scA);
    if (a1 && b0) lds_fx_add(hA, lA, base + nb2, t10, scA);
    if (a1 && b1) lds_fx_add(hA, lA, base + nb2 + 1, t11, scA);
    if (wB) {
      const double wb = wB[p];
      const double u00 = wb * S1 * S2, u01 = wb * S1 * (1.0 - S2);
      const double u10 = wb * (1.0 - S1) * S2, u11 = wb * (1.0 - S1) * (1.0 - S2);
      if (a0 && b0) lds_fx_add(hB, lB, base, u00, scB);
      if (a0 && b1) lds_fx_add(hB, lB, base + 1, u01, scB);
      if (a1 && b0) lds_fx_add(hB, lB, base + nb2, u10, scB);
      if (a1 && b1) lds_fx_add(hB, lB, base + nb2 + 1, u11, scB);
    }
  }
  __syncthreads();
  const int64_t nc = (int64_t)nb1 * nb2;
  const int64_t g0 = (int64_t)r0 * nb2;
  for (int e = threadIdx.x; e < tcells; e += kThreads) {
    if (hA[e]) atomicAdd(reinterpret_cast<unsigned long long *>(accA + g0 + e), hA[e]);
    if (lA[e]) atomicAdd(reinterpret_cast<unsigned long long *>(accA + nc + g0 + e), lA[e]);
    if (wB) {
      if (hB[e]) atomicAdd(reinterpret_cast<unsigned long long *>(accB + g0 + e), hB[e]);
      if (lB[e]) atomicAdd(reinterpret_cast<unsigned long long *>(accB + nc + g0 + e), lB[e]);
    }
    if (cell_count && cnt[e]) atomicAdd(cell_count + g0 + e, cnt[e]);
  }
}

__global__ __launch_bounds__(kThreads) void fx_to_double(const int64_t *__restrict__ acc,
                                                         const double *__restrict__ scale,
                                                         int which, int64_t n,
                                                         double *__restrict__ out) {
  double inv = 1.0 / scale[which];
  const uint64_t *lo_w = reinterpret_cast<const uint64_t *>(acc + n);
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t lo = lo_w[i];
    int64_t hi = acc[i] + (int64_t)(lo >> 32);
    double l = (double)(lo & 0xffffffffull);
    out[i] = fma((double)hi, 0x1p32, l) * inv;
  }
}

int grid_for(int64_t n) {
  int64_t b = (n + kThreads - 1) / kThreads;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

int deposit_impl(const char *who, const double *q1, const double *q2, const double *wA,
                 const double *wB, int64_t n, int nb1, double s1, double e1, int nb2, double s2,
                 double e2, double *outA, double *outB, int32_t *cell_index,
                 int32_t *cell_count, void *ws, int64_t ws_bytes, void *stream) {
#pragma clang fp contract(off)
  // particle arrays may be NULL when there are no particles
  if ((n > 0 && (!q1 || !q2)) || !outA || !ws)
    return set_error(DFCSR_EINVAL, "%s: NULL pointer", who);
  if (n < 0 || nb1 < 1 || nb2 < 1) return set_error(DFCSR_EINVAL, "%s: bad sizes", who);
  if (n >= ((int64_t)1 << 31))
    return set_error(DFCSR_ERANGE, "%s: %lld particles >= 2^31", who, (long long)n);
  if ((int64_t)nb1 * nb2 > (int64_t)1 << 28)
    return set_error(DFCSR_ERANGE, "%s: grid %d x %d too large", who, nb1, nb2);
  if (ws_bytes < dfcsr_histogram_workspace_bytes(nb1, nb2))
    return set_error(DFCSR_ENOWS, "%s: workspace too small", who);
  hipStream_t s = as_stream(stream);
  double *hdr = static_cast<double *>(ws);
  int64_t ncell = (int64_t)nb1 * nb2;
  int64_t *accA = reinterpret_cast<int64_t *>(hdr + kHdr);
  int64_t *accB = accA + 2 * ncell;
  // bins: inv_spacing = 1 / ((end - start) / nbins), deposit.py:55-58
  double inv1 = 1.0 / ((e1 - s1) / (double)nb1);
  double inv2 = 1.0 / ((e2 - s2) / (double)nb2);
  int nb = nblocks_for(n);
  const double *wmax = wB ? wB : wA;   // the non-unit weight, if any
  if (wmax) hipLaunchKernelGGL(absmax_partial, dim3(nb), dim3(kThreads), 0, s, wmax, n, hdr + 8);
  hipLaunchKernelGGL(scale_final, dim3(1), dim3(kThreads), 0, s, (const double *)(hdr + 8), nb,
                     n, wmax ? 1 : 0, hdr);
  hipLaunchKernelGGL(zero_i64, dim3(grid_for(4 * ncell)), dim3(kThreads), 0, s, accA, 4 * ncell,
                     cell_count, cell_count ? ncell : 0);
  // A lone weighted deposit (wA) uses the weighted scale[1] for grid A.
  int sA = wA ? 1 : 0;
  // LDS tiles: 2 words per cell per grid (+ an int count) must fit kTileLdsBytes
  const int ngrid = wB ? 2 : 1;
  const int per_cell = 16 * ngrid + (cell_count ? 4 : 0);
  const int rows = (int)(kTileLdsBytes / per_cell) / nb2;
  const int ntile = rows > 0 ? (nb1 + rows - 1) / rows : 0;
  if (rows >= 4 && ntile <= 16 && n > 0) {
    // enough particles per workgroup that the flush stays small
    int64_t chunks = (n + 16383) / 16384;
    if (chunks < 1) chunks = 1;
    if (chunks * ntile < 512) chunks = (512 + ntile - 1) / ntile;
    int64_t per = (n + chunks - 1) / chunks;
    chunks = (n + per - 1) / per;
    const size_t lds = (size_t)rows * nb2 * per_cell + 16;
    hipLaunchKernelGGL(cic_tile_kernel, dim3((unsigned)chunks, ntile), dim3(kThreads), lds, s, q1,
                       q2, wA, wB, n, per, nb1, s1, inv1, nb2, s2, inv2, rows,
                       (const double *)hdr, sA, accA, accB, cell_index, cell_count);
  } else {
    hipLaunchKernelGGL(cic_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, q1, q2, wA, wB, n,
                       nb1, s1, inv1, nb2, s2, inv2, (const double *)hdr, sA, accA, accB,
                       cell_index, cell_count);
  }
  hipLaunchKernelGGL(fx_to_double, dim3(grid_for(ncell)), dim3(kThreads), 0, s,
                     (const int64_t *)accA, (const double *)hdr, sA, ncell, outA);
  if (outB)   // zero-filled when there are no particles
    hipLaunchKernelGGL(fx_to_double, dim3(grid_for(ncell)), dim3(kThreads), 0, s,
                       (const int64_t *)accB, (const double *)hdr, 1, ncell, outB);
  DFCSR_LAUNCH_CHECK(who);
  return DFCSR_OK;
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_histogram_workspace_bytes(int nbins_1, int nbins_2) {
  if (nbins_1 < 1 || nbins_2 < 1) return 0;
  return (int64_t)kHdr * 8 + 4 * (int64_t)nbins_1 * nbins_2 * 8;
}

extern "C" int dfcsr_histogram_cic_2d(const double *q1, const double *q2, const double *w,
                                      int64_t n, int nbins_1, double bins_start_1,
                                      double bins_end_1, int nbins_2, double bins_start_2,
                                      double bins_end_2, double *hist, int32_t *cell_index,
                                      int32_t *cell_count, void *ws, int64_t ws_bytes,
                                      void *stream) {
  return deposit_impl("histogram_cic_2d", q1, q2, w, nullptr, n, nbins_1, bins_start_1,
                      bins_end_1, nbins_2, bins_start_2, bins_end_2, hist, nullptr, cell_index,
                      cell_count, ws, ws_bytes, stream);
}

extern "C" int dfcsr_deposit_density_vx(const double *x, const double *z, const double *px,
                                        int64_t n, int nbins_x, double x_start, double x_end,
                                        int nbins_z, double z_start, double z_end,
                                        double *density, double *vx, int32_t *cell_index,
                                        int32_t *cell_count, void *ws, int64_t ws_bytes,
                                        void *stream) {
  if ((n > 0 && !px) || !vx) return set_error(DFCSR_EINVAL, "deposit_density_vx: NULL pointer");
  return deposit_impl("deposit_density_vx", x, z, nullptr, px, n, nbins_x, x_start, x_end,
                      nbins_z, z_start, z_end, density, vx, cell_index, cell_count, ws,
                      ws_bytes, stream);
}
