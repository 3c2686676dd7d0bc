// a1: beam moments with a fixed reduction tree (deterministic).
//
// Replaces the numpy statistics the CSR step reads: Beam.update_status and
// its properties (beams.py:88-98, 137-215), get_CSR_mesh's std/mean of
// x_transform (CSR.py:368-374) and get_DF's slice std (deposit.py:147-159).
// Two-pass (mean, then centred sums) like numpy's std; the least-squares
// slope is cov(x, z)/var(z), intercept mean_x - slope*mean_z.
#include "common.h"

namespace dfcsr {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kNSum = 4;   // partial sums per block per pass

int nblocks_for(int64_t n) {
  int64_t b = (n + kThreads * 8 - 1) / (kThreads * 8);
  if (b < 1) b = 1;
  if (b > kMaxBlocks) b = kMaxBlocks;
  return (int)b;
}

// stats layout in workspace (doubles):
//   [0..9] results (same order as the public output)
//   [16 ...] block partials: kMaxBlocks * kNSum
enum { S_MX = 0, S_MZ, S_SX, S_SZ, S_P0, S_P1, S_MXT, S_SXT, S_SSL, S_NSL, S_MSL = 10 };
constexpr int kPartOff = 16;

// pass 0: sum x, sum z
// pass 1: sum (x-mx)^2, (z-mz)^2, (x-mx)(z-mz)
// pass 2: sum xt, slice count, slice sum x
// pass 3: sum (xt-mxt)^2, slice sum (x-msl)^2
__global__ __launch_bounds__(kThreads) void moments_partial(const double *__restrict__ x,
                                                            const double *__restrict__ z,
                                                            int64_t n, int pass,
                                                            double *__restrict__ st) {
  __shared__ double red[kThreads / 64];
  double a[kNSum] = {0.0, 0.0, 0.0, 0.0};
  const double mx = st[S_MX], mz = st[S_MZ], p0 = st[S_P0], p1 = st[S_P1];
  const double mxt = st[S_MXT], msl = st[S_MSL], cut = 0.1 * st[S_SZ];
  int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    double xv = x[i], zv = z[i];
    if (pass == 0) {
      a[0] += xv;
      a[1] += zv;
    } else if (pass == 1) {
      double dx = xv - mx, dz = zv - mz;
      a[0] += dx * dx;
      a[1] += dz * dz;
      a[2] += dx * dz;
    } else {
      double xt = xv - (p0 * zv + p1);
      bool in = fabs(zv) < cut;
      if (pass == 2) {
        a[0] += xt;
        a[1] += in ? 1.0 : 0.0;
        a[2] += in ? xv : 0.0;
      } else {
        double d = xt - mxt, e = xv - msl;
        a[0] += d * d;
        a[1] += in ? e * e : 0.0;
      }
    }
  }
  for (int k = 0; k < kNSum; ++k) {
    double v = block_sum(a[k], red);
    if (threadIdx.x == 0) st[kPartOff + blockIdx.x * kNSum + k] = v;
  }
}

__global__ __launch_bounds__(kThreads) void moments_final(int64_t n, int pass, int nblk,
                                                          double *__restrict__ st,
                                                          double *__restrict__ out) {
  __shared__ double red[kThreads / 64];
  double tot[kNSum];
  for (int k = 0; k < kNSum; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) v += st[kPartOff + b * kNSum + k];
    tot[k] = block_sum(v, red);
  }
  if (threadIdx.x != 0) return;
  double dn = (double)n;
  if (pass == 0) {
    st[S_MX] = tot[0] / dn;
    st[S_MZ] = tot[1] / dn;
  } else if (pass == 1) {
    st[S_SX] = sqrt(tot[0] / dn);
    st[S_SZ] = sqrt(tot[1] / dn);
    double p0 = tot[2] / tot[1];
    st[S_P0] = p0;
    st[S_P1] = st[S_MX] - p0 * st[S_MZ];
  } else if (pass == 2) {
    st[S_MXT] = tot[0] / dn;
    st[S_NSL] = tot[1];
    st[S_MSL] = tot[2] / tot[1];   // NaN for an empty slice, as np.std([])
  } else {
    st[S_SXT] = sqrt(tot[0] / dn);
    st[S_SSL] = sqrt(tot[1] / st[S_NSL]);
    for (int k = 0; k < DFCSR_MOMENTS_COUNT; ++k) out[k] = st[k];
  }
}

// ---- per-step beam statistics (CSR.py:837-859, twiss.py:2-70) -------------
//
// Means and the sample covariance (np.cov, ddof = 1) of the six coordinates
// (x, px, y, py, z, pz) in one pass over the particles: sums of d = v - c
// and of the products d_p d_q, each a fixed reduction tree, with c the mean
// of the first (up to) 256 particles -- within a small fraction of a sigma
// of the bunch's mean, so the shifted sums lose nothing to cancellation
// beyond a few ulps.  The final single-workgroup pass forms
// mean = c + S_d / n and cov = (S_pq - S_p S_q / n) / (n - 1), then the
// Twiss and dispersion parameters of (x, px, pz) and (y, py, pz) and the
// energy statistics.
constexpr int kNCov = 27;          // 6 shifted sums, then the 21 products (upper triangle)
constexpr int kStatOff = 32;       // block partials start here (doubles)

__device__ inline int tri_index(int i, int j) {   // i <= j < 6, row-major upper triangle
  return i * 6 - i * (i - 1) / 2 + (j - i);
}

__global__ __launch_bounds__(kThreads) void stats_partial(
    const double *__restrict__ x, const double *__restrict__ px, const double *__restrict__ y,
    const double *__restrict__ py, const double *__restrict__ z, const double *__restrict__ pz,
    int64_t n, double *__restrict__ st) {
  __shared__ double red[kThreads / 64];
  const double *c[6] = {x, px, y, py, z, pz};
  double a[kNCov];
#pragma unroll
  for (int k = 0; k < kNCov; ++k) a[k] = 0.0;
  // the shift: every workgroup forms the same sum in the same order
  double m[6];
  const double m0 = (double)(n < kThreads ? n : kThreads);
  for (int k = 0; k < 6; ++k)
    m[k] = block_sum((int64_t)threadIdx.x < n ? c[k][threadIdx.x] : 0.0, red) / m0;
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) st[k] = m[k];
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = c[k][i] - m[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) a[k] += v[k];
#pragma unroll
    for (int p = 0; p < 6; ++p)
#pragma unroll
      for (int q = p; q < 6; ++q) a[6 + tri_index(p, q)] += v[p] * v[q];
  }
  // unrolled: constant indices keep a[] in registers (a loop index would
  // place the array in scratch memory)
#pragma unroll
  for (int k = 0; k < kNCov; ++k) {
    const double v = block_sum(a[k], red);
    if (threadIdx.x == 0) st[kStatOff + blockIdx.x * kNCov + k] = v;
  }
}

// Twiss and dispersion of one plane from its 3 x 3 covariance of (u, pu,
// delta) (twiss.py:30-70): out = alpha, beta, gamma, emit, eta, etap, norm_emit.
__device__ inline void twiss_plane(double s00, double s01, double s02, double s11, double s12,
                                   double s22, double p0c, double mc2, double *out) {
  const double delta2 = s22;
  const double eb = s00 - s02 * s02 / delta2;
  const double eg = s11 - s12 * s12 / delta2;
  const double ea = -s01 + s02 * s12 / delta2;
  const double emit = sqrt(eb * eg - ea * ea);
  out[0] = ea / emit;
  out[1] = eb / emit;
  out[2] = eg / emit;
  out[3] = emit;
  out[4] = s02 / delta2;
  out[5] = s12 / delta2;
  out[6] = emit * p0c / mc2;
}

__global__ __launch_bounds__(kThreads) void stats_final(int64_t n, int nblk, double p0c,
                                                        double mc2, const double *__restrict__ st,
                                                        double *__restrict__ out) {
  __shared__ double red[kThreads / 64];
  double tot[kNCov];   // constant indices throughout (unrolled loops): registers
#pragma unroll
  for (int k = 0; k < kNCov; ++k) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kThreads) v += st[kStatOff + b * kNCov + k];
    tot[k] = block_sum(v, red);
  }
  if (threadIdx.x != 0) return;
  const double dn = (double)n;
  const double c[6] = {st[0], st[1], st[2], st[3], st[4], st[5]};
  double *cov = out + DFCSR_STAT_COV;
#pragma unroll
  for (int k = 0; k < 6; ++k) out[DFCSR_STAT_MEAN + k] = c[k] + tot[k] / dn;
#pragma unroll
  for (int p = 0; p < 6; ++p)
#pragma unroll
    for (int q = p; q < 6; ++q) {
      const int t = tri_index(p, q);
      cov[t] = (tot[6 + t] - tot[p] * tot[q] / dn) / (dn - 1.0);
    }
  // planes (x, px, pz) and (y, py, pz): coordinates 0, 1, 5 and 2, 3, 5
  twiss_plane(cov[tri_index(0, 0)], cov[tri_index(0, 1)], cov[tri_index(0, 5)],
              cov[tri_index(1, 1)], cov[tri_index(1, 5)], cov[tri_index(5, 5)], p0c, mc2,
              out + DFCSR_STAT_TWISS_X);
  twiss_plane(cov[tri_index(2, 2)], cov[tri_index(2, 3)], cov[tri_index(2, 5)],
              cov[tri_index(3, 3)], cov[tri_index(3, 5)], cov[tri_index(5, 5)], p0c, mc2,
              out + DFCSR_STAT_TWISS_Y);
  // energy = (pz + 1) p0c (beams.py:166-181): its mean and population std
  out[DFCSR_STAT_MEAN_ENERGY] = (out[DFCSR_STAT_MEAN + 5] + 1.0) * p0c;
  out[DFCSR_STAT_SIGMA_ENERGY] = p0c * sqrt(cov[tri_index(5, 5)] * ((dn - 1.0) / dn));
}

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

extern "C" int64_t dfcsr_beam_statistics_workspace_bytes(int64_t n) {
  (void)n;
  return (int64_t)(kStatOff + kMaxBlocks * kNCov) * (int64_t)sizeof(double);
}

extern "C" int dfcsr_beam_statistics(const double *x, const double *px, const double *y,
                                     const double *py, const double *z, const double *pz,
                                     int64_t n, double p0c, double mc2, void *ws,
                                     int64_t ws_bytes, double *out, void *stream) {
  if (!x || !px || !y || !py || !z || !pz || !out || !ws)
    return set_error(DFCSR_EINVAL, "beam_statistics: NULL pointer");
  if (n < 2)
    return set_error(DFCSR_EINVAL, "beam_statistics: need n >= 2 particles (n=%lld)",
                     (long long)n);
  if (ws_bytes < dfcsr_beam_statistics_workspace_bytes(n))
    return set_error(DFCSR_ENOWS, "beam_statistics: workspace %lld < %lld bytes",
                     (long long)ws_bytes, (long long)dfcsr_beam_statistics_workspace_bytes(n));
  hipStream_t s = as_stream(stream);
  double *st = static_cast<double *>(ws);
  const int nb = nblocks_for(n);
  hipLaunchKernelGGL(stats_partial, dim3(nb), dim3(kThreads), 0, s, x, px, y, py, z, pz, n, st);
  hipLaunchKernelGGL(stats_final, dim3(1), dim3(kThreads), 0, s, n, nb, p0c, mc2,
                     (const double *)st, out);
  DFCSR_LAUNCH_CHECK("beam_statistics");
  return DFCSR_OK;
}

extern "C" int64_t dfcsr_moments_workspace_bytes(int64_t n) {
  (void)n;
  return (int64_t)(kPartOff + kMaxBlocks * kNSum) * (int64_t)sizeof(double);
}

extern "C" int dfcsr_beam_moments(const double *x, const double *z, int64_t n, void *ws,
                                  int64_t ws_bytes, double *out, void *stream) {
  if (!x || !z || !out || !ws) return set_error(DFCSR_EINVAL, "beam_moments: NULL pointer");
  if (n < 2) return set_error(DFCSR_EINVAL, "beam_moments: need n >= 2 particles (n=%lld)",
                              (long long)n);
  if (ws_bytes < dfcsr_moments_workspace_bytes(n))
    return set_error(DFCSR_ENOWS, "beam_moments: workspace %lld < %lld bytes",
                     (long long)ws_bytes, (long long)dfcsr_moments_workspace_bytes(n));
  hipStream_t s = as_stream(stream);
  double *st = static_cast<double *>(ws);
  int nb = nblocks_for(n);
  for (int pass = 0; pass < 4; ++pass) {
    hipLaunchKernelGGL(moments_partial, dim3(nb), dim3(kThreads), 0, s, x, z, n, pass, st);
    hipLaunchKernelGGL(moments_final, dim3(1), dim3(kThreads), 0, s, n, pass, nb, st, out);
  }
  DFCSR_LAUNCH_CHECK("beam_moments");
  return DFCSR_OK;
}
