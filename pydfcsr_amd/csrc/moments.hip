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

}  // namespace
}  // namespace dfcsr

using namespace dfcsr;

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
