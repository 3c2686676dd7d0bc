/* dfcsr.h -- C ABI of the MI355X (gfx950) 2D CSR wake step.
 *
 * Every entry point is stream-ordered and asynchronous on the HIP stream it
 * is given (NULL = the null stream), takes caller-owned DEVICE pointers
 * (contiguous, C order) and returns 0 on success or a negative DFCSR_E*
 * code; dfcsr_last_error() then holds a message for the calling thread.
 * No entry point allocates device memory: scratch comes in as a workspace
 * pointer whose size the matching *_workspace_bytes() call reports.
 * Calls that share a workspace must not overlap on different streams.
 *
 * Array conventions follow the reference (pyDFCSR, paths relative to
 * pyDFCSR_2D/): 2D grids are [x_index][z_index]; the wake mesh is
 * [xbins][zbins] flattened with x outer (CSR.py:384-387).
 */
#ifndef DFCSR_H
#define DFCSR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFCSR_ABI_VERSION 8

enum {
  DFCSR_OK = 0,
  DFCSR_EINVAL = -1,   /* bad argument (shape, size, NULL pointer) */
  DFCSR_EHIP = -2,     /* a HIP runtime call failed */
  DFCSR_ERANGE = -3,   /* a size exceeds a documented maximum */
  DFCSR_ENOWS = -4     /* workspace too small */
};

enum { DFCSR_PREC_FP64 = 0, DFCSR_PREC_MIXED = 1 };

/* Number of doubles written by dfcsr_beam_moments. */
#define DFCSR_MOMENTS_COUNT 10
/* Order of the five history fields in every slice record. */
enum { DFCSR_F_DENSITY = 0, DFCSR_F_DENSITY_X = 1, DFCSR_F_DENSITY_Z = 2,
       DFCSR_F_VX = 3, DFCSR_F_VX_X = 4 };
/* History records are corner-packed: the record of cell (x, z) holds, for
 * each field f, the values at (x, z), (x, z1), (x1, z), (x1, z1) with
 * x1 = min(x + 1, Xi - 1) and z1 = min(z + 1, Zi - 1) -- the 2 x 2
 * neighbourhood a trilinear lookup reads in one time plane, with the
 * reference's last-plane clamp (interp3D.py:31-46) built in.  Element
 * [f * 4 + c], c = 0..3 in that order: 20 values per record.
 * FP64 ring: whole records, time slot innermost, [x][z][slot][20] (slot =
 * ring slot, see DfcsrHistory): the two time planes of a lookup are
 * neighbouring records, and the far-region lookups of one wavefront (many
 * retarded times, nearby z) stay within one x, z block.
 * FP32 rings (mixed precision), two layouts of the same values:
 *   chunked: records cut into 5 chunks of 16 bytes (chunk f = the four
 *   corners of field f) stored z fastest, [x][slot][f][z][4], so the
 *   near-region lookups of one wavefront (one time plane, consecutive z)
 *   read consecutive chunks;
 *   time-innermost: whole records [x][z][slot][20] like the FP64 ring, so a
 *   far-region lookup (two neighbouring planes) is one 160-byte span. */
#define DFCSR_HIST_REC64 20
#define DFCSR_HIST_REC32 20
#define DFCSR_HIST_CHUNKS32 5
#define DFCSR_MAX_ELEMENTS 64
#define DFCSR_MAX_SG_WINDOW 21

int dfcsr_abi_version(void);
const char *dfcsr_last_error(void);

/* ---- a1: beam moments -----------------------------------------------------
 * Replaces the numpy statistics read by the step: Beam.update_status /
 * sigma_x / sigma_z / mean_x / mean_z / slope / x_transform
 * (beams.py:88-98, 137-215), the std/mean of x_transform in get_CSR_mesh
 * (CSR.py:368-374) and the |z| < 0.1 sigma_z slice std of get_DF
 * (deposit.py:147-159).
 * out[0..9] = mean_x, mean_z, sigma_x, sigma_z, slope0, slope1,
 *             mean(x_transform), std(x_transform), slice std(x), slice count.
 * Population (ddof=0) statistics; slope is the least-squares line x(z)
 * (np.polyfit(z, x, 1)).  Deterministic (fixed reduction tree). */
int64_t dfcsr_moments_workspace_bytes(int64_t n);
int dfcsr_beam_moments(const double *x, const double *z, int64_t n, void *ws, int64_t ws_bytes,
                       double *out, void *stream);

/* Per-step beam statistics on the device: replaces CSR2D.update_statistics'
 * reads of Beam.twiss / sigma_energy / mean_energy (CSR.py:837-859,
 * twiss.py:2-70, beams.py:166-181).  out[DFCSR_STATS_COUNT] (device):
 *   [DFCSR_STAT_MEAN + k]  mean of coordinate k (x, px, y, py, z, pz)
 *   [DFCSR_STAT_COV + t]   sample covariance (np.cov, ddof = 1), row-major
 *                          upper triangle: t = 0 (x,x), 1 (x,px), ... 20 (pz,pz)
 *   [DFCSR_STAT_TWISS_X + k], [DFCSR_STAT_TWISS_Y + k]  twiss_dispersion_calc
 *                          of (x, px, pz) / (y, py, pz): alpha, beta, gamma,
 *                          emit, eta, etap, norm_emit (= emit * p0c / mc2)
 *   [DFCSR_STAT_MEAN_ENERGY], [DFCSR_STAT_SIGMA_ENERGY]  of (pz + 1) * p0c
 *                          (population std).
 * Deterministic (fixed reduction tree); no host synchronisation. */
#define DFCSR_STAT_MEAN 0
#define DFCSR_STAT_COV 6
#define DFCSR_STAT_TWISS_X 27
#define DFCSR_STAT_TWISS_Y 34
#define DFCSR_STAT_MEAN_ENERGY 41
#define DFCSR_STAT_SIGMA_ENERGY 42
#define DFCSR_STATS_COUNT 43
int64_t dfcsr_beam_statistics_workspace_bytes(int64_t n);
int dfcsr_beam_statistics(const double *x, const double *px, const double *y, const double *py,
                          const double *z, const double *pz, int64_t n, double p0c, double mc2,
                          void *ws, int64_t ws_bytes, double *out, void *stream);

/* ---- a3: cloud-in-cell deposition ----------------------------------------
 * Replaces histogram_cic_2d (deposit.py:42-87): hist[i1][i2] += w*S1*S2 over
 * the four CIC corners with cell = (q - start) * (1/((end - start)/nbins)).
 * w == NULL means unit weights.  Sums are accumulated as 64-bit fixed point
 * (scale 2^k, k chosen from n * max|w|) so the result is bit-reproducible.
 * Optional outputs (NULL to skip): cell_index[2*p + {0,1}] = the particle's
 * lower cell (i1, i2) (floor, may lie off the grid); cell_count[i1][i2] =
 * number of particles whose lower cell is (i1, i2) on the grid. */
int64_t dfcsr_histogram_workspace_bytes(int nbins_1, int nbins_2);
int dfcsr_histogram_cic_2d(const double *q1, const double *q2, const double *w, int64_t n,
                           int nbins_1, double bins_start_1, double bins_end_1,
                           int nbins_2, double bins_start_2, double bins_end_2,
                           double *hist, int32_t *cell_index, int32_t *cell_count,
                           void *ws, int64_t ws_bytes, void *stream);
/* The two depositions of DF_tracker.get_DF in one pass over the particles
 * (deposit.py:172-182): density (w = 1) and vx sum (w = px). */
int dfcsr_deposit_density_vx(const double *x, const double *z, const double *px, int64_t n,
                             int nbins_x, double x_start, double x_end,
                             int nbins_z, double z_start, double z_end,
                             double *density, double *vx, int32_t *cell_index,
                             int32_t *cell_count, void *ws, int64_t ws_bytes, void *stream);

/* ---- a4-a6: density-function post-processing -------------------------------
 * Replaces deposit.py:183-235 on deposited grids (in place):
 *   velocity threshold division, separable Savitzky-Golay smoothing
 *   (scipy.signal.savgol_filter, mode='interp', axis 0 then 1), trapezoid
 *   normalisation over linspace grids, np.gradient (non-uniform spacing form),
 *   smoothing of the derivatives and the vx_x fill rule.
 * density / vx: in = raw deposits, out = final fields; density_x,
 * density_z, vx_x: out.  Grid axes are np.linspace(x_start, x_end, nbins_x)
 * and np.linspace(z_start, z_end, nbins_z).  window odd, order < window,
 * window <= min(nbins) and <= DFCSR_MAX_SG_WINDOW.
 * info (device, may be NULL) receives {max raw density, dsum, max final
 * density, vx_x fill value}. */
int64_t dfcsr_df_workspace_bytes(int nbins_x, int nbins_z);
int dfcsr_df_process(double *density, double *vx, double *density_x, double *density_z,
                     double *vx_x, int nbins_x, double x_start, double x_end,
                     int nbins_z, double z_start, double z_end,
                     double velocity_threshold, int window, int order, double *info,
                     void *ws, int64_t ws_bytes, void *stream);

/* ---- a7: 2D Savitzky-Golay filter -------------------------------------------
 * Replaces SGolay_filter.sgolay2d(z, window_size, order, derivative)
 * (SGolay_filter.py:3-81; get_DF's commented alternative to a5/a6,
 * deposit.py:187, 194, 227, 232).  z: row-major [nx][nz] float64 on the
 * device.  derivative: DFCSR_SG2D_NONE (the smoothed grid into out),
 * _COL (the slope along the first axis, per index step), _ROW (along the
 * second), _BOTH (col into out, row into out2) -- the reference's None,
 * 'col', 'row', 'both'.  The grid is padded as the reference pads it
 * (mirrored bands about the edge values) and convolved with the rows of the
 * fit's pseudo-inverse; the sum runs in a fixed order where the reference
 * uses an FFT (results agree to rounding).  Limits: window_size odd,
 * 3 <= window_size <= DFCSR_SG2D_MAX_WINDOW (else DFCSR_ERANGE);
 * 0 <= order < window_size (the reference also accepts rank-deficient fits,
 * order >= window_size); order >= 1 for a derivative; nx, nz >=
 * window_size / 2 + 1 (the reference's padding slices need as much).
 * One launch on `stream`; no allocation, no workspace. */
#define DFCSR_SG2D_NONE 0
#define DFCSR_SG2D_COL 1
#define DFCSR_SG2D_ROW 2
#define DFCSR_SG2D_BOTH 3
#define DFCSR_SG2D_MAX_WINDOW 13
int dfcsr_sgolay2d(const double *z, int nx, int nz, int window_size, int order, int derivative,
                   double *out, double *out2, void *stream);
/* Host-side (no device work): the [window_size][window_size] kernels
 * dfcsr_sgolay2d convolves with -- pinv row 0, or the negated rows 1 / 2
 * (SGolay_filter.py:69-81) -- into k0 (and k1 for _BOTH). */
int dfcsr_sgolay2d_kernels(int window_size, int order, int derivative, double *k0, double *k1);

/* ---- a8-a10: history re-grid ------------------------------------------------
 * Replaces DF_tracker.DF_interp / append_interpolant's re-gridding
 * (deposit.py:296-309, 328-337, 379-390): bilinear RegularGridInterpolator
 * (fill 0; vx_x fill = mean(vx_x) of that slice) of slices on their own
 * deposition grids onto the common interp grid
 * np.linspace(xi0, xi1, Xi) x np.linspace(zi0, zi1, Zi).
 * src: raw slices, src_slot_stride doubles apart, each holding the five
 * fields as [field][nbx][nbz] (field order DFCSR_F_*).
 * jobs: HOST array of njobs descriptors; hist64 (FP64), hist32 (FP32
 * chunked) and hist32t (FP32 time-innermost) are rings of cap slots in the
 * layouts above (dst_slot < cap), cells and cells32 the cell-major rings of
 * DfcsrHistory; each may be NULL, but not all five. */
typedef struct {
  int32_t src_slot;   /* raw slice index in src */
  int32_t dst_slot;   /* ring slot in the history */
  int32_t nbx, nbz;   /* deposition grid of that slice */
  double x0, x1, z0, z1;   /* its linspace end points */
} DfcsrRegridJob;
int64_t dfcsr_regrid_workspace_bytes(int njobs);
int dfcsr_regrid(const double *src, int64_t src_slot_stride, const DfcsrRegridJob *jobs,
                 int njobs, double *hist64, float *hist32, float *hist32t, double *cells,
                 float *cells32, int Xi, int Zi, int cap, double xi0, double xi1, double zi0,
                 double zi1, void *ws, int64_t ws_bytes, void *stream);

/* ---- a11/a12: lookup primitives (exposed for parity tests) -----------------
 * History descriptor: ring of cap slots, logical slice i lives in slot
 * (head + i) % cap; axes as DF_tracker.build_interpolant (deposit.py:416-421):
 * min_t = oldest time, dt = (t_last - t_first)/(nt - 1), min_x = xi0,
 * dx = (xi1 - xi0)/(Xi - 1), same for z.  The FP64 ring is needed by the
 * lookup primitives, the integrand and FP64 wakes; a mixed-precision wake
 * needs at least one FP32 ring (with both, its region-0 phases read the
 * time-innermost one and the other phases the chunked one).
 * The cell-major rings (optional) hold the same slices plane by plane:
 * cell (slot, x, z) at ((slot Xi + x) DFCSR_CELL_ROW(Zi) + z) DFCSR_CELL,
 * the five fields of that grid node (40 B in float64, 20 B in float32,
 * against 160 B / 80 B corner-packed records).  Each x row holds Zi + 1
 * cells: z = 0 .. Zi - 1 and, at z = Zi, a copy of cell Zi - 1, so that the
 * z pair (z, z + 1) a lookup reads is interpolate3D's (clamped at the last
 * z) everywhere.  With them the near-region phases of
 * a wake read cells (FP64 wakes: cells, mixed: cells32), whose neighbouring
 * z lie in neighbouring cells; results are identical with or without them.
 * One plane must span at most 4 GiB (the wakes address cells within a plane
 * with 32-bit offsets).
 * dcells32 (optional, with cells32) holds the float32 time differences of
 * cells32 in the same layout: cell (slot of slice i, x, z) = cells32 of
 * slice i + 1 minus that of slice i, rounded to float32, and 0 for the
 * newest slice (dfcsr_cells32_tdiff fills it).  A mixed-precision wake reads
 * the cell-major rings only when both are present: the time lerp
 * v0 + td (v1 - v0) then takes v1 - v0 from dcells32, the same float32
 * value, so the results are those of the float32 rings. */
typedef struct {
  const double *data;    /* [Xi][Zi][cap][DFCSR_HIST_REC64] or NULL */
  const float *data32;   /* [Xi][cap][DFCSR_HIST_CHUNKS32][Zi][4] or NULL */
  const float *data32t;  /* [Xi][Zi][cap][DFCSR_HIST_REC32] or NULL */
  const double *cells;   /* [cap][Xi][Zi + 1][DFCSR_CELL] or NULL */
  const float *cells32;  /* [cap][Xi][Zi + 1][DFCSR_CELL] or NULL */
  int32_t cap, head, nt, Xi, Zi;
  double min_t, min_x, min_z, dt, dx, dz;
  const float *dcells32; /* [cap][Xi][Zi + 1][DFCSR_CELL] or NULL */
  /* zspan (optional, Zi <= 65535): [Xi][cap] z spans, zlo | zhi << 16, from
   * dfcsr_history_zspan.  Entry (x, slot of slice i) bounds the z of every
   * cell of rows x and min(x + 1, Xi - 1), in slice i and (when i < nt - 1)
   * slice i + 1, that can contribute: density, density_x or density_z
   * nonzero, or vx or vx_x not finite (zlo > zhi: none).  A lookup whose
   * corners lie outside the span of its x row and first time plane gathers
   * fields whose integrand is exactly 0 (every term carries a factor
   * density, density_x or density_z, and 0 times a finite value is 0): the
   * wakes skip it, as they skip lookups outside the history (the near regions
   * always, region 0 on the chirp branch, |slope| > 1). */
  const uint32_t *zspan;
} DfcsrHistory;

/* Fills zspan (layout of DfcsrHistory) for the logical slices first .. nt-1
 * of a ring of cap slots (slice i in slot (head + i) % cap) from its
 * float64 cell-major ring cells, or when that is NULL from the float32 one
 * cells32.  From cells, vx and vx_x beyond the float32 range count as not
 * finite, so the spans hold for wakes of either precision; from cells32
 * they hold for mixed-precision wakes only (a float64 value may round to a
 * float32 zero).  Call after dfcsr_regrid with first as for
 * dfcsr_cells32_tdiff.  Zi <= 65535.  Has no counterpart in the reference:
 * it lets the wakes skip lookups whose integrand is exactly 0. */
int dfcsr_history_zspan(const double *cells, const float *cells32, uint32_t *zspan, int Xi, int Zi,
                        int cap, int head, int nt, int first, void *stream);

/* Fills dcells32 (layout of DfcsrHistory) for the logical slices
 * first .. nt-1 of a ring of cap slots whose slice i lives in slot
 * (head + i) % cap: the cells of slice i + 1 minus those of slice i in
 * float32, and 0 for slice nt - 1.  Called after a re-grid: of the new
 * slice alone (append_interpolant's common case) with first = nt - 2, of
 * every slice with first = 0.  Has no counterpart in the reference: it keeps
 * the derived ring of the mixed-precision wake current. */
int dfcsr_cells32_tdiff(const float *cells32, float *dcells32, int Xi, int Zi, int cap, int head,
                        int nt, int first, void *stream);
#define DFCSR_CELL 5
/* cells per x row of the cell-major rings: the grid's Zi and a copy of the last */
#define DFCSR_CELL_ROW(Zi) ((Zi) + 1)

/* Lattice tables (lattice.py:4-110, 136-150): table[k][0..5] =
 * coords_x, coords_y, n_x, n_y, tau_x, tau_y at s = min_s + k*ds, k < ns;
 * distance[], rho[] per element (device pointers). */
typedef struct {
  const double *table;
  int32_t ns;
  int32_t nelem;
  double min_s, ds;
  const double *distance;
  const double *rho;
} DfcsrLattice;

/* interpolate3D (interp3D.py:18-66) on one history field at n points
 * q[3*p + {0,1,2}] = (t, x, z). */
int dfcsr_interpolate3d(const DfcsrHistory *hist, int field, const double *q, int64_t n,
                        double *out, void *stream);
/* interpolate1D (interp1D.py:13-36) on lattice table column col (0..5). */
int dfcsr_interpolate1d(const DfcsrLattice *lat, int col, const double *s, int64_t n,
                        double *out, void *stream);

/* ---- a13-a15: retarded integral on the wake mesh ---------------------------- */
typedef struct {
  double t;                      /* beam position (CSR.py:456) */
  double sigma_x, sigma_z;       /* beam._sigma_x, beam._sigma_z */
  double slope0, slope1;         /* beam._slope (np.polyfit(z, x, 1)) */
  double mean_x;                 /* beam._mean_x */
  double formation_length, n_formation_length;
  double csr_scaling;            /* 8.98755e3 * charge (CSR.py:80) */
  int32_t Nx, Nz;                /* CSR_integration xbins, zbins */
  int32_t mesh_xbins, mesh_zbins;
  double xr0, xr1, zr0, zr1;     /* linspace ends of xrange, zrange (CSR.py:380-381) */
  /* Optional DEVICE array [M][2] of observers (s - t, x); when non-NULL it
   * replaces the mesh formula, e.g. for get_CSR_wake(s, x) at arbitrary
   * points (CSR.py:454). */
  const double *observers;
  /* Optional DEVICE counter: when non-NULL the wake adds to it the number of
   * integrand points it evaluated (the points whose retarded lookup lies in
   * the history; the rest contribute exactly 0 and are skipped).  A
   * diagnostic: it does not change the wake.  No counterpart in the
   * reference, which evaluates every point. */
  unsigned long long *points_evaluated;
} DfcsrWakeParams;

/* get_CSR_integrand (CSR.py:605-782) on an explicit (xp, sp) grid of one
 * observer (s, x, t = p->t): Iz, Ix are [nxp][nsp]. */
int dfcsr_csr_integrand(const DfcsrHistory *hist, const DfcsrLattice *lat, double s, double x,
                        double t, const double *xp, int nxp, const double *sp, int nsp,
                        double *Iz, double *Ix, void *stream);

/* get_CSR_mesh + calculate_2D_CSR + get_CSR_wake (CSR.py:361-602):
 * xmesh, zmesh (M = mesh_xbins*mesh_zbins, both or neither may be NULL),
 * dE_dct, x_kick ([mesh_xbins][mesh_zbins], MeV/m).  precision:
 * DFCSR_PREC_FP64 or DFCSR_PREC_MIXED (FP32 integrand with compensated FP32
 * accumulation; FP64 retardation geometry and final sum; needs an FP32
 * ring; vx at the observer comes from the FP64 ring when present, else
 * from an FP32 ring).  The workspace holds per-phase partial sums (a phase is one
 * integration region and one chunk of its x' rows): its size depends on the
 * mesh and integration bins and the chirp branch of p. */
int64_t dfcsr_wake_workspace_bytes(const DfcsrWakeParams *p);
int dfcsr_wake_mesh(const DfcsrHistory *hist, const DfcsrLattice *lat, const DfcsrWakeParams *p,
                    int precision, double *xmesh, double *zmesh, double *dE_dct, double *x_kick,
                    void *ws, int64_t ws_bytes, void *stream);

/* dfcsr_wake_mesh without a workspace argument, for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The partial sums and
 * observer data live in a static device array of DFCSR_WAKE_SCRATCH_BYTES
 * per GPU, held until the process ends; a mesh whose workspace
 * (dfcsr_wake_workspace_bytes) exceeds it is rejected with DFCSR_ERANGE
 * (32 MiB: up to a 128 x 128 mesh on either branch at 200 x 200
 * integration).  Calls on one GPU share that array: calls that overlap on
 * one GPU (concurrent streams or threads) race.  Calls no allocating
 * routine. */
#define DFCSR_WAKE_SCRATCH_BYTES ((int64_t)32 << 20)
int dfcsr_wake_mesh_noalloc(const DfcsrHistory *hist, const DfcsrLattice *lat,
                            const DfcsrWakeParams *p, int precision, double *xmesh,
                            double *zmesh, double *dE_dct, double *x_kick, void *stream);

/* Batched form for serving: k independent simulations (each its own history,
 * lattice and parameters) in one launch, 1 <= k <= DFCSR_WAKE_MAX_TASKS.  All
 * descriptor arrays are HOST arrays of length k; output pointers per
 * simulation.  The task table is copied into the workspace on the stream and
 * the call waits for that copy before it returns (the kernels stay
 * asynchronous). */
#define DFCSR_WAKE_MAX_TASKS 4096
int dfcsr_wake_mesh_batched(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                            const DfcsrWakeParams *p, int precision, double *const *dE_dct,
                            double *const *x_kick, void *ws, int64_t ws_bytes, void *stream);
int64_t dfcsr_wake_batched_workspace_bytes(int k, const DfcsrWakeParams *p);
/* The batched form in two halves, so the launch can be captured in a HIP
 * graph and replayed with new task contents: dfcsr_wake_tasks_pack writes
 * the k task descriptors into host memory `tasks` (at least
 * dfcsr_wake_task_table_bytes(k) bytes) and the launch shape into *batch;
 * the caller copies those bytes to device memory and
 * dfcsr_wake_batched_launch runs them with a device buffer of at least
 * batch->partial_bytes for the partial sums (no host-device copy inside). */
typedef struct {
  int32_t tasks;          /* k */
  int32_t blocks_far;     /* workgroups of the region-0 launch */
  int32_t blocks_near;    /* workgroups of the launch over the other regions */
  int32_t observers;      /* mesh points over all tasks */
  int64_t partial_bytes;  /* partial-sum buffer the launch needs */
  int32_t precision;      /* the precision the tasks were checked for; the
                             launch must use the same */
} DfcsrWakeBatch;
int64_t dfcsr_wake_task_table_bytes(int k);
int dfcsr_wake_tasks_pack(int k, const DfcsrHistory *hist, const DfcsrLattice *lat,
                          const DfcsrWakeParams *p, int precision, double *const *dE_dct,
                          double *const *x_kick, void *tasks, int64_t tasks_bytes,
                          DfcsrWakeBatch *batch);
int dfcsr_wake_batched_launch(const void *dev_tasks, const DfcsrWakeBatch *batch, int precision,
                              void *partials, int64_t partials_bytes, void *stream);

/* ---- tracking: one lattice-element piece -----------------------------------
 * Replaces bmadx.track_element on the element CSR2D.get_bmadx_element builds
 * for a step piece (CSR.py:146-199, 227-294).  Bmad's canonical coordinates
 * (x, px, y, py, z, pz) relative to the reference momentum p0c [eV]; mc2 the
 * particle rest energy [eV].  Drift: exact.  SBend: exact motion in its
 * uniform field of curvature g, hard-edge pole faces e1 (entrance) and e2
 * (exit) with linear edge kicks (px += x g tan e, py -= y g tan(e - psi),
 * psi = FINT 2 HGAP g (1 + sin^2 e) / cos e); the field's reference momentum is p0c (the
 * reference always passes P0C = beam.init_energy).  Quadrupole: energy-
 * scaled linear map per step with second-order z terms and the low-energy z
 * correction.  Sextupole: drift - thin kick - drift per step.  Offsets and
 * tilt (quadrupole, sextupole) as X_OFFSET / Y_OFFSET / TILT.  In place. */
enum { DFCSR_ELE_DRIFT = 0, DFCSR_ELE_SBEND = 1, DFCSR_ELE_QUAD = 2, DFCSR_ELE_SEXTUPOLE = 3 };
typedef struct {
  int32_t type;        /* DFCSR_ELE_* */
  int32_t num_steps;   /* quadrupole / sextupole steps (>= 1) */
  double L;            /* length of this piece [m] */
  double g;            /* sbend: G = angle / L of the element [1/m] */
  double e1, e2;       /* sbend: face angles at this piece (0 where it has no face) */
  double fint, hgap;   /* sbend entrance fringe: FINT, HGAP (vertical edge focusing) */
  double fintx, hgapx; /* sbend exit fringe: FINTX, HGAPX */
  double k1;           /* quadrupole K1 [1/m^2] */
  double k2;           /* sextupole K2 [1/m^3] */
  double x_offset, y_offset, tilt;
} DfcsrElement;
int dfcsr_track_element(double *x, double *px, double *y, double *py, double *z, double *pz,
                        int64_t n, const DfcsrElement *ele, double p0c, double mc2,
                        void *stream);

/* ---- a16: mesh -> particle kick --------------------------------------------
 * Replaces Beam.apply_wakes (beams.py:108-131): bilinear
 * RegularGridInterpolator((xrange, zrange), step*W*1e6/E0, fill 0) at
 * (x - (slope0*z + slope1), z); pz += dE part; px += x part if transverse_on. */
int dfcsr_apply_wakes(double *px, double *pz, const double *x, const double *z, int64_t n,
                      const double *dE_dct, const double *x_kick, int xbins, int zbins,
                      double xr0, double xr1, double zr0, double zr1, double slope0,
                      double slope1, double step_size, double init_energy, int transverse_on,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif
