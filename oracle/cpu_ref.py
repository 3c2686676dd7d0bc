"""CPU oracle for the 2D CSR wake step -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product path (``pydfcsr_amd``) never imports anything under
``oracle/`` and fails loudly when its HIP library is missing.

It is a float64 numpy restatement of pyDFCSR's hot path, written from the
reference's source text (paths relative to the reference root
``pyDFCSR_2D/``).  Each function cites the lines it follows.  numpy/scipy
routines that the reference itself calls (``savgol_filter``, ``np.gradient``,
``np.trapz``, ``RegularGridInterpolator``, ``np.polyfit``) are called here the
same way, so the oracle inherits their exact arithmetic.

Pinning (what checks each function, see tests/test_oracle_*.py):
  histogram_cic_2d      brute force on hand-built particles, serial-loop form
                        bit-exact, golden vectors from the reference module.
  interpolate3D/1D      exact on multilinear functions, truncation/clamp/zero
                        special cases, golden vectors from the reference module.
  reference_trajectory  closed-form arcs, unit/orthogonal tau and n.
  get_DF                scipy savgol_filter / np.gradient composition; golden
                        vectors from the reference module.
  DFTracker             RegularGridInterpolator re-grid, window pop rule,
                        golden vectors from the reference module.
  csr_integrand/wake    zero-density and linearity invariants, scalar loop
                        form, golden vectors from the reference module.
  apply_wakes           RegularGridInterpolator closed form on bilinear fields.
  track_*               bmadx's Drift / SBend / Quadrupole single-particle results
                        held by the reference's test notebook (1e-12); the
                        symplectic condition of the maps' Jacobians; the linear
                        limit against the reference's r_gen6 matrices (golden
                        vectors: pole faces, the fringe term at FINT = 1/2,
                        tilt, the z row).  Sextupole: symplectic, feed-down to
                        the quadrupole map, second-order convergence to its
                        field equations; the choice of symmetric splitting
                        and FINT other than 1/2 are parity unpinned.
  sgolay2d              golden vectors from the reference module (every
                        derivative mode, minimal grids); the notebook's ramp
                        (test/test.ipynb cell 23); exact on polynomials of
                        the fit's degree away from the border; padding of a
                        ramp is its linear continuation.
  run_loop              golden vectors of the reference's own CSR2D.run on six
                        lattices (leading drift, nsep 2 and 5, elements shorter
                        than a step incl. the first and a one-piece bend, the
                        chicane workload), bmadx replaced by a deterministic
                        stand-in map: pieces, sigma_z, positions, formation
                        lengths and the history window bit-exact.
  beam_statistics       twiss: a beam built with known Twiss/dispersion (exact
                        sample moments), beta*gamma - alpha^2 = 1, emittance =
                        sqrt(det) of the dispersion-free matrix; energy
                        statistics against the definitions.
"""
from __future__ import annotations

import math
from collections import deque

import numpy as np
from scipy.interpolate import RegularGridInterpolator
from scipy.signal import savgol_filter

_trapz = getattr(np, "trapezoid", None) or np.trapz

# CSR_scaling = 8.98755e3 * charge[C]   (CSR.py:80)
CSR_SCALING_PER_COULOMB = 8.98755e3


# ---------------------------------------------------------------------------
# a3  histogram_cic_2d   (deposit.py:42-87)
# ---------------------------------------------------------------------------
def cic_cells(q1, q2, nbins_1, bins_start_1, bins_end_1, nbins_2, bins_start_2, bins_end_2):
    """Per-particle lower-cell indices and CIC fractions (deposit.py:55-75).

    The FP64 expressions are the reference's: ``cell = (q - start) * inv``
    with ``inv = 1 / ((end - start) / nbins)`` and ``i = floor(cell)``.
    """
    q1 = np.asarray(q1, dtype=np.float64)
    q2 = np.asarray(q2, dtype=np.float64)
    inv1 = 1.0 / ((bins_end_1 - bins_start_1) / nbins_1)
    inv2 = 1.0 / ((bins_end_2 - bins_start_2) / nbins_2)
    c1 = (q1 - bins_start_1) * inv1
    c2 = (q2 - bins_start_2) * inv2
    f1 = np.floor(c1)
    f2 = np.floor(c2)
    S1 = 1.0 - (c1 - f1)
    S2 = 1.0 - (c2 - f2)
    return f1.astype(np.int64), f2.astype(np.int64), S1, S2


def histogram_cic_2d(q1, q2, w, nbins_1, bins_start_1, bins_end_1,
                     nbins_2, bins_start_2, bins_end_2):
    """CIC deposition with the reference's serial accumulation order.

    deposit.py:65-85 adds, for each particle in turn, its four corner
    contributions ``w*S1*S2`` etc.  ``np.add.at`` with particle-major,
    corner-minor ordering reproduces the same FP64 sum order exactly.
    """
    i1, i2, S1, S2 = cic_cells(q1, q2, nbins_1, bins_start_1, bins_end_1,
                               nbins_2, bins_start_2, bins_end_2)
    w = np.asarray(w, dtype=np.float64)
    n = w.shape[0]
    vals = np.empty((n, 4))
    vals[:, 0] = w * S1 * S2
    vals[:, 1] = w * S1 * (1.0 - S2)
    vals[:, 2] = w * (1.0 - S1) * S2
    vals[:, 3] = w * (1.0 - S1) * (1.0 - S2)
    a = np.stack([i1, i1, i1 + 1, i1 + 1], axis=1)
    b = np.stack([i2, i2 + 1, i2, i2 + 1], axis=1)
    ok = (a >= 0) & (a < nbins_1) & (b >= 0) & (b < nbins_2)
    hist = np.zeros((nbins_1, nbins_2), dtype=np.float64)
    flat = (a * nbins_2 + b)[ok]
    np.add.at(hist.reshape(-1), flat, vals[ok])
    return hist


def histogram_cic_2d_loop(q1, q2, w, nbins_1, bins_start_1, bins_end_1,
                          nbins_2, bins_start_2, bins_end_2):
    """Literal scalar-loop form of deposit.py:54-87, for tiny inputs."""
    bin_spacing_1 = (bins_end_1 - bins_start_1) / nbins_1
    inv_spacing_1 = 1. / bin_spacing_1
    bin_spacing_2 = (bins_end_2 - bins_start_2) / nbins_2
    inv_spacing_2 = 1. / bin_spacing_2
    hist = np.zeros((nbins_1, nbins_2))
    for i in range(len(w)):
        c1 = (q1[i] - bins_start_1) * inv_spacing_1
        c2 = (q2[i] - bins_start_2) * inv_spacing_2
        i1 = int(math.floor(c1))
        i2 = int(math.floor(c2))
        S1 = 1. - (c1 - i1)
        S2 = 1. - (c2 - i2)
        if 0 <= i1 < nbins_1:
            if 0 <= i2 < nbins_2:
                hist[i1, i2] += w[i] * S1 * S2
            if 0 <= i2 + 1 < nbins_2:
                hist[i1, i2 + 1] += w[i] * S1 * (1. - S2)
        if 0 <= i1 + 1 < nbins_1:
            if 0 <= i2 < nbins_2:
                hist[i1 + 1, i2] += w[i] * (1. - S1) * S2
            if 0 <= i2 + 1 < nbins_2:
                hist[i1 + 1, i2 + 1] += w[i] * (1. - S1) * (1. - S2)
    return hist


def cell_counts(i1, i2, nbins_1, nbins_2):
    """Integer count of particles whose lower cell (i1, i2) is on the grid."""
    ok = (i1 >= 0) & (i1 < nbins_1) & (i2 >= 0) & (i2 < nbins_2)
    cnt = np.zeros(nbins_1 * nbins_2, dtype=np.int64)
    np.add.at(cnt, (i1 * nbins_2 + i2)[ok], 1)
    return cnt.reshape(nbins_1, nbins_2)


# ---------------------------------------------------------------------------
# a1  beam moments   (beams.py:88-98, 137-215; deposit.py:147-159)
# ---------------------------------------------------------------------------
def beam_moments(x, z):
    """sigma/mean/slope/x_transform statistics the step reads.

    beams.py:201-204 ``slope = np.polyfit(z, x, 1)``; beams.py:206-211
    ``x_transform = x - np.polyval(slope, z)``; CSR.py:371-372 std/mean of
    x_transform; deposit.py:157-159 the |z| < 0.1 sigma_z slice std.
    """
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    sigma_x = np.std(x)
    sigma_z = np.std(z)
    p = np.polyfit(z, x, deg=1)
    xt = x - np.polyval(p, z)
    slice_ind = np.argwhere(np.abs(z) < 0.1 * sigma_z)
    with np.errstate(invalid="ignore", divide="ignore"):
        slice_sig = np.std(x[slice_ind]) if slice_ind.size else np.nan
    return dict(mean_x=np.mean(x), mean_z=np.mean(z), sigma_x=sigma_x, sigma_z=sigma_z,
                slope=p, mean_xt=np.mean(xt), sigma_xt=np.std(xt), slice_sigma_x=slice_sig,
                n_slice=int(slice_ind.shape[0]))


# ---------------------------------------------------------------------------
# per-step statistics   (CSR.py:837-859; twiss.py:2-70; beams.py:166-181)
# ---------------------------------------------------------------------------
def twiss_dispersion_calc(sigma3):
    """Twiss and dispersion of one plane from the 3 x 3 covariance of
    (u, pu, delta), twiss.py:30-70: the dispersion-free 2 x 2 matrix
    [[eb, -ea], [-ea, eg]] = Sigma_uu - eta eta^T delta2 and its emittance."""
    delta2 = sigma3[2, 2]
    xd = sigma3[0, 2]
    pd = sigma3[1, 2]
    eb = sigma3[0, 0] - xd ** 2 / delta2
    eg = sigma3[1, 1] - pd ** 2 / delta2
    ea = -sigma3[0, 1] + xd * pd / delta2
    emit = np.sqrt(eb * eg - ea ** 2)
    return dict(alpha=ea / emit, beta=eb / emit, gamma=eg / emit, emit=emit, eta=xd / delta2,
                etap=pd / delta2)


def beam_statistics(x, px, y, py, z, pz, p0c, mc2):
    """The per-step record of CSR2D.update_statistics (CSR.py:837-859):
    twiss_from_bmadx_particles (twiss.py:2-27: np.cov of (x, px, pz) and
    (y, py, pz), norm_emit = emit p0c / mc2) and the energy (pz + 1) p0c's
    mean and np.std (beams.py:166-181)."""
    out = {}
    for plane, (u, pu) in (("x", (x, px)), ("y", (y, py))):
        tw = twiss_dispersion_calc(np.cov([u, pu, pz]))
        tw["norm_emit"] = tw["emit"] * p0c / mc2
        for k, v in tw.items():
            out[f"{k}_{plane}"] = v
    energy = (np.asarray(pz) + 1) * p0c
    return dict(twiss=out, mean_energy=np.mean(energy), sigma_energy=np.std(energy),
                cov=np.cov([x, px, y, py, z, pz]))


# ---------------------------------------------------------------------------
# tracking: bmadx.track_element as CSR2D calls it (CSR.py:146-199, 227-294)
# ---------------------------------------------------------------------------
# bmadx (Bmad-X, git+https://github.com/bmad-sim/Bmad-X/, environment.yml:24,
# no version pinned) is absent.  These are the published Bmad maps for the
# element types the reference builds, in Bmad's canonical coordinates
# (x, px = Px/P0, y, py, z = -beta c (t - t_ref), pz = dP/P0), restated from the
# physics, and pinned on the Drift / SBend / Quadrupole single-particle results
# the reference's test notebook holds (test/test_BmadX_tracking.ipynb cells
# 25, 28, 31; tests/golden/ref_bmadx_tracking.json) and on the reference's
# r_gen6 matrices in the linear limit (tests/golden/ref_r_gen6.npz).
# Sextupole: no reference result exists; its field equations pin it up to
# the choice of symmetric splitting (parity unpinned there).
def sqrt_one(v):
    """sqrt(1 + v) - 1 without cancellation."""
    return v / (np.sqrt(1 + v) + 1)


def low_energy_z_correction(pz, p0c, mc2, ds):
    """ds (beta/beta0 - 1): the z slip of a particle slower than the
    reference; its pz series where mc2 (beta0 pz)^2 < 3e-7 E_tot."""
    pz = np.asarray(pz, dtype=np.float64)
    beta = (1 + pz) * p0c / np.sqrt(((1 + pz) * p0c) ** 2 + mc2 ** 2)
    beta0 = p0c / np.sqrt(p0c ** 2 + mc2 ** 2)
    e_tot = np.sqrt(p0c ** 2 + mc2 ** 2)
    series = ds * pz * (1 - 3 * (pz * beta0 ** 2) / 2 + pz ** 2 * beta0 ** 2 *
                        (2 * beta0 ** 2 - (mc2 / e_tot) ** 2 / 2)) * (mc2 / e_tot) ** 2
    exact = ds * (beta - beta0) / beta0
    return np.where(mc2 * (beta0 * pz) ** 2 < 3e-7 * e_tot, series, exact)


def track_drift(c, L, p0c, mc2):
    """Exact field-free drift of length L; c = (x, px, y, py, z, pz)."""
    x, px, y, py, z, pz = (np.asarray(v, dtype=np.float64) for v in c)
    P = 1 + pz
    Px, Py = px / P, py / P
    pxy2 = Px ** 2 + Py ** 2
    Pl = np.sqrt(1 - pxy2)
    x = x + L * Px / Pl
    y = y + L * Py / Pl
    # z + L (beta/beta0 - 1/Pl), both differences without cancellation
    dz = L * (sqrt_one((mc2 ** 2 * (2 * pz + pz ** 2)) / ((p0c * P) ** 2 + mc2 ** 2)) +
              sqrt_one(-pxy2) / Pl)
    return np.stack([x, px, y, py, z + dz, pz])


def _quad_plane(k, L, P):
    """One plane of a quadrupole, u'' = -k u (k > 0 focusing), u' = pu/P:
    the 2 x 2 map and the path-length coefficients of -1/2 int u'^2 ds in
    (u^2, u pu, pu^2)."""
    sk = np.sqrt(np.abs(k))
    th = sk * L
    with np.errstate(invalid="ignore", divide="ignore"):
        cx = np.where(k > 0, np.cos(th), np.where(k < 0, np.cosh(th), 1.0))
        sx = np.where(k > 0, np.sin(th) / sk, np.where(k < 0, np.sinh(th) / sk, L))
    m = (cx, sx / P, -k * sx * P, cx)
    d = (k * (cx * sx - L) / 4, k * sx ** 2 / (2 * P), -(L + cx * sx) / (4 * P ** 2))
    return m, d


def _offset_set(c, x_off, y_off, tilt):
    x, px, y, py, z, pz = c
    x, y = x - x_off, y - y_off
    ct, st = np.cos(tilt), np.sin(tilt)
    return (x * ct + y * st, px * ct + py * st, -x * st + y * ct, -px * st + py * ct, z, pz)


def _offset_unset(c, x_off, y_off, tilt):
    x, px, y, py, z, pz = c
    ct, st = np.cos(tilt), np.sin(tilt)
    x, px, y, py = x * ct - y * st, px * ct - py * st, x * st + y * ct, px * st + py * ct
    return (x + x_off, px, y + y_off, py, z, pz)


def track_quadrupole(c, L, K1, p0c, mc2, num_steps=1, x_off=0.0, y_off=0.0, tilt=0.0):
    """Quadrupole: energy-scaled linear map per step, second-order z terms
    of the path length, low-energy z correction."""
    x, px, y, py, z, pz = _offset_set(tuple(np.asarray(v, dtype=np.float64) for v in c),
                                      x_off, y_off, tilt)
    ds = L / num_steps
    for _ in range(num_steps):
        P = 1 + pz
        k = K1 / P
        tx, dx = _quad_plane(k, ds, P)
        ty, dy = _quad_plane(-k, ds, P)
        z = (z + dx[0] * x ** 2 + dx[1] * x * px + dx[2] * px ** 2 +
             dy[0] * y ** 2 + dy[1] * y * py + dy[2] * py ** 2)
        x, px = tx[0] * x + tx[1] * px, tx[2] * x + tx[3] * px
        y, py = ty[0] * y + ty[1] * py, ty[2] * y + ty[3] * py
        z = z + low_energy_z_correction(pz, p0c, mc2, ds)
    return np.stack(_offset_unset((x, px, y, py, z, pz), x_off, y_off, tilt))


def _bend_edge(c, g, e, fint=0.0, hgap=0.0):
    """Hard-edge pole face rotated by e (linear edge kicks); the fringe
    integral softens the vertical focusing by psi = fint (2 hgap) g
    (1 + sin^2 e) / cos e (r_gen6.py:49-56 at FINT = 1/2; other FINT values:
    parity unpinned)."""
    x, px, y, py, z, pz = c
    psi = fint * (2 * hgap) * g * (1 + np.sin(e) ** 2) / np.cos(e)
    if e == 0 and psi == 0:
        return (x, px, y, py, z, pz)
    return (x, px + x * g * np.tan(e), y, py - y * g * np.tan(e - psi), z, pz)


def track_sbend(c, L, g, e1, e2, p0c, mc2, fint=0.0, hgap=0.0, fintx=0.0, hgapx=0.0):
    """Sector bend of curvature g (= angle / L) in its uniform field, exact:
    the particle's circle of radius P_h / g meets the exit face at angle
    g L; entrance and exit faces rotated by e1, e2.  Written without the
    cancellations of r - 1/g: px1 = px cos + (p_s - 1 - g x) sin,
    x1 = x cos + ((p_s1 - 1) - cos (p_s - 1) + px sin) / g,
    phi = g L + asin(px / P_h) - asin(px1 / P_h), path = phi P / g."""
    c = tuple(np.asarray(v, dtype=np.float64) for v in c)
    if g == 0:
        return np.stack(_bend_edge(tuple(track_drift(_bend_edge(c, g, e1, fint, hgap), L, p0c,
                                                     mc2)), g, e2, fintx, hgapx))
    x, px, y, py, z, pz = _bend_edge(c, g, e1, fint, hgap)
    P = 1 + pz
    th = g * L
    ct, st = np.cos(th), np.sin(th)
    pxy2 = px ** 2 + py ** 2
    ps_1 = sqrt_one(2 * pz + pz ** 2 - pxy2)
    Ph = np.sqrt(P ** 2 - py ** 2)
    px1 = px * ct + (ps_1 - g * x) * st
    ps1_1 = sqrt_one(2 * pz + pz ** 2 - py ** 2 - px1 ** 2)
    x1 = x * ct + (ps1_1 - ct * ps_1 + px * st) / g
    dphi = np.arcsin(px / Ph) - np.arcsin(px1 / Ph)
    y1 = y + py * (L + dphi / g)
    beta = P * p0c / np.sqrt((P * p0c) ** 2 + mc2 ** 2)
    beta0 = p0c / np.sqrt(p0c ** 2 + mc2 ** 2)
    z1 = z + L * beta / beta0 - P * (L + dphi / g)
    out = _bend_edge((x1, px1, y1, py, z1, P - 1), g, e2, fintx, hgapx)
    return np.stack(out)


def track_sextupole(c, L, K2, p0c, mc2, num_steps=1, x_off=0.0, y_off=0.0, tilt=0.0):
    """Sextupole as drift - thin kick - drift per step (the kick of K2 ds:
    dpx = -K2 ds (x^2 - y^2) / 2, dpy = K2 ds x y).  Pinned to its field
    equations; drift-kick-drift rather than kick-drift-kick is parity
    unpinned."""
    cc = _offset_set(tuple(np.asarray(v, dtype=np.float64) for v in c), x_off, y_off, tilt)
    ds = L / num_steps
    for _ in range(num_steps):
        cc = tuple(track_drift(cc, ds / 2, p0c, mc2))
        x, px, y, py, z, pz = cc
        cc = (x, px - K2 * ds * (x * x - y * y) / 2, y, py + K2 * ds * x * y, z, pz)
        cc = tuple(track_drift(cc, ds / 2, p0c, mc2))
    return np.stack(_offset_unset(cc, x_off, y_off, tilt))


# ---------------------------------------------------------------------------
# linear maps of element pieces (r_gen6.py:4-122), for linear-optics checks
# ---------------------------------------------------------------------------
def transfer_matrix(L, angle=0.0, k1=0.0, E1=0.0, E2=0.0, hgap=0.0):
    """6x6 linear map (x, px, y, py, z, pz) of a bend/quad/drift piece: the
    linear optics r_gen6.py:4-122 computes, written from the textbook
    sector-bend and quadrupole matrices with pole-face kicks (the vertical
    one corrected by the fringe term, FINT = 1/2)."""
    R = np.eye(6)
    if L == 0:
        return R
    h = angle / L
    kx2 = k1 + h * h
    ky2 = -k1

    def plane(k2):
        k = math.sqrt(abs(k2))
        ph = k * L
        if abs(ph) < 1e-12:
            return np.array([[1.0, L], [0.0, 1.0]]), None, None, None
        if k2 > 0:
            co, si = math.cos(ph), math.sin(ph)
            return np.array([[co, si / k], [-k * si, co]]), co, si, k
        co, si = math.cosh(ph), math.sinh(ph)
        return np.array([[co, si / k], [k * si, co]]), co, si, k

    Rx, co, si, kx = plane(kx2)
    Ry = plane(ky2)[0]
    R[0:2, 0:2] = Rx
    R[2:4, 2:4] = Ry
    if co is not None:
        D0 = h * (1 - co) / kx2
        D1 = h * si / kx
        R[0, 5] = D0
        R[1, 5] = D1
        R[4, 0] = -D1
        R[4, 1] = -D0
        R[4, 5] = -(h ** 2) * (kx * L - kx * Rx[0, 1]) / kx ** 3
    if angle:
        def face(e):
            F = np.eye(6)
            psi = h * hgap * (1 + math.sin(e) ** 2) / math.cos(e)
            F[1, 0] = math.tan(e) * h
            F[3, 2] = -math.tan(e - psi) * h
            return F
        R = face(E2) @ R @ face(E1)
    return R



# ---------------------------------------------------------------------------
# a4-a6  DF_tracker.get_DF   (deposit.py:145-245)
# ---------------------------------------------------------------------------
def _sg2(a, window, order):
    """deposit.py:195-196: savgol along axis 0, then axis 1 (mode='interp')."""
    return savgol_filter(savgol_filter(a, window_length=window, polyorder=order, axis=0),
                         window_length=window, polyorder=order, axis=1)


def get_DF(x, z, px, t, xbins=100, zbins=100, xlim=5, zlim=5, filter_order=0,
           filter_window=0, velocity_threhold=5, **_unused):
    """Deposition + density-function post-processing (deposit.py:145-245)."""
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    px = np.asarray(px, dtype=np.float64)
    sigma_x = np.std(x)
    sigma_z = np.std(z)
    xmean = np.mean(x)
    zmean = np.mean(z)
    slice_ind = np.argwhere(np.abs(z) < 0.1 * sigma_z)
    with np.errstate(invalid="ignore", divide="ignore"):
        slice_sigX = np.std(x[slice_ind]) if slice_ind.size else np.nan
        frac = sigma_x / slice_sigX
    if frac > 5:                                    # deposit.py:160-167
        xb, zb, window = xbins, zbins, filter_window
    else:
        xb, zb, window = 100, 100, 5
    x0, x1 = xmean - xlim * sigma_x, xmean + xlim * sigma_x
    z0, z1 = zmean - zlim * sigma_z, zmean + zlim * sigma_z
    x_grids = np.linspace(x0, x1, xb)
    z_grids = np.linspace(z0, z1, zb)
    density = histogram_cic_2d(x, z, np.ones(x.shape), xb, x0, x1, zb, z0, z1)
    vx = histogram_cic_2d(x, z, px, xb, x0, x1, zb, z0, z1)
    threshold = np.max(density) / velocity_threhold
    m = density > threshold
    vx[m] /= density[m]
    density = _sg2(density, window, filter_order)
    vx = _sg2(vx, window, filter_order)
    dsum = _trapz(_trapz(density, x_grids, axis=0), z_grids)
    density /= dsum
    vx[density <= threshold] = 0                    # raw-count threshold (deposit.py:204)
    density_x, density_z = np.gradient(density, x_grids, z_grids)
    vx_x, _vx_z = np.gradient(vx, x_grids, z_grids)
    density_x = _sg2(density_x, window, filter_order)
    density_z = _sg2(density_z, window, filter_order)
    vx_x = _sg2(vx_x, window, filter_order)
    threshold = np.max(density) / velocity_threhold * 8
    with np.errstate(invalid="ignore"):
        fill = np.mean(vx_x[density > threshold]) if np.any(density > threshold) else np.nan
    vx_x[density < threshold] = fill
    return dict(x_grids=x_grids, z_grids=z_grids, density=density, vx=vx,
                density_x=density_x, density_z=density_z, vx_x=vx_x, t=t,
                sigma_x=sigma_x, sigma_z=sigma_z, xmean=xmean, zmean=zmean,
                frac=frac, nbins=(xb, zb), window=window,
                x_range=(x0, x1), z_range=(z0, z1))


FIELDS = ("density", "density_x", "density_z", "vx", "vx_x")


# ---------------------------------------------------------------------------
# a7  SGolay_filter.sgolay2d   (SGolay_filter.py:3-81; get_DF's commented
#     alternative, deposit.py:187, 194, 227, 232)
# ---------------------------------------------------------------------------
def sgolay2d_pad(z, h):
    """The grid padded by h on every side (SGolay_filter.py:36-66), by
    distance k = 1..h from each edge: top/left bands mirror row/column k about
    the edge value from below (e - |z_k - e|), bottom/right bands from above
    (e + |z_k - e|); the top-left and bottom-right corners mirror about the
    corner value; the top-right corner mirrors the right band about its first
    row, the bottom-left corner the bottom band about its first column."""
    z = np.asarray(z, dtype=np.float64)
    nx, nz = z.shape
    Z = np.zeros((nx + 2 * h, nz + 2 * h))
    Z[h:h + nx, h:h + nz] = z
    for k in range(1, h + 1):
        Z[h - k, h:h + nz] = z[0] - np.abs(z[k] - z[0])
        Z[h + nx - 1 + k, h:h + nz] = z[-1] + np.abs(z[nx - 1 - k] - z[-1])
        Z[h:h + nx, h - k] = z[:, 0] - np.abs(z[:, k] - z[:, 0])
        Z[h:h + nx, h + nz - 1 + k] = z[:, -1] + np.abs(z[:, nz - 1 - k] - z[:, -1])
    for k1 in range(1, h + 1):
        for k2 in range(1, h + 1):
            e = z[0, 0]
            Z[h - k1, h - k2] = e - abs(z[k1, k2] - e)
            e = z[-1, -1]
            Z[h + nx - 1 + k1, h + nz - 1 + k2] = e + abs(z[nx - 1 - k1, nz - 1 - k2] - e)
            col = h + nz - 1 + k2                  # right band, real rows 0 and k1
            e = Z[h, col]
            Z[h - k1, col] = e - abs(Z[h + k1, col] - e)
            row = h + nx - 1 + k1                  # bottom band, real columns 0 and k2
            e = Z[row, h]
            Z[row, h - k2] = e - abs(Z[row, h + k2] - e)
    return Z


def sgolay2d_fit_terms(order):
    """(x power, y power) of the fit's monomials in the reference's order
    (SGolay_filter.py:24): by degree, x power falling within a degree."""
    return [(deg - n, n) for deg in range(order + 1) for n in range(deg + 1)]


def sgolay2d(z, window_size, order, derivative=None):
    """SGolay_filter.sgolay2d: least-squares fit of a degree-`order` 2D
    polynomial over each window, evaluated as convolutions of the padded
    grid with rows of pinv(A), A[p, t] = dx_p^ex_t dy_p^ey_t at window
    offsets dx (first axis, slow) and dy; 'col' / 'row' convolve with the
    negated slope rows 1 / 2, 'both' returns (col, row)."""
    if window_size % 2 == 0:
        raise ValueError("window_size must be odd")
    terms = sgolay2d_fit_terms(order)
    if window_size ** 2 < len(terms):
        raise ValueError("order is too high for the window size")
    from scipy.signal import fftconvolve
    h = window_size // 2
    off = np.arange(-h, h + 1, dtype=np.float64)
    dx, dy = np.meshgrid(off, off, indexing="ij")
    A = np.stack([dx.ravel() ** ex * dy.ravel() ** ey for ex, ey in terms], axis=1)
    Ainv = np.linalg.pinv(A)
    Z = sgolay2d_pad(z, h)
    kern = lambda t, sign: sign * Ainv[t].reshape(window_size, window_size)
    if derivative is None:
        return fftconvolve(Z, kern(0, 1.0), mode="valid")
    if derivative == "col":
        return fftconvolve(Z, kern(1, -1.0), mode="valid")
    if derivative == "row":
        return fftconvolve(Z, kern(2, -1.0), mode="valid")
    if derivative == "both":
        return (fftconvolve(Z, kern(1, -1.0), mode="valid"),
                fftconvolve(Z, kern(2, -1.0), mode="valid"))
    raise ValueError(f"unknown derivative {derivative!r}")


# ---------------------------------------------------------------------------
# a8-a10  DF history   (deposit.py:247-426)
# ---------------------------------------------------------------------------
class DFTracker:
    """History of DF slices, re-gridding and 3D stacking (deposit.py:90-426)."""

    def __init__(self, xbins=100, zbins=100, xlim=5, zlim=5, filter_order=0,
                 filter_window=0, velocity_threhold=5, upper_limit=None):
        self.cfg = dict(xbins=xbins, zbins=zbins, xlim=xlim, zlim=zlim,
                        filter_order=filter_order, filter_window=filter_window,
                        velocity_threhold=velocity_threhold)
        self.upper_limit = upper_limit
        self.start_time = 0.0
        self.interp_start = 0
        self.DF_log = deque()
        self.time_log = deque()
        self.sigma_x_log = deque()
        self.sigma_z_log = deque()
        self.time_interp = deque()
        self.interp = {f: deque() for f in FIELDS}
        self.sigma_x_interp = None
        self.sigma_z_interp = None
        self.x_grid_interp = None
        self.z_grid_interp = None
        self.cur = None
        self.n_regrid = 0

    def get_DF(self, x, z, px, t):
        self.cur = get_DF(x, z, px, t, **self.cfg)
        return self.cur

    def append_DF(self):                                 # deposit.py:247-256
        c = self.cur
        self.DF_log.append(c)
        self.time_log.append(c["t"])
        self.sigma_x_log.append(c["sigma_x"])
        self.sigma_z_log.append(c["sigma_z"])
        self.end_time = c["t"]

    def pop_left_DF(self, new_start_time):               # deposit.py:258-280
        while self.start_time < new_start_time:
            self.DF_log.popleft()
            self.time_log.popleft()
            self.sigma_x_log.popleft()
            self.sigma_z_log.popleft()
            self.start_time = self.time_log[0]
        while self.interp_start < new_start_time:
            for f in FIELDS:
                self.interp[f].popleft()
            self.time_interp.popleft()
            self.interp_start = self.time_interp[0]

    def DF_interp(self, DF, x_grids, z_grids, fill_value=0.0):   # deposit.py:296-309
        X, Z = np.meshgrid(self.x_grid_interp, self.z_grid_interp, indexing="ij")
        f = RegularGridInterpolator((x_grids, z_grids), DF, method="linear",
                                    fill_value=fill_value, bounds_error=False)
        return f((X, Z))

    def _interp_slice(self, c):
        out = {}
        for f in FIELDS:
            fill = np.mean(c["vx_x"]) if f == "vx_x" else 0.0
            out[f] = self.DF_interp(c[f], c["x_grids"], c["z_grids"], fill_value=fill)
        return out

    def append_interpolant(self, formation_length, n_formation_length):   # deposit.py:312-392
        start_point = np.amax(a=(0, self.end_time - n_formation_length * formation_length))
        self.pop_left_DF(new_start_time=start_point)
        c = self.cur
        if (self.sigma_x_interp and self.sigma_z_interp
                and 2 > c["sigma_x"] / self.sigma_x_interp > 1 / 2
                and 2 > c["sigma_z"] / self.sigma_z_interp > 1 / 2):
            self.time_interp.append(c["t"])
            sl = self._interp_slice(c)
            for f in FIELDS:
                self.interp[f].append(sl[f])
            return False
        max_sx, min_sx = np.max(self.sigma_x_log), np.min(self.sigma_x_log)
        max_sz, min_sz = np.max(self.sigma_z_log), np.min(self.sigma_z_log)
        xbins = int(500 * (max_sx / min_sx))
        zbins = int(500 * (max_sz / min_sz))
        if isinstance(self.upper_limit, int):
            xbins = min(xbins, self.upper_limit)
            zbins = min(zbins, self.upper_limit)
        self.sigma_x_interp = max_sx
        self.sigma_z_interp = max_sz
        self.x_grid_interp = np.linspace(c["xmean"] - 5 * max_sx, c["xmean"] + 5 * max_sx, xbins)
        self.z_grid_interp = np.linspace(c["zmean"] - 5 * max_sz, c["zmean"] + 5 * max_sz, zbins)
        self.interp = {f: deque() for f in FIELDS}
        self.time_interp = self.time_log.copy()
        for d in self.DF_log:
            sl = self._interp_slice(d)
            for f in FIELDS:
                self.interp[f].append(sl[f])
        self.n_regrid += 1
        return True

    def build_interpolant(self):                         # deposit.py:395-426
        self.min_x, self.max_x = self.time_interp[0], self.time_interp[-1]
        self.min_y, self.max_y = self.x_grid_interp[0], self.x_grid_interp[-1]
        self.min_z, self.max_z = self.z_grid_interp[0], self.z_grid_interp[-1]
        with np.errstate(invalid="ignore", divide="ignore"):
            self.delta_x = (self.max_x - self.min_x) / (len(self.time_interp) - 1)
        self.delta_y = (self.max_y - self.min_y) / (self.x_grid_interp.shape[0] - 1)
        self.delta_z = (self.max_z - self.min_z) / (self.z_grid_interp.shape[0] - 1)
        self.data = {f: np.array(self.interp[f]) for f in FIELDS}

    def axes(self):
        return dict(min_t=self.min_x, min_x=self.min_y, min_z=self.min_z,
                    dt=self.delta_x, dx=self.delta_y, dz=self.delta_z)


# ---------------------------------------------------------------------------
# a11  interpolate3D   (interp3D.py:18-66)   a12  interpolate1D (interp1D.py:13-36)
# ---------------------------------------------------------------------------
def _trunc_index(u, size):
    """int(u) truncation with the reference's validity rule.

    interp3D.py:30-52: ``i0 = int(u)``, ``i1 = i0`` at the last plane else
    ``i0 + 1``; valid iff ``i0 >= 0 and i1 < size``.  NaN is invalid.
    """
    with np.errstate(invalid="ignore"):
        f = np.trunc(u)
        ok = (f >= 0) & (f <= size - 1)
    i0 = np.where(ok, f, 0).astype(np.int64)
    i1 = np.where(i0 == size - 1, i0, i0 + 1)
    d = np.where(ok, u - i0, 0.0)
    return i0, i1, d, ok


def interpolate3D(xval, yval, zval, data, min_x, min_y, min_z, delta_x, delta_y, delta_z):
    """Vectorised interp3D.py:18-66 (same arithmetic, same order)."""
    xval = np.asarray(xval, dtype=np.float64)
    nx, ny, nz = data.shape
    with np.errstate(invalid="ignore", divide="ignore"):
        u = (xval - min_x) / delta_x
        v = (np.asarray(yval, dtype=np.float64) - min_y) / delta_y
        w = (np.asarray(zval, dtype=np.float64) - min_z) / delta_z
    x0, x1, xd, okx = _trunc_index(u, nx)
    y0, y1, yd, oky = _trunc_index(v, ny)
    z0, z1, zd, okz = _trunc_index(w, nz)
    ok = okx & oky & okz
    c00 = data[x0, y0, z0] * (1 - xd) + data[x1, y0, z0] * xd
    c01 = data[x0, y0, z1] * (1 - xd) + data[x1, y0, z1] * xd
    c10 = data[x0, y1, z0] * (1 - xd) + data[x1, y1, z0] * xd
    c11 = data[x0, y1, z1] * (1 - xd) + data[x1, y1, z1] * xd
    c0 = c00 * (1 - yd) + c10 * yd
    c1 = c01 * (1 - yd) + c11 * yd
    return np.where(ok, c0 * (1 - zd) + c1 * zd, 0.0)


def interpolate3D_loop(xval, yval, zval, data, min_x, min_y, min_z, delta_x, delta_y, delta_z):
    """Literal scalar-loop form of interp3D.py:18-66, for tiny inputs."""
    result = np.zeros(len(xval))
    xs, ys, zs = data.shape
    for i in range(len(xval)):
        x = (xval[i] - min_x) / delta_x
        y = (yval[i] - min_y) / delta_y
        z = (zval[i] - min_z) / delta_z
        if not (np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
            continue
        x0 = int(x); x1 = x0 if x0 == xs - 1 else x0 + 1
        y0 = int(y); y1 = y0 if y0 == ys - 1 else y0 + 1
        z0 = int(z); z1 = z0 if z0 == zs - 1 else z0 + 1
        xd, yd, zd = x - x0, y - y0, z - z0
        if x0 >= 0 and y0 >= 0 and z0 >= 0 and x1 < xs and y1 < ys and z1 < zs:
            c00 = data[x0, y0, z0] * (1 - xd) + data[x1, y0, z0] * xd
            c01 = data[x0, y0, z1] * (1 - xd) + data[x1, y0, z1] * xd
            c10 = data[x0, y1, z0] * (1 - xd) + data[x1, y1, z0] * xd
            c11 = data[x0, y1, z1] * (1 - xd) + data[x1, y1, z1] * xd
            c0 = c00 * (1 - yd) + c10 * yd
            c1 = c01 * (1 - yd) + c11 * yd
            result[i] = c0 * (1 - zd) + c1 * zd
    return result


def interpolate1D(xval, data, min_x, delta_x):
    """Vectorised interp1D.py:13-36 (out-of-range -> 0)."""
    u = (np.asarray(xval, dtype=np.float64) - min_x) / delta_x
    x0, x1, xd, ok = _trunc_index(u, data.shape[0])
    return np.where(ok, data[x0] * (1 - xd) + data[x1] * xd, 0.0)


# ---------------------------------------------------------------------------
# a12  lattice tables   (lattice.py:4-110, 136-173)
# ---------------------------------------------------------------------------
def reference_trajectory(elements, Nsample=2000):
    """Reference orbit tables; ``elements`` is the ordered list of element
    dicts (the lattice YAML without its leading ``step_size`` key).

    Returns s, rho, distance, nsep, coords, n_vec, tau_vec (lattice.py:4-110).
    """
    Nel = len(elements)
    distance = np.zeros(Nel)
    rho = np.zeros(Nel)
    nsep = np.zeros(Nel)
    for k, e in enumerate(elements):
        L = e["L"]
        nsep[k] = e["nsep"]
        if e["type"] == "dipole":
            rho[k] = e["angle"] / L
        distance[k] = L if k == 0 else L + distance[k - 1]
    s = np.linspace(0, distance[-1], Nsample)
    coords = np.zeros((Nsample, 2))
    tau = np.zeros((Nsample, 2))
    nv = np.zeros((Nsample, 2))
    th = 0.0
    tau[0] = (np.cos(th), np.sin(th))
    nv[0] = (np.sin(th), -1 * np.cos(th))
    ei = 0
    for cnt in range(1, Nsample):
        st = s[cnt]
        if st > distance[ei]:
            ei += 1
        ds = s[cnt] - s[cnt - 1]
        e = elements[ei]
        if e["type"] == "dipole":
            L, ang = e["L"], e["angle"]
            phi = ds / L * ang
            r = L / ang
            x0 = coords[cnt - 1, 0] - r * np.sin(th)
            y0 = coords[cnt - 1, 1] + r * np.cos(th)
            coords[cnt, 0] = x0 + r * np.sin(phi + th)
            coords[cnt, 1] = y0 - r * np.cos(phi + th)
        else:
            phi = 0
            coords[cnt, 0] = coords[cnt - 1, 0] + ds * np.cos(th)
            coords[cnt, 1] = coords[cnt - 1, 1] + ds * np.sin(th)
        th += phi
        tau[cnt] = (np.cos(th), np.sin(th))
        nv[cnt] = (np.sin(th), -1 * np.cos(th))
    return s, rho, distance, nsep, coords, nv, tau


class LatticeTables:
    """Lattice tables + 1D interpolant axis (lattice.py:136-150)."""

    def __init__(self, elements, Nsample=2000):
        (self.s, self.rho, self.distance, self.nsep, self.coords, self.n_vec,
         self.tau_vec) = reference_trajectory(elements, Nsample)
        self.Nelement = len(elements)
        self.min_x, self.max_x = self.s[0], self.s[-1]
        self.delta_x = (self.max_x - self.min_x) / (self.s.shape[0] - 1)


# ---------------------------------------------------------------------------
# a13  get_CSR_integrand   (CSR.py:605-782)
# ---------------------------------------------------------------------------
def csr_integrand(s, x, t, sp, xp, lat, hist):
    """Integrand grids (I_z, I_x) for one observer (CSR.py:605-782).

    ``hist`` is a DFTracker after build_interpolant(); ``lat`` a LatticeTables.
    """
    a = hist.axes()

    def i3(f, tv, xv, zv):
        return interpolate3D(tv, xv, zv, hist.data[f], a["min_t"], a["min_x"], a["min_z"],
                             a["dt"], a["dx"], a["dz"])

    vx = i3("vx", np.array([t]), np.array([x]), np.array([s - t]))[0]
    spf = sp.ravel()
    xpf = xp.ravel()

    def i1(col, v):
        return interpolate1D(v, col, lat.min_x, lat.delta_x)

    sv = np.array([s])
    X0_s, X0_sp = i1(lat.coords[:, 0], sv)[0], i1(lat.coords[:, 0], spf)
    Y0_s, Y0_sp = i1(lat.coords[:, 1], sv)[0], i1(lat.coords[:, 1], spf)
    nsx, nspx = i1(lat.n_vec[:, 0], sv)[0], i1(lat.n_vec[:, 0], spf)
    nsy, nspy = i1(lat.n_vec[:, 1], sv)[0], i1(lat.n_vec[:, 1], spf)
    tsx, tspx = i1(lat.tau_vec[:, 0], sv)[0], i1(lat.tau_vec[:, 0], spf)
    tsy, tspy = i1(lat.tau_vec[:, 1], sv)[0], i1(lat.tau_vec[:, 1], spf)

    rx = X0_s - X0_sp + x * nsx - xpf * nspx
    ry = Y0_s - Y0_sp + x * nsy - xpf * nspy
    R = np.sqrt(rx ** 2 + ry ** 2)

    rho_sp = np.zeros(spf.shape)                     # CSR.py:651-656
    for k in range(lat.Nelement):
        if k == 0:
            rho_sp[spf < lat.distance[k]] = lat.rho[k]
        else:
            rho_sp[(spf < lat.distance[k]) & (spf >= lat.distance[k - 1])] = lat.rho[k]

    t_ret = t - R
    zq = spf - t_ret
    dens = i3("density", t_ret, xpf, zq)
    dens_x = i3("density_x", t_ret, xpf, zq)
    dens_z = i3("density_z", t_ret, xpf, zq)
    vx_ret = i3("vx", t_ret, xpf, zq)
    vx_x_ret = i3("vx_x", t_ret, xpf, zq)

    vs, vs_ret, vs_s_ret = 1, 1, 0
    scale = 1 + xpf * rho_sp
    vel_x = vs * tsx + vx * nsx
    vel_y = vs * tsy + vx * nsy
    vr_x = vs_ret * tspx + vx_ret * nspx
    vr_y = vs_ret * tspy + vx_ret * nspy
    with np.errstate(invalid="ignore", divide="ignore"):
        nab_x = dens_x * nspx + dens_z / scale * tspx
        nab_y = dens_x * nspy + dens_z / scale * tspy
        div_v = vs_s_ret + vx_x_ret
        part1 = vel_x * vr_x + vel_y * vr_y
        num1 = scale * ((vel_x - part1 * vr_x) * nab_x + (vel_y - part1 * vr_y) * nab_y)
        num2 = -scale * part1 * dens * div_v
        Iz = num1 / R + num2 / R
        nmx = nsx - nspx
        nmy = nsy - nspy
        p1 = rx * nmx + ry * nmy
        p2 = nsx * tspx + nsy * tspy
        pd = -(vr_x * nab_x + vr_y * nab_y) - dens * div_v
        W1 = scale * p1 / (R * R * R) * dens
        W2 = scale * p1 / (R * R) * pd
        W3 = -scale * p2 / R * pd
        Ix = W1 + W2 + W3
    return Iz.reshape(xp.shape), Ix.reshape(xp.shape)


# ---------------------------------------------------------------------------
# a14  get_CSR_wake   (CSR.py:454-602)
# ---------------------------------------------------------------------------
def wake_regions(s, x, t, sigma_z, sigma_x, slope0, xmean, formation_length,
                 n_formation_length, Nx, Nz):
    """Integration regions (xp axis, sp axis) for one observer (CSR.py:456-553).

    Returns a list of (xp_1d, sp_1d) in the reference's summation order.
    """
    tan_theta = slope0
    x0 = (s - t) * slope0
    if np.abs(tan_theta) <= 1:
        s2 = s - 500 * sigma_z
        s3 = s - 20 * sigma_z
        s4 = s + 5 * sigma_z
        s1 = np.max((0, s2 - n_formation_length * formation_length))
        sp1 = np.linspace(s1, s2, Nz)
        sp2 = np.linspace(s2, s3, Nz)
        sp3 = np.linspace(s3, s4, Nz)
        xp_w = np.linspace(x0 - 20 * sigma_x, x0 + 20 * sigma_x, 2 * Nx)
        xp_n = np.linspace(x0 - 10 * sigma_x, x0 + 10 * sigma_x, Nx)
        return [(xp_w, sp1), (xp_n, sp2), (xp_n, sp3)]
    if tan_theta > 0:
        tan_alpha = -2 * tan_theta / (1 - tan_theta ** 2)
        d = (10 * sigma_x + xmean - x) / tan_alpha
        x1_l, x1_r = x + 0.1 * sigma_x, x + 10 * sigma_x
        x2_l, x2_r = x - 3 * sigma_x, x1_l
    else:
        tan_alpha = 2 * tan_theta / (1 - tan_theta ** 2)
        d = -(xmean - x - 10 * sigma_x) / tan_alpha
        x1_l, x1_r = x - 10 * sigma_x, x - 1 * sigma_x
        x2_l, x2_r = x1_r, x + 3 * sigma_x
    s4 = s + 3 * sigma_z
    s3 = np.max((0, s - d))
    s2 = s3 - 200 * sigma_z
    x3_l, x3_r = x0 - 5 * sigma_x, x0 + 5 * sigma_x
    x4_l, x4_r = x0 - 20 * sigma_x, x0 + 20 * sigma_x
    s1 = np.max((0, s2 - n_formation_length * formation_length))
    sp1 = np.linspace(s1, s2, Nz)
    sp2 = np.linspace(s2, s3, Nz)
    sp3 = np.linspace(s3, s4, Nz)
    xp1 = np.linspace(x1_l, x1_r, Nx)
    xp2 = np.linspace(x2_l, x2_r, Nx)
    xp3 = np.linspace(x3_l, x3_r, Nx)
    xp4 = np.linspace(x4_l, x4_r, 2 * Nx)
    return [(xp4, sp1), (xp3, sp2), (xp1, sp3), (xp2, sp3)]


def csr_wake(s, x, t, beam, formation_length, n_formation_length, Nx, Nz, lat, hist,
             csr_scaling):
    """(dE_dct, x_kick) at one observer: CSR.py:454-602 with nested trapz."""
    regs = wake_regions(s, x, t, beam["sigma_z"], beam["sigma_x"], beam["slope"][0],
                        beam["mean_x"], formation_length, n_formation_length, Nx, Nz)
    dE = 0.0
    xk = 0.0
    for xp1, sp1 in regs:
        xpm, spm = np.meshgrid(xp1, sp1, indexing="ij")
        Iz, Ix = csr_integrand(s, x, t, spm, xpm, lat, hist)
        dE = dE + (-csr_scaling * _trapz(_trapz(Iz, xp1, axis=0), sp1))
        xk = xk + csr_scaling * _trapz(_trapz(Ix, xp1, axis=0), sp1)
    return dE, xk


# ---------------------------------------------------------------------------
# a2  get_CSR_mesh (CSR.py:361-394)   a15  calculate_2D_CSR (CSR.py:397-418)
# ---------------------------------------------------------------------------
def csr_mesh(x, z, xbins, zbins, xlim, zlim):
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    p = np.polyfit(z, x, deg=1)
    xt = x - np.polyval(p, z)
    sig_x = np.std(xt)
    mean_x = np.mean(xt)
    sig_z = np.std(z)
    mean_z = np.mean(z)
    zrange = np.linspace(mean_z - zlim * sig_z, mean_z + zlim * sig_z, zbins)
    xrange = np.linspace(mean_x - xlim * sig_x, mean_x + xlim * sig_x, xbins)
    xm_t, zm = np.meshgrid(xrange, zrange, indexing="ij")
    xm_t = xm_t.flatten()
    zm = zm.flatten()
    xmesh = xm_t + np.polyval(p, zm)
    return dict(xmesh=xmesh, zmesh=zm, zrange=zrange, xrange=xrange, slope=p)


def calculate_2D_CSR(position, mesh, beam, formation_length, n_formation_length, Nx, Nz,
                     lat, hist, csr_scaling, xbins, zbins, rows=None):
    """Loop over mesh points (CSR.py:397-418).  ``rows`` restricts the loop
    to a subset of flattened mesh indices (others are left as NaN)."""
    Mtot = xbins * zbins
    dE = np.full(Mtot, np.nan)
    xk = np.full(Mtot, np.nan)
    idx = range(Mtot) if rows is None else rows
    for i in idx:
        s = position + mesh["zmesh"][i]
        x = mesh["xmesh"][i]
        dE[i], xk[i] = csr_wake(s, x, position, beam, formation_length, n_formation_length,
                                Nx, Nz, lat, hist, csr_scaling)
    return dE.reshape(xbins, zbins), xk.reshape(xbins, zbins)


# ---------------------------------------------------------------------------
# a16  Beam.apply_wakes   (beams.py:108-131)
# ---------------------------------------------------------------------------
def apply_wakes(x, z, px, pz, dE_dct, x_kick, xrange, zrange, step_size, init_energy,
                transverse_on):
    """Returns new (px, pz) after the CSR kick (beams.py:108-131)."""
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z, dtype=np.float64)
    p = np.polyfit(z, x, deg=1)
    xt = x - np.polyval(p, z)
    dE_E1 = step_size * dE_dct * 1e6 / init_energy
    f = RegularGridInterpolator((xrange, zrange), dE_E1, fill_value=0.0, bounds_error=False)
    pz_new = pz + f(np.array([xt, z]).T)
    if transverse_on:
        dxp = step_size * x_kick * 1e6 / init_energy
        f = RegularGridInterpolator((xrange, zrange), dxp, fill_value=0.0, bounds_error=False)
        px_new = px + f(np.array([xt, z]).T)
    else:
        px_new = np.array(px, dtype=np.float64, copy=True)
    return px_new, pz_new


# ---------------------------------------------------------------------------
# a17  run loop: step schedule, element pieces, formation length
#      (CSR.py:140-144, 146-199, 202-358; lattice.py:152-173)
# ---------------------------------------------------------------------------
def lattice_steps(distance, step_size):
    """(positions_record, steps_per_element) of Lattice.get_steps
    (lattice.py:152-173): the step positions 0, ds, 2 ds, ... up to the
    lattice end (arange with a half-step margin), and per element the number
    of step positions in (distance[k-1], distance[k]] (position 0 excluded
    from the first element)."""
    positions = np.arange(0, distance[-1] + step_size / 2, step_size)
    spe = np.zeros(len(distance), dtype=int)
    prev = 0
    for k, d in enumerate(distance):
        ind = np.searchsorted(positions, d, side="right")
        spe[k] = ind - prev - 1 if k == 0 else ind - prev
        prev = ind
    return positions, spe


def formation_length(R, sigma_z):
    """CSR.py:140-142 (in-bend form; the loop passes 5 * the beam's sigma_z)."""
    return (24 * (R ** 2) * sigma_z) ** (1 / 3)


def element_piece(ele, DL, entrance=False, exit=False):
    """The bmadx element a piece of length DL is tracked as (CSR.py:146-199),
    as a dict: kind, L and, for a bend, G (the G key, else angle / L), E1,
    E2 and FRINGE_AT -- the element's own E1/E2/FRINGE_AT when the piece is
    the whole element, E1 and 'entrance_end' at an entrance piece, E2 and
    'exit_end' at an exit piece, neither ('no_end') inside; a quadrupole's
    K1, a sextupole's K2.  A bend tracked in one piece without a FRINGE_AT
    key fails, as in the reference (CSR.py:177 reads it unbound)."""
    typ = ele["type"]
    if typ == "dipole":
        G = ele["G"] if "G" in ele else ele["angle"] / ele["L"]
        E1, E2 = ele.get("E1", 0), ele.get("E2", 0)
        if entrance and exit:
            if "FRINGE_AT" not in ele:
                raise KeyError("FRINGE_AT: a bend tracked in one piece needs it (CSR.py:177)")
            return dict(kind="SBend", L=DL, G=G, E1=E1, E2=E2, FRINGE_AT=ele["FRINGE_AT"])
        if entrance:
            return dict(kind="SBend", L=DL, G=G, E1=E1, E2=0.0, FRINGE_AT="entrance_end")
        if exit:
            return dict(kind="SBend", L=DL, G=G, E1=0.0, E2=E2, FRINGE_AT="exit_end")
        return dict(kind="SBend", L=DL, G=G, E1=0.0, E2=0.0, FRINGE_AT="no_end")
    if typ == "drift":
        return dict(kind="Drift", L=DL)
    if typ == "quad":
        return dict(kind="Quadrupole", L=DL, K1=ele["K1"])
    if typ == "sextupole":
        return dict(kind="Sextupole", L=DL, K2=ele["K2"])
    raise ValueError(f"element type {typ!r}")


def run_loop(lattice_config, track, sigma_z, stop_time=None):
    """The element/step loop of CSR2D.run (CSR.py:202-358) without the CSR
    work: a generator yielding one dict per step (step, element, position,
    formation_length -- the value append_interpolant receives --, csr: whether
    the step computes the wake, step % nsep == 0).

    lattice_config: ordered mapping, step_size first, then the elements.
    track(piece, DL, update_step): moves the caller's beam through one piece
    (a dict of element_piece); sigma_z(): the beam's sigma_z now.

    Per element: the rest of the step that crossed into it is tracked in
    the previous element (exit piece) unless that element was taken whole;
    an element holding no step position is tracked whole (entrance and
    exit); the formation length is set from 5 sigma_z in a bend and in
    every element after the first bend (with the last bend's radius), and
    accumulates element lengths before it; the first step in an element is
    the part of the step past the element boundary (entrance piece)."""
    items = list(lattice_config.items())
    DL = items[0][1]
    assert items[0][0] == "step_size"
    names = [k for k, _ in items[1:]]
    eles = [dict(v) for _, v in items[1:]]
    distance = np.zeros(len(eles))
    for k, e in enumerate(eles):
        distance[k] = e["L"] if k == 0 else e["L"] + distance[k - 1]
    positions, spe = lattice_steps(distance, DL)
    nsep = np.array([float(e["nsep"]) for e in eles])
    position = 0
    step_count = 1
    skip_ele = False
    afterbend = False
    R_rec = None
    fl = 0.0
    prev = None
    for k, (name, e) in enumerate(zip(names, eles)):
        L = e["L"]
        steps = spe[k]
        if (not skip_ele) and k > 0:
            DL_1 = distance[k - 1] - position
            if DL_1 > 1.0e-6:
                track(element_piece(prev, DL_1, exit=True), DL_1, False)
                position += DL_1
        if steps == 0:
            skip_ele = True
            track(element_piece(e, L, entrance=True, exit=True), L, False)
            position += L
        if e["type"] == "dipole":
            R = L / e["angle"]
            afterbend = True
            R_rec = R
            fl = formation_length(R, 5 * sigma_z())
        elif afterbend:
            fl = formation_length(R_rec, 5 * sigma_z())
        else:
            fl += L
        for step in range(steps):
            if step == 0 and k > 0:
                DL_2 = positions[step_count] - distance[k - 1]
                track(element_piece(e, DL_2, entrance=True), DL_2, True)
                position += DL_2
                skip_ele = False
            else:
                track(element_piece(e, DL), DL, True)
                position += DL
            yield dict(step=step_count, element=name, position=position, formation_length=fl,
                       csr=bool(step % nsep[k] == 0))
            step_count += 1
            if stop_time and position > stop_time:
                return
        prev = e
