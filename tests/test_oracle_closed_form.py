"""Pins the CPU oracle against mathematics that fixes its answers.

Closed forms, invariants, brute force on tiny inputs and library routines,
each chosen so that a plausible slip (a dropped term, a wrong sign or index,
a transposed operand) fails at least one of them.
"""
import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator

from oracle import cpu_ref as O


# -- a3 CIC deposition ------------------------------------------------------
def test_cic_particle_on_node_single_cell():
    # grid 4 x 5 cells over [0, 4) x [0, 10): a particle exactly on a node
    h = O.histogram_cic_2d(np.array([2.0]), np.array([4.0]), np.array([3.0]), 4, 0.0, 4.0, 5, 0.0, 10.0)
    exp = np.zeros((4, 5))
    exp[2, 2] = 3.0
    assert np.array_equal(h, exp)


def test_cic_off_node_split_and_edges():
    # cell (1.25, 0.5) -> S1 = 0.75, S2 = 0.5
    h = O.histogram_cic_2d(np.array([1.25]), np.array([1.0]), np.array([2.0]), 4, 0.0, 4.0, 5, 0.0, 10.0)
    exp = np.zeros((4, 5))
    exp[1, 0] = 2 * 0.75 * 0.5
    exp[1, 1] = 2 * 0.75 * 0.5
    exp[2, 0] = 2 * 0.25 * 0.5
    exp[2, 1] = 2 * 0.25 * 0.5
    assert np.allclose(h, exp, rtol=0, atol=1e-15)
    # particle in the last cell loses its (i+1) neighbours; one below the
    # grid keeps only its upper neighbour
    h = O.histogram_cic_2d(np.array([3.5, -0.25]), np.array([4.0, 4.0]), np.array([1.0, 1.0]),
                           4, 0.0, 4.0, 5, 0.0, 10.0)
    assert h[3, 2] == 0.5 and h[0, 2] == 0.75 and np.isclose(h.sum(), 1.25)


def test_cic_total_weight_and_loop_form():
    rng = np.random.default_rng(1)
    q1 = rng.uniform(0.2, 2.95, 500)      # i+1 stays on the grid
    q2 = rng.uniform(0.3, 7.9, 500)
    w = rng.normal(size=500)
    h = O.histogram_cic_2d(q1, q2, w, 4, 0.0, 4.0, 5, 0.0, 10.0)
    # interior particles keep all their weight
    assert np.isclose(h.sum(), w.sum(), rtol=1e-12)
    assert np.array_equal(h, O.histogram_cic_2d_loop(q1, q2, w, 4, 0.0, 4.0, 5, 0.0, 10.0))
    i1, i2, _, _ = O.cic_cells(q1, q2, 4, 0.0, 4.0, 5, 0.0, 10.0)
    assert np.array_equal(i1, np.floor(q1).astype(int)) and np.array_equal(i2, np.floor(q2 / 2).astype(int))
    c = O.cell_counts(i1, i2, 4, 5)
    assert c.sum() == 500


# -- a11 / a12 interpolation --------------------------------------------------
def test_interpolate3d_exact_on_trilinear():
    t = np.linspace(0.0, 2.0, 5)
    x = np.linspace(-1.0, 1.0, 7)
    z = np.linspace(0.5, 3.5, 9)
    T, X, Z = np.meshgrid(t, x, z, indexing="ij")
    f = lambda a, b, c: 1.5 + 2 * a - 3 * b + 0.5 * c + a * b - 2 * b * c + 0.25 * a * b * c
    data = f(T, X, Z)
    rng = np.random.default_rng(2)
    q = np.stack([rng.uniform(0, 2, 200), rng.uniform(-1, 1, 200), rng.uniform(0.5, 3.5, 200)])
    r = O.interpolate3D(*q, data, 0.0, -1.0, 0.5, 0.5, 1 / 3, 0.375)
    assert np.allclose(r, f(*q), rtol=0, atol=1e-12)


def test_interpolate3d_truncation_clamp_and_zero():
    data = np.arange(2 * 3 * 4, dtype=float).reshape(2, 3, 4)
    mn = (0.0, 0.0, 0.0)
    dl = (1.0, 1.0, 1.0)
    # u in (-1, 0): trunc -> 0, negative fraction extrapolates
    r = O.interpolate3D(np.array([-0.5]), np.array([1.0]), np.array([1.0]), data, *mn, *dl)
    assert np.isclose(r[0], data[0, 1, 1] - 0.5 * (data[1, 1, 1] - data[0, 1, 1]))
    # u in [size-1, size): clamps to the last plane
    r = O.interpolate3D(np.array([1.7]), np.array([1.0]), np.array([1.0]), data, *mn, *dl)
    assert np.isclose(r[0], data[1, 1, 1])
    # u <= -1 or u >= size: zero
    r = O.interpolate3D(np.array([-1.0, 2.0, 0.5]), np.array([1.0, 1.0, 3.0]),
                        np.array([1.0, 1.0, 1.0]), data, *mn, *dl)
    assert np.array_equal(r, np.zeros(3))
    # NaN query: zero
    assert O.interpolate3D(np.array([np.nan]), np.array([1.0]), np.array([1.0]), data, *mn, *dl)[0] == 0
    q = np.random.default_rng(3).uniform(-1.5, 4.5, size=(3, 300))
    assert np.array_equal(O.interpolate3D(*q, data, *mn, *dl), O.interpolate3D_loop(*q, data, *mn, *dl))


def test_interpolate1d_linear_and_out_of_range():
    d = 3.0 - 2.0 * np.linspace(0, 1, 11)
    xs = np.array([-0.2, 0.0, 0.33, 0.999, 1.0, 1.09, 1.1, -0.05, -0.1])
    r = O.interpolate1D(xs, d, 0.0, 0.1)
    exp = np.array([0.0, 3.0, 3 - 0.66, 3 - 1.998, 1.0, 1.0, 0.0, 3.0 + 0.1, 0.0])
    assert np.allclose(r, exp, atol=1e-12)


# -- a12 lattice tables ------------------------------------------------------
def test_reference_trajectory_closed_form():
    R = 6.0
    ang = 0.3 / R
    els = [dict(type="dipole", L=0.3, angle=ang, nsep=1), dict(type="drift", L=0.5, nsep=1)]
    s, rho, dist, nsep, coords, nv, tau = O.reference_trajectory(els, Nsample=401)
    assert np.allclose(rho, [1 / R, 0.0]) and np.allclose(dist, [0.3, 0.8])
    inb = s <= 0.3 - 1e-12
    th = s[inb] / R
    # arc of a leading dipole: x = R sin(th), y = R (1 - cos th)
    assert np.allclose(coords[inb, 0], R * np.sin(th), atol=1e-12)
    assert np.allclose(coords[inb, 1], R * (1 - np.cos(th)), atol=1e-12)
    assert np.allclose(np.hypot(*tau.T), 1) and np.allclose(np.hypot(*nv.T), 1)
    assert np.allclose((tau * nv).sum(1), 0, atol=1e-15)
    # n = tau rotated by -90 degrees
    assert np.allclose(nv[:, 0], tau[:, 1]) and np.allclose(nv[:, 1], -tau[:, 0])


# -- a5 / a6 library composition (savgol 'interp', gradient) -----------------
def test_get_df_uniform_threshold_quirks():
    rng = np.random.default_rng(4)
    n = 20000
    z = rng.normal(0, 50e-6, n)
    x = rng.normal(0, 30e-6, n)
    px = 0.01 * x
    r = O.get_DF(x, z, px, 0.0, xbins=32, zbins=32, filter_order=2, filter_window=5,
                 velocity_threhold=1000)
    assert r["nbins"] == (100, 100)          # frac <= 5 fallback (deposit.py:160-167)
    dsum = np.trapezoid(np.trapezoid(r["density"], r["x_grids"], axis=0), r["z_grids"])
    assert np.isclose(dsum, 1.0, rtol=1e-12)
    # vx = px/x ratio -> 0.01 * x where density is significant
    core = r["density"] > 0.5 * r["density"].max()
    X = np.broadcast_to(r["x_grids"][:, None], r["density"].shape)
    assert np.allclose(r["vx"][core], 0.01 * X[core], rtol=0.05, atol=2e-8)


# -- a16 kick ------------------------------------------------------------------
def test_apply_wakes_bilinear_closed_form():
    xr = np.linspace(-1, 1, 5)
    zr = np.linspace(-2, 2, 7)
    X, Z = np.meshgrid(xr, zr, indexing="ij")
    dE = 3 + X - 2 * Z + 0.5 * X * Z
    xk = -1 + 2 * X + Z
    rng = np.random.default_rng(5)
    z = rng.uniform(-2.5, 2.5, 300)
    x = 0.3 * z + rng.uniform(-1.2, 1.2, 300)
    px0 = np.zeros(300)
    pz0 = np.zeros(300)
    px1, pz1 = O.apply_wakes(x, z, px0, pz0, dE, xk, xr, zr, 0.1, 1e9, 1)
    p = np.polyfit(z, x, 1)
    xt = x - np.polyval(p, z)
    inside = (np.abs(xt) <= 1) & (np.abs(z) <= 2)
    c = 0.1 * 1e6 / 1e9
    assert np.allclose(pz1[inside], c * (3 + xt - 2 * z + 0.5 * xt * z)[inside], rtol=1e-12, atol=1e-18)
    assert np.allclose(px1[inside], c * (-1 + 2 * xt + z)[inside], rtol=1e-12, atol=1e-18)
    assert np.all(pz1[~inside] == 0) and np.all(px1[~inside] == 0)


# -- a13 integrand invariants --------------------------------------------------
def _toy_history(scale=1.0):
    T = O.DFTracker(xbins=24, zbins=24, filter_order=2, filter_window=5, velocity_threhold=1000)
    rng = np.random.default_rng(6)
    for k, t in enumerate([0.0, 0.05, 0.1]):
        z = rng.normal(0, 40e-6, 8000)
        x = rng.normal(0, 20e-6, 8000) + 3.0 * z
        px = 1e-3 * rng.normal(size=8000)
        T.get_DF(x, z, px, t)
        T.append_DF()
        T.append_interpolant(float("inf") if k == 0 else 1.0, 1.0)
    T.build_interpolant()
    if scale != 1.0:
        for f in ("density", "density_x", "density_z"):
            T.data[f] = T.data[f] * scale
    return T


def test_integrand_zero_density_and_linearity():
    L = O.LatticeTables([dict(type="dipole", L=0.5, angle=0.05, nsep=1)])
    T = _toy_history()
    s, x, t = 0.1 + 5e-5, 1e-5, 0.1
    xp, sp = np.meshgrid(np.linspace(-1e-4, 1e-4, 9), np.linspace(0.06, 0.1002, 11), indexing="ij")
    Iz, Ix = O.csr_integrand(s, x, t, sp, xp, L, T)
    assert np.all(np.isfinite(Iz)) and np.any(Iz != 0)
    T2 = _toy_history(scale=2.5)
    Iz2, Ix2 = O.csr_integrand(s, x, t, sp, xp, L, T2)
    # linear in the density fields (W1+W2+W3 cancels partly: compare to the max)
    assert np.abs(Iz2 - 2.5 * Iz).max() <= 1e-12 * np.abs(Iz).max()
    assert np.abs(Ix2 - 2.5 * Ix).max() <= 1e-12 * np.abs(Ix).max()
    T0 = _toy_history(scale=0.0)
    Iz0, Ix0 = O.csr_integrand(s, x, t, sp, xp, L, T0)
    assert np.all(Iz0 == 0) and np.all(Ix0 == 0)


def test_rgi_restatement_matches_scipy():
    # the product re-grid and kick kernels restate RegularGridInterpolator's
    # linear rule: interval search + (1-y0)(1-y1) corner weights
    xg = np.linspace(-1, 1, 9)
    zg = np.linspace(0, 2, 13)
    V = np.random.default_rng(7).normal(size=(9, 13))
    q = np.random.default_rng(8).uniform(-1.2, 2.2, size=(400, 2))
    ref = RegularGridInterpolator((xg, zg), V, fill_value=0.0, bounds_error=False)(q)
    i = np.clip(np.searchsorted(xg, q[:, 0], side="right") - 1, 0, 7)
    j = np.clip(np.searchsorted(zg, q[:, 1], side="right") - 1, 0, 11)
    y0 = (q[:, 0] - xg[i]) / (xg[i + 1] - xg[i])
    y1 = (q[:, 1] - zg[j]) / (zg[j + 1] - zg[j])
    v = (V[i, j] * (1 - y0) * (1 - y1) + V[i, j + 1] * (1 - y0) * y1
         + V[i + 1, j] * y0 * (1 - y1) + V[i + 1, j + 1] * y0 * y1)
    oob = (q[:, 0] < xg[0]) | (q[:, 0] > xg[-1]) | (q[:, 1] < zg[0]) | (q[:, 1] > zg[-1])
    v[oob] = 0.0
    assert np.array_equal(v, ref)


# -- per-step statistics (twiss.py:2-70) -------------------------------------------
def _beam_with_covariance(cov, n=4000, seed=0):
    """Particles whose sample covariance (ddof = 1) is exactly cov (to rounding):
    centre random samples, whiten with their own Cholesky factor, colour with
    cov's."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((cov.shape[0], n))
    X -= X.mean(axis=1, keepdims=True)
    Lc = np.linalg.cholesky(np.cov(X))
    Lt = np.linalg.cholesky(cov)
    return Lt @ np.linalg.solve(Lc, X)


@pytest.mark.parametrize("alpha,beta,emit,eta,etap,sd", [
    (-1.3, 7.5, 2e-9, 0.12, -0.03, 1e-3), (0.4, 0.8, 5e-7, 0.0, 0.0, 2e-4),
    (2.0, 25.0, 1e-10, -0.5, 0.2, 5e-5)])
def test_twiss_recovers_constructed_beam(alpha, beta, emit, eta, etap, sd):
    gamma = (1 + alpha ** 2) / beta
    d2 = sd ** 2
    cov = np.zeros((3, 3))
    cov[:2, :2] = emit * np.array([[beta, -alpha], [-alpha, gamma]]) + \
        d2 * np.outer([eta, etap], [eta, etap])
    cov[:2, 2] = cov[2, :2] = d2 * np.array([eta, etap])
    cov[2, 2] = d2
    u, pu, d = _beam_with_covariance(cov)
    tw = O.twiss_dispersion_calc(np.cov([u, pu, d]))
    for k, v in dict(alpha=alpha, beta=beta, gamma=gamma, emit=emit, eta=eta, etap=etap).items():
        assert abs(tw[k] - v) <= 1e-9 * max(abs(v), 1.0) * (1 if k != "emit" else emit), (k, tw[k], v)
    # invariant of any beam
    assert abs(tw["beta"] * tw["gamma"] - tw["alpha"] ** 2 - 1) < 1e-12


def test_beam_statistics_planes_and_energy():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((6, 6)) * np.array([1e-4, 1e-5, 2e-4, 3e-5, 5e-5, 1e-3])[:, None]
    C = A @ A.T
    x, px, y, py, z, pz = _beam_with_covariance(C, n=5000, seed=4) + \
        np.array([1e-5, 0, -2e-5, 0, 3e-6, 2e-3])[:, None]
    p0c, mc2 = 1e9, 510998.95069
    st = O.beam_statistics(x, px, y, py, z, pz, p0c, mc2)
    Xc = np.stack([x, px, y, py, z, pz])
    Xc = Xc - Xc.mean(axis=1, keepdims=True)
    assert np.allclose(st["cov"], Xc @ Xc.T / (Xc.shape[1] - 1), rtol=1e-12, atol=0)
    # each plane uses (u, pu, pz): the x plane must not see y and vice versa
    for plane, idx in (("x", [0, 1, 5]), ("y", [2, 3, 5])):
        s3 = C[np.ix_(idx, idx)]
        d2 = s3[2, 2]
        m = s3[:2, :2] - np.outer(s3[:2, 2], s3[:2, 2]) / d2
        emit = np.sqrt(np.linalg.det(m))
        assert abs(st["twiss"][f"emit_{plane}"] / emit - 1) < 1e-9
        assert abs(st["twiss"][f"beta_{plane}"] - m[0, 0] / emit) < 1e-9 * abs(m[0, 0] / emit)
        assert abs(st["twiss"][f"alpha_{plane}"] + m[0, 1] / emit) < 1e-9 * max(abs(m[0, 1] / emit), 1)
        assert abs(st["twiss"][f"eta_{plane}"] - s3[0, 2] / d2) < 1e-9 * abs(s3[0, 2] / d2)
        assert abs(st["twiss"][f"norm_emit_{plane}"] - emit * p0c / mc2) < 1e-9 * emit * p0c / mc2
    assert abs(st["mean_energy"] - p0c * (1 + pz.mean())) < 1e-15 * p0c * 4
    assert abs(st["sigma_energy"] - p0c * np.sqrt(np.mean((pz - pz.mean()) ** 2))) < \
        1e-12 * st["sigma_energy"]


# -- tracking (bmadx.track_element as CSR.py:146-199 builds it) -------------------
def _gold_tracking():
    import json
    import os
    d = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "ref_bmadx_tracking.json")))
    p = d["particle_in"]
    return d, [p[k] for k in ("x", "px", "y", "py", "z", "pz")], p["p0c"], p["mc2"]


_KEYS = ("x", "px", "y", "py", "z", "pz")


def _track_case(c, case, p0c, mc2):
    e = case["element"]
    if e == "Drift":
        return O.track_drift(c, case["L"], p0c, mc2)
    if e == "SBend":
        return O.track_sbend(c, case["L"], case["G"], case["E1"], case["E2"], p0c, mc2)
    return O.track_quadrupole(c, case["L"], case["K1"], p0c, mc2)


def test_tracking_matches_bmadx_single_particles():
    """The reference notebook's bmadx results (test_BmadX_tracking.ipynb
    cells 25, 28, 31) within 1e-12 relative per coordinate."""
    d, c, p0c, mc2 = _gold_tracking()
    for case in d["cases"]:
        r = _track_case(c, case, p0c, mc2)
        g = np.array([case["out"][k] for k in _KEYS])
        assert np.all(np.abs(r - g) <= 1e-12 * np.abs(g)), (case["element"], r - g)


def _jacobian(f, c0, h=1e-7):
    c0 = np.asarray(c0, dtype=np.float64)
    J = np.zeros((6, 6))
    for j in range(6):
        e = np.zeros(6)
        e[j] = h
        J[:, j] = (np.asarray(f(c0 + e)) - np.asarray(f(c0 - e))) / (2 * h)
    return J


@pytest.mark.parametrize("name", ["drift", "quad_f", "quad_d", "bend", "bend_edges", "sext",
                                  "quad_tilt"])
def test_tracking_maps_are_symplectic(name):
    """J^T S J = S for the Jacobian at a generic point: ties the z update to
    the energy dependence of the transverse map, so a wrong sign or a dropped
    path-length term fails."""
    p0c, mc2 = 1e9, 510998.95069
    maps = {
        "drift": lambda c: O.track_drift(c, 0.7, p0c, mc2),
        "quad_f": lambda c: O.track_quadrupole(c, 0.3, 4.0, p0c, mc2),
        "quad_d": lambda c: O.track_quadrupole(c, 0.3, -6.0, p0c, mc2, num_steps=3),
        "bend": lambda c: O.track_sbend(c, 0.3, 0.05 / 0.3, 0.0, 0.0, p0c, mc2),
        "bend_edges": lambda c: O.track_sbend(c, 0.3, -0.08 / 0.3, 0.02, -0.08, p0c, mc2,
                                              fint=0.6, hgap=0.017, fintx=0.5, hgapx=0.02),
        "sext": lambda c: O.track_sextupole(c, 0.2, 30.0, p0c, mc2, num_steps=2),
        "quad_tilt": lambda c: O.track_quadrupole(c, 0.3, 4.0, p0c, mc2, x_off=1e-4,
                                                  y_off=-2e-4, tilt=0.3),
    }
    c0 = np.array([1.2e-3, -0.7e-3, 0.8e-3, 0.5e-3, 0.3e-3, 2e-3])
    J = _jacobian(maps[name], c0)
    S = np.kron(np.eye(3), np.array([[0.0, 1.0], [-1.0, 0.0]]))
    assert np.abs(J.T @ S @ J - S).max() < 1e-7, np.abs(J.T @ S @ J - S).max()


def _r_gen6_vectors():
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_r_gen6.npz"))
    return d["cases"], d["R"]


def test_transfer_matrix_matches_r_gen6_vectors():
    """The oracle's linear map equals the reference's r_gen6 (r_gen6.py:4-122)
    on every unrolled case of tests/golden/ref_r_gen6.npz, pole faces and the
    fringe term included."""
    from oracle.cpu_ref import transfer_matrix
    cases, mats = _r_gen6_vectors()
    n = 0
    for (L, angle, k1, roll, e1, e2, hgap), R in zip(cases, mats):
        if roll:
            continue
        T = transfer_matrix(L, angle=angle, k1=k1, E1=e1, E2=e2, hgap=hgap)
        assert np.allclose(T, R, rtol=0, atol=1e-15), (L, angle, k1, e1, e2, hgap)
        n += 1
    assert n == 10


@pytest.mark.parametrize("case", [0, 1, 2, 3, 4, 5, 6, 7, 8, 10])
def test_tracking_linear_limit_is_r_gen6(case):
    """At small amplitude the maps equal the reference's r_gen6 matrices
    (tests/golden/ref_r_gen6.npz): the transverse rows exactly, so the pole
    faces and the fringe term at r_gen6's FINT = 1/2 (r_gen6.py:49-56) are
    pinned, and the z row up to the relativistic slip L / (beta gamma)^2 that
    r_gen6 leaves out (R56 of an ultrarelativistic beam).  Case 10 is a
    quadrupole rolled by pi/4 (r_gen6's roll is the element's tilt); cases 9
    (a combined-function bend) and 11 (a bare rotation) are no tracked
    element."""
    cases, mats = _r_gen6_vectors()
    L, angle, k1, roll, e1, e2, hgap = cases[case]
    R = mats[case].copy()
    p0c, mc2 = 1e10, 510998.95069
    if angle:
        f = lambda c: O.track_sbend(c, L, angle / L, e1, e2, p0c, mc2, fint=0.5, hgap=hgap,
                                    fintx=0.5, hgapx=hgap)
    elif k1:
        f = lambda c: O.track_quadrupole(c, L, k1, p0c, mc2, tilt=roll)
    else:
        f = lambda c: O.track_drift(c, L, p0c, mc2)
    J = _jacobian(f, np.zeros(6), h=1e-6)
    assert np.abs(J[:4] - R[:4]).max() < 1e-12
    R[4, 5] += L * (mc2 / p0c) ** 2
    assert np.abs(J[4] - R[4]).max() < 1e-10
    assert np.array_equal(np.round(J[5], 6), [0, 0, 0, 0, 0, 1])


def test_sextupole_feed_down_is_a_quadrupole():
    """Off axis by x0 a short sextupole focuses like a quadrupole of
    K1 = K2 x0 (the feed-down of (x + iy)^2): its linear map about the orbit
    equals the quadrupole map (pinned on bmadx results) to third order in L.
    A flipped sign, a lost 1/2 or swapped planes change the focusing term
    K1 L = 1.2e-3 by its own size."""
    p0c, mc2 = 1e9, 510998.95069
    L, K2, x0 = 0.02, 30.0, 2e-3
    c0 = np.array([x0, 0.0, 0.0, 0.0, 0.0, 0.0])
    Js = _jacobian(lambda c: O.track_sextupole(c, L, K2, p0c, mc2), c0)
    Jq = _jacobian(lambda c: O.track_quadrupole(c, L, K2 * x0, p0c, mc2), c0)
    assert np.abs(Js[:4, :4] - Jq[:4, :4]).max() < 1e-7
    assert Js[1, 0] < -1e-3 and Js[3, 2] > 1e-3


def test_sextupole_converges_to_its_field_equations():
    """With many steps the sextupole map is the flow of x' = px/ps,
    px' = -K2 (x^2 - y^2)/2, y' = py/ps, py' = K2 x y,
    z' = beta/beta0 - (1 + pz)/ps (ps = sqrt((1+pz)^2 - px^2 - py^2)),
    integrated here by scipy's DOP853; the error falls as the square of the
    step (a symmetric splitting).  Which symmetric splitting bmadx uses
    (drift-kick-drift or kick-drift-kick) no reference result fixes."""
    from scipy.integrate import solve_ivp
    p0c, mc2, L, K2 = 1e9, 510998.95069, 0.2, 300.0
    b0 = p0c / np.hypot(p0c, mc2)

    def rhs(s, u):
        x, px, y, py, z, pz = u
        P = 1 + pz
        ps = np.sqrt(P * P - px * px - py * py)
        b = P * p0c / np.hypot(P * p0c, mc2)
        return [px / ps, -K2 * (x * x - y * y) / 2, py / ps, K2 * x * y, b / b0 - P / ps, 0.0]

    u0 = np.array([2e-3, 1e-3, 1e-3, -5e-4, 1e-4, 1e-3])
    exact = solve_ivp(rhs, (0.0, L), u0, method="DOP853", rtol=1e-13, atol=1e-16).y[:, -1]
    e10 = np.abs(O.track_sextupole(u0, L, K2, p0c, mc2, num_steps=10) - exact).max()
    e100 = np.abs(O.track_sextupole(u0, L, K2, p0c, mc2, num_steps=100) - exact).max()
    e1000 = np.abs(O.track_sextupole(u0, L, K2, p0c, mc2, num_steps=1000) - exact).max()
    assert e1000 < 2e-12
    assert 80 < e10 / e100 < 120 and 80 < e100 / e1000 < 120
    # the field matters at this size: a plain drift misses by 1e-4
    assert np.abs(O.track_drift(u0, L, p0c, mc2) - exact).max() > 1e-4
