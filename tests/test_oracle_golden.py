"""Pins the CPU oracle against the reference's own functions.

The fixtures under tests/golden/ref_*.npz were written by
tests/golden/make_reference_vectors.py, which ran pyDFCSR's hot-path
functions (deposit.py, interp3D.py, interp1D.py, lattice.py, CSR.py,
beams.py) on seeded inputs.  Every comparison here is bit-exact.
"""
import os

import numpy as np
import pytest

from oracle import cpu_ref as O

LAT_ELEMENTS = [dict(type="drift", L=0.1, nsep=1), dict(type="dipole", L=0.3, angle=0.05, nsep=1),
                dict(type="drift", L=0.4, nsep=2), dict(type="dipole", L=0.3, angle=-0.05, nsep=1)]


def _g(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def test_histogram_cic_2d_golden(golden_dir):
    d = _g(golden_dir, "ref_histogram_cic_2d.npz")
    a = d["args"]
    args = (int(a[0]), a[1], a[2], int(a[3]), a[4], a[5])
    assert np.array_equal(O.histogram_cic_2d(d["q1"], d["q2"], np.ones(d["q1"].shape), *args), d["h_ones"])
    assert np.array_equal(O.histogram_cic_2d(d["q1"], d["q2"], d["w"], *args), d["h_w"])


def test_interpolate_golden(golden_dir):
    d = _g(golden_dir, "ref_interpolate.npz")
    q = d["q3"]
    r = O.interpolate3D(q[:, 0], q[:, 1], q[:, 2], d["data3"], *d["mins"], *d["dels"])
    assert np.array_equal(r, d["r3"])
    assert np.array_equal(O.interpolate1D(d["q1"], d["data1"], float(d["min1"]), float(d["del1"])), d["r1"])


def test_lattice_golden(golden_dir):
    d = _g(golden_dir, "ref_lattice.npz")
    L = O.LatticeTables(LAT_ELEMENTS)
    for k in ("s", "rho", "distance", "nsep", "coords", "n_vec", "tau_vec"):
        assert np.array_equal(getattr(L, k), d[k]), k
    assert L.delta_x == d["delta_x"] and L.min_x == d["min_x"]


@pytest.mark.parametrize("tag,expect_bins", [("corr", (40, 48)), ("fallback", (100, 100))])
def test_get_df_golden(golden_dir, tag, expect_bins):
    d = _g(golden_dir, f"ref_get_df_{tag}.npz")
    r = O.get_DF(d["x"], d["z"], d["px"], 0.25, xbins=40, zbins=48, filter_order=2,
                 filter_window=7, velocity_threhold=1000.0)
    assert r["nbins"] == expect_bins
    for k in ("x_grids", "z_grids", "density", "vx", "density_x", "density_z", "vx_x"):
        assert np.array_equal(r[k], d[k]), k


def _history(d):
    T = O.DFTracker(xbins=40, zbins=48, filter_order=2, filter_window=7,
                    velocity_threhold=1000.0, upper_limit=40)
    flags = []
    for k in range(len(d["times"])):
        x, z, px = d["inputs"][k]
        T.get_DF(x, z, px, float(d["times"][k]))
        T.append_DF()
        flags.append(int(T.append_interpolant(float(d["fls"][k]), float(d["n_fl"]))))
        if k >= 1:
            T.build_interpolant()
    return T, flags


def test_history_golden(golden_dir):
    d = _g(golden_dir, "ref_history.npz")
    T, flags = _history(d)
    assert flags == list(d["regrid_flags"])
    assert np.array_equal(np.array(T.time_interp), d["time_interp"])
    assert np.array_equal(T.x_grid_interp, d["x_grid_interp"])
    assert np.array_equal(T.z_grid_interp, d["z_grid_interp"])
    ax = np.array([T.min_x, T.min_y, T.min_z, T.delta_x, T.delta_y, T.delta_z])
    assert np.array_equal(ax, d["axes"])
    for f in O.FIELDS:
        assert np.array_equal(T.data[f], d[f]), f


def test_wake_golden(golden_dir):
    h = _g(golden_dir, "ref_history.npz")
    T, _ = _history(h)
    w = _g(golden_dir, "ref_wake.npz")
    L = O.LatticeTables(LAT_ELEMENTS)
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]), slope=w["slope"],
                mean_x=float(w["mean_x"]))
    m = O.csr_mesh(w["beam_x"], w["beam_z"], 3, 4, 3, 3)
    for k in ("xmesh", "zmesh", "xrange", "zrange"):
        assert np.array_equal(m[k], w[k]), k
    dE, xk = O.calculate_2D_CSR(0.3, m, beam, 0.05, 1.0, 12, 16, L, T, float(w["csr_scaling"]), 3, 4)
    assert np.array_equal(dE, w["dE_dct"]) and np.array_equal(xk, w["x_kick"])
    for sl, key in ((2.5, "chirp_pos"), (-2.5, "chirp_neg")):
        b = dict(beam, slope=np.array([sl, w["slope"][1]]))
        res = np.array([O.csr_wake(0.3 + w["zmesh"][i], w["xmesh"][i], 0.3, b, 0.05, 1.0, 12, 16,
                                   L, T, float(w["csr_scaling"])) for i in w["chirp_rows"]])
        assert np.array_equal(res, w[key]), key
    # debug integrand grids of one observer (CSR.py:599-600)
    i = 6
    s, x = 0.3 + w["zmesh"][i], w["xmesh"][i]
    regs = O.wake_regions(s, x, 0.3, beam["sigma_z"], beam["sigma_x"], beam["slope"][0],
                          beam["mean_x"], 0.05, 1.0, 12, 16)
    for r, (xp, sp) in enumerate(regs, start=1):
        xpm, spm = np.meshgrid(xp, sp, indexing="ij")
        Iz, Ix = O.csr_integrand(s, x, 0.3, spm, xpm, L, T)
        assert np.array_equal(Iz, w[f"dbg_Iz{r}"]) and np.array_equal(Ix, w[f"dbg_Ix{r}"])


def test_apply_wakes_golden(golden_dir):
    a = _g(golden_dir, "ref_apply_wakes.npz")
    px1, pz1 = O.apply_wakes(a["x"], a["z"], a["px0"], a["pz0"], a["dE_dct"], a["x_kick"],
                             a["xrange"], a["zrange"], float(a["step_size"]), float(a["energy"]), 1)
    assert np.array_equal(px1, a["px1"]) and np.array_equal(pz1, a["pz1"])


def test_c_oracle_matches_golden_wake(golden_dir):
    """The oracle's C restatement (CPU baseline) against the reference's wakes."""
    from oracle import csr_ref_c as C
    h = _g(golden_dir, "ref_history.npz")
    T, _ = _history(h)
    w = _g(golden_dir, "ref_wake.npz")
    L = O.LatticeTables(LAT_ELEMENTS)
    keep = C._Keep()
    H = C.hist_from_tracker(T, keep)
    Lr = C.lat_from_tables(L, keep)
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]), slope=w["slope"],
                mean_x=float(w["mean_x"]))
    obs = np.stack([w["zmesh"], w["xmesh"]], axis=1)
    dE, xk = C.wake(H, Lr, 0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), obs, 2)
    for got, ref in ((dE, w["dE_dct"].ravel()), (xk, w["x_kick"].ravel())):
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    for sl, key in ((2.5, "chirp_pos"), (-2.5, "chirp_neg")):
        b = dict(beam, slope=np.array([sl, w["slope"][1]]))
        rows = w["chirp_rows"]
        dE, xk = C.wake(H, Lr, 0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), obs[rows], 2)
        res = np.stack([dE, xk], axis=1)
        assert np.abs(res - w[key]).max() <= 1e-12 * np.abs(w[key]).max()
    # the device ring layout ([Xi][Zi][slot][20] corner-packed records, field
    # f at element 4 f, the ring wrapping around) reads the same history
    data = np.stack([T.data[f] for f in O.FIELDS], axis=-1)
    nt, Xi, Zi = data.shape[:3]
    cap, head = nt + 3, 5
    rec = np.zeros((Xi, Zi, cap, 20))
    for i in range(nt):
        rec[:, :, (head + i) % cap, 0::4] = data[i]
    a = T.axes()
    H2 = C.hist_from_records(rec, head, nt, (a["min_t"], a["min_x"], a["min_z"], a["dt"],
                                             a["dx"], a["dz"]), keep)
    dE2, xk2 = C.wake(H2, Lr, 0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), obs, 1)
    dE1, xk1 = C.wake(H, Lr, 0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), obs, 1)
    assert np.array_equal(dE1, dE2) and np.array_equal(xk1, xk2)
