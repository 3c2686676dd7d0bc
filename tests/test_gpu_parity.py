"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden vectors, at sizes the oracle finishes in seconds.

Tolerances (float64 throughout):
  * per-particle cell indices and per-cell counts: bit-exact;
  * interpolate3D / interpolate1D / the integrand grids: bit-exact (same IEEE
    operation sequence, no fused multiply-add);
  * deposited grids: |d| <= 1e-13 * max (exact fixed-point sums against the
    reference's serial float64 sums);
  * density-function fields, history and kicks: <= 1e-9 * max;
  * wakes (dE_dct, x_kick): <= 1e-10 * max|W| (summation order differs).
"""
import os

import numpy as np
import pytest
import torch

from oracle import cpu_ref as O
from tests.gpu_util import DeviceHistory, DeviceLattice, rel_max, wake_params

pytestmark = pytest.mark.gpu

DEV = "cuda"
LAT_ELEMENTS = [dict(type="drift", L=0.1, nsep=1), dict(type="dipole", L=0.3, angle=0.05, nsep=1),
                dict(type="drift", L=0.4, nsep=2), dict(type="dipole", L=0.3, angle=-0.05, nsep=1)]


def _g(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(DEV)


@pytest.fixture(scope="module")
def K():
    from pydfcsr_amd import _lib, kernels
    _lib.load()
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return kernels


@pytest.fixture(scope="module")
def ws():
    from pydfcsr_amd import _lib
    return _lib.Workspace(DEV)


@pytest.fixture(scope="module")
def golden_history(golden_dir):
    d = _g(golden_dir, "ref_history.npz")
    T = O.DFTracker(xbins=40, zbins=48, filter_order=2, filter_window=7,
                    velocity_threhold=1000.0, upper_limit=40)
    for k in range(len(d["times"])):
        x, z, px = d["inputs"][k]
        T.get_DF(x, z, px, float(d["times"][k]))
        T.append_DF()
        T.append_interpolant(float(d["fls"][k]), float(d["n_fl"]))
        if k >= 1:
            T.build_interpolant()
    return d, T


# -- a1 ------------------------------------------------------------------------------
def test_beam_moments(K, ws):
    rng = np.random.default_rng(11)
    z = rng.normal(3e-5, 60e-6, 200003)
    x = 0.7 * z + rng.normal(-1e-5, 20e-6, z.size)
    m = K.beam_moments(_t(x), _t(z), ws)
    ref = O.beam_moments(x, z)
    for k in ("mean_x", "mean_z", "sigma_x", "sigma_z", "mean_xt", "sigma_xt", "slice_sigma_x"):
        assert abs(m[k] - ref[k]) <= 1e-12 * max(abs(ref[k]), ref["sigma_x"]), k
    assert np.allclose(m["slope"], ref["slope"], rtol=1e-11, atol=1e-17)
    assert m["n_slice"] == ref["n_slice"]
    m2 = K.beam_moments(_t(x), _t(z), ws)
    assert all(m2[k] == m[k] for k in m if k != "slope")   # deterministic


def test_beam_statistics(K, ws):
    """Per-step record (CSR.py:837-859): Twiss of x and y from np.cov of
    (x, px, pz) / (y, py, pz), energy mean and spread, the 6 x 6 covariance;
    a correlated beam with dispersion, 300 007 particles, against the oracle
    within 1e-11 relative; deterministic."""
    rng = np.random.default_rng(12)
    n = 300007
    A = rng.standard_normal((6, 6)) * np.array([1e-4, 1e-5, 2e-4, 3e-5, 5e-5, 1e-3])[:, None]
    X = A @ rng.standard_normal((6, n)) + np.array([1e-5, 2e-6, -2e-5, 0, 3e-6, 2e-3])[:, None]
    p0c, mc2 = 1.0e9, 510998.95069
    coords = _t(X)
    v = K.beam_statistics(coords, p0c, mc2, ws)
    d = K.statistics_dict(v)
    ref = O.beam_statistics(*X, p0c, mc2)
    for k, r in ref["twiss"].items():
        assert abs(d["twiss"][k] - r) <= 1e-11 * abs(r), (k, d["twiss"][k], r)
    for k in ("mean_energy", "sigma_energy"):
        assert abs(d[k] - ref[k]) <= 1e-12 * abs(ref[k]), k
    assert np.allclose(d["cov"], ref["cov"], rtol=1e-11, atol=0)
    assert np.allclose(d["mean"], X.mean(axis=1), rtol=1e-12, atol=1e-20)
    v2 = K.beam_statistics(coords, p0c, mc2, ws)
    assert torch.equal(v, v2)


# -- a3 ------------------------------------------------------------------------------
def test_histogram_golden_and_cells(K, ws, golden_dir):
    d = _g(golden_dir, "ref_histogram_cic_2d.npz")
    a = d["args"]
    args = (int(a[0]), float(a[1]), float(a[2]), int(a[3]), float(a[4]), float(a[5]))
    q1, q2 = _t(d["q1"]), _t(d["q2"])
    h1, cells, counts = K.histogram_cic_2d(q1, q2, None, *args, ws, want_cells=True)
    h2 = K.histogram_cic_2d(q1, q2, _t(d["w"]), *args, ws)
    assert rel_max(h1.cpu(), d["h_ones"]) <= 1e-13
    assert rel_max(h2.cpu(), d["h_w"]) <= 1e-13
    i1, i2, _, _ = O.cic_cells(d["q1"], d["q2"], *args)
    c = cells.cpu().numpy()
    assert np.array_equal(c[:, 0], i1) and np.array_equal(c[:, 1], i2)
    assert np.array_equal(counts.cpu().numpy(), O.cell_counts(i1, i2, args[0], args[3]))


def test_histogram_edges_and_determinism(K, ws):
    # particles on nodes, on the far edge, below the grid, zero weights
    q1 = np.array([0.0, 1.0, 4.0, -0.25, 3.5, 2.0, 3.999999999])
    q2 = np.array([0.0, 4.0, 4.0, 4.0, 9.0, 10.0, 9.999999999])
    w = np.array([1.0, 2.0, 3.0, 1.0, 0.0, 5.0, 1.0])
    h = K.histogram_cic_2d(_t(q1), _t(q2), _t(w), 4, 0.0, 4.0, 5, 0.0, 10.0, ws).cpu().numpy()
    ref = O.histogram_cic_2d_loop(q1, q2, w, 4, 0.0, 4.0, 5, 0.0, 10.0)
    assert np.allclose(h, ref, rtol=0, atol=1e-15 * np.abs(ref).max())
    # empty input
    e = K.histogram_cic_2d(_t(np.zeros(0)), _t(np.zeros(0)), None, 3, 0.0, 1.0, 3, 0.0, 1.0, ws)
    assert torch.count_nonzero(e) == 0
    # 2^20 particles: bit-identical on repeat
    rng = np.random.default_rng(3)
    x = _t(rng.normal(size=1 << 20))
    z = _t(rng.normal(size=1 << 20))
    px = _t(rng.normal(size=1 << 20) * 1e-3)
    d1, v1 = K.deposit_density_vx(x, z, px, 64, -5.0, 5.0, 64, -5.0, 5.0, ws)
    d1, v1 = d1.clone(), v1.clone()
    d2, v2 = K.deposit_density_vx(x, z, px, 64, -5.0, 5.0, 64, -5.0, 5.0, ws)
    assert torch.equal(d1, d2) and torch.equal(v1, v2)
    xr, zr, pr = x.cpu().numpy(), z.cpu().numpy(), px.cpu().numpy()
    assert rel_max(d1.cpu(), O.histogram_cic_2d(xr, zr, np.ones(xr.size), 64, -5.0, 5.0, 64, -5.0, 5.0)) <= 1e-13
    assert rel_max(v1.cpu(), O.histogram_cic_2d(xr, zr, pr, 64, -5.0, 5.0, 64, -5.0, 5.0)) <= 1e-13


# -- a11 / a12 ----------------------------------------------------------------------
def test_interpolate_golden_bitexact(K, golden_dir):
    from pydfcsr_amd import _lib
    d = _g(golden_dir, "ref_interpolate.npz")
    data = d["data3"]
    nt, nx, nz = data.shape
    from tests.gpu_util import corner_records, ring64
    stack = np.zeros((nt, nx, nz, 5))
    stack[..., 0] = data
    h = _t(ring64(corner_records(stack), nt, 0))
    H = _lib.History(h.data_ptr(), None, None, None, None, nt, 0, nt, nx, nz,
                     *[float(v) for v in d["mins"]],
                     *[float(v) for v in d["dels"]])
    r = K.interpolate3D(H, 0, _t(d["q3"]), DEV).cpu().numpy()
    assert np.array_equal(r, d["r3"])
    tab = np.zeros((d["data1"].size, 6))
    tab[:, 3] = d["data1"]
    tt = _t(tab)
    dist = _t(np.array([1.0]))
    L = _lib.Lattice(tt.data_ptr(), tab.shape[0], 1, float(d["min1"]), float(d["del1"]),
                     dist.data_ptr(), dist.data_ptr())
    r1 = K.interpolate1D(L, 3, _t(d["q1"])).cpu().numpy()
    assert np.array_equal(r1, d["r1"])


# -- a4-a6 ---------------------------------------------------------------------------
@pytest.mark.parametrize("tag,bins", [("corr", (40, 48)), ("fallback", (100, 100))])
def test_get_df_golden(golden_dir, tag, bins):
    from pydfcsr_amd.deposit import DF_tracker
    d = _g(golden_dir, f"ref_get_df_{tag}.npz")
    tr = DF_tracker(dict(xbins=40, zbins=48, filter_order=2, filter_window=7,
                         velocity_threhold=1000.0), device=DEV)
    m = tr.get_DF(_t(d["x"]), _t(d["z"]), _t(d["px"]), 0.25)
    assert (m.nbx, m.nbz) == bins
    f = tr.current_fields()
    assert np.allclose(f["x_grids"], d["x_grids"], rtol=1e-13, atol=0)
    for k in ("density", "vx", "density_x", "density_z", "vx_x"):
        assert rel_max(f[k].cpu(), d[k]) <= 1e-9, k


@pytest.mark.parametrize("bins,window,order", [((7, 9), 7, 2), ((33, 65), 9, 3),
                                                ((65, 33), 5, 2), ((17, 40), 13, 4)])
def test_get_df_ragged_grids_vs_oracle(bins, window, order):
    """The smoothing runs as tiles of 8 x 32 outputs; grids that end inside a
    tile, are narrower than a tile, or whose edge windows (the first / last
    `window` samples) reach past a partial tile's halo give the oracle's
    get_DF (deposit.py:145-245) on a correlated beam (configured grid)."""
    from pydfcsr_amd.deposit import DF_tracker
    rng = np.random.default_rng(sum(bins) + window)
    n = 20000
    z = rng.normal(0.0, 60e-6, n)
    x = rng.normal(0.0, 20e-6, n) + 4.0 * z
    px = rng.normal(0.0, 1e-4, n) + 0.3 * z / 60e-6 * 1e-4
    cfg = dict(xbins=bins[0], zbins=bins[1], filter_order=order, filter_window=window,
               velocity_threhold=1000.0)
    ref = O.get_DF(x, z, px, 0.1, **cfg)
    assert ref["nbins"] == bins
    tr = DF_tracker(cfg, device=DEV)
    m = tr.get_DF(_t(x), _t(z), _t(px), 0.1)
    assert (m.nbx, m.nbz) == bins
    f = tr.current_fields()
    for k in ("density", "vx", "density_x", "density_z", "vx_x"):
        a, b = f[k].cpu().numpy(), ref[k]
        assert np.array_equal(np.isnan(a), np.isnan(b)), k
        assert rel_max(np.nan_to_num(a), np.nan_to_num(b)) <= 1e-9, k


# -- a8-a10 --------------------------------------------------------------------------
def test_history_golden(golden_dir):
    """The device history against the reference's: re-grid flags, window,
    axes and fields.  Through the same full re-grids, window pops and ring
    growth, the float32 time-difference ring stays slice i + 1 minus slice i
    of the float32 cell-major ring (bit-exact), 0 for the newest slice, the
    span table the one the float64 cell-major ring gives, and each x row of
    the cell-major rings ends with a copy of its last cell; the float64 cells
    equal the corner-packed ring's node values."""
    from pydfcsr_amd.deposit import DF_tracker
    from tests.gpu_util import zspan_table
    d = _g(golden_dir, "ref_history.npz")
    tr = DF_tracker(dict(xbins=40, zbins=48, filter_order=2, filter_window=7,
                         velocity_threhold=1000.0, upper_limit=40), device=DEV, capacity=3)
    flags = []
    for k in range(len(d["times"])):
        x, z, px = d["inputs"][k]
        tr.get_DF(_t(x), _t(z), _t(px), float(d["times"][k]))
        tr.append_DF()
        flags.append(int(tr.append_interpolant(float(d["fls"][k]), float(d["n_fl"]))))
        idx = [tr._slot(i) for i in range(tr.nt)]
        c32 = tr.cells32.cpu().numpy()[idx]
        want = np.zeros_like(c32)
        want[:-1] = c32[1:] - c32[:-1]
        assert np.array_equal(tr.dcells32.cpu().numpy()[idx], want), k
        c64 = tr.cells.cpu().numpy()[idx]
        # each x row ends with a copy of its last cell
        assert np.array_equal(c64[:, :, -1], c64[:, :, -2]), k
        assert np.array_equal(c32[:, :, -1], c32[:, :, -2]), k
        zs = tr.zspan.cpu().numpy()[:, idx]
        assert np.array_equal(zs, zspan_table(c64[:, :, :-1], len(idx), 0)), k
        if k >= 1:
            tr.build_interpolant()
    assert flags == list(d["regrid_flags"])
    assert np.array_equal(np.array([m.t for m in tr.log]), d["time_interp"])
    ax = np.array([tr.min_x, tr.min_y, tr.min_z, tr.delta_x, tr.delta_y, tr.delta_z])
    assert np.allclose(ax, d["axes"], rtol=1e-13, atol=0)
    hf = tr.history_fields()
    for f in ("density", "density_x", "density_z", "vx", "vx_x"):
        assert rel_max(hf[f].cpu(), d[f]) <= 1e-9, f
    # the float64 cells are the corner-packed ring's grid-node values, exactly
    c64 = tr.cells.cpu().numpy()[[tr._slot(i) for i in range(tr.nt)]]
    for k, f in enumerate(("density", "density_x", "density_z", "vx", "vx_x")):
        assert np.array_equal(c64[:, :, :-1, k], hf[f].cpu().numpy()), f


# -- a13 -----------------------------------------------------------------------------
def test_integrand_golden_bitexact(K, golden_dir, golden_history):
    _, T = golden_history
    w = _g(golden_dir, "ref_wake.npz")
    L = O.LatticeTables(LAT_ELEMENTS)
    DH = DeviceHistory(T, DEV, wrap=2)
    DL = DeviceLattice(L, DEV)
    i = 6
    s, x = 0.3 + w["zmesh"][i], w["xmesh"][i]
    xps = [w["dbg_xp_w"], w["dbg_xp_n"], w["dbg_xp_n"]]
    sps = [w["dbg_sp1"], w["dbg_sp2"], w["dbg_sp3"]]
    for r in range(3):
        Iz, Ix = K.csr_integrand(DH.desc, DL.desc, s, x, 0.3, _t(xps[r]), _t(sps[r]))
        assert np.array_equal(Iz.cpu().numpy(), w[f"dbg_Iz{r + 1}"])
        assert np.array_equal(Ix.cpu().numpy(), w[f"dbg_Ix{r + 1}"])


# -- a13-a15 -------------------------------------------------------------------------
def _golden_wake_inputs(golden_dir, golden_history, wrap=0):
    _, T = golden_history
    w = _g(golden_dir, "ref_wake.npz")
    L = O.LatticeTables(LAT_ELEMENTS)
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]), slope=w["slope"],
                mean_x=float(w["mean_x"]))
    xr, zr = w["xrange"], w["zrange"]
    P = wake_params(0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                    (xr[0], xr[-1], xr.size), (zr[0], zr[-1], zr.size))
    return w, beam, DeviceHistory(T, DEV, wrap=wrap), DeviceLattice(L, DEV), P


def test_wake_golden(K, ws, golden_dir, golden_history):
    w, beam, DH, DL, P = _golden_wake_inputs(golden_dir, golden_history)
    out = K.wake_mesh(DH.desc, DL.desc, P, 0, DEV, ws)
    assert np.allclose(out["xmesh"].cpu().numpy(), w["xmesh"], rtol=0, atol=1e-18)
    assert np.array_equal(out["zmesh"].cpu().numpy(), w["zmesh"])
    assert rel_max(out["dE_dct"].cpu().view(3, 4), w["dE_dct"]) <= 1e-10
    assert rel_max(out["x_kick"].cpu().view(3, 4), w["x_kick"]) <= 1e-10
    # deterministic, and independent of where the ring wraps
    out2 = K.wake_mesh(DH.desc, DL.desc, P, 0, DEV, ws)
    assert torch.equal(out["dE_dct"], out2["dE_dct"]) and torch.equal(out["x_kick"], out2["x_kick"])
    _, _, DHw, _, _ = _golden_wake_inputs(golden_dir, golden_history, wrap=3)
    out3 = K.wake_mesh(DHw.desc, DL.desc, P, 0, DEV, ws)
    assert torch.equal(out["dE_dct"], out3["dE_dct"]) and torch.equal(out["x_kick"], out3["x_kick"])
    # with the cell-major ring (its slots without a slice hold NaN: a read of
    # the wrong slot shows), wrapped or not: identical
    _, T = golden_history
    for wrap in (0, 1, 3):
        DHc = DeviceHistory(T, DEV, wrap=wrap, rings=("fp64", "fp32", "fp32t", "cells"))
        o = K.wake_mesh(DHc.desc, DL.desc, P, 0, DEV, ws)
        assert torch.equal(out["dE_dct"], o["dE_dct"]), wrap
        assert torch.equal(out["x_kick"], o["x_kick"]), wrap


def _last_z_lookups(T, L, beam, t, obs):
    """Valid near-region lookups (regions 2 and 3 of the non-chirp branch,
    CSR.py:461-485) of the observers obs[k] = (s - t, x) whose retarded z'
    falls in the history's last z cell, where interpolate3D clamps z1 to z0;
    from the oracle's region grids and lattice tables."""
    a = T.axes()
    nt, Zi = T.data["density"].shape[0], T.data["density"].shape[2]

    def i1(col, v):
        return O.interpolate1D(v, col, L.min_x, L.delta_x)

    hits = 0
    for z, x in obs:
        s = t + z
        regs = O.wake_regions(s, x, t, beam["sigma_z"], beam["sigma_x"], beam["slope"][0],
                              beam["mean_x"], 0.05, 1.0, 12, 16)
        for xp1, sp1 in regs[1:]:
            xp, sp = [g.ravel() for g in np.meshgrid(xp1, sp1, indexing="ij")]
            sv = np.array([s])
            rx = (i1(L.coords[:, 0], sv)[0] - i1(L.coords[:, 0], sp)
                  + x * i1(L.n_vec[:, 0], sv)[0] - xp * i1(L.n_vec[:, 0], sp))
            ry = (i1(L.coords[:, 1], sv)[0] - i1(L.coords[:, 1], sp)
                  + x * i1(L.n_vec[:, 1], sv)[0] - xp * i1(L.n_vec[:, 1], sp))
            R = np.sqrt(rx ** 2 + ry ** 2)
            ut = (t - R - a["min_t"]) / a["dt"]
            ux = (xp - a["min_x"]) / a["dx"]
            uz = (sp - (t - R) - a["min_z"]) / a["dz"]
            ok = (ut > -1) & (ut < nt) & (ux > -1) & (ux < T.data["density"].shape[1])
            hits += int(np.count_nonzero(ok & (uz >= Zi - 1) & (uz < Zi)))
    return hits


class _RolledHistory:
    """A history's fields rolled by k cells along z (its axes unchanged): the
    golden bunch's core moved to the top z edge of the interp grid."""

    def __init__(self, T, k):
        self.data = {f: np.roll(v, k, axis=2) for f, v in T.data.items()}
        self._axes = T.axes()

    def axes(self):
        return self._axes


@pytest.mark.parametrize("precision", [0, 1])
def test_wake_cells_last_z_bitexact(K, ws, golden_dir, golden_history, precision):
    """Observers at the top z edge of a history whose bunch core lies there,
    so that near-region lookups fall in the grid's last z cell on nonzero
    fields (interpolate3D clamps z1 to z0 there; the cell-major rings read
    the pair (z - 1, z) and give its upper cell to both z corners): the
    wakes through the cell-major rings equal those through the
    corner-packed rings bit for bit, FP64 (float64 cells) and mixed (float32
    cells and their time differences), and match the oracle within 1e-10
    (FP64) / 1e-4 (mixed) of max|W|."""
    w, beam, _, DL, _ = _golden_wake_inputs(golden_dir, golden_history)
    _, T0 = golden_history
    Zi = T0.data["density"].shape[2]
    T = _RolledHistory(T0, Zi // 2)
    L = O.LatticeTables(LAT_ELEMENTS)
    a = T.axes()
    max_z = a["min_z"] + (Zi - 1) * a["dz"]
    obs_np = np.array([(max_z - f * a["dz"], beam["mean_x"] + g * beam["sigma_x"])
                       for f in (0.05, 0.3, 0.7, 1.5) for g in (-1.0, 0.0, 1.0)])
    assert _last_z_lookups(T, L, beam, 0.3, obs_np) > 0
    obs = _t(obs_np)   # the parameters hold its device address: keep it alive
    P = wake_params(0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), (0.0, 0.0, 1),
                    (0.0, 0.0, len(obs_np)), observers=obs)
    rings = ("fp64",) if precision == 0 else ("fp32", "fp32t")
    cells = ("cells",) if precision == 0 else ("cells32",)
    res = []
    for r in (rings, rings + cells):
        DH = DeviceHistory(T, DEV, wrap=2, rings=r)
        o = K.wake_mesh(DH.desc, DL.desc, P, precision, DEV, ws)
        res.append(np.stack([o["dE_dct"].cpu().numpy(), o["x_kick"].cpu().numpy()]))
    assert np.array_equal(res[0], res[1])
    ref = np.array([O.csr_wake(0.3 + z, x, 0.3, beam, 0.05, 1.0, 12, 16, L, T,
                               float(w["csr_scaling"])) for z, x in obs_np]).T
    assert np.abs(ref[0]).max() > 0
    tol = 1e-10 if precision == 0 else 1e-4
    for k in range(2):
        assert rel_max(res[0][k], ref[k]) <= tol, k


@pytest.mark.parametrize("slope0,key", [(2.5, "chirp_pos"), (-2.5, "chirp_neg")])
def test_wake_chirp_branch_golden(K, ws, golden_dir, golden_history, slope0, key):
    w, beam, DH, DL, _ = _golden_wake_inputs(golden_dir, golden_history)
    b = dict(beam, slope=np.array([slope0, w["slope"][1]]))
    rows = w["chirp_rows"]
    obs = _t(np.stack([w["zmesh"][rows], w["xmesh"][rows]], axis=1))
    P = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), (0.0, 0.0, 1),
                    (0.0, 0.0, len(rows)), observers=obs)
    out = K.wake_mesh(DH.desc, DL.desc, P, 0, DEV, ws)
    res = np.stack([out["dE_dct"].cpu().numpy(), out["x_kick"].cpu().numpy()], axis=1)
    assert rel_max(res[:, 0], w[key][:, 0]) <= 1e-10
    assert rel_max(res[:, 1], w[key][:, 1]) <= 1e-10


# -- a16 -----------------------------------------------------------------------------
def test_apply_wakes_golden(K, ws, golden_dir):
    a = _g(golden_dir, "ref_apply_wakes.npz")
    x, z = _t(a["x"]), _t(a["z"])
    px, pz = _t(a["px0"]), _t(a["pz0"])
    m = K.beam_moments(x, z, ws)
    xr, zr = a["xrange"], a["zrange"]
    K.apply_wakes(px, pz, x, z, _t(a["dE_dct"]), _t(a["x_kick"]), xr[0], xr[-1], zr[0], zr[-1],
                  m["slope0"], m["slope1"], float(a["step_size"]), float(a["energy"]), 1,
                  xr.size, zr.size)
    dpz_ref = a["pz1"] - a["pz0"]
    dpx_ref = a["px1"] - a["px0"]
    assert rel_max(pz.cpu().numpy() - a["pz0"], dpz_ref) <= 1e-9
    assert rel_max(px.cpu().numpy() - a["px0"], dpx_ref) <= 1e-9


def test_wake_mixed_precision_golden(K, ws, golden_dir, golden_history):
    """Mixed form (FP32 integrand, FP64 retardation and sums) against the
    reference's float64 wakes: max |err| <= 1e-4 * max|W| for both wakes."""
    w, beam, DH, DL, P = _golden_wake_inputs(golden_dir, golden_history)
    out = K.wake_mesh(DH.desc, DL.desc, P, 1, DEV, ws)
    e_z = rel_max(out["dE_dct"].cpu().view(3, 4), w["dE_dct"])
    e_x = rel_max(out["x_kick"].cpu().view(3, 4), w["x_kick"])
    assert e_z <= 1e-4 and e_x <= 1e-4, (e_z, e_x)
    out2 = K.wake_mesh(DH.desc, DL.desc, P, 1, DEV, ws)
    assert torch.equal(out["dE_dct"], out2["dE_dct"])   # deterministic
    for slope0, key in ((2.5, "chirp_pos"), (-2.5, "chirp_neg")):
        b = dict(beam, slope=np.array([slope0, w["slope"][1]]))
        rows = w["chirp_rows"]
        obs = _t(np.stack([w["zmesh"][rows], w["xmesh"][rows]], axis=1))
        Pc = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), (0.0, 0.0, 1),
                         (0.0, 0.0, len(rows)), observers=obs)
        o = K.wake_mesh(DH.desc, DL.desc, Pc, 1, DEV, ws)
        assert rel_max(o["dE_dct"].cpu().numpy(), w[key][:, 0]) <= 1e-4
        assert rel_max(o["x_kick"].cpu().numpy(), w[key][:, 1]) <= 1e-4


@pytest.mark.parametrize("Xi,Zi", [(6, 8), (5, 7)])   # float4 and scalar plane paths
def test_cells32_tdiff_matches_numpy(K, Xi, Zi):
    """The time differences of the float32 cell-major ring: for each logical
    slice i of a wrapped window (first .. nt-1), slice i + 1 minus slice i in
    float32 (numpy's float32 subtraction, bit-exact), 0 for the newest; no
    other slot is written."""
    rng = np.random.default_rng(Xi * 31 + Zi)
    cap, head, nt = 9, 6, 7
    scale = 10.0 ** rng.uniform(-3, 3, (cap, 1, 1, 5))
    c = (rng.standard_normal((cap, Xi, Zi, 5)) * scale).astype(np.float32)
    sentinel = 12345.0
    cd = torch.from_numpy(c).to(DEV)
    for first in (0, 3, nt - 1, nt):
        d = torch.full((cap, Xi, Zi, 5), sentinel, dtype=torch.float32, device=DEV)
        K.cells32_tdiff(cd, d, head, nt, first)
        want = np.full_like(c, sentinel)
        for i in range(first, nt):
            s0, s1 = (head + i) % cap, (head + i + 1) % cap
            want[s0] = 0.0 if i == nt - 1 else c[s1] - c[s0]
        assert np.array_equal(d.cpu().numpy(), want), first


@pytest.mark.parametrize("fp32", [False, True])
def test_history_zspan_matches_numpy(K, fp32):
    """The span table from the float64 (or float32) cell-major ring: per x
    row and slot of each slice first .. nt-1 of a wrapped window, the z span
    of cells of rows x, x + 1 in slices i, i + 1 with a nonzero density or
    gradient or a velocity that is not finite; no other slot written.  Slices
    with rows of zeros, a NaN vx and values that are float32-exact."""
    from tests.gpu_util import zspan_table, cell_ring
    rng = np.random.default_rng(7 + fp32)
    cap, head, nt, Xi, Zi = 9, 6, 7, 37, 53
    data = np.zeros((nt, Xi, Zi, 5))
    for i in range(nt):   # a band of contributing cells that moves, some rows empty
        z0, z1 = 3 + 4 * i, 20 + 3 * i
        data[i, 2:30, z0:z1, :3] = rng.standard_normal((28, z1 - z0, 3))
        data[i, :, :, 3:] = rng.standard_normal((Xi, Zi, 2))
    data[2, 5:9] = 0.0
    data[3, 11, 40, 3] = np.nan       # vx not finite outside the band
    data[4, 1:3, 50, 1] = 1.0         # lone gradients past the band
    data = data.astype(np.float32).astype(np.float64)
    dt = np.float32 if fp32 else np.float64
    ring = torch.from_numpy(cell_ring(data.astype(dt), cap, head, dt)).to(DEV)
    sentinel = np.int32(0x12345)
    for first in (0, 3, nt - 1, nt):
        zs = torch.full((Xi, cap), int(sentinel), dtype=torch.int32, device=DEV)
        K.history_zspan(None if fp32 else ring, ring if fp32 else None, zs, head, nt, first)
        want = zspan_table(data, cap, head)
        for i in range(first):
            want[:, (head + i) % cap] = sentinel
        for s in range(cap):
            if (s - head) % cap >= nt:
                want[:, s] = sentinel
        assert np.array_equal(zs.cpu().numpy(), want), first


def test_wake_zspan_skips_exact_zeros(K, ws, golden_dir, golden_history):
    """With the span table the wakes skip the lookups whose gathered density
    and gradients are all 0 (an integrand of exactly 0): fewer points are
    evaluated and the wakes are bit-identical, FP64 (corner-packed and
    cell-major rings) and mixed (corner-packed and cell-major float32), on
    the golden mesh and on the explicit observers of the chirp branch."""
    w, beam, _, DL, P = _golden_wake_inputs(golden_dir, golden_history)
    _, T = golden_history
    b = dict(beam, slope=np.array([2.5, w["slope"][1]]))
    rows = w["chirp_rows"]
    obs = _t(np.stack([w["zmesh"][rows], w["xmesh"][rows]], axis=1))
    Pc = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), (0.0, 0.0, 1),
                     (0.0, 0.0, len(rows)), observers=obs)
    cases = [(0, ("fp64",)), (0, ("fp64", "cells")), (1, ("fp32", "fp32t")),
             (1, ("fp32", "fp32t", "cells32"))]
    for prec, rings in cases:
        for params in (P, Pc):
            res, ev = [], []
            for extra in ((), ("zspan",)):
                DH = DeviceHistory(T, DEV, wrap=2, rings=rings + extra)
                o = K.wake_mesh(DH.desc, DL.desc, params, prec, DEV, ws)
                res.append(np.stack([o["dE_dct"].cpu().numpy(), o["x_kick"].cpu().numpy()]))
                ev.append(K.wake_points_evaluated(DH.desc, DL.desc, params, prec, DEV, ws))
            assert np.array_equal(res[0], res[1]), (prec, rings)
            assert ev[1] < ev[0], (prec, rings, ev)


def test_wake_far_zspan_skip_chirp_mesh(K, ws, golden_dir, golden_history):
    """On the chirp branch the observer-major region 0 skips, on the span
    table, the lookups that gather zero density and gradients (an integrand of
    exactly 0): the wakes are bit-identical with and without the table (FP64
    and mixed), and fewer points are evaluated than by the same observers
    given explicitly (region 0 then runs by row chunks, with no such skip;
    the golden history holds 335 such points of region 0 there), whose
    wakes agree to 1e-12 (FP64) / 1e-10 (mixed: the same float32 points,
    summed in float64 in another order).  The batched launch, beside a task
    off the chirp branch, is bit-identical with and without the table too."""
    w, beam, _, DL, _ = _golden_wake_inputs(golden_dir, golden_history)
    _, T = golden_history
    b = dict(beam, slope=np.array([2.5, w["slope"][1]]))
    xr, zr = w["xrange"], w["zrange"]
    P = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                    (xr[0], xr[-1], xr.size), (zr[0], zr[-1], zr.size))
    for prec, rings, tol in ((0, ("fp64", "cells"), 1e-12),
                             (1, ("fp32", "fp32t", "cells32"), 1e-10)):
        res, ev, hs = [], [], []
        for extra in ((), ("zspan",)):
            DH = DeviceHistory(T, DEV, wrap=2, rings=rings + extra)
            hs.append(DH)
            o = K.wake_mesh(DH.desc, DL.desc, P, prec, DEV, ws)
            res.append(np.stack([o["dE_dct"].cpu().numpy(), o["x_kick"].cpu().numpy()]))
            ev.append(K.wake_points_evaluated(DH.desc, DL.desc, P, prec, DEV, ws))
        assert np.array_equal(res[0], res[1]), (prec, rings)
        assert np.isfinite(res[1]).all() and np.abs(res[1]).max() > 0
        # the batched launch (serving) makes the same choice per task, here beside a task
        # off the chirp branch (the golden slope): bit-identical with and without the
        # table, and equal to the single launch to rounding (another kernel's code)
        P0 = wake_params(0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                         (xr[0], xr[-1], xr.size), (zr[0], zr[-1], zr.size))
        bat = []
        for H in hs:
            ob = [{k: torch.empty(xr.size * zr.size, dtype=torch.float64, device=DEV)
                   for k in ("dE_dct", "x_kick")} for _ in range(2)]
            K.wake_mesh_batched([H.desc, H.desc], [DL.desc, DL.desc], [P, P0], prec, ob, DEV, ws)
            bat.append([[ob[j][k].cpu().numpy() for k in ("dE_dct", "x_kick")] for j in (0, 1)])
        o0 = K.wake_mesh(DH.desc, DL.desc, P0, prec, DEV, ws)
        for f, k in enumerate(("dE_dct", "x_kick")):
            for j in (0, 1):
                assert np.array_equal(bat[0][j][f], bat[1][j][f]), (prec, j, k)
            assert rel_max(bat[1][0][f], res[1][f]) <= 1e-13, (prec, k)
            assert rel_max(bat[1][1][f], o0[k].cpu().numpy()) <= 1e-13, (prec, k)
        obs = torch.stack([o["zmesh"], o["xmesh"]], dim=1).contiguous()
        Po = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]), (0.0, 0.0, 1),
                         (0.0, 0.0, obs.shape[0]), observers=obs)
        oo = K.wake_mesh(DH.desc, DL.desc, Po, prec, DEV, ws)
        ev_obs = K.wake_points_evaluated(DH.desc, DL.desc, Po, prec, DEV, ws)
        assert ev[1] < ev_obs <= ev[0], (prec, ev, ev_obs)
        for f, k in enumerate(("dE_dct", "x_kick")):
            assert rel_max(oo[k].cpu().numpy(), res[1][f]) <= tol, (prec, k)


def test_wake_mixed_precision_ring_layouts(K, ws, golden_dir, golden_history):
    """The mixed form reads either float32 ring layout (chunked, time-innermost,
    or both: region 0 the time-innermost one, the later regions the chunked
    one), with or without the float32 cell-major ring and its time
    differences for the later regions (slots without a slice hold NaN), with
    the same arithmetic, so the wakes are bit-identical across layouts;
    without the float64 ring, vx at the observer comes from a float32 ring,
    within 1e-4 of max|W| of the reference's float64 wakes."""
    w, beam, _, DL, P = _golden_wake_inputs(golden_dir, golden_history)
    _, T = golden_history
    res = {}
    for rings in (("fp64", "fp32", "fp32t"), ("fp64", "fp32"), ("fp64", "fp32t"),
                  ("fp32", "fp32t"), ("fp32",), ("fp32t",)):
        for cells in ((), ("cells32",)):
            DH = DeviceHistory(T, DEV, wrap=2, rings=rings + cells)
            out = K.wake_mesh(DH.desc, DL.desc, P, 1, DEV, ws)
            res[rings + cells] = np.stack([out["dE_dct"].cpu().numpy(),
                                           out["x_kick"].cpu().numpy()])
            e_z = rel_max(res[rings + cells][0].reshape(3, 4), w["dE_dct"])
            e_x = rel_max(res[rings + cells][1].reshape(3, 4), w["x_kick"])
            assert e_z <= 1e-4 and e_x <= 1e-4, (rings, cells, e_z, e_x)
    with64 = [r for r in res if "fp64" in r]
    without = [r for r in res if "fp64" not in r]
    for group in (with64, without):
        for r in group[1:]:
            assert np.array_equal(res[r], res[group[0]]), (r, group[0])
    # FP64 precision refuses a history without the float64 ring
    DH = DeviceHistory(T, DEV, rings=("fp32", "fp32t"))
    with pytest.raises(Exception, match="FP64 history ring"):
        K.wake_mesh(DH.desc, DL.desc, P, 0, DEV, ws)


# -- tracking (bmadx.track_element, CSR.py:146-199) ----------------------------------
def _elem(**kw):
    from pydfcsr_amd import _lib
    e = _lib.Element()
    e.num_steps = 1
    for k, v in kw.items():
        setattr(e, k, v)
    return e


def test_track_golden_single_particles(golden_dir):
    """The reference notebook's bmadx results (Drift, SBend, Quadrupole on one
    particle at 1e-3 in every coordinate, p0c 4e7 eV) through the device
    kernel, within 1e-12 relative per coordinate."""
    import json
    from pydfcsr_amd import _lib
    from pydfcsr_amd import kernels as K
    d = json.load(open(os.path.join(golden_dir, "ref_bmadx_tracking.json")))
    p = d["particle_in"]
    keys = ("x", "px", "y", "py", "z", "pz")
    for case in d["cases"]:
        c = _t(np.array([[p[k]] for k in keys]))
        if case["element"] == "Drift":
            e = _elem(type=_lib.DFCSR_ELE_DRIFT, L=case["L"])
        elif case["element"] == "SBend":
            e = _elem(type=_lib.DFCSR_ELE_SBEND, L=case["L"], g=case["G"], e1=case["E1"],
                      e2=case["E2"])
        else:
            e = _elem(type=_lib.DFCSR_ELE_QUAD, L=case["L"], k1=case["K1"])
        K.track_element(c, e, p["p0c"], p["mc2"])
        r = c.cpu().numpy()[:, 0]
        g = np.array([case["out"][k] for k in keys])
        assert np.all(np.abs(r - g) <= 1e-12 * np.abs(g)), (case["element"], r - g)


@pytest.mark.parametrize("kind", ["drift", "sbend", "sbend_faces", "quad", "quad_offset",
                                  "sextupole"])
def test_track_beam_vs_oracle(kind):
    """A 6D Gaussian beam (100 003 particles, amplitudes ~1e-3) through one
    piece of each element type: device kernel against the oracle's maps,
    within 1e-12 of each coordinate's largest value."""
    from pydfcsr_amd import _lib
    from pydfcsr_amd import kernels as K
    rng = np.random.default_rng(21)
    X = rng.normal(0.0, 1.0, (6, 100003)) * np.array([1e-3, 5e-4, 8e-4, 3e-4, 1e-3, 2e-3])[:, None]
    p0c, mc2 = 1e9, 510998.95069
    if kind == "drift":
        e = _elem(type=_lib.DFCSR_ELE_DRIFT, L=0.7)
        ref = O.track_drift(X, 0.7, p0c, mc2)
    elif kind == "sbend":
        e = _elem(type=_lib.DFCSR_ELE_SBEND, L=0.3, g=0.05 / 0.3)
        ref = O.track_sbend(X, 0.3, 0.05 / 0.3, 0.0, 0.0, p0c, mc2)
    elif kind == "sbend_faces":
        e = _elem(type=_lib.DFCSR_ELE_SBEND, L=0.3, g=-0.08 / 0.3, e1=0.02, e2=-0.08, fint=0.6,
                  hgap=0.017, fintx=0.5, hgapx=0.02)
        ref = O.track_sbend(X, 0.3, -0.08 / 0.3, 0.02, -0.08, p0c, mc2, fint=0.6, hgap=0.017,
                            fintx=0.5, hgapx=0.02)
    elif kind == "quad":
        e = _elem(type=_lib.DFCSR_ELE_QUAD, L=0.3, k1=-6.0, num_steps=3)
        ref = O.track_quadrupole(X, 0.3, -6.0, p0c, mc2, num_steps=3)
    elif kind == "quad_offset":
        e = _elem(type=_lib.DFCSR_ELE_QUAD, L=0.3, k1=4.0, x_offset=1e-4, y_offset=-2e-4,
                  tilt=0.3)
        ref = O.track_quadrupole(X, 0.3, 4.0, p0c, mc2, x_off=1e-4, y_off=-2e-4, tilt=0.3)
    else:
        e = _elem(type=_lib.DFCSR_ELE_SEXTUPOLE, L=0.2, k2=30.0, num_steps=2)
        ref = O.track_sextupole(X, 0.2, 30.0, p0c, mc2, num_steps=2)
    c = _t(X)
    K.track_element(c, e, p0c, mc2)
    r = c.cpu().numpy()
    scale = np.abs(ref).max(axis=1, keepdims=True)
    err = (np.abs(r - ref) / scale).max()
    assert err <= 1e-12, (kind, err)


# -- a7 sgolay2d (SGolay_filter.py:3-81) -------------------------------------
# Tolerance: |d| <= 1e-12 * max|z| (the device sums the window in a fixed
# order; the reference and the oracle convolve by FFT).
def _sg_run(K, z, w, order, mode):
    print(f"sgolay2d {z.shape} window {w} order {order} {mode}", flush=True)
    r = K.sgolay2d(_t(z), w, order, mode)
    torch.cuda.synchronize()
    return tuple(x.cpu().numpy() for x in r) if mode == "both" else r.cpu().numpy()


def test_sgolay2d_golden(K, golden_dir):
    d = _g(golden_dir, "ref_sgolay2d.npz")
    for i, (nx, nz, w, order) in enumerate(d["cases"]):
        z = d[f"z{i}"]
        tol = 1e-12 * np.abs(z).max()
        for mode in (None, "col", "row"):
            got = _sg_run(K, z, int(w), int(order), mode)
            assert np.abs(got - d[f"{mode or 'none'}{i}"]).max() <= tol, (i, mode)
        col, row = _sg_run(K, z, int(w), int(order), "both")
        assert np.abs(col - d[f"bothcol{i}"]).max() <= tol, (i, "both")
        assert np.abs(row - d[f"bothrow{i}"]).max() <= tol, (i, "both")


@pytest.mark.parametrize("shape,w,order", [((100, 37), 3, 1), ((257, 130), 5, 2),
                                           ((16, 16), 7, 3), ((17, 15), 13, 4),
                                           ((7, 40), 13, 2), ((33, 64), 11, 4)])
def test_sgolay2d_vs_oracle(K, shape, w, order):
    """Ragged tiles (sizes off multiples of 16), the largest window, grids
    at the smallest size the padding allows, an offset and scale far from
    1."""
    rng = np.random.default_rng(w * 100 + order)
    nx, nz = shape
    xg, zg = np.meshgrid(np.linspace(-2, 2, nx), np.linspace(-1, 1, nz), indexing="ij")
    z = 1e6 * (3.0 + np.sin(2 * xg) * np.cos(3 * zg) + 0.1 * rng.normal(size=shape))
    tol = 1e-12 * np.abs(z).max()
    for mode in (None, "col", "row", "both"):
        got = _sg_run(K, z, w, order, mode)
        ref = O.sgolay2d(z, w, order, mode)
        if mode == "both":
            assert max(np.abs(got[0] - ref[0]).max(), np.abs(got[1] - ref[1]).max()) <= tol
        else:
            assert np.abs(got - ref).max() <= tol, mode


def test_sgolay2d_notebook_ramp(K):
    """test/test.ipynb cell 23: 'row' of the 3 x 3 ramp is 1 everywhere."""
    got = _sg_run(K, np.arange(1.0, 10.0).reshape(3, 3), 3, 1, "row")
    assert np.allclose(got, 1.0, rtol=0, atol=1e-14)
