"""Typed wrappers over the C ABI, one per reference operator on the hot path.

Each takes/returns torch float64 tensors on the HIP device and launches on
torch's current stream.  Reference functions they replace (paths relative to
pyDFCSR_2D/):

  beam_moments        beams.py:88-98,137-215; CSR.py:368-374; deposit.py:147-159
  histogram_cic_2d    deposit.py:42-87
  deposit_density_vx  deposit.py:172-182
  df_process          deposit.py:183-235
  regrid              deposit.py:296-309, 328-337, 379-390
  interpolate3D       interp3D.py:18-66
  interpolate1D       interp1D.py:13-36
  csr_integrand       CSR.py:605-782
  wake_mesh           CSR.py:361-602
  apply_wakes         beams.py:108-131
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import DfcsrError, check, ptr, require_device_tensor, stream_handle

MOMENT_KEYS = ("mean_x", "mean_z", "sigma_x", "sigma_z", "slope0", "slope1", "mean_xt",
                "sigma_xt", "slice_sigma_x", "n_slice")


def _f64(t, name):
    return require_device_tensor(t, torch.float64, name)


def beam_moments(x, z, ws: _lib.Workspace, out: torch.Tensor | None = None, host=True):
    lib = _lib.load()
    _f64(x, "x")
    _f64(z, "z")
    n = x.numel()
    if z.numel() != n:
        raise DfcsrError("beam_moments: x and z differ in length")
    if out is None:
        out = torch.empty(_lib.MOMENTS_COUNT, dtype=torch.float64, device=x.device)
    b = ws.get("moments", lib.dfcsr_moments_workspace_bytes(n))
    check(lib.dfcsr_beam_moments(ptr(x), ptr(z), n, ptr(b), b.numel(), ptr(out),
                                 stream_handle(x.device)), "beam_moments")
    if not host:
        return out
    return moments_dict(out)


def moments_dict(out: torch.Tensor):
    """The moments vector as a dict of floats, plus "slope" = [slope0, slope1]."""
    v = out.cpu().numpy()
    d = dict(zip(MOMENT_KEYS, (float(a) for a in v)))
    d["slope"] = np.array([d["slope0"], d["slope1"]])
    return d


def beam_statistics(coords: torch.Tensor, p0c, mc2, ws: _lib.Workspace,
                    out: torch.Tensor | None = None):
    """Device statistics of the (6, N) coordinates (x, px, y, py, z, pz):
    means, np.cov triangle, Twiss x / y and energy statistics, written into
    ``out`` (a device float64 view of DFCSR_STATS_COUNT values) without a
    host synchronisation."""
    lib = _lib.load()
    _f64(coords, "coords")
    if coords.dim() != 2 or coords.shape[0] != 6:
        raise DfcsrError(f"beam_statistics: coords must be (6, N), got {tuple(coords.shape)}")
    n = coords.shape[1]
    if out is None:
        out = torch.empty(_lib.STATS_COUNT, dtype=torch.float64, device=coords.device)
    if (out.dtype != torch.float64 or out.numel() != _lib.STATS_COUNT or not out.is_contiguous()
            or out.device != coords.device):
        raise DfcsrError("beam_statistics: out must be a contiguous float64 device vector of "
                         f"{_lib.STATS_COUNT}")
    b = ws.get("statistics", lib.dfcsr_beam_statistics_workspace_bytes(n))
    rows = [coords[k] for k in range(6)]
    check(lib.dfcsr_beam_statistics(*[ptr(r) for r in rows], n, float(p0c), float(mc2), ptr(b),
                                    b.numel(), ptr(out), stream_handle(coords.device)),
          "beam_statistics")
    return out


def statistics_dict(v) -> dict:
    """Host dict of one statistics vector: twiss (the reference's keys
    alpha_x ... norm_emit_y), mean_energy, sigma_energy, mean, cov (6 x 6)."""
    v = np.asarray(v.cpu() if isinstance(v, torch.Tensor) else v, dtype=np.float64)
    tw = {}
    for plane, off in (("x", _lib.STAT_TWISS_X), ("y", _lib.STAT_TWISS_Y)):
        for k, key in enumerate(_lib.TWISS_KEYS):
            tw[f"{key}_{plane}"] = float(v[off + k])
    cov = np.zeros((6, 6))
    t = _lib.STAT_COV
    for i in range(6):
        for j in range(i, 6):
            cov[i, j] = cov[j, i] = v[t]
            t += 1
    return dict(twiss=tw, mean_energy=float(v[_lib.STAT_MEAN_ENERGY]),
                sigma_energy=float(v[_lib.STAT_SIGMA_ENERGY]),
                mean=v[_lib.STAT_MEAN:_lib.STAT_MEAN + 6].copy(), cov=cov)


def track_element(coords: torch.Tensor, ele: _lib.Element, p0c, mc2):
    """Tracks the (6, N) device coordinates through one element piece, in
    place (dfcsr_track_element)."""
    lib = _lib.load()
    _f64(coords, "coords")
    if coords.dim() != 2 or coords.shape[0] != 6:
        raise DfcsrError(f"track_element: coords must be (6, N), got {tuple(coords.shape)}")
    rows = [ptr(coords[k]) for k in range(6)]
    check(lib.dfcsr_track_element(*rows, coords.shape[1], ctypes.byref(ele), float(p0c),
                                  float(mc2), stream_handle(coords.device)), "track_element")
    return coords


def histogram_cic_2d(q1, q2, w, nbins_1, bins_start_1, bins_end_1, nbins_2, bins_start_2,
                     bins_end_2, ws: _lib.Workspace, want_cells=False):
    """Device histogram_cic_2d; w=None means unit weights.  With want_cells,
    also returns per-particle lower-cell indices (n, 2) and per-cell counts."""
    lib = _lib.load()
    _f64(q1, "q1")
    _f64(q2, "q2")
    if w is not None:
        _f64(w, "w")
    n = q1.numel()
    dev = q1.device
    hist = torch.empty((nbins_1, nbins_2), dtype=torch.float64, device=dev)
    cells = torch.empty((n, 2), dtype=torch.int32, device=dev) if want_cells else None
    counts = torch.empty((nbins_1, nbins_2), dtype=torch.int32, device=dev) if want_cells else None
    b = ws.get("hist", lib.dfcsr_histogram_workspace_bytes(nbins_1, nbins_2))
    check(lib.dfcsr_histogram_cic_2d(ptr(q1), ptr(q2), ptr(w), n, nbins_1, bins_start_1,
                                     bins_end_1, nbins_2, bins_start_2, bins_end_2, ptr(hist),
                                     ptr(cells), ptr(counts), ptr(b), b.numel(),
                                     stream_handle(dev)), "histogram_cic_2d")
    if want_cells:
        return hist, cells, counts
    return hist


def deposit_density_vx(x, z, px, nbx, x0, x1, nbz, z0, z1, ws: _lib.Workspace, density=None,
                       vx=None, cells=None, counts=None):
    lib = _lib.load()
    for t, nm in ((x, "x"), (z, "z"), (px, "px")):
        _f64(t, nm)
    n = x.numel()
    dev = x.device
    if density is None:
        density = torch.empty((nbx, nbz), dtype=torch.float64, device=dev)
    if vx is None:
        vx = torch.empty((nbx, nbz), dtype=torch.float64, device=dev)
    b = ws.get("hist", lib.dfcsr_histogram_workspace_bytes(nbx, nbz))
    check(lib.dfcsr_deposit_density_vx(ptr(x), ptr(z), ptr(px), n, nbx, x0, x1, nbz, z0, z1,
                                       ptr(density), ptr(vx), ptr(cells), ptr(counts), ptr(b),
                                       b.numel(), stream_handle(dev)), "deposit_density_vx")
    return density, vx


def df_process(density, vx, x0, x1, z0, z1, velocity_threshold, window, order,
               ws: _lib.Workspace, out=None, info=None):
    """In place on density/vx; returns (density, vx, density_x, density_z, vx_x)."""
    lib = _lib.load()
    nbx, nbz = density.shape
    dev = density.device
    if out is None:
        out = [torch.empty((nbx, nbz), dtype=torch.float64, device=dev) for _ in range(3)]
    b = ws.get("df", lib.dfcsr_df_workspace_bytes(nbx, nbz))
    check(lib.dfcsr_df_process(ptr(density), ptr(vx), ptr(out[0]), ptr(out[1]), ptr(out[2]), nbx,
                               x0, x1, nbz, z0, z1, float(velocity_threshold), int(window),
                               int(order), ptr(info), ptr(b), b.numel(), stream_handle(dev)),
          "df_process")
    return density, vx, out[0], out[1], out[2]


def savgol_operator(window, order):
    """Host-side Savitzky-Golay operator: (interior stencil, left-edge rows)."""
    lib = _lib.load()
    c = np.zeros(window)
    e = np.zeros((max(window // 2, 1), window))
    check(lib.dfcsr_savgol_operator(window, order, c.ctypes.data, e.ctypes.data), "savgol_operator")
    return c, e[: window // 2]


def sgolay2d(z: torch.Tensor, window_size: int, order: int, derivative=None):
    """SGolay_filter.sgolay2d (SGolay_filter.py:3-81) on the device: z a
    float64 (nx, nz) tensor; derivative None (smoothed grid), 'col' (slope
    along the first axis, per index step), 'row' (along the second) or
    'both' (the tuple (col, row)), as the reference returns them."""
    lib = _lib.load()
    if derivative not in _lib.SG2D_DERIVATIVE:
        raise ValueError(f"derivative must be one of None, 'col', 'row', 'both', not {derivative!r}")
    if not isinstance(z, torch.Tensor) or z.dim() != 2:
        raise DfcsrError("sgolay2d: z must be a 2-D tensor")
    z = require_device_tensor(z.contiguous(), torch.float64, "z")
    nx, nz = z.shape
    out = torch.empty_like(z)
    out2 = torch.empty_like(z) if derivative == "both" else None
    check(lib.dfcsr_sgolay2d(ptr(z), nx, nz, int(window_size), int(order),
                             _lib.SG2D_DERIVATIVE[derivative], ptr(out), ptr(out2),
                             stream_handle(z.device)), "sgolay2d")
    return (out, out2) if derivative == "both" else out


def sgolay2d_kernels(window_size: int, order: int, derivative=None):
    """Host-side: the (window, window) kernels sgolay2d convolves with (one,
    or two for 'both')."""
    lib = _lib.load()
    w = int(window_size)
    k0 = np.zeros((max(w, 1), max(w, 1)))
    k1 = np.zeros_like(k0)
    check(lib.dfcsr_sgolay2d_kernels(w, int(order), _lib.SG2D_DERIVATIVE[derivative],
                                     k0.ctypes.data, k1.ctypes.data), "sgolay2d_kernels")
    return (k0, k1) if derivative == "both" else k0


def regrid(src: torch.Tensor, slot_stride: int, jobs, hist64, hist32, Xi, Zi, xi0, xi1, zi0, zi1,
           ws: _lib.Workspace, hist32t=None, cells=None, cells32=None):
    """jobs: list of (src_slot, dst_slot, nbx, nbz, x0, x1, z0, z1).  Rings (any
    may be None, not all): hist64 float64 (Xi, Zi, cap, 20), hist32 float32
    chunked (Xi, cap, 5, Zi, 4), hist32t float32 time-innermost (Xi, Zi, cap, 20),
    cells / cells32 cell-major float64 / float32 (cap, Xi, Zi + 1, 5): each x
    row's Zi cells and a copy of its last."""
    lib = _lib.load()
    n = len(jobs)
    if n == 0:
        return
    arr = (_lib.RegridJob * n)(*[_lib.RegridJob(*j) for j in jobs])
    b = ws.get("regrid", lib.dfcsr_regrid_workspace_bytes(n))
    rings = [r for r in (hist64, hist32t) if r is not None]
    if rings:
        cap = int(rings[0].shape[2])
    elif hist32 is not None:
        cap = int(hist32.shape[1])
    elif cells is not None or cells32 is not None:
        cap = int((cells if cells is not None else cells32).shape[0])
    else:
        raise DfcsrError("regrid: no history ring")
    want = {id(hist64): ((Xi, Zi, cap, _lib.HIST_REC64), torch.float64),
            id(hist32t): ((Xi, Zi, cap, _lib.HIST_REC32), torch.float32),
            id(hist32): ((Xi, cap, _lib.HIST_CHUNKS32, Zi, 4), torch.float32),
            id(cells): ((cap, Xi, Zi + 1, _lib.CELL), torch.float64),
            id(cells32): ((cap, Xi, Zi + 1, _lib.CELL), torch.float32)}
    for r in (hist64, hist32, hist32t, cells, cells32):
        if r is None:
            continue
        shape, dtype = want[id(r)]
        if tuple(r.shape) != shape or r.dtype != dtype or not r.is_contiguous():
            raise DfcsrError(f"regrid: history ring {tuple(r.shape)} {r.dtype} does not match "
                             f"{shape} {dtype}")
    check(lib.dfcsr_regrid(ptr(src), int(slot_stride), arr, n, ptr(hist64), ptr(hist32),
                           ptr(hist32t), ptr(cells), ptr(cells32), int(Xi), int(Zi), cap,
                           float(xi0), float(xi1), float(zi0), float(zi1), ptr(b), b.numel(),
                           stream_handle(src.device)), "regrid")


def cells32_tdiff(cells32: torch.Tensor, dcells32: torch.Tensor, head: int, nt: int,
                  first: int = 0):
    """Fills dcells32 (cap, Xi, Zi + 1, 5) float32 with the time differences of
    the float32 cell-major ring cells32 for logical slices first .. nt-1 (slice i
    in slot (head + i) % cap): slice i + 1 minus slice i, 0 for the newest."""
    lib = _lib.load()
    shape = tuple(cells32.shape)
    if (len(shape) != 4 or shape[3] != _lib.CELL or cells32.dtype != torch.float32
            or tuple(dcells32.shape) != shape or dcells32.dtype != torch.float32
            or not cells32.is_contiguous() or not dcells32.is_contiguous()):
        raise DfcsrError(f"cells32_tdiff: rings {shape} {cells32.dtype} / "
                         f"{tuple(dcells32.shape)} {dcells32.dtype} are not two contiguous "
                         f"float32 (cap, Xi, Zi + 1, {_lib.CELL}) rings")
    cap, Xi, Zi = shape[0], shape[1], shape[2] - 1
    check(lib.dfcsr_cells32_tdiff(ptr(cells32), ptr(dcells32), int(Xi), int(Zi), int(cap),
                                  int(head), int(nt), int(first), stream_handle(cells32.device)),
          "cells32_tdiff")


def history_zspan(cells, cells32, zspan: torch.Tensor, head: int, nt: int, first: int = 0):
    """Fills zspan (Xi, cap) int32 -- read as uint32 zlo | zhi << 16 -- for the
    logical slices first .. nt-1 from the float64 cell-major ring cells (cap,
    Xi, Zi + 1, 5), or when that is None from the float32 one cells32: the z span
    of the cells of rows x, x + 1 in slices i, i + 1 whose density or
    gradients are nonzero or whose velocities are not finite."""
    lib = _lib.load()
    ring = cells if cells is not None else cells32
    if ring is None:
        raise DfcsrError("history_zspan: no cell-major ring")
    want = torch.float64 if cells is not None else torch.float32
    if (ring.dim() != 4 or ring.shape[3] != _lib.CELL or ring.dtype != want
            or not ring.is_contiguous()):
        raise DfcsrError(f"history_zspan: ring {tuple(ring.shape)} {ring.dtype} is not a "
                         f"contiguous (cap, Xi, Zi + 1, {_lib.CELL}) {want} ring")
    cap, Xi, Zi = ring.shape[0], ring.shape[1], ring.shape[2] - 1
    if (tuple(zspan.shape) != (Xi, cap) or zspan.dtype != torch.int32
            or not zspan.is_contiguous()):
        raise DfcsrError(f"history_zspan: table {tuple(zspan.shape)} {zspan.dtype} is not a "
                         f"contiguous ({Xi}, {cap}) int32 table")
    check(lib.dfcsr_history_zspan(ptr(cells), ptr(cells32) if cells is None else None,
                                  ptr(zspan), int(Xi), int(Zi), int(cap), int(head), int(nt),
                                  int(first), stream_handle(ring.device)), "history_zspan")


def interpolate3D(hist: _lib.History, field: int, q: torch.Tensor, device):
    lib = _lib.load()
    out = torch.empty(q.shape[0], dtype=torch.float64, device=device)
    check(lib.dfcsr_interpolate3d(ctypes.byref(hist), int(field), ptr(q), q.shape[0], ptr(out),
                                  stream_handle(device)), "interpolate3d")
    return out


def interpolate1D(lat: _lib.Lattice, col: int, s: torch.Tensor):
    lib = _lib.load()
    out = torch.empty(s.numel(), dtype=torch.float64, device=s.device)
    check(lib.dfcsr_interpolate1d(ctypes.byref(lat), int(col), ptr(s), s.numel(), ptr(out),
                                  stream_handle(s.device)), "interpolate1d")
    return out


def csr_integrand(hist: _lib.History, lat: _lib.Lattice, s, x, t, xp: torch.Tensor,
                  sp: torch.Tensor):
    """Integrand grids (Iz, Ix), shape (len(xp), len(sp)), for one observer."""
    lib = _lib.load()
    dev = xp.device
    Iz = torch.empty((xp.numel(), sp.numel()), dtype=torch.float64, device=dev)
    Ix = torch.empty_like(Iz)
    check(lib.dfcsr_csr_integrand(ctypes.byref(hist), ctypes.byref(lat), float(s), float(x),
                                  float(t), ptr(xp), xp.numel(), ptr(sp), sp.numel(), ptr(Iz),
                                  ptr(Ix), stream_handle(dev)), "csr_integrand")
    return Iz, Ix


def wake_mesh(hist: _lib.History, lat: _lib.Lattice, p: _lib.WakeParams, precision, device,
              ws: _lib.Workspace, out=None, want_mesh=True):
    lib = _lib.load()
    M = p.mesh_xbins * p.mesh_zbins
    if out is None:
        out = {k: torch.empty(M, dtype=torch.float64, device=device)
               for k in ("xmesh", "zmesh", "dE_dct", "x_kick")}
    b = ws.get("wake", lib.dfcsr_wake_workspace_bytes(ctypes.byref(p)))
    check(lib.dfcsr_wake_mesh(ctypes.byref(hist), ctypes.byref(lat), ctypes.byref(p),
                              int(precision), ptr(out["xmesh"]) if want_mesh else None,
                              ptr(out["zmesh"]) if want_mesh else None, ptr(out["dE_dct"]),
                              ptr(out["x_kick"]), ptr(b), b.numel(), stream_handle(device)),
          "wake_mesh")
    return out


def wake_points_evaluated(hist: _lib.History, lat: _lib.Lattice, p: _lib.WakeParams, precision,
                          device, ws: _lib.Workspace):
    """Integrand points a wake with these inputs evaluates (the ones whose
    retarded lookup lies in the history): one extra wake with the device
    counter of DfcsrWakeParams.points_evaluated."""
    cnt = torch.zeros(1, dtype=torch.int64, device=device)
    q = _lib.WakeParams.from_buffer_copy(p)
    q.points_evaluated = cnt.data_ptr()
    wake_mesh(hist, lat, q, precision, device, ws, want_mesh=False)
    return int(cnt.item())


def wake_mesh_batched(hists, lats, params, precision, outs, device, ws: _lib.Workspace):
    lib = _lib.load()
    k = len(hists)
    H = (_lib.History * k)(*hists)
    L = (_lib.Lattice * k)(*lats)
    P = (_lib.WakeParams * k)(*params)
    dE = (_lib.c_dp * k)(*[o["dE_dct"].data_ptr() for o in outs])
    xk = (_lib.c_dp * k)(*[o["x_kick"].data_ptr() for o in outs])
    b = ws.get("wake_batched", lib.dfcsr_wake_batched_workspace_bytes(k, P))
    check(lib.dfcsr_wake_mesh_batched(k, H, L, P, int(precision), dE, xk, ptr(b), b.numel(),
                                      stream_handle(device)), "wake_mesh_batched")


def apply_wakes(px, pz, x, z, dE_dct, x_kick, xr0, xr1, zr0, zr1, slope0, slope1, step_size,
                init_energy, transverse_on, xbins, zbins):
    lib = _lib.load()
    check(lib.dfcsr_apply_wakes(ptr(px), ptr(pz), ptr(x), ptr(z), x.numel(), ptr(dE_dct),
                                ptr(x_kick), int(xbins), int(zbins), float(xr0), float(xr1),
                                float(zr0), float(zr1), float(slope0), float(slope1),
                                float(step_size), float(init_energy), int(bool(transverse_on)),
                                stream_handle(x.device)), "apply_wakes")
