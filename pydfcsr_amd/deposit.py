"""Device-resident density-function history (DF_tracker, deposit.py:90-426).

Per step: get_DF deposits and post-processes the beam into the five fields
on the deposition grid; append_DF logs them; append_interpolant drops slices
older than the retarded-time window and re-grids either the new slice or,
when the beam size left the [1/2, 2] band, every logged slice onto a new
interp grid; build_interpolant fixes the axes the wake kernel reads.

Storage is rings with one slot per logged slice, on the device for the
whole run:
  raw    [cap][5][nbmax*nbmax] float64  fields on the slice's own grid
  hist   [Xi][Zi][cap][20]     float64  re-gridded corner-packed records
  hist32t[Xi][Zi][cap][20]     float32  the same records (mixed precision)
  hist32 [Xi][cap][5][Zi][4]   float32  the same, chunked (optional)
  cells  [cap][Xi][Zi+1][5]    float64  cell-major: the five fields at (x, z), and
                                        per x row a copy of its last cell
  cells32[cap][Xi][Zi+1][5]    float32  the same in float32
  dcells32[cap][Xi][Zi+1][5]   float32  cells32 of the next slice minus this one's (0
                                        for the newest): the mixed wake's time lerp
  zspan  [Xi][cap]             int32    per x row and slot the z span of the cells that
                                        can contribute (rows x, x+1; slices i, i+1):
                                        the wakes skip lookups outside it (integrand 0)
(record (x, z) holds every field at (x, z), (x, z+1), (x+1, z), (x+1, z+1),
clamped to the last plane: one record per time plane per lookup.  hist and
hist32t keep the slot innermost (the two planes of a lookup are neighbours,
one lookup reads one contiguous span: what the far region's lookups want);
the cell-major rings keep a plane's cells together, neighbouring z in
neighbouring cells: what the near regions' lookups want.  hist32 stores each field's four corners as one
16-byte chunk, z fastest; it serves the near regions when the cell-major
rings are not kept.  keep_fp64 chooses hist (and cells), keep_fp32 hist32t
(and cells32, dcells32), keep_cells the cell-major rings, keep_chunked hist32.)
Appending writes slot (head + nt) % cap; popping advances head.  Nothing is
copied per step: the reference's np.array(deque) stacking of the whole
history (deposit.py:422-426) becomes a descriptor update.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, kernels
from ._lib import DfcsrError
from .params import Deposition_params


@dataclass
class SliceMeta:
    t: float
    sigma_x: float
    sigma_z: float
    xmean: float
    zmean: float
    nbx: int
    nbz: int
    x0: float
    x1: float
    z0: float
    z1: float
    frac: float
    window: int


class DF_tracker:
    def __init__(self, input_dic=None, device="cuda", keep_fp32=True, capacity=16,
                 keep_fp64=True, keep_cells=True, keep_chunked=None):
        """keep_fp64: hold the float64 history rings (FP64 wakes, lookups);
        keep_fp32: hold the float32 rings (mixed-precision wakes);
        keep_cells: with them the cell-major rings the wakes' near regions
        read; keep_chunked: the chunked float32 ring (default: when the
        cell-major rings are not kept)."""
        self.params = Deposition_params(input_dic)
        p = self.params
        self.xbins, self.zbins, self.xlim, self.zlim = p.xbins, p.zbins, p.xlim, p.zlim
        self.filter_order, self.filter_window = p.filter_order, p.filter_window
        self.velocity_threhold, self.upper_limit = p.velocity_threhold, p.upper_limit
        self.device = torch.device(device)
        self.keep_fp32 = keep_fp32
        self.keep_fp64 = keep_fp64
        self.keep_cells = bool(keep_cells)
        self.keep_chunked = (not self.keep_cells) if keep_chunked is None else bool(keep_chunked)
        if not (keep_fp32 or keep_fp64):
            raise DfcsrError("DF_tracker needs the float64 or the float32 history rings")
        self.ws = _lib.Workspace(self.device)
        self.nbmax = max(int(self.xbins), int(self.zbins), 100)
        self.cap = 0
        self.head = 0
        self.log: list[SliceMeta] = []   # logical order, oldest first
        self.start_time = 0.0
        self.end_time = 0.0
        self.sigma_x_interp = None
        self.sigma_z_interp = None
        self.Xi = self.Zi = 0
        self.xi = self.zi = None          # (start, stop) of the interp grid
        self.raw = None
        self.hist = None      # float64 (Xi, Zi, cap, 20)
        self.hist32 = None    # float32 chunked (Xi, cap, 5, Zi, 4)
        self.hist32t = None   # float32 time-innermost (Xi, Zi, cap, 20)
        self.cells = None     # float64 cell-major (cap, Xi, Zi + 1, 5)
        self.cells32 = None   # float32 cell-major (cap, Xi, Zi + 1, 5)
        self.dcells32 = None  # its time differences (cap, Xi, Zi + 1, 5)
        self.zspan = None     # z spans of contributing cells (Xi, cap), uint32 as int32
        self.cur = None
        self.n_regrid = 0
        # largest interp grid side the rings may take (None: unbounded, as the
        # reference); a re-grid past it raises (serving admits by it)
        self.max_grid = None
        self.info = torch.zeros(4, dtype=torch.float64, device=self.device)
        self._ensure_capacity(capacity)

    # -- storage -------------------------------------------------------------------
    @property
    def nt(self):
        return len(self.log)

    def _slot(self, i):
        return (self.head + i) % self.cap

    def _raw_stride(self):
        return 5 * self.nbmax * self.nbmax

    def _ensure_capacity(self, need):
        if need <= self.cap:
            return
        new_cap = max(need, 2 * self.cap, 16)
        raw = torch.empty((new_cap, self._raw_stride()), dtype=torch.float64, device=self.device)
        rings = (None,) * 7
        if self.Xi:
            rings = self._rings(self.Xi, self.Zi, new_cap)
        hist, hist32, hist32t, cells, cells32, dcells32, zspan = rings
        if self.cap:
            # keep the ring's logical order, pending slot included (index nt)
            for i in range(min(self.nt + 1, self.cap)):
                s = self._slot(i)
                raw[i].copy_(self.raw[s])
                if hist is not None and self.hist is not None:
                    hist[:, :, i].copy_(self.hist[:, :, s])
                if hist32t is not None and self.hist32t is not None:
                    hist32t[:, :, i].copy_(self.hist32t[:, :, s])
                if hist32 is not None and self.hist32 is not None:
                    hist32[:, i].copy_(self.hist32[:, s])
                for new, old in ((cells, self.cells), (cells32, self.cells32),
                                 (dcells32, self.dcells32)):
                    if new is not None and old is not None:
                        new[i].copy_(old[s])
                if zspan is not None and self.zspan is not None:
                    zspan[:, i].copy_(self.zspan[:, s])
        self.raw, self.hist, self.hist32, self.hist32t = raw, hist, hist32, hist32t
        self.cells, self.cells32, self.dcells32 = cells, cells32, dcells32
        self.zspan = zspan
        self.cap = new_cap
        self.head = 0

    def _rings(self, Xi, Zi, cap):
        """(hist, hist32, hist32t, cells, cells32, dcells32, zspan) for an Xi x
        Zi grid.  The wakes address cells within a plane with 32-bit offsets: a
        cell-major ring whose plane would pass 4 GiB is not kept (the rings
        serve).  The span table comes with a cell-major ring (it is derived
        from one) when z indices fit 16 bits."""
        hist = hist32 = hist32t = cells = cells32 = dcells32 = zspan = None
        dev = self.device
        if self.keep_fp64:
            hist = torch.empty((Xi, Zi, cap, _lib.HIST_REC64), dtype=torch.float64, device=dev)
            if self.keep_cells and Xi * (Zi + 1) * _lib.CELL * 8 <= (1 << 32):
                cells = torch.empty((cap, Xi, Zi + 1, _lib.CELL), dtype=torch.float64,
                                    device=dev)
        if self.keep_fp32:
            hist32t = torch.empty((Xi, Zi, cap, _lib.HIST_REC32), dtype=torch.float32, device=dev)
            if self.keep_cells and Xi * (Zi + 1) * _lib.CELL * 4 <= (1 << 32):
                cells32 = torch.empty((cap, Xi, Zi + 1, _lib.CELL), dtype=torch.float32,
                                      device=dev)
                dcells32 = torch.empty_like(cells32)
            if self.keep_chunked:
                hist32 = torch.empty((Xi, cap, _lib.HIST_CHUNKS32, Zi, 4), dtype=torch.float32,
                                     device=dev)
        # spans from float32 cells hold for mixed wakes only (a float64 value
        # may round to a float32 zero): from cells32 only without float64 rings
        if (cells is not None or (cells32 is not None and not self.keep_fp64)) and Zi <= 0xffff:
            zspan = torch.empty((Xi, cap), dtype=torch.int32, device=dev)
        return hist, hist32, hist32t, cells, cells32, dcells32, zspan

    def _alloc_hist(self, Xi, Zi):
        if self.Xi != Xi or self.Zi != Zi or (self.hist is None and self.hist32t is None):
            # free first
            self.hist = self.hist32 = self.hist32t = None
            self.cells = self.cells32 = self.dcells32 = self.zspan = None
            (self.hist, self.hist32, self.hist32t, self.cells, self.cells32,
             self.dcells32, self.zspan) = self._rings(Xi, Zi, self.cap)
        self.Xi, self.Zi = Xi, Zi

    def raw_fields(self, i=None):
        """Five fields of logical slice i (default: the current one) as
        (nbx, nbz) views, in the order density, density_x, density_z, vx, vx_x."""
        if i is None:
            slot, m = self._slot(self.nt), self.cur
        else:
            slot, m = self._slot(i), self.log[i]
        n = m.nbx * m.nbz
        base = self.raw[slot]
        return [base[f * n:(f + 1) * n].view(m.nbx, m.nbz) for f in range(5)]

    # -- deposit.py:145-245 ----------------------------------------------------------
    def get_DF(self, x, z, px, t, moments=None):
        """Deposits and post-processes the beam into the pending slot.
        ``moments`` (from Beam.update_status) saves one reduction."""
        m = moments if moments is not None else kernels.beam_moments(x, z, self.ws)
        sigma_x, sigma_z = m["sigma_x"], m["sigma_z"]
        xmean, zmean = m["mean_x"], m["mean_z"]
        with np.errstate(invalid="ignore", divide="ignore"):
            frac = np.float64(sigma_x) / np.float64(m["slice_sigma_x"])
        if frac > 5:
            nbx, nbz, window = int(self.xbins), int(self.zbins), int(self.filter_window)
        else:
            nbx, nbz, window = 100, 100, 5
        if max(nbx, nbz) > self.nbmax:
            raise DfcsrError(f"deposition grid {nbx}x{nbz} above the allocated {self.nbmax}")
        x0 = xmean - self.xlim * sigma_x
        x1 = xmean + self.xlim * sigma_x
        z0 = zmean - self.zlim * sigma_z
        z1 = zmean + self.zlim * sigma_z
        self._ensure_capacity(self.nt + 1)
        self.cur = SliceMeta(float(t), sigma_x, sigma_z, xmean, zmean, nbx, nbz, x0, x1, z0, z1,
                             float(frac), window)
        d, dx, dz, vx, vxx = self.raw_fields()
        kernels.deposit_density_vx(x, z, px, nbx, x0, x1, nbz, z0, z1, self.ws, density=d, vx=vx)
        kernels.df_process(d, vx, x0, x1, z0, z1, self.velocity_threhold, window,
                           int(self.filter_order), self.ws, out=[dx, dz, vxx], info=self.info)
        return self.cur

    def current_fields(self):
        d, dx, dz, vx, vxx = self.raw_fields()
        c = self.cur
        return dict(density=d, density_x=dx, density_z=dz, vx=vx, vx_x=vxx,
                    x_grids=np.linspace(c.x0, c.x1, c.nbx), z_grids=np.linspace(c.z0, c.z1, c.nbz))

    def append_DF(self):   # deposit.py:247-256
        self.log.append(self.cur)
        self.end_time = self.cur.t

    def pop_left_DF(self, new_start_time):   # deposit.py:258-280
        # The reference pops its interp deque (which does not yet hold the new
        # slice) with the same rule and fails with IndexError when that empties
        # it, i.e. when even the newest earlier slice is older than the window.
        if self.nt >= 2 and self.start_time < new_start_time and self.log[-2].t < new_start_time:
            raise DfcsrError("history window n_formation_length*formation_length is shorter "
                             "than one step: every earlier slice would be dropped")
        while self.start_time < new_start_time:
            self.log.pop(0)
            self.head = (self.head + 1) % self.cap
            self.start_time = self.log[0].t

    def append_interpolant(self, formation_length, n_formation_length):   # deposit.py:312-392
        start_point = max(0.0, self.end_time - n_formation_length * formation_length)
        self.pop_left_DF(new_start_time=start_point)
        c = self.cur
        if (self.sigma_x_interp and self.sigma_z_interp
                and 2 > c.sigma_x / self.sigma_x_interp > 1 / 2
                and 2 > c.sigma_z / self.sigma_z_interp > 1 / 2):
            self._regrid([self.nt - 1])
            return False
        sx = [m.sigma_x for m in self.log]
        sz = [m.sigma_z for m in self.log]
        max_sx, min_sx, max_sz, min_sz = max(sx), min(sx), max(sz), min(sz)
        xbins = int(500 * (max_sx / min_sx))
        zbins = int(500 * (max_sz / min_sz))
        if isinstance(self.upper_limit, int):
            xbins = min(xbins, self.upper_limit)
            zbins = min(zbins, self.upper_limit)
        if self.max_grid is not None and max(xbins, zbins) > self.max_grid:
            raise DfcsrError(f"re-grid to {xbins} x {zbins} exceeds the {self.max_grid} per side "
                             "this history was admitted for")
        self.sigma_x_interp = max_sx
        self.sigma_z_interp = max_sz
        self.xi = (c.xmean - 5 * max_sx, c.xmean + 5 * max_sx)
        self.zi = (c.zmean - 5 * max_sz, c.zmean + 5 * max_sz)
        self._alloc_hist(xbins, zbins)
        self._regrid(list(range(self.nt)))
        self.n_regrid += 1
        return True

    def _regrid(self, idx):
        jobs = []
        for i in idx:
            m = self.log[i]
            s = self._slot(i)
            jobs.append((s, s, m.nbx, m.nbz, m.x0, m.x1, m.z0, m.z1))
        for k in range(0, len(jobs), 4096):
            kernels.regrid(self.raw, self._raw_stride(), jobs[k:k + 4096], self.hist, self.hist32,
                           self.Xi, self.Zi, self.xi[0], self.xi[1], self.zi[0], self.zi[1],
                           self.ws, hist32t=self.hist32t, cells=self.cells,
                           cells32=self.cells32)
        # the re-gridded slices and each one's predecessor change differences
        # and spans (a span covers the slice and the next)
        first = max(0, min(idx) - 1) if idx else self.nt
        if self.dcells32 is not None:
            kernels.cells32_tdiff(self.cells32, self.dcells32, self.head, self.nt, first=first)
        if self.zspan is not None:
            kernels.history_zspan(self.cells, self.cells32, self.zspan, self.head, self.nt,
                                  first=first)

    def build_interpolant(self):   # deposit.py:395-426
        nt = self.nt
        self.min_x, self.max_x = self.log[0].t, self.log[-1].t
        self.min_y, self.max_y = self.xi
        self.min_z, self.max_z = self.zi
        with np.errstate(invalid="ignore", divide="ignore"):
            self.delta_x = (np.float64(self.max_x) - self.min_x) / np.float64(nt - 1)
        self.delta_y = (self.max_y - self.min_y) / (self.Xi - 1)
        self.delta_z = (self.max_z - self.min_z) / (self.Zi - 1)

    def descriptor(self):
        """ctypes history descriptor for the kernels (after build_interpolant)."""
        def p(t):
            return t.data_ptr() if t is not None else None
        return _lib.History(p(self.hist), p(self.hist32), p(self.hist32t), p(self.cells),
                            p(self.cells32), self.cap, self.head, self.nt, self.Xi, self.Zi,
                            float(self.min_x), float(self.min_y), float(self.min_z),
                            float(self.delta_x), float(self.delta_y), float(self.delta_z),
                            p(self.dcells32), p(self.zspan))

    def history_fields(self):
        """Stacked (nt, Xi, Zi) copies of the five fields in logical order
        (the reference's data_*_interp arrays; for inspection and tests)."""
        idx = torch.tensor([self._slot(i) for i in range(self.nt)], device=self.device)
        ring = self.hist if self.hist is not None else self.hist32t
        h = ring[..., 0::4].index_select(2, idx).double()   # corner (x, z) of each field
        return {f: h[..., k].permute(2, 0, 1).contiguous() for k, f in enumerate(_lib.FIELDS)}

    def footprint_bytes(self):
        b = self.raw.numel() * 8 if self.raw is not None else 0
        if self.hist is not None:
            b += self.hist.numel() * 8
        for r in (self.hist32, self.hist32t, self.cells, self.cells32, self.dcells32,
                  self.zspan):
            if r is not None:
                b += r.numel() * r.element_size()
        return b
