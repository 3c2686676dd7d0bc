"""Lattice: element list, reference-orbit tables and the step schedule.

Mirrors Lattice / get_referece_traj of the reference (lattice.py:4-204).
The orbit tables are built once on the host and uploaded to the device as
one interleaved (ns, 6) table [coords_x, coords_y, n_x, n_y, tau_x, tau_y]
(the layout the wake kernels read), plus per-element ``distance`` and
``rho``.  The sample recurrence below is the reference's, operation for
operation, so the tables are bit-identical to its arrays.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .yaml_parser import parse_yaml


def reference_tables(elements, nsample=2000):
    """(s, rho, distance, nsep, coords, n_vec, tau_vec) of lattice.py:4-110.

    elements: ordered list of element dicts (type, L, nsep, angle...)."""
    nel = len(elements)
    lengths = np.array([float(e["L"]) for e in elements])
    nsep = np.array([float(e["nsep"]) for e in elements])
    rho = np.zeros(nel)
    for k, e in enumerate(elements):
        if e["type"] == "dipole":
            rho[k] = e["angle"] / e["L"]
    distance = np.zeros(nel)
    acc = 0.0
    for k in range(nel):
        acc = lengths[k] if k == 0 else lengths[k] + acc
        distance[k] = acc
    s = np.linspace(0, distance[-1], nsample)
    coords = np.zeros((nsample, 2))
    tau = np.zeros((nsample, 2))
    nvec = np.zeros((nsample, 2))
    theta = 0
    tau[0] = (np.cos(theta), np.sin(theta))
    nvec[0] = (np.sin(theta), -1 * np.cos(theta))
    ele = 0
    for i in range(1, nsample):
        if s[i] > distance[ele]:
            ele += 1
        ds = s[i] - s[i - 1]
        e = elements[ele]
        if e["type"] == "dipole":
            L, ang = e["L"], e["angle"]
            phi = ds / L * ang
            r = L / ang
            # centre of curvature of this arc, then the point at angle theta+phi
            cx = coords[i - 1, 0] - r * np.sin(theta)
            cy = coords[i - 1, 1] + r * np.cos(theta)
            coords[i, 0] = cx + r * np.sin(phi + theta)
            coords[i, 1] = cy - r * np.cos(phi + theta)
        else:
            phi = 0
            coords[i, 0] = coords[i - 1, 0] + ds * np.cos(theta)
            coords[i, 1] = coords[i - 1, 1] + ds * np.sin(theta)
        theta += phi
        tau[i] = (np.cos(theta), np.sin(theta))
        nvec[i] = (np.sin(theta), -1 * np.cos(theta))
    return s, rho, distance, nsep, coords, nvec, tau


class Lattice:
    """Lattice from ``{'lattice_input_file': path}`` (lattice.py:112-204)."""

    def __init__(self, input_lattice, device=None, base_dir=None, nsample=2000):
        assert "lattice_input_file" in input_lattice, \
            "Error in parsing lattice: must include the keyword <lattice_input_file>"
        self.lattice_input_file = input_lattice["lattice_input_file"]
        cfg = input_lattice.get("_config")
        if cfg is None:
            cfg = parse_yaml(self.lattice_input_file, base_dir=base_dir)
        assert "step_size" in cfg, \
            "Required input parameter step_size to Lattice.__init__(**kwargs) was not found."
        self.lattice_config = cfg
        self.element_names = list(cfg.keys())[1:]
        self.elements = [dict(cfg[k]) for k in self.element_names]
        self._Nelement = len(self.elements)
        (self.s, self.rho, self.distance, self.nsep, self.coords, self.n_vec,
         self.tau_vec) = reference_tables(self.elements, nsample)
        self._lattice_length = self.distance[-1]
        self.get_steps()
        self.min_x, self.max_x = self.s[0], self.s[-1]
        self.delta_x = (self.max_x - self.min_x) / (self.s.shape[0] - 1)
        self.current_element = None
        self.device = torch.device(device) if device is not None else None
        self._dev = None

    def get_steps(self):   # lattice.py:152-173
        self.step_size = self.lattice_config["step_size"]
        self._positions_record = np.arange(0, self.lattice_length + self.step_size / 2,
                                           self.step_size)
        self._total_steps = len(self._positions_record)
        self.steps_per_element = np.zeros((self.Nelement,), dtype=int)
        csr_idx = []
        prev = 0
        for k, d in enumerate(self.distance):
            ind = np.searchsorted(self._positions_record, d, side="right")
            csr_idx.extend(np.arange(prev, ind, self.nsep[k]).tolist())
            self.steps_per_element[k] = ind - prev - 1 if k == 0 else ind - prev
            prev = ind
        self._CSR_steps_index = np.array(csr_idx, dtype=float)

    @property
    def lattice_length(self):
        return self._lattice_length

    @property
    def total_steps(self):
        return self._total_steps

    @property
    def steps_record(self):
        return self._positions_record

    @property
    def CSR_steps_index(self):
        return self._CSR_steps_index

    @property
    def Nelement(self):
        return self._Nelement

    def update(self, ele_name):
        self.current_element = ele_name

    # -- device side ---------------------------------------------------------
    def device_tables(self, device=None):
        """Uploads the tables once; returns the ctypes descriptor."""
        device = torch.device(device) if device is not None else self.device
        if self._dev is None or self._dev[0] != device:
            tab = np.stack([self.coords[:, 0], self.coords[:, 1], self.n_vec[:, 0],
                            self.n_vec[:, 1], self.tau_vec[:, 0], self.tau_vec[:, 1]], axis=1)
            t = torch.from_numpy(np.ascontiguousarray(tab)).to(device)
            dist = torch.from_numpy(self.distance.copy()).to(device)
            rho = torch.from_numpy(self.rho.copy()).to(device)
            desc = _lib.Lattice(t.data_ptr(), tab.shape[0], self.Nelement, float(self.min_x),
                                float(self.delta_x), dist.data_ptr(), rho.data_ptr())
            self._dev = (device, t, dist, rho, desc)
        return self._dev[4]
