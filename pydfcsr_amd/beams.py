"""Beam state on the device and its per-step statistics.

Mirrors the reference's Beam (beams.py:12-229): particle coordinates in Bmad
order (x, px, y, py, z, pz) relative to the constant reference momentum
``init_energy`` (eV), the cached moments of update_status, the CSR kick of
apply_wakes and tracking through lattice elements.

Statistics (sigma, mean, slope = polyfit(z, x, 1), x_transform moments and
the |z| < 0.1 sigma_z slice std) come from one deterministic device
reduction; the per-step record (Twiss of x and y, energy mean and spread)
from another (dfcsr_beam_statistics); the kick is the gfx950 mesh->particle
kernel.

Tracking replaces bmadx's track_element (absent here) with the gfx950
element kernel (dfcsr_track_element): exact drift, exact sector bend with
hard-edge faces, quadrupole with second-order path length and low-energy z
terms, sextupole drift-kick-drift; the drift, bend and quadrupole maps
reproduce the reference notebook's bmadx results.
"""
from __future__ import annotations


import numpy as np
import torch

from . import _lib, h5io, kernels

# electron rest energy [eV]: scipy.constants' 'electron mass energy equivalent
# in MeV' (CODATA 2018) x 1e6, as physical_constants.py:11
MC2 = 510998.95069


def bmadx_element(ele: dict, DL: float, entrance=False, exit=False) -> _lib.Element:
    """The element of a step piece of length DL, as CSR2D.get_bmadx_element
    builds it for bmadx (CSR.py:146-199): a dipole's G (G key, else angle/L),
    its faces E1 / E2 and their fringe (FINT, HGAP / FINTX, HGAPX) only at the
    piece that holds them (FRINGE_AT entrance_end / exit_end); a dipole taken
    in one piece keeps the faces its own FRINGE_AT names (both_ends,
    entrance_end, exit_end or no_end) and, as in the reference, must have
    that key; a quad's K1, a sextupole's K2, with X_OFFSET, Y_OFFSET, TILT
    and NUM_STEPS when given."""
    typ = ele["type"]
    e = _lib.Element()
    e.L = float(DL)
    e.num_steps = int(ele.get("NUM_STEPS", 1))
    if typ == "dipole":
        e.type = _lib.DFCSR_ELE_SBEND
        e.g = float(ele["G"]) if "G" in ele else float(ele["angle"]) / float(ele["L"])
        if entrance and exit:
            if "FRINGE_AT" not in ele:
                raise ValueError("a dipole tracked in one piece needs FRINGE_AT (the reference "
                                 "reads it unbound otherwise, CSR.py:177)")
            at = ele["FRINGE_AT"]
            if at not in ("both_ends", "entrance_end", "exit_end", "no_end"):
                raise ValueError(f"FRINGE_AT {at!r}")
            entrance = at in ("both_ends", "entrance_end")
            exit = at in ("both_ends", "exit_end")
        if entrance:
            e.e1 = float(ele.get("E1", 0.0))
            e.fint = float(ele.get("FINT", 0.0) or 0.0)
            e.hgap = float(ele.get("HGAP", 0.0) or 0.0)
        if exit:
            e.e2 = float(ele.get("E2", 0.0))
            e.fintx = float(ele.get("FINTX", ele.get("FINT", 0.0)) or 0.0)
            e.hgapx = float(ele.get("HGAPX", ele.get("HGAP", 0.0)) or 0.0)
    elif typ == "drift":
        e.type = _lib.DFCSR_ELE_DRIFT
    elif typ in ("quad", "sextupole"):
        e.type = _lib.DFCSR_ELE_QUAD if typ == "quad" else _lib.DFCSR_ELE_SEXTUPOLE
        if typ == "quad":
            e.k1 = float(ele["K1"])
        else:
            e.k2 = float(ele["K2"])
        e.x_offset = float(ele.get("X_OFFSET", 0.0))
        e.y_offset = float(ele.get("Y_OFFSET", 0.0))
        e.tilt = float(ele.get("TILT", 0.0))
    else:
        raise ValueError(f"element type {typ!r}: the reference builds dipole, drift, quad and "
                         "sextupole pieces only (CSR.py:153-197)")
    return e


class Beam:
    """Beam from an ``input_beam`` config (beams.py:16-86).

    Styles: ``from_file`` (6 text columns x px y py z pz, plus charge [C] and
    energy [eV]) and ``gaussian`` (seeded synthetic beam: n_particle, charge,
    energy, sigma_x, sigma_z, chirp, sigma_delta, optional two_gaussian).
    ``distgen`` and ``ParticleGroup`` need packages this build does not ship.
    """

    def __init__(self, input_beam, device="cuda", base_dir=None):
        self.check_inputs(input_beam)
        self.input_beam_config = input_beam
        self.style = input_beam["style"]
        self.device = torch.device(device)
        if self.style == "from_file":
            coords = np.loadtxt(input_beam["beamfile"])
            assert coords.shape[1] == 6, \
                f"Error: input beam must have 6 dimension, but get {coords.shape[1]} instead"
            self._charge = input_beam["charge"]
            self._init_energy = input_beam["energy"]
            arr = coords.T.copy()
        elif self.style == "gaussian":
            arr = gaussian_beam(**{k: v for k, v in input_beam.items() if k != "style"})
            self._charge = float(input_beam["charge"])
            self._init_energy = float(input_beam["energy"])
        else:
            raise NotImplementedError(
                f"input_beam style {self.style!r} needs distgen/openPMD, which this build "
                "does not include; use 'from_file' or 'gaussian'")
        self.coords = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64)).to(self.device)
        self._init_gamma = self._init_energy / MC2
        self.position = 0
        self.step = 0
        self.ws = _lib.Workspace(self.device)
        self.update_status()

    @staticmethod
    def from_arrays(x, px, z, pz, charge, energy, y=None, py=None, device="cuda"):
        b = Beam.__new__(Beam)
        b.style = "arrays"
        b.device = torch.device(device)
        n = len(x)
        arr = np.zeros((6, n))
        arr[0], arr[1], arr[4], arr[5] = x, px, z, pz
        if y is not None:
            arr[2] = y
        if py is not None:
            arr[3] = py
        b.coords = torch.from_numpy(arr).to(b.device)
        b._charge = float(charge)
        b._init_energy = float(energy)
        b._init_gamma = b._init_energy / MC2
        b.position = 0
        b.step = 0
        b.ws = _lib.Workspace(b.device)
        b.update_status()
        return b

    @staticmethod
    def check_inputs(input_beam):
        assert "style" in input_beam, "ERROR: input_beam must have keyword <style>"
        req = {"from_file": ["style", "beamfile", "charge", "energy"],
               "distgen": ["style", "distgen_input_file"],
               "ParticleGroup": ["style", "ParticleGroup_h5"],
               "gaussian": ["style", "n_particle", "charge", "energy"]}
        if input_beam["style"] not in req:
            raise Exception("input beam parsing Error: invalid input style")
        if input_beam["style"] != "gaussian":
            allowed = req[input_beam["style"]] + ["verbose"]
            for p in input_beam:
                assert p in allowed, f"Incorrect param given to Beam.__init__(**kwargs): {p}\n" \
                                     f"Allowed params: {allowed}"
        for r in req[input_beam["style"]]:
            assert r in input_beam, \
                f"Required input parameter {r} to Beam.__init__(**kwargs) was not found."

    # -- statistics (beams.py:88-98, 137-215) ---------------------------------
    def update_status(self):
        """Launches the moments of the current coordinates; they stay on the
        device (moments_device) until a host value is read."""
        self._moments_dev = kernels.beam_moments(self.x, self.z, self.ws, host=False)
        self._moments_host = None

    @property
    def moments_device(self):
        """Device float64 vector of the DFCSR_MOMENTS_* values."""
        return self._moments_dev

    @property
    def moments(self):
        """The moments as a dict of floats (copied from the device on first read)."""
        if self._moments_host is None:
            self._moments_host = kernels.moments_dict(self._moments_dev)
        return self._moments_host

    _sigma_x = property(lambda self: self.moments["sigma_x"])
    _sigma_z = property(lambda self: self.moments["sigma_z"])
    _slope = property(lambda self: self.moments["slope"])
    _mean_x = property(lambda self: self.moments["mean_x"])
    _mean_z = property(lambda self: self.moments["mean_z"])

    def track(self, element, step_size, update_step=True):   # beams.py:101-106
        """Tracks through one element piece (an Element from bmadx_element)
        on the device, in place."""
        self.coords = self.coords.contiguous()
        kernels.track_element(self.coords, element, self.init_energy, MC2)
        self.position += step_size
        if update_step:
            self.step += 1
        self.update_status()

    def apply_wakes(self, dE_dct, x_kick, xrange, zrange, step_size, transverse_on):
        """beams.py:108-131.  xrange/zrange: the mesh linspace axes, given as
        arrays or as (start, stop, num) triples."""
        xr = _axis(xrange)
        zr = _axis(zrange)
        dE = torch.as_tensor(dE_dct, dtype=torch.float64, device=self.device).contiguous()
        xk = torch.as_tensor(x_kick, dtype=torch.float64, device=self.device).contiguous()
        self.coords = self.coords.contiguous()
        m = self.moments
        kernels.apply_wakes(self.coords[1], self.coords[5], self.coords[0], self.coords[4], dE, xk,
                            xr[0], xr[1], zr[0], zr[1], m["slope0"], m["slope1"], step_size,
                            self.init_energy, transverse_on, xr[2], zr[2])
        # the reference recomputes the status here (beams.py:131); the kick
        # changes px and pz only, and the status is moments of x and z, so
        # the current ones stand

    # -- properties --------------------------------------------------------------
    x = property(lambda self: self.coords[0])
    px = property(lambda self: self.coords[1])
    y = property(lambda self: self.coords[2])
    py = property(lambda self: self.coords[3])
    z = property(lambda self: self.coords[4])
    pz = property(lambda self: self.coords[5])
    charge = property(lambda self: self._charge)
    init_energy = property(lambda self: self._init_energy)
    init_gamma = property(lambda self: self._init_gamma)
    sigma_x = property(lambda self: self._sigma_x)
    sigma_z = property(lambda self: self._sigma_z)
    mean_x = property(lambda self: self._mean_x)
    mean_z = property(lambda self: self._mean_z)
    slope = property(lambda self: self._slope)

    @property
    def x_transform(self):
        return self.x - (self._slope[0] * self.z + self._slope[1])

    @property
    def energy(self):
        return (self.pz + 1) * self.init_energy

    def statistics(self, out=None):
        """Device vector of the per-step statistics (dfcsr_beam_statistics):
        means, covariance, Twiss x / y, energy mean and spread; the particle
        momentum reference p0c is init_energy (beams.py:33)."""
        self.coords = self.coords.contiguous()
        return kernels.beam_statistics(self.coords, self.init_energy, MC2, self.ws, out)

    @property
    def particle_group(self):   # beams.py:226-230
        """The beam as openPMD particle data (the dict of
        bmadx_particles_to_openpmd, interfaces.py:65-121), copied to the host."""
        c = self.coords.cpu().numpy()
        return h5io.openpmd_particle_data(*c, self.position, self.init_energy, MC2, self.charge)

    @property
    def twiss(self):   # beams.py:222-224 -> twiss.py:2-27
        return kernels.statistics_dict(self.statistics())["twiss"]

    @property
    def mean_energy(self):   # beams.py:171-173
        return kernels.statistics_dict(self.statistics())["mean_energy"]

    @property
    def sigma_energy(self):   # beams.py:179-181 (np.std: population)
        return kernels.statistics_dict(self.statistics())["sigma_energy"]


def _axis(a):
    if isinstance(a, (tuple, list)) and len(a) == 3:
        return float(a[0]), float(a[1]), int(a[2])
    a = np.asarray(a.cpu() if isinstance(a, torch.Tensor) else a, dtype=np.float64)
    return float(a[0]), float(a[-1]), int(a.shape[0])


def gaussian_beam(n_particle, sigma_x=50e-6, sigma_z=50e-6, chirp=0.0, sigma_delta=0.0,
                  seed=0, two_gaussian=None, charge=None, energy=None, sigma_px=0.0, **_):
    """Seeded synthetic beam, returns (6, n) coordinates.

    two_gaussian: dict(sigmas=(s1, s2), separation=d, ratio=(q1, q2)) puts the
    fraction q1 at -d/2 with sigma s1 and q2 at +d/2 with sigma s2 in z."""
    rng = np.random.default_rng(seed)
    n = int(n_particle)
    arr = np.zeros((6, n))
    if two_gaussian:
        s1, s2 = two_gaussian["sigmas"]
        d = two_gaussian["separation"]
        q1, _q2 = two_gaussian["ratio"]
        n1 = int(round(n * q1))
        arr[4, :n1] = rng.normal(-d / 2, s1, n1)
        arr[4, n1:] = rng.normal(d / 2, s2, n - n1)
    else:
        arr[4] = rng.normal(0.0, sigma_z, n)
    arr[0] = rng.normal(0.0, sigma_x, n)
    if sigma_px:
        arr[1] = rng.normal(0.0, sigma_px, n)
    arr[5] = chirp * arr[4]
    if sigma_delta:
        arr[5] += rng.normal(0.0, sigma_delta, n)
    return arr
