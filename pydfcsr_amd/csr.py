"""CSR2D: the reference's driver (CSR.py:27-879) over the gfx950 step.

Same YAML input, same run loop (element boundaries split a step, the
formation length follows the last bend, CSR every nsep steps), same result
attributes: dE_dct and x_kick on the (xbins, zbins) wake mesh in MeV/m,
CSR_xmesh / CSR_zmesh / CSR_zrange / CSR_xrange_transformed, and the kicked
particles.  The CSR step itself (deposition -> density functions -> history
-> retarded integral -> kicks) runs on the device through libdfcsr.

``precision`` selects the integral's arithmetic: 'fp64' (plain) or 'mixed'
(FP32 integrand with compensated FP32 accumulation, FP64 retardation
geometry and final sum).
"""
from __future__ import annotations

import os
import time

import numpy as np
import torch

from . import _lib, h5io, kernels
from .beams import Beam, bmadx_element
from .deposit import DF_tracker
from .lattice import Lattice
from .params import CSR_params, Integration_params
from .yaml_parser import parse_yaml

CSR_SCALING_PER_COULOMB = 8.98755e3   # MeV/m per (C/m^2), CSR.py:80


class CSR2D:
    def __init__(self, input_file=None, parallel=False, device=None, precision="fp64",
                 config=None, verbose=False, keep_fp32=None, history_capacity=16,
                 keep_fp64=None):
        """input_file: YAML path (or a parsed dict via ``config``).  parallel is
        accepted for interface parity: the reference's MPI split over mesh
        points is the kernel's launch grid here.  precision "fp64" or "mixed";
        the history keeps the rings that precision reads unless keep_fp64 /
        keep_fp32 ask for more (e.g. to compare both precisions on one state)."""
        if device is None:
            device = "cuda"
        self.device = torch.device(device)
        self.precision = precision
        self.keep_fp32 = (precision == "mixed") if keep_fp32 is None else bool(keep_fp32)
        self.keep_fp64 = (precision == "fp64") if keep_fp64 is None else bool(keep_fp64)
        self.history_capacity = int(history_capacity)
        self.verbose = verbose
        self.timestamp = time.strftime("%Y-%m-%dT%H_%M_%S")
        self.parallel = False
        self.formation_length = None
        self.ws = _lib.Workspace(self.device)
        self.timing = False        # record HIP events around each CSR step
        self.timing_events = []    # (start, wake_start, wake_end, end, points, M)
        if input_file is not None or config is not None:
            self.parse_input(input_file, config)
            self.input_file = input_file
        self.initialization()
        self.prefix = f"{self.CSR_params.write_name}-{self.timestamp}"

    # -- CSR.py:48-68, 127-138 ---------------------------------------------------
    def parse_input(self, input_file, config=None):
        base = None
        if config is None:
            config = parse_yaml(input_file)
            base = os.path.dirname(os.path.abspath(input_file))
        self.check_input_consistency(config)
        self.input = config
        self.beam = Beam(config["input_beam"], device=self.device, base_dir=base)
        self.lattice = Lattice(config["input_lattice"], device=self.device, base_dir=base)
        self.DF_tracker = DF_tracker(config.get("particle_deposition") or {}, device=self.device,
                                     keep_fp32=self.keep_fp32,
                                     capacity=self.history_capacity,
                                     keep_fp64=self.keep_fp64)
        self.integration_params = Integration_params(config.get("CSR_integration") or {})
        self.CSR_params = CSR_params(config.get("CSR_computation") or {})

    def check_input_consistency(self, input):
        self.required_inputs = ["input_beam", "input_lattice"]
        allowed = self.required_inputs + ["particle_deposition", "distribution_interpolation",
                                          "CSR_integration", "CSR_computation"]
        for p in input:
            assert p in allowed, (f"Incorrect param given to {self.__class__.__name__}.__init__"
                                  f"(**kwargs): {p}\nAllowed params: {allowed}")
        for r in self.required_inputs:
            assert r in input, (f"Required input parameter {r} to "
                                f"{self.__class__.__name__}.__init__(**kwargs) was not found.")

    # -- CSR.py:70-81 -----------------------------------------------------------
    def initialization(self):
        b = self.beam
        self.DF_tracker.get_DF(b.x, b.z, b.px, t=b.position, moments=b.moments)
        self.DF_tracker.append_DF()
        self.DF_tracker.append_interpolant(formation_length=float("inf"),
                                           n_formation_length=self.integration_params.n_formation_length)
        self.CSR_scaling = CSR_SCALING_PER_COULOMB * self.beam.charge
        self.init_statistics()

    def init_statistics(self):   # CSR.py:82-108
        n = self.lattice.total_steps
        st = {"twiss": {f"{key}_{plane}": np.zeros(n) for plane in ("x", "y")
                        for key in _lib.TWISS_KEYS}}
        st["slope"] = np.zeros((n, 2))
        for k in ("sigma_x", "sigma_z", "sigma_energy", "mean_x", "mean_z", "mean_energy"):
            st[k] = np.zeros(n)
        self._statistics = st
        # moments, Twiss and energy rows stay on the device until the dict is read
        self._stats_dev = torch.zeros((n, _lib.STATS_COUNT), dtype=torch.float64,
                                      device=self.device)
        self._moments_rows = torch.zeros((n, _lib.MOMENTS_COUNT), dtype=torch.float64,
                                         device=self.device)
        self._stats_dirty = False
        self.update_statistics(step=0)
        self.inbend = False
        self.afterbend = False
        self.R_rec = None
        self.phi_rec = None

    def update_statistics(self, step):   # CSR.py:837-859
        if step >= len(self._statistics["sigma_x"]):
            return
        b = self.beam
        self._moments_rows[step].copy_(b.moments_device)   # device rows, no sync
        b.statistics(out=self._stats_dev[step])
        self._stats_dirty = True

    @property
    def statistics(self):
        """The reference's statistics dict (CSR.py:82-106): per step twiss
        (alpha/beta/gamma/emit/eta/etap/norm_emit of x and y), slope, sigma_x,
        sigma_z, sigma_energy, mean_x, mean_z, mean_energy.  Reading it copies
        the device-side rows once."""
        if self._stats_dirty:
            m = self._moments_rows.cpu().numpy()
            st = self._statistics
            col = kernels.MOMENT_KEYS.index
            st["slope"][:, 0] = m[:, col("slope0")]
            st["slope"][:, 1] = m[:, col("slope1")]
            for k in ("sigma_x", "sigma_z", "mean_x", "mean_z"):
                st[k][:] = m[:, col(k)]
            v = self._stats_dev.cpu().numpy()
            tw = self._statistics["twiss"]
            for plane, off in (("x", _lib.STAT_TWISS_X), ("y", _lib.STAT_TWISS_Y)):
                for k, key in enumerate(_lib.TWISS_KEYS):
                    tw[f"{key}_{plane}"][:] = v[:, off + k]
            self._statistics["mean_energy"][:] = v[:, _lib.STAT_MEAN_ENERGY]
            self._statistics["sigma_energy"][:] = v[:, _lib.STAT_SIGMA_ENERGY]
            self._stats_dirty = False
        return self._statistics

    def get_formation_length(self, R, sigma_z, phi=0.0, inbend=True):   # CSR.py:140-144
        if inbend:
            self.formation_length = (24 * (R ** 2) * sigma_z) ** (1 / 3)
        else:
            self.formation_length = (3 * R ** 2 * phi ** 4) / (4 * (-6 * sigma_z + R * phi ** 3))

    def get_bmadx_element(self, ele, DL, entrance=False, exit=False):   # CSR.py:146-199
        return bmadx_element(self.lattice.lattice_config[ele], DL, entrance=entrance, exit=exit)

    def _track(self, ele_name, DL, entrance=False, exit=False, update_step=True):
        element = self.get_bmadx_element(ele_name, DL, entrance=entrance, exit=exit)
        self.beam.track(element, DL, update_step=update_step)

    # -- CSR.py:202-358 ---------------------------------------------------------------
    def run(self, stop_time=None, debug=False, max_steps=None, on_step=None):
        """The reference's element/step loop.  ``max_steps`` stops after that
        many steps; ``on_step(self, step_count)`` is called after each step."""
        for step_count in self.iter_steps(stop_time=stop_time, debug=debug):
            if on_step is not None:
                on_step(self, step_count)
            if max_steps is not None and step_count >= max_steps:
                return
        if self._stopped:
            return
        self.dump_beam(label="end")
        self.write_statistics()

    def iter_steps(self, stop_time=None, debug=False):
        """Generator form of run(): yields the step count after each step."""
        for kind, step in self.iter_phases(stop_time=stop_time, debug=debug, external_wake=False):
            yield step

    def iter_phases(self, stop_time=None, debug=False, external_wake=False):
        """Step generator yielding ("done", step) after each step.  With
        external_wake, a step that needs a wake first yields ("wake", step)
        with the mesh prepared (prepare_wake()); the caller computes the wake
        into self.wake_out (e.g. batched with other simulations) and resumes,
        and the step then applies the kicks."""
        step_count = 1
        self._stopped = False
        DL = self.lattice.step_size
        ele_count = 0
        skip_ele = False
        self.inbend = False
        self.afterbend = False
        self.formation_length = 0.0
        ele_prev = None
        for ele in self.lattice.element_names:
            self.lattice.update(ele)
            cfg = self.lattice.lattice_config[ele]
            L = cfg["L"]
            typ = cfg["type"]
            steps = self.lattice.steps_per_element[ele_count]
            if (not skip_ele) and ele_count > 0:
                # the part of the step still in the previous element (CSR.py:228-236)
                DL_1 = self.lattice.distance[ele_count - 1] - self.beam.position
                if DL_1 > 1.0e-6:
                    self._track(ele_prev, DL_1, exit=True, update_step=False)
            if steps == 0:
                skip_ele = True
                self._track(ele, L, entrance=True, exit=True, update_step=False)
            if typ == "dipole":
                angle = cfg["angle"]
                R = L / angle
                self.inbend = True
                self.afterbend = True
                self.R_rec = R
                self.phi_rec = angle
                self.get_formation_length(R=R, sigma_z=5 * self.beam.sigma_z, inbend=True)
            else:
                self.inbend = False
                if self.afterbend:
                    self.get_formation_length(R=self.R_rec, sigma_z=5 * self.beam.sigma_z,
                                              inbend=True)
                else:
                    self.formation_length += L
            for step in range(steps):
                t0 = time.time()
                if step == 0 and ele_count > 0:
                    DL_2 = self.lattice._positions_record[step_count] - \
                        self.lattice.distance[ele_count - 1]
                    self._track(ele, DL_2, entrance=True)
                    skip_ele = False
                else:
                    self._track(ele, DL)
                ev = self._event() if self.timing else None
                if debug or self.CSR_params.compute_CSR:
                    self.update_history()
                if self.CSR_params.compute_CSR and step % self.lattice.nsep[ele_count] == 0:
                    self.get_CSR_mesh()
                    ev_w0 = self._event() if self.timing else None
                    if external_wake:
                        yield ("wake", step_count)
                        self.finish_wake()
                    else:
                        self.calculate_2D_CSR()
                    ev_w1 = self._event() if self.timing else None
                    pts = self.integrand_points()
                    if self.CSR_params.apply_CSR:
                        self.beam.apply_wakes(self.dE_dct, self.x_kick, self.mesh_x_axis,
                                              self.mesh_z_axis, DL * self.lattice.nsep[ele_count],
                                              self.CSR_params.transverse_on)
                    if self.timing:
                        self.timing_events.append((ev, ev_w0, ev_w1, self._event(), pts,
                                                   self._mesh_M()))
                    wb = self.CSR_params.write_beam
                    if wb == "all" or (isinstance(wb, list) and step_count in wb):
                        self.dump_beam(label=step_count)
                    if self.CSR_params.write_wakes:
                        self.write_wakes()
                self.update_statistics(step=step_count)
                if self.verbose:
                    print(f"Finish step {step_count}, s = {self.beam.position},  in "
                          f"{time.time() - t0} seconds", flush=True)
                yield ("done", step_count)
                step_count += 1
                if stop_time and self.beam.position > stop_time:
                    self._stopped = True
                    return
            ele_prev = ele
            ele_count += 1

    # -- one CSR step (the accelerated path) ------------------------------------------
    def update_history(self):
        """get_DF -> append_DF -> append_interpolant -> build_interpolant (CSR.py:297-307)."""
        b = self.beam
        tr = self.DF_tracker
        tr.get_DF(b.x, b.z, b.px, t=b.position, moments=b.moments)
        tr.append_DF()
        tr.append_interpolant(formation_length=self.formation_length,
                              n_formation_length=self.integration_params.n_formation_length)
        tr.build_interpolant()

    def csr_step(self, step_size):
        """History update, wake and kick for the current beam (one CSR step)."""
        self.update_history()
        self.get_CSR_mesh()
        self.calculate_2D_CSR()
        if self.CSR_params.apply_CSR:
            self.beam.apply_wakes(self.dE_dct, self.x_kick, self.mesh_x_axis, self.mesh_z_axis,
                                  step_size, self.CSR_params.transverse_on)

    # -- CSR.py:361-394 -------------------------------------------------------------
    def get_CSR_mesh(self):
        m = self.beam.moments
        sig_x, mean_x = m["sigma_xt"], m["mean_xt"]
        sig_z, mean_z = m["sigma_z"], m["mean_z"]
        p = self.CSR_params
        self.mesh_z_axis = (mean_z - p.zlim * sig_z, mean_z + p.zlim * sig_z, int(p.zbins))
        self.mesh_x_axis = (mean_x - p.xlim * sig_x, mean_x + p.xlim * sig_x, int(p.xbins))
        self.CSR_zrange = np.linspace(*self.mesh_z_axis)
        self.CSR_xrange_transformed = np.linspace(*self.mesh_x_axis)
        self._mesh_slope = np.array(m["slope"])

    def wake_params(self):
        b = self.beam
        p = self.CSR_params
        ip = self.integration_params
        return _lib.WakeParams(
            float(b.position), float(b._sigma_x), float(b._sigma_z), float(b._slope[0]),
            float(self._mesh_slope[1]), float(b._mean_x), float(self.formation_length),
            float(ip.n_formation_length), float(self.CSR_scaling), int(ip.xbins), int(ip.zbins),
            int(p.xbins), int(p.zbins), float(self.mesh_x_axis[0]), float(self.mesh_x_axis[1]),
            float(self.mesh_z_axis[0]), float(self.mesh_z_axis[1]), None)

    # -- CSR.py:397-418 ----------------------------------------------------------------
    def calculate_2D_CSR(self):
        prec = _lib.DFCSR_PREC_MIXED if self.precision == "mixed" else _lib.DFCSR_PREC_FP64
        self._wake_out = kernels.wake_mesh(self.DF_tracker.descriptor(),
                                           self.lattice.device_tables(self.device),
                                           self.wake_params(), prec, self.device, self.ws,
                                           out=getattr(self, "_wake_out", None) if
                                           getattr(self, "_wake_M", None) == self._mesh_M() else None)
        self._wake_M = self._mesh_M()
        p = self.CSR_params
        self.dE_dct = self._wake_out["dE_dct"].view(p.xbins, p.zbins)
        self.x_kick = self._wake_out["x_kick"].view(p.xbins, p.zbins)
        self.CSR_xmesh = self._wake_out["xmesh"]
        self.CSR_zmesh = self._wake_out["zmesh"]

    def prepare_wake(self):
        """(history, lattice, params, outputs) of the pending wake."""
        M = self._mesh_M()
        if getattr(self, "wake_out", None) is None or self.wake_out["dE_dct"].numel() != M:
            self.wake_out = {k: torch.empty(M, dtype=torch.float64, device=self.device)
                             for k in ("dE_dct", "x_kick")}
        return (self.DF_tracker.descriptor(), self.lattice.device_tables(self.device),
                self.wake_params(), self.wake_out)

    def finish_wake(self):
        p = self.CSR_params
        self.dE_dct = self.wake_out["dE_dct"].view(p.xbins, p.zbins)
        self.x_kick = self.wake_out["x_kick"].view(p.xbins, p.zbins)
        self.CSR_xmesh, self.CSR_zmesh = self.host_mesh()

    def host_mesh(self):
        """(CSR_xmesh, CSR_zmesh) of the current mesh computed on the host
        with the reference's expressions (CSR.py:380-389)."""
        xt, zm = np.meshgrid(self.CSR_xrange_transformed, self.CSR_zrange, indexing="ij")
        zm = zm.flatten()
        return xt.flatten() + np.polyval(self._mesh_slope, zm), zm

    # -- output files (CSR.py:784-879) ------------------------------------------------
    def _output_file(self, what):
        return os.path.join(self.CSR_params.workdir, f"{self.prefix}-{what}.h5")

    def dump_beam(self, label):   # CSR.py:784-797
        filename = self._output_file(f"particles-{label}")
        if self.verbose:
            print("Beam at position {} is written to {}".format(self.beam.position, filename))
        h5io.write_particles_file(filename, self.beam.particle_group, verbose=self.verbose)

    def write_wakes(self):   # CSR.py:799-835
        b = self.beam
        xm, zm = self.CSR_xmesh, self.CSR_zmesh
        xm = xm.cpu().numpy() if isinstance(xm, torch.Tensor) else np.asarray(xm)
        zm = zm.cpu().numpy() if isinstance(zm, torch.Tensor) else np.asarray(zm)
        h5io.write_wakes_file(self._output_file("wakes"), b.step, b.position, b.init_gamma,
                              b.init_energy, self.lattice.current_element, b.charge, xm, zm,
                              self.dE_dct.cpu().numpy(), self.x_kick.cpu().numpy(),
                              verbose=self.verbose)

    def write_statistics(self):   # CSR.py:860-879
        lat = self.lattice
        h5io.write_statistics_file(self._output_file("statistics"), lat.steps_record,
                                   lat.coords, lat.n_vec, lat.tau_vec, self.statistics,
                                   verbose=self.verbose)

    def _event(self):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        return e

    def integrand_points(self):
        """Integrand evaluations of the last wake: M * P with P = 4 Nx Nz
        (three regions) or 5 Nx Nz (four, chirp branch), CSR.py:541-597."""
        ip = self.integration_params
        per = (5 if abs(self.beam._slope[0]) > 1 else 4) * ip.xbins * ip.zbins
        return self._mesh_M() * per

    def _mesh_M(self):
        return int(self.CSR_params.xbins) * int(self.CSR_params.zbins)

    # -- CSR.py:454-602 (per-observer diagnostic entry) -------------------------------
    def _regions(self, s, x):
        b = self.beam
        t = b.position
        Nx, Nz = self.integration_params.xbins, self.integration_params.zbins
        sz, sx, tan_theta, xmean = b._sigma_z, b._sigma_x, b._slope[0], b._mean_x
        x0 = (s - t) * b._slope[0]
        nfl = self.integration_params.n_formation_length * self.formation_length
        if np.abs(tan_theta) <= 1:
            s2, s3, s4 = s - 500 * sz, s - 20 * sz, s + 5 * sz
            s1 = np.max((0, s2 - nfl))
            xw = np.linspace(x0 - 20 * sx, x0 + 20 * sx, 2 * Nx)
            xn = np.linspace(x0 - 10 * sx, x0 + 10 * sx, Nx)
            return [(xw, np.linspace(s1, s2, Nz)), (xn, np.linspace(s2, s3, Nz)),
                    (xn, np.linspace(s3, s4, Nz))]
        if tan_theta > 0:
            ta = -2 * tan_theta / (1 - tan_theta ** 2)
            d = (10 * sx + xmean - x) / ta
            x1 = (x + 0.1 * sx, x + 10 * sx)
            x2 = (x - 3 * sx, x1[0])
        else:
            ta = 2 * tan_theta / (1 - tan_theta ** 2)
            d = -(xmean - x - 10 * sx) / ta
            x1 = (x - 10 * sx, x - 1 * sx)
            x2 = (x1[1], x + 3 * sx)
        s4 = s + 3 * sz
        s3 = np.max((0, s - d))
        s2 = s3 - 200 * sz
        s1 = np.max((0, s2 - nfl))
        sp1, sp2, sp3 = (np.linspace(s1, s2, Nz), np.linspace(s2, s3, Nz), np.linspace(s3, s4, Nz))
        return [(np.linspace(x0 - 20 * sx, x0 + 20 * sx, 2 * Nx), sp1),
                (np.linspace(x0 - 5 * sx, x0 + 5 * sx, Nx), sp2),
                (np.linspace(*x1, Nx), sp3), (np.linspace(*x2, Nx), sp3)]

    def get_CSR_wake(self, s, x, debug=False):
        """(dE_dct, x_kick) at one observer; debug=True returns the region
        axes and integrand grids instead (CSR.py:571-572, 599-600)."""
        regs = self._regions(s, x)
        t = self.beam.position
        grids = []
        dE = 0.0
        xk = 0.0
        trapz = getattr(np, "trapezoid", None) or np.trapz
        for xp, sp in regs:
            xpm, spm = np.meshgrid(xp, sp, indexing="ij")
            Iz, Ix = self.get_CSR_integrand(s, x, t, spm, xpm)
            grids.append((Iz, Ix))
            dE = dE + (-self.CSR_scaling * trapz(trapz(Iz, xp, axis=0), sp))
            xk = xk + self.CSR_scaling * trapz(trapz(Ix, xp, axis=0), sp)
        if debug:
            return regs, grids
        return dE, xk

    def get_CSR_integrand(self, s, x, t, sp, xp, ignore_vx=False):
        """Integrand on meshgrid(xp, sp, indexing='ij') arrays (CSR.py:605-782)."""
        if ignore_vx:
            raise NotImplementedError("ignore_vx is never set by the reference (CSR.py:471-474)")
        xp1 = torch.as_tensor(np.ascontiguousarray(np.asarray(xp)[:, 0]), device=self.device)
        sp1 = torch.as_tensor(np.ascontiguousarray(np.asarray(sp)[0, :]), device=self.device)
        Iz, Ix = kernels.csr_integrand(self.DF_tracker.descriptor(),
                                       self.lattice.device_tables(self.device), s, x, t, xp1, sp1)
        return Iz.cpu().numpy(), Ix.cpu().numpy()
