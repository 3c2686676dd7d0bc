"""Write golden input/output vectors from the reference's own hot-path functions.

Run ONLY in the build container, where the reference source tree is mounted:

    python tests/golden/make_reference_vectors.py /root/reference/pyDFCSR_2D [r_gen6|sgolay2d|run_loop]

(with ``r_gen6`` / ``sgolay2d`` / ``run_loop`` only ``ref_r_gen6.npz`` /
``ref_sgolay2d.npz`` / ``ref_run_loop.npz`` is written)

The reference package cannot be imported as shipped (numba, bmadx, mpi4py,
h5py, distgen and pmd_beamphysics are absent).  The functions captured here
are plain numpy/scipy code once numba's decorators are identities, so this
script registers import-only placeholders for those absent modules before
loading the reference modules from their files.  None of the placeholders is
called by a captured function except numba's ``jit``/``jitclass``/``prange``
(identity / ``range``), whose pure-Python semantics equal the JIT's for these
functions on finite inputs.

Nothing of the reference travels: the outputs are plain ``.npz`` arrays
(inputs and expected outputs) under ``tests/golden/``.
"""
from __future__ import annotations

import collections
import importlib.util
import os
import sys
import tempfile
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _install_placeholders():
    def jit(*a, **k):
        if a and callable(a[0]):
            return a[0]
        return lambda f: f

    numba = types.ModuleType("numba")
    numba.jit = jit
    numba.njit = jit
    numba.prange = range
    class _Type:                      # numba type objects are only subscripted in specs
        def __getitem__(self, item):
            return self

    numba.double = _Type()
    exp = types.ModuleType("numba.experimental")
    exp.jitclass = lambda spec=None: (lambda c: c)
    numba.experimental = exp
    sys.modules["numba"] = numba
    sys.modules["numba.experimental"] = exp

    bmadx = types.ModuleType("bmadx")
    bmadx.Particle = collections.namedtuple("Particle", "x px y py z pz s p0c mc2")
    bmadx.M_ELECTRON = 510998.95
    for name in ("Drift", "SBend", "Quadrupole", "Sextupole", "track_element"):
        setattr(bmadx, name, None)
    bmadx.__path__ = []
    consts = types.ModuleType("bmadx.constants")
    consts.C_LIGHT, consts.M_ELECTRON, consts.E_CHARGE = 299792458.0, 510998.95, 1.602176634e-19
    bmadx.constants = consts
    sys.modules["bmadx"] = bmadx
    sys.modules["bmadx.constants"] = consts
    for name in ("mpi4py", "h5py", "distgen", "pmd_beamphysics"):
        m = types.ModuleType(name)
        sys.modules[name] = m
    sys.modules["mpi4py"].MPI = None
    sys.modules["distgen"].Generator = None
    sys.modules["pmd_beamphysics"].ParticleGroup = None


def _load_reference(ref_dir):
    _install_placeholders()
    pkg = types.ModuleType("pyDFCSR_2D")
    pkg.__path__ = [ref_dir]
    sys.modules["pyDFCSR_2D"] = pkg
    # beams.py imports the openPMD converters of interfaces.py (bmadx/torch
    # glue for beam I/O); none of the captured functions uses them.
    itf = types.ModuleType("pyDFCSR_2D.interfaces")
    itf.openpmd_to_bmadx_particles = itf.bmadx_particles_to_openpmd = None
    sys.modules["pyDFCSR_2D.interfaces"] = itf
    mods = {}
    for name in ("tools", "yaml_parser", "params", "interp1D", "interp3D", "deposit",
                 "lattice", "twiss", "r_gen6", "twiss_R", "physical_constants",
                 "beams", "CSR"):
        spec = importlib.util.spec_from_file_location(f"pyDFCSR_2D.{name}",
                                                      os.path.join(ref_dir, f"{name}.py"))
        mod = importlib.util.module_from_spec(spec)
        sys.modules[f"pyDFCSR_2D.{name}"] = mod
        spec.loader.exec_module(mod)
        mods[name] = mod
    return mods


LATTICE_YAML = """step_size: 0.05
D0:
  type: drift
  L: 0.1
  nsep: 1
B1:
  type: dipole
  L: 0.3
  angle: 0.05
  nsep: 1
D1:
  type: drift
  L: 0.4
  nsep: 2
B2:
  type: dipole
  L: 0.3
  angle: -0.05
  nsep: 1
"""


def make_beam(rng, n, sx, sz, chirp=0.0, corr=0.0):
    z = rng.normal(0.0, sz, n)
    x = rng.normal(0.0, sx, n) + corr * z
    px = rng.normal(0.0, 1e-4, n) + 0.02 * x
    pz = chirp * z + rng.normal(0.0, 1e-4, n)
    return x, px, z, pz


def main(ref_dir):
    R = _load_reference(ref_dir)
    dep, i3, i1, lat_mod, CSR, params, beams = (R["deposit"], R["interp3D"], R["interp1D"],
                                                R["lattice"], R["CSR"], R["params"], R["beams"])
    rng = np.random.default_rng(12345)
    out = {}

    # a3 histogram_cic_2d (deposit.py:42-87), particles straddling the grid edges
    q1 = rng.normal(0.0, 1.0, 3000)
    q2 = rng.normal(0.0, 2.0, 3000)
    w = rng.normal(0.5, 0.3, 3000)
    h1 = dep.histogram_cic_2d(q1, q2, np.ones(q1.shape), 16, -2.5, 2.5, 12, -5.0, 5.0)
    h2 = dep.histogram_cic_2d(q1, q2, w, 16, -2.5, 2.5, 12, -5.0, 5.0)
    np.savez(os.path.join(HERE, "ref_histogram_cic_2d.npz"), q1=q1, q2=q2, w=w,
             args=np.array([16, -2.5, 2.5, 12, -5.0, 5.0]), h_ones=h1, h_w=h2)

    # a11 interpolate3D (interp3D.py:18-66) incl. truncation / clamp / out-of-range
    data = rng.normal(size=(5, 7, 9))
    mins = np.array([0.1, -0.3, 0.2])
    dels = np.array([0.5, 0.25, 0.125])
    sizes = np.array(data.shape)
    u = rng.uniform(-1.6, sizes[None, :] + 0.6, size=(600, 3))
    u[:40, 0] = rng.uniform(-0.99, -0.01, 40)          # (-1, 0): negative fraction
    u[40:80, 1] = rng.uniform(sizes[1] - 1, sizes[1] - 0.01, 40)   # clamp plane
    vals = mins[None, :] + u * dels[None, :]
    r3 = i3.interpolate3D(vals[:, 0].copy(), vals[:, 1].copy(), vals[:, 2].copy(), data,
                          mins[0], mins[1], mins[2], dels[0], dels[1], dels[2])
    d1 = rng.normal(size=23)
    u1 = rng.uniform(-1.5, 23.5, 300)
    v1 = -0.4 + u1 * 0.05
    r1 = i1.interpolate1D(v1.copy(), d1, -0.4, 0.05)
    np.savez(os.path.join(HERE, "ref_interpolate.npz"), data3=data, mins=mins, dels=dels,
             q3=vals, r3=r3, data1=d1, q1=v1, r1=r1, min1=-0.4, del1=0.05)

    # a12 lattice tables (lattice.py:4-173)
    tmpdir = tempfile.mkdtemp()
    lat_path = os.path.join(tmpdir, "lat.yaml")
    with open(lat_path, "w") as f:
        f.write(LATTICE_YAML)
    lat = lat_mod.Lattice({"lattice_input_file": lat_path})
    np.savez(os.path.join(HERE, "ref_lattice.npz"), s=lat.s, rho=lat.rho, distance=lat.distance,
             nsep=lat.nsep, coords=lat.coords, n_vec=lat.n_vec, tau_vec=lat.tau_vec,
             positions_record=lat._positions_record, steps_per_element=lat.steps_per_element,
             min_x=lat.min_x, delta_x=lat.delta_x)

    # a4 get_DF (deposit.py:145-245): correlated beam (configured grid) and
    # uncorrelated beam (100x100 / window 5 fallback)
    cfg = dict(xbins=40, zbins=48, xlim=5, zlim=5, filter_order=2, filter_window=7,
               velocity_threhold=1000)
    for tag, corr in (("corr", 4.0), ("fallback", 0.0)):
        x, px, z, pz = make_beam(rng, 12000, 30e-6, 60e-6, corr=corr)
        tr = dep.DF_tracker(dict(cfg))
        tr.get_DF(x=x, z=z, px=px, t=0.25)
        np.savez(os.path.join(HERE, f"ref_get_df_{tag}.npz"), x=x, z=z, px=px, t=0.25,
                 cfg_xbins=40, cfg_zbins=48, cfg_order=2, cfg_window=7, cfg_vt=1000.0,
                 x_grids=tr.x_grids, z_grids=tr.z_grids, density=tr.density, vx=tr.vx,
                 density_x=tr.density_x, density_z=tr.density_z, vx_x=tr.vx_x)

    # a8-a10 history: 7 slices with a sigma jump (forces a full re-grid), a
    # window pop, capped interp grid (upper_limit) to keep the fixture small
    tr = dep.DF_tracker(dict(cfg, upper_limit=40))
    hist_inputs = []
    times = [0.0, 0.05, 0.1, 0.15, 0.2, 0.25, 0.3]
    sig = [30e-6, 31e-6, 32e-6, 90e-6, 91e-6, 92e-6, 93e-6]
    fls = [float("inf"), 0.2, 0.2, 0.2, 0.06, 0.06, 0.06]
    regrid_flags = []
    for k, (t, sx, fl) in enumerate(zip(times, sig, fls)):
        x, px, z, pz = make_beam(rng, 3000, sx, 50e-6 + 2e-6 * k, corr=0.5)
        hist_inputs.append(np.stack([x, z, px]))
        tr.get_DF(x=x, z=z, px=px, t=t)
        tr.append_DF()
        grid_before = tr.x_grid_interp
        tr.append_interpolant(formation_length=fl, n_formation_length=1.0)
        regrid_flags.append(int(tr.x_grid_interp is not grid_before))   # new interp grid
        if k >= 1:          # one slice: deposit.py:419 divides by zero
            tr.build_interpolant()
    np.savez(os.path.join(HERE, "ref_history.npz"), inputs=np.stack(hist_inputs),
             times=np.array(times), fls=np.array(fls), n_fl=1.0,
             cfg_xbins=40, cfg_zbins=48, cfg_order=2, cfg_window=7, cfg_vt=1000.0,
             upper_limit=40, time_interp=np.array(tr.time_interp),
             regrid_flags=np.array(regrid_flags),
             x_grid_interp=tr.x_grid_interp, z_grid_interp=tr.z_grid_interp,
             axes=np.array([tr.min_x, tr.min_y, tr.min_z, tr.delta_x, tr.delta_y, tr.delta_z]),
             density=tr.data_density_interp, density_x=tr.data_density_x_interp,
             density_z=tr.data_density_z_interp, vx=tr.data_vx_interp,
             vx_x=tr.data_vx_x_interp)

    # a13-a15 integrand, wake and mesh loop (CSR.py:361-782) on that history
    class _Beam:
        pass

    xl, pxl, zl, pzl = hist_inputs[-1][0], hist_inputs[-1][2], hist_inputs[-1][1], None
    obj = CSR.CSR2D.__new__(CSR.CSR2D)
    obj.DF_tracker = tr
    obj.lattice = lat
    obj.integration_params = params.Integration_params({"n_formation_length": 1.0,
                                                        "zbins": 16, "xbins": 12})
    obj.CSR_params = params.CSR_params({"xbins": 3, "zbins": 4, "xlim": 3, "zlim": 3,
                                        "workdir": tmpdir})
    obj.CSR_scaling = 8.98755e3 * 1e-9
    obj.formation_length = 0.05
    Bm = beams.Beam.__new__(beams.Beam)
    Particle = sys.modules["bmadx"].Particle
    Bm.particle = Particle(xl, pxl, np.zeros_like(xl), np.zeros_like(xl), zl,
                           np.zeros_like(xl), 0.0, 2e9, 510998.95)
    Bm._init_energy = 2e9
    Bm.position = 0.3
    Bm.update_status()
    obj.beam = Bm
    obj.get_CSR_mesh()
    obj.calculate_2D_CSR()
    chirp_runs = []
    for slope_override in (2.5, -2.5):
        Bm._slope = np.array([slope_override, Bm._slope[1]])
        res = [obj.get_CSR_wake(0.3 + obj.CSR_zmesh[i], obj.CSR_xmesh[i]) for i in (0, 5, 11)]
        chirp_runs.append(np.array(res))
    Bm.update_status()
    dbg = obj.get_CSR_wake(0.3 + obj.CSR_zmesh[6], obj.CSR_xmesh[6], debug=True)
    np.savez(os.path.join(HERE, "ref_wake.npz"),
             beam_x=xl, beam_z=zl, beam_px=pxl, position=0.3,
             sigma_x=Bm._sigma_x, sigma_z=Bm._sigma_z, slope=Bm._slope, mean_x=Bm._mean_x,
             mean_z=Bm._mean_z, formation_length=0.05, n_fl=1.0, Nx=12, Nz=16,
             csr_scaling=obj.CSR_scaling, mesh_xbins=3, mesh_zbins=4, mesh_xlim=3, mesh_zlim=3,
             xmesh=obj.CSR_xmesh, zmesh=obj.CSR_zmesh, zrange=obj.CSR_zrange,
             xrange=obj.CSR_xrange_transformed, dE_dct=obj.dE_dct, x_kick=obj.x_kick,
             chirp_pos=chirp_runs[0], chirp_neg=chirp_runs[1], chirp_rows=np.array([0, 5, 11]),
             dbg_xp_w=dbg[0], dbg_xp_n=dbg[1], dbg_sp1=dbg[2], dbg_sp2=dbg[3], dbg_sp3=dbg[4],
             dbg_Iz1=dbg[5], dbg_Ix1=dbg[6], dbg_Iz2=dbg[7], dbg_Ix2=dbg[8],
             dbg_Iz3=dbg[9], dbg_Ix3=dbg[10])

    # a16 apply_wakes (beams.py:108-131)
    pz0 = rng.normal(0, 1e-4, xl.shape[0])
    px0 = np.array(pxl)
    Bm.particle = Particle(xl, px0, np.zeros_like(xl), np.zeros_like(xl), zl, pz0, 0.0, 2e9,
                           510998.95)
    Bm.apply_wakes(obj.dE_dct, obj.x_kick, obj.CSR_xrange_transformed, obj.CSR_zrange,
                   0.05, 1)
    np.savez(os.path.join(HERE, "ref_apply_wakes.npz"), x=xl, z=zl, px0=px0, pz0=pz0,
             dE_dct=obj.dE_dct, x_kick=obj.x_kick, xrange=obj.CSR_xrange_transformed,
             zrange=obj.CSR_zrange, step_size=0.05, energy=2e9, px1=Bm.particle.px,
             pz1=Bm.particle.pz)
    print("golden vectors written to", HERE)


# (L, angle, k1, roll, E1, E2, hgap): drift, quadrupoles, sector and
# combined-function bends, pole faces, the fringe term (hgap, FINT = 1/2)
# and rolls
R_GEN6_CASES = [
    (0.3, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0),
    (0.3, 0.0, 5.0, 0.0, 0.0, 0.0, 0.0),
    (0.3, 0.0, -5.0, 0.0, 0.0, 0.0, 0.0),
    (0.3, 0.05, 0.0, 0.0, 0.0, 0.0, 0.0),
    (0.3, 0.05, 0.0, 0.0, 0.05, 0.0, 0.0),
    (0.3, -0.05, 0.0, 0.0, 0.0, -0.05, 0.0),
    (0.3, 0.05, 0.0, 0.0, 0.02, -0.03, 0.017),
    (0.3, -0.08, 0.0, 0.0, 0.0, 0.0, 0.02),
    (0.5, 0.2, 0.0, 0.0, 0.1, 0.1, 0.025),
    (0.3, 0.05, 2.0, 0.0, 0.0, 0.0, 0.0),
    (0.3, 0.0, 5.0, np.pi / 4, 0.0, 0.0, 0.0),
    (0.0, 0.0, 0.0, 0.3, 0.0, 0.0, 0.0),
]


def transfer_matrices(ref_dir):
    """r_gen6 (r_gen6.py:4-122) on R_GEN6_CASES -> ref_r_gen6.npz."""
    R = _load_reference(ref_dir)
    rg = R["r_gen6"].r_gen6
    cases = np.array(R_GEN6_CASES, dtype=np.float64)
    mats = np.stack([rg(L, angle=a, k1=k, roll=r, E1=e1, E2=e2, hgap=hg)
                     for (L, a, k, r, e1, e2, hg) in R_GEN6_CASES])
    np.savez(os.path.join(HERE, "ref_r_gen6.npz"), cases=cases, R=mats)
    print("r_gen6 vectors written to", HERE)


# (nx, nz, window_size, order): grids with a smooth part and noise; the last
# two at the smallest size the padding allows (window_size // 2 + 1)
SGOLAY_CASES = [(40, 50, 3, 1), (40, 50, 5, 2), (33, 17, 7, 3), (24, 29, 9, 2),
                (4, 4, 7, 2), (6, 5, 9, 3)]
SGOLAY_MODES = (None, "col", "row", "both")


def sgolay_vectors(ref_dir):
    """SGolay_filter.sgolay2d (SGolay_filter.py:3-81) on SGOLAY_CASES, every
    derivative mode -> ref_sgolay2d.npz (keys z{i}, {mode}{i}; 'both' as
    bothcol{i}, bothrow{i})."""
    spec = importlib.util.spec_from_file_location("SGolay_filter",
                                                  os.path.join(ref_dir, "SGolay_filter.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    rng = np.random.default_rng(2718)
    out = {"cases": np.array(SGOLAY_CASES, dtype=np.int64)}
    for i, (nx, nz, w, order) in enumerate(SGOLAY_CASES):
        xg, zg = np.meshgrid(np.linspace(-1, 1, nx), np.linspace(-1, 1, nz), indexing="ij")
        z = np.exp(-(xg ** 2 + 2 * zg ** 2)) + 0.3 * xg * zg + 0.05 * rng.normal(size=(nx, nz))
        out[f"z{i}"] = z
        for mode in SGOLAY_MODES:
            r = mod.sgolay2d(z, w, order, derivative=mode)
            if mode == "both":
                out[f"bothcol{i}"], out[f"bothrow{i}"] = r
            else:
                out[f"{mode or 'none'}{i}"] = r
    np.savez(os.path.join(HERE, "ref_sgolay2d.npz"), **out)
    print("sgolay2d vectors written to", HERE)


SHORT_LATTICE_YAML = """step_size: 0.05
B0:
  type: dipole
  L: 0.03
  angle: 0.01
  E1: 0.01
  E2: 0.0
  FRINGE_AT: both_ends
  nsep: 1
D1:
  type: drift
  L: 0.2
  nsep: 3
Q1:
  type: quad
  L: 0.02
  K1: 2.0
  nsep: 1
B1:
  type: dipole
  L: 0.25
  angle: -0.03
  E1: -0.01
  E2: -0.02
  nsep: 2
S1:
  type: sextupole
  L: 0.04
  K2: 1.0
  nsep: 1
D2:
  type: drift
  L: 0.31
  nsep: 1
"""


def _c2_chicane_yaml():
    """The 4-bend chicane of the bench's chicane workload: 0.3 m bends at
    +, -, -, + 0.05 rad, drifts 2/1/2 m, 400 steps."""
    rows = [("B1", "dipole", 0.3, 0.05, 0.0, 0.05), ("D1", "drift", 2.0),
            ("B2", "dipole", 0.3, -0.05, -0.05, 0.0), ("D2", "drift", 1.0),
            ("B3", "dipole", 0.3, -0.05, 0.0, -0.05), ("D3", "drift", 2.0),
            ("B4", "dipole", 0.3, 0.05, 0.05, 0.0)]
    out = [f"step_size: {6.2 / 400!r}"]
    for r in rows:
        out += [f"{r[0]}:", f"  type: {r[1]}", f"  L: {r[2]!r}", "  nsep: 1"]
        if r[1] == "dipole":
            out += [f"  angle: {r[3]!r}", f"  E1: {r[4]!r}", f"  E2: {r[5]!r}"]
    return "\n".join(out) + "\n"


# (tag, lattice source, n_formation_length): a leading drift and nsep 2; the
# reference's FODO example (a first element and two consecutive elements
# shorter than one step, no bend: the formation length accumulates every
# element); the reference's chicane example (leading drift, nsep 5,
# n_formation_length 1); the chicane workload (400 steps); a lattice with a
# bend shorter than one step (tracked in one piece with its FRINGE_AT) and a
# short quadrupole and sextupole; the reference's dipole example (a 1 rad
# bend with optional fringe keys).
RUN_LOOP_CASES = [("drift_first", LATTICE_YAML, 1.0),
                  ("fodo", "example/input/fodo_lattice.yaml", 4),
                  ("chicane_example", "example/input/chicane_lattice.yaml", 1),
                  ("chicane_c2", None, 4),
                  ("short", SHORT_LATTICE_YAML, 2),
                  ("dipole_example", "example/input/dipole_lattice.yaml", 4)]


def run_loop_vectors(ref_dir):
    """The reference's own run loop (CSR.py:202-358) on RUN_LOOP_CASES with
    bmadx's element classes replaced by records and track_element by the
    stand-in map of run_loop_tracker.py; the wake itself (calculate_2D_CSR)
    replaced by a recorder and apply_CSR off, so only the loop's decisions
    are captured.  Per lattice: every tracked piece (kind, L, G, E1, E2,
    FRINGE_AT, K1, K2) with sigma_z after it; per step: position, the
    formation length passed to append_interpolant, the history window's
    first time and length after it, whether CSR ran, pieces so far.
    -> ref_run_loop.npz"""
    import contextlib
    import io
    import json
    sys.path.insert(0, HERE)
    from run_loop_tracker import make_beam, piece_row, stand_in_map
    R = _load_reference(ref_dir)
    CSR, beams, ypar = R["CSR"], R["beams"], R["yaml_parser"]
    Particle = sys.modules["bmadx"].Particle
    rec = {}

    def element(kind):
        def make(**kw):
            d = dict(kind=kind, **kw)
            d.pop("P0C", None)
            return d
        return make

    def track_element(p, ele):
        new = stand_in_map((p.x, p.px, p.y, p.py, p.z, p.pz), ele)
        rec["pieces"].append(piece_row(ele))
        rec["sigz"].append(np.std(new[4]))
        return Particle(*new, p.s, p.p0c, p.mc2)

    for kind in ("Drift", "SBend", "Quadrupole", "Sextupole"):
        setattr(CSR, kind, element(kind))
    beams.track_element = track_element

    tmp = tempfile.mkdtemp()
    beam = make_beam()
    beam_file = os.path.join(tmp, "beam.txt")
    np.savetxt(beam_file, beam.T, fmt="%.17g")
    beam = np.loadtxt(beam_file).T.copy()
    out = {"beam": beam}
    for tag, src, nfl in RUN_LOOP_CASES:
        if src is None:
            src = _c2_chicane_yaml()
        lat_file = os.path.join(tmp, f"{tag}.yaml")
        if src.endswith(".yaml"):
            with open(os.path.join(ref_dir, src)) as f:
                src = f.read()
        with open(lat_file, "w") as f:
            f.write(src)
        lat_cfg = ypar.parse_yaml(lat_file)
        cfg_file = os.path.join(tmp, f"{tag}_cfg.yaml")
        with open(cfg_file, "w") as f:
            f.write(f"""input_beam:
  style: from_file
  beamfile: {beam_file}
  charge: 1.0e-9
  energy: 1.0e+9
input_lattice:
  lattice_input_file: {lat_file}
particle_deposition:
  xbins: 32
  zbins: 32
  xlim: 5
  zlim: 5
  filter_order: 2
  filter_window: 5
  velocity_threhold: 1000
  upper_limit: 48
CSR_integration:
  n_formation_length: {nfl}
  zbins: 8
  xbins: 8
CSR_computation:
  compute_CSR: 1
  apply_CSR: 0
  transverse_on: 1
  xbins: 4
  zbins: 4
  write_wakes: False
  workdir: {tmp}
""")
        rec.update(pieces=[], sigz=[], steps=[], fl=None, csr=False)
        obj = CSR.CSR2D.__new__(CSR.CSR2D)
        obj.timestamp = "fixture"

        def append_interpolant(formation_length, n_formation_length, _orig=None):
            rec["fl"] = formation_length
            return _orig(formation_length=formation_length,
                         n_formation_length=n_formation_length)

        def calculate_2D_CSR(self=obj):
            rec["csr"] = True
            self.dE_dct = np.zeros((self.CSR_params.xbins, self.CSR_params.zbins))
            self.x_kick = np.zeros_like(self.dE_dct)

        def update_statistics(step, self=obj):
            tr = self.DF_tracker
            rec["steps"].append([step, float(self.beam.position),
                                 np.nan if rec["fl"] is None else rec["fl"],
                                 tr.time_interp[0], len(tr.time_interp), float(rec["csr"]),
                                 len(rec["pieces"])])
            rec["fl"] = None
            rec["csr"] = False

        obj.update_statistics = update_statistics
        obj.calculate_2D_CSR = calculate_2D_CSR
        obj.dump_beam = lambda label: None
        obj.write_statistics = lambda: None
        with contextlib.redirect_stdout(io.StringIO()):
            obj.parse_input(cfg_file)
            tr = obj.DF_tracker
            tr.append_interpolant = (
                lambda formation_length, n_formation_length, _o=tr.append_interpolant:
                append_interpolant(formation_length, n_formation_length, _o))
            obj.formation_length = None
            obj.initialization()
            obj.prefix = "fixture"
            obj.parallel = False
            obj.run()
        lat = obj.lattice
        out[f"{tag}_lattice_json"] = np.array(json.dumps(
            [["step_size", lat_cfg["step_size"]]] +
            [[k, dict(v)] for k, v in list(lat_cfg.items())[1:]]))
        out[f"{tag}_nfl"] = np.float64(nfl)
        out[f"{tag}_pieces"] = np.array(rec["pieces"], dtype=np.float64)
        out[f"{tag}_sigz"] = np.array(rec["sigz"], dtype=np.float64)
        out[f"{tag}_steps"] = np.array(rec["steps"], dtype=np.float64)
        out[f"{tag}_positions_record"] = lat._positions_record
        out[f"{tag}_steps_per_element"] = lat.steps_per_element
        print(tag, len(rec["pieces"]), "pieces", len(rec["steps"]), "step records")
    np.savez_compressed(os.path.join(HERE, "ref_run_loop.npz"), **out)
    print("run-loop vectors written to", HERE)


if __name__ == "__main__":
    # make_reference_vectors.py [REF_DIR] [r_gen6 | sgolay2d | run_loop]:
    # every fixture, or only the transfer matrices / the 2D filter's / the
    # run loop's
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/pyDFCSR_2D"
    if sys.argv[2:] == ["r_gen6"]:
        transfer_matrices(ref)
    elif sys.argv[2:] == ["sgolay2d"]:
        sgolay_vectors(ref)
    elif sys.argv[2:] == ["run_loop"]:
        run_loop_vectors(ref)
    else:
        main(ref)
        transfer_matrices(ref)
        sgolay_vectors(ref)
        run_loop_vectors(ref)
