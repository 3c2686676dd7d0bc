"""CPU tests of the host-side mirror of the reference interface (no device)."""
import os
from collections import OrderedDict

import numpy as np
import pytest

from pydfcsr_amd.lattice import Lattice, reference_tables
from pydfcsr_amd.params import CSR_params, Deposition_params, Integration_params


def _golden_lattice_cfg():
    cfg = OrderedDict(step_size=0.05)
    cfg["D0"] = dict(type="drift", L=0.1, nsep=1)
    cfg["B1"] = dict(type="dipole", L=0.3, angle=0.05, nsep=1)
    cfg["D1"] = dict(type="drift", L=0.4, nsep=2)
    cfg["B2"] = dict(type="dipole", L=0.3, angle=-0.05, nsep=1)
    return cfg


def test_lattice_tables_bitexact_with_reference(golden_dir):
    d = np.load(os.path.join(golden_dir, "ref_lattice.npz"))
    L = Lattice({"lattice_input_file": "<inline>", "_config": _golden_lattice_cfg()})
    for k in ("s", "rho", "distance", "nsep", "coords", "n_vec", "tau_vec"):
        assert np.array_equal(getattr(L, k), d[k]), k
    assert np.array_equal(L._positions_record, d["positions_record"])
    assert np.array_equal(L.steps_per_element, d["steps_per_element"])
    assert L.delta_x == d["delta_x"]


def test_lattice_from_yaml_file(tmp_path):
    p = tmp_path / "lat.yaml"
    p.write_text("step_size: 0.1\nD0:\n  type: drift\n  L: 0.1\n  nsep: 1\n"
                 "B1:\n  type: dipole\n  L: 1.0\n  angle: 0.1\n  nsep: 1\n")
    L = Lattice({"lattice_input_file": str(p)})
    assert L.element_names == ["D0", "B1"] and L.Nelement == 2
    assert np.isclose(L.lattice_length, 1.1)
    with pytest.raises(AssertionError):
        Lattice({"wrong_key": str(p)})


def test_params_defaults_and_unknown_keys():
    ip = Integration_params()
    assert (ip.n_formation_length, ip.zbins, ip.xbins) == (4, 200, 200)
    cp = CSR_params({"xbins": 10})
    assert (cp.xbins, cp.zbins, cp.xlim, cp.zlim, cp.transverse_on) == (10, 30, 5, 5, 1)
    dp = Deposition_params({"velocity_threhold": 1000, "upper_limit": 2000})
    assert dp.velocity_threhold == 1000 and dp.upper_limit == 2000 and dp.xbins == 100
    with pytest.raises(TypeError):
        Integration_params({"nbins": 3})
    with pytest.raises(TypeError):
        Deposition_params({"velocity_threshold": 3})   # the reference's key is misspelt


def test_reference_tables_closed_form_arc():
    els = [dict(type="dipole", L=0.3, angle=0.05, nsep=1)]
    s, rho, dist, nsep, coords, nv, tau = reference_tables(els, 301)
    R = 0.3 / 0.05
    th = s / R
    assert np.allclose(coords[:, 0], R * np.sin(th), atol=1e-12)
    assert np.allclose(coords[:, 1], R * (1 - np.cos(th)), atol=1e-12)


@pytest.mark.parametrize("slope0", [0.3, -0.8, 2.5, -2.5, 6.0])
def test_near_planes_covers_the_near_regions_reach(slope0):
    """CSR2D.near_planes (the depth of the wakes' recent-plane copy) covers
    every near-region lookup: for mesh observers, the oldest retarded plane
    their regions 1.. reach (get_CSR_wake's bounds, oracle.wake_regions, and
    |r - r'| from the lattice tables) lies within the copied planes, unless
    the copy is capped (then older lookups read the ring, as designed)."""
    import math
    from types import SimpleNamespace
    from oracle import cpu_ref as O
    from pydfcsr_amd.csr import CSR2D
    elements = [dict(type="drift", L=0.3, nsep=1), dict(type="dipole", L=0.5, angle=0.05, nsep=1),
                dict(type="drift", L=1.0, nsep=1)]
    lat = O.LatticeTables(elements)
    sz, sx, mean_x, slope1 = 60e-6, 40e-6, 3e-6, -1e-6
    t, dt, nt, cap = 0.7, 5e-3, 140, 16
    zr = (-5 * sz, 5 * sz, 9)
    xr = (-5 * sx, 5 * sx, 9)
    fake = SimpleNamespace(
        DF_tracker=SimpleNamespace(nt=nt, delta_x=dt, recent_planes=cap),
        beam=SimpleNamespace(_sigma_z=sz, _sigma_x=sx, _slope=np.array([slope0, slope1]),
                             _mean_x=mean_x),
        mesh_z_axis=zr, mesh_x_axis=xr, _mesh_slope=np.array([slope0, slope1]))
    K = CSR2D.near_planes(fake)
    assert 1 <= K <= min(cap, nt)
    min_t = t - (nt - 1) * dt

    def col(v, c):
        return O.interpolate1D(np.atleast_1d(v), lat.coords[:, c] if c < 2 else lat.n_vec[:, c - 2],
                               lat.min_x, lat.delta_x)

    need = 0
    for zm in np.linspace(*zr):
        for xt in np.linspace(*xr):
            s, x = t + zm, xt + slope0 * zm + slope1
            regs = O.wake_regions(s, x, t, sz, sx, slope0, mean_x, 1.0, 4, 24, 24)
            X0, Y0, nx, ny = (col(s, c)[0] for c in range(4))
            for xp, sp in regs[1:]:
                sp = sp[(sp >= lat.min_x) & (sp <= lat.max_x)]
                if sp.size == 0:
                    continue
                XP, SP = np.meshgrid(xp, sp, indexing="ij")
                rx = X0 - col(SP.ravel(), 0) + x * nx - XP.ravel() * col(SP.ravel(), 2)
                ry = Y0 - col(SP.ravel(), 1) + x * ny - XP.ravel() * col(SP.ravel(), 3)
                R = np.sqrt(rx * rx + ry * ry).max()
                i0 = math.trunc((t - R - min_t) / dt)
                need = max(need, nt - max(i0, 0))
    assert K >= min(need, cap), (K, need)
