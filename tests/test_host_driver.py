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

