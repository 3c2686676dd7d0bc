"""The driver's run loop (CSR2D.iter_phases, CSR.py:202-358) against the
reference's own, on the six lattices of tests/golden/ref_run_loop.npz.

The fixture was written by running the reference's CSR2D.run with bmadx
replaced by the deterministic stand-in map of
tests/golden/run_loop_tracker.py; here the driver runs the same lattices
from the same beam with its tracking replaced by the same map (applied on
the host to the device particles), and its whole CSR step on the device.
Compared, in order: every piece the driver tracks (element kind, length bit
for bit, bend strength, faces and fringe where FRINGE_AT puts them,
quadrupole and sextupole strengths), sigma_z after each piece (the device
moments against numpy's std: 1e-13), and per step the position (bit for
bit), the formation length (1e-13: it is computed from sigma_z), the first
time and length of the history window (bit for bit) and whether the step
computed a wake.
"""
import json
import os
from collections import OrderedDict

import numpy as np
import pytest

from tests.golden.run_loop_tracker import FRINGE, KINDS, stand_in_map

pytestmark = pytest.mark.gpu

DEV = "cuda"
TAGS = ("drift_first", "fodo", "chicane_example", "chicane_c2", "short", "dipole_example")


@pytest.fixture(scope="module")
def ref(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "ref_run_loop.npz")))


def _piece_of(e):
    from pydfcsr_amd import _lib
    kind = {_lib.DFCSR_ELE_DRIFT: "Drift", _lib.DFCSR_ELE_SBEND: "SBend",
            _lib.DFCSR_ELE_QUAD: "Quadrupole", _lib.DFCSR_ELE_SEXTUPOLE: "Sextupole"}[e.type]
    return dict(kind=kind, L=e.L, G=e.g, K1=e.k1, K2=e.k2)


@pytest.mark.parametrize("tag", TAGS)
def test_driver_loop_matches_reference(ref, tag, tmp_path):
    import torch
    from pydfcsr_amd.csr import CSR2D
    lat = OrderedDict((k, v) for k, v in json.loads(str(ref[f"{tag}_lattice_json"])))
    beam_file = tmp_path / "beam.txt"
    np.savetxt(beam_file, ref["beam"].T, fmt="%.17g")
    cfg = {
        "input_beam": dict(style="from_file", beamfile=str(beam_file), charge=1e-9,
                           energy=1e9),
        "input_lattice": dict(lattice_input_file=f"<{tag}>", _config=lat),
        "particle_deposition": dict(xbins=32, zbins=32, xlim=5, zlim=5, filter_order=2,
                                    filter_window=5, velocity_threhold=1000, upper_limit=48),
        "CSR_integration": dict(n_formation_length=float(ref[f"{tag}_nfl"]), zbins=8, xbins=8),
        "CSR_computation": dict(compute_CSR=1, apply_CSR=0, transverse_on=1, xbins=4, zbins=4,
                                write_wakes=False),
    }
    sim = CSR2D(config=cfg, device=DEV, precision="fp64")
    assert np.array_equal(sim.lattice.steps_record, ref[f"{tag}_positions_record"])
    assert np.array_equal(sim.lattice.steps_per_element, ref[f"{tag}_steps_per_element"])
    elements, sigz = [], []

    def track(ele_name, DL, entrance=False, exit=False, update_step=True):
        e = sim.get_bmadx_element(ele_name, DL, entrance=entrance, exit=exit)
        elements.append(e)
        b = sim.beam
        new = stand_in_map(b.coords.cpu().numpy(), _piece_of(e))
        b.coords = torch.from_numpy(np.ascontiguousarray(new)).to(b.device)
        b.position += DL
        if update_step:
            b.step += 1
        b.update_status()
        sigz.append(b._sigma_z)

    sim._track = track
    flag = {"csr": False}
    orig_wake = sim.calculate_2D_CSR

    def wake():
        flag["csr"] = True
        orig_wake()

    sim.calculate_2D_CSR = wake
    tr = sim.DF_tracker
    rows = [[0, 0.0, np.inf, tr.log[0].t, tr.nt, 0.0, 0]]
    for step in sim.iter_steps():
        rows.append([step, sim.beam.position, sim.formation_length, tr.log[0].t, tr.nt,
                     float(flag["csr"]), len(elements)])
        flag["csr"] = False
    torch.cuda.synchronize()

    P = ref[f"{tag}_pieces"]
    assert len(elements) == P.shape[0], (len(elements), P.shape[0])
    for k, (e, r) in enumerate(zip(elements, P)):
        kind, L, G, E1, E2, fr, K1, K2 = r
        p = _piece_of(e)
        fringe = FRINGE[int(fr)] if fr >= 0 else None
        want_e1 = E1 if fringe in ("entrance_end", "both_ends") else 0.0
        want_e2 = E2 if fringe in ("exit_end", "both_ends") else 0.0
        assert p["kind"] == KINDS[int(kind)] and p["L"] == L, (k, p, r)
        assert (p["G"], p["K1"], p["K2"]) == (G, K1, K2), (k, p, r)
        assert (e.e1, e.e2) == (want_e1, want_e2), (k, e.e1, e.e2, r)
    assert np.allclose(sigz, ref[f"{tag}_sigz"], rtol=1e-13, atol=0)
    rows = np.array(rows, dtype=np.float64)
    S = ref[f"{tag}_steps"]
    assert rows.shape == S.shape, (rows.shape, S.shape)
    exact = [0, 1, 3, 4, 5, 6]
    assert np.array_equal(rows[:, exact], S[:, exact]), \
        np.argwhere(rows[:, exact] != S[:, exact])[:5]
    assert np.allclose(rows[:, 2], S[:, 2], rtol=1e-13, atol=0)
