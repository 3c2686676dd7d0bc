"""The oracle's run loop (CSR.py:202-358) against the reference's own.

tests/golden/ref_run_loop.npz holds what the reference's CSR2D.run did on
six lattices with bmadx replaced by the deterministic stand-in map of
tests/golden/run_loop_tracker.py (make_reference_vectors.py run_loop): every
tracked piece with the beam's sigma_z after it, and per step the position,
the formation length handed to append_interpolant, the history window's
first time and length after it and whether the wake was computed.  The
oracle's run_loop, driven by the same map from the same beam, must
reproduce all of it bit for bit, and its history (DFTracker) the window.
Planted mistakes in the oracle (the formation length's exponent, the first
element's step count, a bend's faces swapped, the leading-drift
accumulation) must each fail on at least one lattice.
"""
import json
from collections import OrderedDict

import numpy as np
import pytest

from tests.golden.run_loop_tracker import piece_row, stand_in_map

TAGS = ("drift_first", "fodo", "chicane_example", "chicane_c2", "short", "dipole_example")


@pytest.fixture(scope="module")
def ref(golden_dir):
    import os
    return dict(np.load(os.path.join(golden_dir, "ref_run_loop.npz")))


def lattice_config(ref, tag):
    rows = json.loads(str(ref[f"{tag}_lattice_json"]))
    return OrderedDict((k, v) for k, v in rows)


def replay(O, ref, tag, with_history=True):
    """Runs the oracle's loop on the fixture's beam; returns (pieces, sigma_z
    after each, step rows) in the fixture's layout."""
    beam = ref["beam"].copy()
    state = {"c": beam}
    pieces, sigz = [], []

    def track(piece, DL, update_step):
        assert piece["L"] == DL
        state["c"] = stand_in_map(state["c"], piece)
        pieces.append(piece_row(piece))
        sigz.append(np.std(state["c"][4]))

    def sigma_z():
        return np.std(state["c"][4])

    nfl = float(ref[f"{tag}_nfl"])
    tr = O.DFTracker(xbins=32, zbins=32, xlim=5, zlim=5, filter_order=2, filter_window=5,
                     velocity_threhold=1000, upper_limit=48)
    c = state["c"]
    rows = []
    if with_history:
        tr.get_DF(c[0], c[4], c[1], 0)
        tr.append_DF()
        tr.append_interpolant(float("inf"), nfl)
    rows.append([0, 0.0, np.inf, tr.time_interp[0] if with_history else 0.0,
                 len(tr.time_interp) if with_history else 1, 0.0, 0])
    for r in O.run_loop(lattice_config(ref, tag), track, sigma_z):
        c = state["c"]
        if with_history:
            tr.get_DF(c[0], c[4], c[1], r["position"])
            tr.append_DF()
            tr.append_interpolant(r["formation_length"], nfl)
        rows.append([r["step"], r["position"], r["formation_length"],
                     tr.time_interp[0] if with_history else np.nan,
                     len(tr.time_interp) if with_history else np.nan, float(r["csr"]),
                     len(pieces)])
    return np.array(pieces), np.array(sigz), np.array(rows)


@pytest.mark.parametrize("tag", TAGS)
def test_run_loop_matches_reference(ref, tag):
    from oracle import cpu_ref as O
    lat = lattice_config(ref, tag)
    eles = list(lat.values())[1:]
    dist = np.cumsum([e["L"] for e in eles])
    # the reference accumulates distance[k] = L + distance[k-1], as cumsum
    positions, spe = O.lattice_steps(dist, lat["step_size"])
    assert np.array_equal(positions, ref[f"{tag}_positions_record"])
    assert np.array_equal(spe, ref[f"{tag}_steps_per_element"])
    pieces, sigz, rows = replay(O, ref, tag)
    assert pieces.shape == ref[f"{tag}_pieces"].shape, (pieces.shape, ref[f"{tag}_pieces"].shape)
    assert np.array_equal(pieces, ref[f"{tag}_pieces"])
    assert np.array_equal(sigz, ref[f"{tag}_sigz"])
    assert np.array_equal(rows, ref[f"{tag}_steps"])


def test_fixture_covers_the_loop_branches(ref):
    """The lattices exercise every branch of the loop: elements holding no
    step position (first element, consecutive ones, a one-piece bend with
    its FRINGE_AT), exit pieces, nsep 2 and 5 (steps without a wake), the
    formation length accumulating before the first bend and following the
    last bend after it, and history windows that drop slices."""
    spe = {t: ref[f"{t}_steps_per_element"] for t in TAGS}
    assert spe["fodo"][0] == 0 and (spe["fodo"][2] == 0)
    assert (spe["short"] == 0).sum() == 2
    one_piece_bend = ref["short_pieces"][0]
    assert one_piece_bend[0] == 1 and one_piece_bend[5] == 3    # SBend, both_ends
    fr = np.concatenate([ref[f"{t}_pieces"][:, 5] for t in TAGS])
    assert {0.0, 1.0, 2.0, 3.0} <= set(fr.tolist())
    csr = np.concatenate([ref[f"{t}_steps"][1:, 5] for t in TAGS])
    assert (csr == 0).any() and (csr == 1).any()
    st = ref["drift_first_steps"]
    assert st[1, 2] == 0.1 and st[-1, 3] > 0      # leading drift, window moved
    st = ref["chicane_c2_steps"]
    # set once per element from that element's entrance sigma_z: 7 values
    assert len(np.unique(st[1:, 2])) == 7
    assert st[-1, 3] > 0


def _fails(O, ref, tags=TAGS):
    for tag in tags:
        try:
            pieces, sigz, rows = replay(O, ref, tag, with_history=False)
        except Exception:
            return True
        same = (pieces.shape == ref[f"{tag}_pieces"].shape
                and np.array_equal(pieces, ref[f"{tag}_pieces"])
                and np.array_equal(rows[:, [0, 1, 2, 5, 6]],
                                   ref[f"{tag}_steps"][:, [0, 1, 2, 5, 6]]))
        if not same:
            return True
    return False


def test_planted_mistakes_fail(ref, monkeypatch):
    from oracle import cpu_ref as O
    # unplanted, the loop alone (no history) agrees
    assert not _fails(O, ref)
    with monkeypatch.context() as m:
        m.setattr(O, "formation_length", lambda R, s: (24 * R * R * s) ** 0.5)
        assert _fails(O, ref)
    with monkeypatch.context() as m:
        orig = O.lattice_steps

        def steps_no_first_offset(distance, step_size):
            p, s = orig(distance, step_size)
            s = s.copy()
            s[0] += 1
            return p, s
        m.setattr(O, "lattice_steps", steps_no_first_offset)
        assert _fails(O, ref)
    with monkeypatch.context() as m:
        orig_piece = O.element_piece

        def swapped(ele, DL, entrance=False, exit=False):
            p = orig_piece(ele, DL, entrance, exit)
            if p["kind"] == "SBend":
                p["E1"], p["E2"] = p["E2"], p["E1"]
            return p
        m.setattr(O, "element_piece", swapped)
        assert _fails(O, ref, tags=("chicane_example", "chicane_c2", "short"))
    with monkeypatch.context() as m:
        # the leading-drift accumulation dropped: no formation length before
        # the first bend
        orig_loop = O.run_loop

        def no_drift_accumulation(cfg, track, sigma_z, stop_time=None):
            bent = False
            for r in orig_loop(cfg, track, sigma_z, stop_time):
                bent = bent or cfg[r["element"]]["type"] == "dipole"
                yield r if bent else dict(r, formation_length=0.0)
        m.setattr(O, "run_loop", no_drift_accumulation)
        assert _fails(O, ref, tags=("drift_first",))
        assert _fails(O, ref, tags=("fodo",))
