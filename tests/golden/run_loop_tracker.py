"""A deterministic stand-in for bmadx's track_element, used to pin the run loop.

The run loop of the reference (CSR.py:202-358) decides which element pieces
are tracked (the part of a step left in the previous element, a whole
element shorter than one step, the part of the first step in a new element)
and computes the formation length from the beam's sigma_z at each element
entrance (CSR.py:246-269).  To record that logic from the reference itself,
make_reference_vectors.py (``run_loop``) runs the reference's CSR2D.run with
bmadx's element classes replaced by plain records and track_element by
``stand_in_map`` below; the tests drive the oracle's run loop and the
driver's with the same map, so the beam (and so sigma_z) evolves the same
way in all three.  The map is not physics: a dipole piece shortens the bunch
(so sigma_z, and with it the formation length, changes from element to
element), the other pieces shear x and px.  Its arithmetic is plain float64
numpy, identical wherever it runs.

Pieces are described by dicts with the keys the reference passes to bmadx
(CSR.py:146-199): kind ('SBend', 'Drift', 'Quadrupole', 'Sextupole'), L, and
for a bend G, E1, E2, FRINGE_AT, for a quadrupole K1, for a sextupole K2,
plus whatever optional keys the lattice file holds for that element.
"""
from __future__ import annotations

import numpy as np

KINDS = ("Drift", "SBend", "Quadrupole", "Sextupole")
FRINGE = ("no_end", "entrance_end", "exit_end", "both_ends")


def stand_in_map(coords, piece):
    """New (6, n) coordinates after one piece; coords rows x px y py z pz."""
    x, px, y, py, z, pz = (np.asarray(c, dtype=np.float64) for c in coords)
    L = float(piece["L"])
    kind = piece["kind"]
    x1 = x + L * px
    if kind == "SBend":
        G = float(piece["G"])
        px1 = px + (L * G) * pz
        z1 = z * (1.0 - 2.0 * L * abs(G)) + (L * G) * x
    elif kind == "Quadrupole":
        px1 = px - (L * float(piece["K1"])) * x
        z1 = z
    elif kind == "Sextupole":
        px1 = px - (L * float(piece["K2"])) * (x * x)
        z1 = z
    else:
        px1 = px
        z1 = z + (1e-3 * L) * pz
    return np.stack([x1, px1, y, py, z1, pz])


def piece_row(piece):
    """Fixed numeric form of a piece: kind, L, G, E1, E2, FRINGE_AT, K1, K2."""
    return [float(KINDS.index(piece["kind"])), float(piece["L"]),
            float(piece.get("G", 0.0)), float(piece.get("E1", 0.0)),
            float(piece.get("E2", 0.0)),
            float(FRINGE.index(piece["FRINGE_AT"])) if "FRINGE_AT" in piece else -1.0,
            float(piece.get("K1", 0.0)), float(piece.get("K2", 0.0))]


def make_beam(n=3000, seed=2024):
    """The run-loop fixture's beam, (6, n): 1 GeV-scale chirped Gaussian."""
    rng = np.random.default_rng(seed)
    z = rng.normal(0.0, 100e-6, n)
    x = rng.normal(0.0, 30e-6, n)
    px = rng.normal(0.0, 1e-5, n)
    pz = -10.0 * z + rng.normal(0.0, 1e-4, n)
    return np.stack([x, px, np.zeros(n), np.zeros(n), z, pz])
