"""Extract the single-particle Bmad-X tracking results that the reference's
own tests hold, as a fixture.

Run ONLY in the build container, where the reference source tree is mounted:

    python tests/golden/extract_bmadx_vectors.py /root/reference/pyDFCSR_2D

pyDFCSR tracks with bmadx (Bmad-X, ``git+https://github.com/bmad-sim/Bmad-X/``,
environment.yml:24, no version pinned), which is absent here.  Its notebook
``test/test_BmadX_tracking.ipynb`` tracked one particle with all six
coordinates 1e-3 (p0c = 4e7 eV, mc2 = M_ELECTRON) through a Drift (cell 25),
an SBend (cell 28) and a Quadrupole (cell 31) and kept the printed outputs.
This script reads the notebook as JSON (it runs nothing of it), parses the
element arguments from the cell sources and the printed Particle tuples from
the outputs, and writes them to ``ref_bmadx_tracking.json``.
"""
from __future__ import annotations

import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
NUM = r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?"


def _particle(text):
    """Particle(x=..., px=..., ..., p0c=..., mc2=...) -> dict of floats."""
    out = {}
    for key in ("x", "px", "y", "py", "z", "pz", "s", "p0c", "mc2"):
        m = re.search(rf"\b{key}=(?:array\()?({NUM})", text)
        out[key] = float(m.group(1))
    return out


def _element(src, ctor):
    m = re.search(rf"{ctor}\(([^)]*)\)", src)
    args = {}
    for part in m.group(1).split(","):
        k, v = (t.strip() for t in part.split("="))
        args[k] = v
    return args


def main(ref_root):
    nb = json.load(open(os.path.join(ref_root, "test", "test_BmadX_tracking.ipynb")))
    cells = nb["cells"]

    def out_text(i):
        for o in cells[i].get("outputs", []):
            if "data" in o and "text/plain" in o["data"]:
                return "".join(o["data"]["text/plain"])
        raise ValueError(f"cell {i} has no printed output")

    src = lambda i: "".join(cells[i]["source"])
    p_in = _particle(out_text(21))
    cases = []
    drift = _element(src(23), "Drift")
    cases.append(dict(element="Drift", L=float(drift["L"]), out=_particle(out_text(25)),
                      cells=[23, 25]))
    bend = _element(src(28), "SBend")
    cases.append(dict(element="SBend", L=float(bend["L"]), G=float(bend["G"]),
                      E1=float(bend["E1"]), E2=float(bend["E2"]),
                      P0C=p_in["p0c"] if bend["P0C"] == "p0c" else float(bend["P0C"]),
                      out=_particle(out_text(28)), cells=[28]))
    quad = _element(src(31), "Quadrupole")
    cases.append(dict(element="Quadrupole", L=float(quad["L"]), K1=float(quad["K1"]),
                      out=_particle(out_text(31)), cells=[31]))
    doc = {"source": "pyDFCSR_2D/test/test_BmadX_tracking.ipynb (printed outputs)",
           "particle_in": p_in, "cases": cases}
    path = os.path.join(HERE, "ref_bmadx_tracking.json")
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/pyDFCSR_2D")
