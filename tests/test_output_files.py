"""Output files (CSR.py:784-879): the wakes, statistics and particle files in
the reference's HDF5 layout.

The files written by ``pydfcsr_amd.h5io`` are dumped by h5py (any
interpreter that has it: ``DFCSR_H5PY_PYTHON``, else this one, else
/opt/conda/bin/python3.9) and compared object by object -- names, attribute
types and values, dataset dtypes, shapes and values, all exact -- with files
that h5py itself writes through the reference's own calls on the same data
(``tests/h5_layout.py``)."""
import subprocess

import numpy as np
import pytest

from pydfcsr_amd import h5io

from tests.h5_readback import PY, dump, run_helper as _run


def _writer_available():
    try:
        h5io.backend_name()
        return True
    except RuntimeError:
        return False


pytestmark = [pytest.mark.skipif(PY is None, reason="no interpreter with h5py to read files back"),
              pytest.mark.skipif(not _writer_available(), reason="no HDF5 writer backend")]


def assert_same_tree(mine, ref):
    assert sorted(mine) == sorted(ref)
    for name in ref:
        assert mine[name] == ref[name], name


def _wake_steps(rng, n=3, shape=(4, 6)):
    steps = []
    for k in range(n):
        steps.append(dict(step=k + 1, position=0.0155 * (k + 1) + 3.255, mean_gamma=1956.95,
                          beam_energy=1e9, element=("B1", "D1", "B2")[k], charge=1e-9,
                          xmesh=rng.normal(size=shape[0] * shape[1]),
                          zmesh=rng.normal(size=shape[0] * shape[1]),
                          dE_dct=rng.normal(size=shape), x_kick=rng.normal(size=shape)))
    steps[1]["dE_dct"][0, 0] = np.nan      # wakes of an empty mesh point are NaN
    return steps


def test_backend_reports_itself():
    name = h5io.backend_name()
    assert name == "h5py" or name.startswith("libhdf5 1.")


def test_wakes_file_matches_the_reference_layout(tmp_path):
    rng = np.random.default_rng(7)
    steps = _wake_steps(rng)
    mine = tmp_path / "run-wakes.h5"
    mine.write_bytes(b"stale")           # step 1 starts the file afresh
    for st in steps:
        h5io.write_wakes_file(str(mine), **st)
    npz = tmp_path / "wakes.npz"
    np.savez(npz, **{k: np.array([st[k] for st in steps]) for k in steps[0]})
    ref = tmp_path / "ref-wakes.h5"
    _run("wakes", str(npz), str(ref))
    a, b = dump(mine), dump(ref)
    assert_same_tree(a, b)
    assert a["step_2"]["attrs"]["element"]["value"] == {"str": "D1"}
    assert a["step_3/longitudinal"]["attrs"]["unit"]["type"]["variable"] is True
    assert a["step_1/transverse/xkicks"]["shape"] == [4, 6]
    # the postprocessor's reads (postprocessor.py:60-105)
    assert np.array_equal(np.array(a["step_3/longitudinal/dE_dct"]["data"]), steps[2]["dE_dct"])
    assert np.allclose(np.array(a["step_1/longitudinal/x_grids"]["data"]).ravel(),
                       steps[0]["xmesh"], rtol=0, atol=0)


def test_wakes_file_appends_after_step_one(tmp_path):
    rng = np.random.default_rng(3)
    steps = _wake_steps(rng, n=2)
    f = tmp_path / "w.h5"
    h5io.write_wakes_file(str(f), **steps[0])
    h5io.write_wakes_file(str(f), **steps[1])
    h5io.write_wakes_file(str(f), **dict(steps[0], step=1))   # a new run starts over
    assert sorted(k for k in dump(f) if "/" not in k and k != "/") == ["step_1"]


def _statistics(rng, n=9):
    keys = ("alpha", "beta", "gamma", "emit", "eta", "etap", "norm_emit")
    st = {"twiss": {f"{k}_{p}": rng.normal(size=n) for p in ("x", "y") for k in keys}}
    st["slope"] = rng.normal(size=(n, 2))
    for k in ("sigma_x", "sigma_z", "sigma_energy", "mean_x", "mean_z", "mean_energy"):
        st[k] = rng.normal(size=n)
    st["mean_energy"][-1] = np.nan
    return st


def test_statistics_file_matches_the_reference_layout(tmp_path):
    rng = np.random.default_rng(11)
    st = _statistics(rng)
    lat = dict(step_positions=np.arange(0, 0.09, 0.01), coords=rng.normal(size=(2000, 2)),
               n_vec=rng.normal(size=(2000, 2)), tau_vec=rng.normal(size=(2000, 2)))
    mine = tmp_path / "run-statistics.h5"
    h5io.write_statistics_file(str(mine), statistics=st, **lat)
    flat = dict(lat)
    flat.update({k: v for k, v in st.items() if k != "twiss"})
    flat.update({"twiss/" + k: v for k, v in st["twiss"].items()})
    npz = tmp_path / "st.npz"
    np.savez(npz, **flat)
    ref = tmp_path / "ref-statistics.h5"
    _run("statistics", str(npz), str(ref))
    a = dump(mine)
    assert_same_tree(a, dump(ref))
    assert len([k for k in a if k.startswith("twiss/")]) == 14
    assert a["slope"]["shape"] == [9, 2]


def test_particle_file_openpmd_layout(tmp_path):
    """openPMD-beamphysics layout (the openPMD 2.0 standard with the
    BeamPhysics extension; the writer library itself is not installed, so
    this checks the standard's records, not a byte-level match)."""
    rng = np.random.default_rng(5)
    n = 1000
    x, px, z, pz = (rng.normal(scale=s, size=n) for s in (3e-5, 1e-5, 1e-4, 1e-3))
    y = py = np.zeros(n)
    p0c, mc2, q, s = 1e9, 510998.95069, 1e-9, 2.5
    d = h5io.openpmd_particle_data(x, px, y, py, z, pz, s, p0c, mc2, q)
    f = tmp_path / "run-particles-end.h5"
    h5io.write_particles_file(str(f), d)
    t = dump(f)
    root = t["/"]["attrs"]
    assert root["openPMD"]["value"] == {"bytes": "2.0.0"}
    assert root["basePath"]["value"] == {"bytes": "/"}
    assert root["particlesPath"]["value"] == {"bytes": "."}
    assert root["speciesType"]["value"] == {"bytes": "electron"}
    assert root["numParticles"]["value"]["num"] == n
    assert root["totalCharge"]["value"]["num"] == pytest.approx(q, rel=1e-12)
    # interfaces.py:96-115: momenta in eV/c, t = (s - z)/(c beta)
    assert np.array_equal(t["position/x"]["data"], x.tolist())
    assert np.allclose(t["momentum/x"]["data"], px * p0c, rtol=1e-15, atol=0)
    pzev = p0c * np.sqrt((1 + pz) ** 2 - px ** 2)
    assert np.allclose(t["momentum/z"]["data"], pzev, rtol=1e-15, atol=0)
    p = (1 + pz) * p0c
    beta = p / np.hypot(p, mc2)
    assert np.allclose(t["time"]["data"], (s - z) / (299792458.0 * beta), rtol=1e-14, atol=0)
    # constant components are groups holding value and shape (y, py, weight, status)
    for name, val in (("position/y", 0.0), ("momentum/y", 0.0), ("particleStatus", 1),
                      ("weight", q / n)):
        assert t[name]["kind"] == "group", name
        assert t[name]["attrs"]["value"]["value"]["num"] == pytest.approx(val, rel=1e-15)
        assert t[name]["attrs"]["shape"]["value"]["num"] == [n]
    assert t["momentum/x"]["attrs"]["unitDimension"]["value"]["num"] == [1, 1, -1, 0, 0, 0, 0]
    assert t["momentum/x"]["attrs"]["unitSI"]["value"]["num"] == pytest.approx(5.344286e-28,
                                                                              rel=1e-6)
    assert t["weight"]["attrs"]["unitDimension"]["value"]["num"] == [0, 0, 1, 1, 0, 0, 0]



_H5PY_WRITER = r"""
import importlib.util, sys
import numpy as np
spec = importlib.util.spec_from_file_location("h5io", sys.argv[1])
h5io = importlib.util.module_from_spec(spec)
spec.loader.exec_module(h5io)
assert h5io.backend_name() == "h5py"
d = np.load(sys.argv[2])
steps = [{k: d[k][i] for k in d.files} for i in range(len(d["step"]))]
for st in steps:
    st["element"] = str(st["element"])
    h5io.write_wakes_file(sys.argv[3], **st)
"""


def test_h5py_backend_writes_the_same_files(tmp_path):
    """Where h5py is importable the writer uses it; its files equal the
    C-library backend's."""
    if h5io.backend_name() == "h5py":
        pytest.skip("this interpreter already writes through h5py")
    steps = _wake_steps(np.random.default_rng(9))
    npz = tmp_path / "wakes.npz"
    np.savez(npz, **{k: np.array([st[k] for st in steps]) for k in steps[0]})
    lib_file, py_file = tmp_path / "lib.h5", tmp_path / "py.h5"
    for st in steps:
        h5io.write_wakes_file(str(lib_file), **st)
    r = subprocess.run([PY, "-c", _H5PY_WRITER, h5io.__file__, str(npz), str(py_file)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert_same_tree(dump(lib_file), dump(py_file))
