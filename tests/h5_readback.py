"""Reading HDF5 files back in tests: an interpreter with h5py (any one:
``DFCSR_H5PY_PYTHON``, else this one, else a conda python3.9) runs
``tests/h5_layout.py`` and returns its JSON."""
import json
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HELPER = os.path.join(HERE, "h5_layout.py")


def _h5py_python():
    cands = [os.environ.get("DFCSR_H5PY_PYTHON"), sys.executable, "/opt/conda/bin/python3.9",
             shutil.which("python3.9")]
    for c in cands:
        if c and os.path.exists(c):
            r = subprocess.run([c, "-c", "import h5py"], capture_output=True)
            if r.returncode == 0:
                return c
    return None


PY = _h5py_python()


def run_helper(*args):
    r = subprocess.run([PY, HELPER, *args], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def dump(path):
    """{object path: {kind, attrs, dtype, shape, data}} of a file."""
    return json.loads(run_helper("dump", str(path)))
