"""ctypes binding of the oracle's C restatement (liboracle_csr.so).

TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's cpu_baseline leg.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_csr.so")

c_dp = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f64 = ctypes.c_double


class RefHist(ctypes.Structure):
    _fields_ = [("base", c_dp * 5), ("st", c_i64), ("sx", c_i64), ("sz", c_i64), ("nt", c_i32),
                ("Xi", c_i32), ("Zi", c_i32), ("cap", c_i32), ("head", c_i32), ("min_t", c_f64),
                ("min_x", c_f64), ("min_z", c_f64), ("dt", c_f64), ("dx", c_f64), ("dz", c_f64)]


class RefLat(ctypes.Structure):
    _fields_ = [("cx", c_dp), ("cy", c_dp), ("nx", c_dp), ("ny", c_dp), ("tx", c_dp),
                ("ty", c_dp), ("ns", c_i32), ("nelem", c_i32), ("min_s", c_f64), ("ds", c_f64),
                ("distance", c_dp), ("rho", c_dp)]


class RefWake(ctypes.Structure):
    _fields_ = [("t", c_f64), ("sigma_x", c_f64), ("sigma_z", c_f64), ("slope0", c_f64),
                ("mean_x", c_f64), ("formation_length", c_f64), ("n_formation_length", c_f64),
                ("csr_scaling", c_f64), ("Nx", c_i32), ("Nz", c_i32)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.run(["make", "-C", HERE], check=True, stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(LIB)
        _lib.csr_ref_wake.argtypes = [ctypes.POINTER(RefHist), ctypes.POINTER(RefLat),
                                      ctypes.POINTER(RefWake), c_dp, c_i32, c_dp, c_dp, c_i32]
        _lib.csr_ref_wake.restype = c_i32
        _lib.csr_ref_threads.restype = c_i32
    return _lib


class _Keep:
    """Holds numpy arrays alive while ctypes structs point at them."""

    def __init__(self):
        self.refs = []

    def p(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.refs.append(a)
        return a.ctypes.data


def hist_from_tracker(T, keep: _Keep):
    """Stacked (Nt, Xi, Zi) arrays of an oracle DFTracker."""
    fields = ("density", "density_x", "density_z", "vx", "vx_x")
    nt, Xi, Zi = T.data["density"].shape
    a = T.axes()
    base = (c_dp * 5)(*[keep.p(T.data[f]) for f in fields])
    return RefHist(base, Xi * Zi, Zi, 1, nt, Xi, Zi, nt, 0, a["min_t"], a["min_x"], a["min_z"],
                   a["dt"], a["dx"], a["dz"])


def hist_from_records(rec: np.ndarray, head, nt, axes, keep: _Keep):
    """The device's float64 ring rec[Xi][Zi][cap][20] of corner-packed
    records; field f of cell (x, z) is element 4 f of its record."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    keep.refs.append(rec)
    Xi, Zi, cap, R = rec.shape
    base = (c_dp * 5)(*[rec.ctypes.data + 8 * 4 * f for f in range(5)])
    return RefHist(base, R, Zi * cap * R, cap * R, nt, Xi, Zi, cap, head, *axes)


def lat_from_tables(lat, keep: _Keep):
    return RefLat(keep.p(lat.coords[:, 0]), keep.p(lat.coords[:, 1]), keep.p(lat.n_vec[:, 0]),
                  keep.p(lat.n_vec[:, 1]), keep.p(lat.tau_vec[:, 0]), keep.p(lat.tau_vec[:, 1]),
                  lat.coords.shape[0], len(lat.distance), float(lat.min_x), float(lat.delta_x),
                  keep.p(lat.distance), keep.p(lat.rho))


def wake(hist: RefHist, lat: RefLat, t, beam, formation_length, n_fl, Nx, Nz, csr_scaling, obs,
         nthreads=0):
    """dE_dct, x_kick at observers obs (M, 2) = (s - t, x)."""
    lib = load()
    obs = np.ascontiguousarray(obs, dtype=np.float64)
    M = obs.shape[0]
    dE = np.zeros(M)
    xk = np.zeros(M)
    P = RefWake(float(t), float(beam["sigma_x"]), float(beam["sigma_z"]), float(beam["slope"][0]),
                float(beam["mean_x"]), float(formation_length), float(n_fl), float(csr_scaling),
                int(Nx), int(Nz))
    lib.csr_ref_wake(ctypes.byref(hist), ctypes.byref(lat), ctypes.byref(P), obs.ctypes.data, M,
                     dE.ctypes.data, xk.ctypes.data, int(nthreads))
    return dE, xk


def threads():
    return load().csr_ref_threads()
