"""HDF5 output files of a run, in the reference's on-disk layout.

Three files, as the reference writes them with h5py:

* ``{prefix}-wakes.h5`` (CSR.py:799-835): one group ``step_{n}`` per CSR
  step with attributes step, position, mean_gamma, beam_energy, element,
  charge, and sub-groups ``longitudinal`` (x_grids, z_grids, dE_dct) and
  ``transverse`` (x_grids, z_grids, xkicks), each with attribute
  unit = 'MeV/m'.  The file is appended to step by step.
* ``{prefix}-statistics.h5`` (CSR.py:860-879, tools.py:69-77): datasets
  step_positions, coords, n_vec, tau_vec and the statistics dict (group
  ``twiss`` with 14 datasets; slope, sigma_x, sigma_z, sigma_energy,
  mean_x, mean_z, mean_energy).
* ``{prefix}-particles-{label}.h5`` (CSR.py:784-797): the beam as an
  openPMD-beamphysics particle file, the data of bmadx_particles_to_openpmd
  (interfaces.py:65-121) written at the file root as
  ``ParticleGroup.write(filename)`` does.

Writing goes through h5py when it is importable (the reference's own
dependency).  Without it, the HDF5 C library is called through ctypes:
``DFCSR_HDF5_LIB`` names it, else it is searched for on the loader path and
under the lib directories of installed Python distributions.  Values are
stored with the types h5py gives the same Python objects: int -> int64,
float -> float64, str -> variable-length UTF-8, bytes -> fixed-length ASCII
(null-padded), numpy arrays with their own dtype (little-endian).
"""
from __future__ import annotations

import ctypes
import ctypes.util
import glob
import os
import sys

import numpy as np

__all__ = ["H5File", "backend_name", "write_wakes_file", "write_statistics_file",
           "write_particles_file", "openpmd_particle_data"]


# --------------------------------------------------------------------------------------
# backends

class _H5pyBackend:
    name = "h5py"

    def __init__(self, h5py):
        self.h5py = h5py

    def open(self, path, mode):
        return self.h5py.File(path, mode)

    def close(self, f):
        f.close()

    def create_group(self, f, name):
        f.create_group(name)

    def create_dataset(self, f, name, data):
        f.create_dataset(name, data=data)

    def set_attr(self, f, obj, key, value):
        f[obj].attrs[key] = value


class _LibBackend:
    """The HDF5 1.10+ C API over ctypes (hid_t is 64-bit from 1.10 on)."""
    name = "libhdf5"
    ACC_RDWR, ACC_TRUNC = 0x0001, 0x0002
    S_SCALAR = 0
    CSET_ASCII, CSET_UTF8 = 0, 1
    STR_NULLTERM, STR_NULLPAD = 0, 1
    VARIABLE = ctypes.c_size_t(-1).value

    def __init__(self, path):
        lib = ctypes.CDLL(path)
        hid, herr, hsize = ctypes.c_int64, ctypes.c_int, ctypes.c_uint64
        cp, vp, u = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_uint
        sig = {
            "H5open": (herr, []),
            "H5get_libversion": (herr, [ctypes.POINTER(u)] * 3),
            "H5Eset_auto2": (herr, [hid, vp, vp]),
            "H5Fcreate": (hid, [cp, u, hid, hid]),
            "H5Fopen": (hid, [cp, u, hid]),
            "H5Fclose": (herr, [hid]),
            "H5Gcreate2": (hid, [hid, cp, hid, hid, hid]),
            "H5Gclose": (herr, [hid]),
            "H5Oopen": (hid, [hid, cp, hid]),
            "H5Oclose": (herr, [hid]),
            "H5Screate": (hid, [ctypes.c_int]),
            "H5Screate_simple": (hid, [ctypes.c_int, ctypes.POINTER(hsize), ctypes.POINTER(hsize)]),
            "H5Sclose": (herr, [hid]),
            "H5Dcreate2": (hid, [hid, cp, hid, hid, hid, hid, hid]),
            "H5Dwrite": (herr, [hid, hid, hid, hid, hid, vp]),
            "H5Dclose": (herr, [hid]),
            "H5Aexists": (ctypes.c_int, [hid, cp]),
            "H5Adelete": (herr, [hid, cp]),
            "H5Acreate2": (hid, [hid, cp, hid, hid, hid, hid]),
            "H5Awrite": (herr, [hid, hid, vp]),
            "H5Aclose": (herr, [hid]),
            "H5Tcopy": (hid, [hid]),
            "H5Tset_size": (herr, [hid, ctypes.c_size_t]),
            "H5Tset_cset": (herr, [hid, ctypes.c_int]),
            "H5Tset_strpad": (herr, [hid, ctypes.c_int]),
            "H5Tclose": (herr, [hid]),
        }
        for fn, (res, args) in sig.items():
            f = getattr(lib, fn)
            f.restype, f.argtypes = res, args
        self.lib = lib
        self._call("H5open")
        ver = [u(), u(), u()]
        self._call("H5get_libversion", *[ctypes.byref(v) for v in ver])
        self.version = tuple(v.value for v in ver)
        if self.version < (1, 10, 0):
            raise RuntimeError(f"{path}: HDF5 {self.version} is older than 1.10 (64-bit hid_t)")
        self.lib.H5Eset_auto2(0, None, None)     # errors come back as return codes
        g = lambda n: ctypes.c_int64.in_dll(lib, n + "_g").value
        self.types = {
            np.dtype("<f8"): g("H5T_IEEE_F64LE"), np.dtype("<f4"): g("H5T_IEEE_F32LE"),
            np.dtype("<i8"): g("H5T_STD_I64LE"), np.dtype("<i4"): g("H5T_STD_I32LE"),
            np.dtype("<i2"): g("H5T_STD_I16LE"), np.dtype("i1"): g("H5T_STD_I8LE"),
            np.dtype("<u8"): g("H5T_STD_U64LE"), np.dtype("<u4"): g("H5T_STD_U32LE"),
            np.dtype("u1"): g("H5T_STD_U8LE"),
        }
        self.c_s1 = g("H5T_C_S1")
        self.path = path

    def _call(self, fn, *args):
        r = getattr(self.lib, fn)(*args)
        if r < 0:
            raise OSError(f"HDF5 {fn} failed ({r})")
        return r

    # files and objects
    def open(self, path, mode):
        p = os.fsencode(path)
        if mode == "w" or (mode == "a" and not os.path.exists(path)):
            return self._call("H5Fcreate", p, self.ACC_TRUNC, 0, 0)
        if mode == "a":
            return self._call("H5Fopen", p, self.ACC_RDWR, 0)
        raise ValueError(f"mode {mode!r}: the writer opens files with 'w' or 'a'")

    def close(self, f):
        self._call("H5Fclose", f)

    def create_group(self, f, name):
        self._call("H5Gclose", self._call("H5Gcreate2", f, name.encode(), 0, 0, 0))

    def _space(self, shape):
        if shape == ():
            return self._call("H5Screate", self.S_SCALAR)
        dims = (ctypes.c_uint64 * len(shape))(*shape)
        return self._call("H5Screate_simple", len(shape), dims, None)

    def _typed(self, value):
        """(HDF5 type id, owned, contiguous buffer, shape) of a value, with
        the type h5py would give it."""
        if isinstance(value, str):
            t = self._call("H5Tcopy", self.c_s1)
            self._call("H5Tset_size", t, self.VARIABLE)
            self._call("H5Tset_cset", t, self.CSET_UTF8)
            self._call("H5Tset_strpad", t, self.STR_NULLTERM)
            s = ctypes.c_char_p(value.encode("utf-8"))
            return t, True, (ctypes.c_char_p * 1)(s), ()
        if isinstance(value, (bytes, np.bytes_)):
            b = bytes(value)
            t = self._call("H5Tcopy", self.c_s1)
            self._call("H5Tset_size", t, max(1, len(b)))
            self._call("H5Tset_strpad", t, self.STR_NULLPAD)
            return t, True, ctypes.create_string_buffer(b, max(1, len(b))), ()
        a = np.asarray(value)
        if a.dtype == np.bool_:
            raise TypeError("boolean values are not written by this writer")
        shape = a.shape                # ascontiguousarray makes a 0-d array 1-d
        a = np.ascontiguousarray(a, dtype=a.dtype.newbyteorder("<"))
        if a.dtype not in self.types:
            raise TypeError(f"unsupported dtype {a.dtype}")
        return self.types[a.dtype], False, a, shape

    def create_dataset(self, f, name, data):
        t, owned, buf, shape = self._typed(np.asarray(data))
        sp = self._space(shape)
        try:
            d = self._call("H5Dcreate2", f, name.encode(), t, sp, 0, 0, 0)
            try:
                self._call("H5Dwrite", d, t, 0, 0, 0, buf.ctypes.data)
            finally:
                self._call("H5Dclose", d)
        finally:
            self._call("H5Sclose", sp)

    def set_attr(self, f, obj, key, value):
        o = self._call("H5Oopen", f, obj.encode(), 0)
        try:
            if self.lib.H5Aexists(o, key.encode()) > 0:
                self._call("H5Adelete", o, key.encode())
            t, owned, buf, shape = self._typed(value)
            sp = self._space(shape)
            try:
                a = self._call("H5Acreate2", o, key.encode(), t, sp, 0, 0)
                try:
                    ptr = buf.ctypes.data if isinstance(buf, np.ndarray) else ctypes.addressof(buf)
                    self._call("H5Awrite", a, t, ptr)
                finally:
                    self._call("H5Aclose", a)
            finally:
                self._call("H5Sclose", sp)
                if owned:
                    self._call("H5Tclose", t)
        finally:
            self._call("H5Oclose", o)


def _library_candidates():
    env = os.environ.get("DFCSR_HDF5_LIB")
    if env:
        return [env]
    out = []
    found = ctypes.util.find_library("hdf5")
    if found:
        out.append(found)
    prefixes = {sys.prefix, sys.base_prefix, "/usr", "/usr/local", "/opt/conda"}
    for p in sorted(prefixes):
        for d in ("lib", "lib64", "lib/x86_64-linux-gnu"):
            out += sorted(glob.glob(os.path.join(p, d, "libhdf5.so*")))
            out += sorted(glob.glob(os.path.join(p, d, "libhdf5_serial.so*")))
    return out


_BACKEND = None


def _backend():
    global _BACKEND
    if _BACKEND is not None:
        return _BACKEND
    if os.environ.get("DFCSR_HDF5_LIB") is None:
        try:
            import h5py
            _BACKEND = _H5pyBackend(h5py)
            return _BACKEND
        except ImportError:
            pass
    errors = []
    for cand in _library_candidates():
        try:
            _BACKEND = _LibBackend(cand)
            return _BACKEND
        except (OSError, RuntimeError, AttributeError) as e:
            errors.append(f"{cand}: {e}")
    raise RuntimeError("HDF5 output needs h5py or the HDF5 C library (>= 1.10); set "
                       "DFCSR_HDF5_LIB to libhdf5.so" + (": " + "; ".join(errors) if errors else ""))


def backend_name():
    """'h5py' or 'libhdf5 <version> (<path>)': what the writer uses here."""
    b = _backend()
    if b.name == "h5py":
        return "h5py"
    return "libhdf5 {}.{}.{} ({})".format(*b.version, b.path)


class H5File:
    """A file being written: groups, datasets and attributes by absolute or
    file-relative path.  Mode 'w' truncates, 'a' appends (creating the file
    if it is missing)."""

    def __init__(self, path, mode="w"):
        self._b = _backend()
        self._f = self._b.open(str(path), mode)

    def create_group(self, name):
        self._b.create_group(self._f, name)
        return name

    def create_dataset(self, name, data):
        self._b.create_dataset(self._f, name, data)
        return name

    def set_attr(self, obj, key, value):
        self._b.set_attr(self._f, obj, key, value)

    def close(self):
        if self._f is not None:
            self._b.close(self._f)
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# --------------------------------------------------------------------------------------
# the three files

def _remove_existing(filename, verbose):
    if os.path.isfile(filename):
        os.remove(filename)
        if verbose:
            print("Existing file " + filename + " deleted.")


def write_wakes_file(filename, step, position, mean_gamma, beam_energy, element, charge,
                     xmesh, zmesh, dE_dct, x_kick, verbose=False):
    """Append the wake of one CSR step (CSR.py:799-835).  Step 1 starts the
    file afresh."""
    dE_dct = np.asarray(dE_dct, dtype=np.float64)
    x_kick = np.asarray(x_kick, dtype=np.float64)
    xg = np.asarray(xmesh, dtype=np.float64).reshape(dE_dct.shape)
    zg = np.asarray(zmesh, dtype=np.float64).reshape(dE_dct.shape)
    if step == 1:
        _remove_existing(filename, verbose)
        if verbose:
            print("Wakes written to ", filename)
    with H5File(filename, "a") as hf:
        g = "step_" + str(step)
        hf.create_group(g)
        hf.set_attr(g, "step", int(step))
        hf.set_attr(g, "position", float(position))
        hf.set_attr(g, "mean_gamma", float(mean_gamma))
        hf.set_attr(g, "beam_energy", float(beam_energy))
        hf.set_attr(g, "element", str(element))
        hf.set_attr(g, "charge", float(charge))
        for sub, name, w in (("longitudinal", "dE_dct", dE_dct), ("transverse", "xkicks", x_kick)):
            p = hf.create_group(g + "/" + sub)
            hf.set_attr(p, "unit", "MeV/m")
            hf.create_dataset(p + "/x_grids", xg)
            hf.create_dataset(p + "/z_grids", zg)
            hf.create_dataset(p + "/" + name, w)


def _dict_to_file(hf, dic, group=None):
    """tools.py:69-77, including its placement of nested dicts: a dict value
    becomes a group at the file root."""
    for key, item in dic.items():
        if not isinstance(item, dict):
            hf.create_dataset((group + "/" + key) if group else key, np.asarray(item))
        else:
            _dict_to_file(hf, item, hf.create_group(key))


def write_statistics_file(filename, step_positions, coords, n_vec, tau_vec, statistics,
                          verbose=False):
    """The per-step statistics and the lattice reference orbit (CSR.py:860-879)."""
    _remove_existing(filename, verbose)
    if verbose:
        print("Statistics written to ", filename)
    with H5File(filename, "w") as hf:
        hf.create_dataset("step_positions", np.asarray(step_positions))
        hf.create_dataset("coords", np.asarray(coords))
        hf.create_dataset("n_vec", np.asarray(n_vec))
        hf.create_dataset("tau_vec", np.asarray(tau_vec))
        _dict_to_file(hf, statistics)


C_LIGHT = 299792458.0
E_CHARGE = 1.602176634e-19      # C; the momentum unit eV/c is E_CHARGE / C_LIGHT kg m/s


def openpmd_particle_data(x, px, y, py, z, pz, s, p0c, mc2, charge):
    """bmadx_particles_to_openpmd (interfaces.py:65-121): Bmad phase space
    (x, px/p0, y, py/p0, z, delta) at position s to openPMD quantities: x, y, z
    in m; px, py, pz in eV/c; t = (s - z)/(c beta); status 1; equal weights
    summing to the bunch charge."""
    x, px, y, py, z, pz = (np.asarray(a, dtype=np.float64) for a in (x, px, y, py, z, pz))
    d = {"x": x, "px": px * p0c, "y": y, "py": py * p0c, "z": z,
         "pz": p0c * ((pz + 1.0) ** 2 - px ** 2 - py ** 2) ** 0.5}
    p = (1 + pz) * p0c
    beta = ((p / mc2) ** 2 / (1 + (p / mc2) ** 2)) ** 0.5
    d["t"] = (s - z) / (C_LIGHT * beta)
    d["status"] = np.ones_like(x, dtype=np.int64)
    d["weight"] = np.ones_like(x) * charge / len(x)
    if not np.isclose(mc2, 510998.95):
        raise ValueError("only electrons are supported as of now")
    d["species"] = "electron"
    return d


# openPMD unit records: (unitSI, unitDimension over L, M, T, I, theta, N, J)
_UNITS = {
    "m": (1.0, (1, 0, 0, 0, 0, 0, 0)),
    "eV/c": (E_CHARGE / C_LIGHT, (1, 1, -1, 0, 0, 0, 0)),
    "s": (1.0, (0, 0, 1, 0, 0, 0, 0)),
    "C": (1.0, (0, 0, 1, 1, 0, 0, 0)),
    "1": (1.0, (0, 0, 0, 0, 0, 0, 0)),
}


def _component(hf, name, data, unit):
    """An openPMD record component: a dataset, or, when every value is the
    same, a group holding the constant as attributes value and shape."""
    data = np.asarray(data)
    if data.size and np.all(data == data.flat[0]):
        hf.create_group(name)
        hf.set_attr(name, "value", data.flat[0])
        hf.set_attr(name, "shape", np.asarray(data.shape, dtype=np.int64))
    else:
        hf.create_dataset(name, data)
    si, dim = _UNITS[unit]
    hf.set_attr(name, "unitSI", float(si))
    hf.set_attr(name, "unitDimension", np.asarray(dim, dtype=np.float64))
    hf.set_attr(name, "unitSymbol", unit.encode())


def write_particles_file(filename, data, verbose=False):
    """An openPMD-beamphysics particle file with the bunch at the root
    (openPMD 2.0.0, extension BeamPhysics;SpeciesType; basePath '/',
    particlesPath '.')."""
    _remove_existing(filename, verbose)
    with H5File(filename, "w") as hf:
        for k, v in (("basePath", "/"), ("dataType", "openPMD"), ("openPMD", "2.0.0"),
                     ("openPMDextension", "BeamPhysics;SpeciesType"), ("particlesPath", ".")):
            hf.set_attr("/", k, v.encode())
        w = data["weight"]
        live = data["status"] == 1
        hf.set_attr("/", "speciesType", data["species"].encode())
        hf.set_attr("/", "numParticles", int(len(w)))
        hf.set_attr("/", "totalCharge", float(np.sum(w)))
        hf.set_attr("/", "chargeLive", float(np.sum(w[live])))
        hf.set_attr("/", "chargeUnitSI", 1.0)
        hf.create_group("position")
        hf.create_group("momentum")
        for c in ("x", "y", "z"):
            _component(hf, "position/" + c, data[c], "m")
            _component(hf, "momentum/" + c, data["p" + c], "eV/c")
        _component(hf, "time", data["t"], "s")
        _component(hf, "weight", data["weight"], "C")
        _component(hf, "particleStatus", data["status"], "1")
