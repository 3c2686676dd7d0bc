"""HDF5 layout helper for the output-file tests, run by an interpreter that
has h5py (it need not be the one running the tests).

    h5_layout.py dump FILE            JSON of every object: kind, attributes
                                      with their HDF5 type, dataset dtype,
                                      shape and values
    h5_layout.py wakes NPZ OUT        OUT as the reference writes its wakes
                                      file with h5py (CSR.py:799-835), one
                                      step per entry of NPZ
    h5_layout.py statistics NPZ OUT   OUT as the reference writes its
                                      statistics file (CSR.py:860-879,
                                      tools.py:69-77)

The two writers restate the reference's h5py calls on the same data, so a
dump of their files is what a file in the reference's layout must give.
"""
import json
import sys

import h5py
import numpy as np


def _value(v):
    if isinstance(v, bytes):
        return {"bytes": v.decode("latin-1")}
    if isinstance(v, str):
        return {"str": v}
    a = np.asarray(v)
    if a.dtype.kind == "S":
        return {"bytes": a.tobytes().decode("latin-1")}
    return {"num": a.tolist(), "dtype": a.dtype.str, "shape": list(a.shape)}


def _type(t):
    if t.get_class() == h5py.h5t.STRING:
        out = {"class": "string", "variable": bool(t.is_variable_str()),
               "cset": int(t.get_cset())}
        if not t.is_variable_str():
            out["size"] = int(t.get_size())
            out["strpad"] = int(t.get_strpad())
        return out
    return {"class": int(t.get_class()), "size": int(t.get_size()),
            "order": int(t.get_order())}


def _attrs(obj):
    out = {}
    for k in obj.attrs:
        a = h5py.h5a.open(obj.id, k.encode())
        out[k] = {"type": _type(a.get_type()), "shape": list(a.shape),
                  "value": _value(obj.attrs[k])}
    return out


def dump(path):
    tree = {}
    with h5py.File(path, "r") as f:
        tree["/"] = {"kind": "group", "attrs": _attrs(f)}

        def visit(name, obj):
            if isinstance(obj, h5py.Dataset):
                tree[name] = {"kind": "dataset", "attrs": _attrs(obj),
                              "type": _type(obj.id.get_type()),
                              "dtype": obj.dtype.str, "shape": list(obj.shape),
                              "data": np.asarray(obj[()]).tolist()}
            else:
                tree[name] = {"kind": "group", "attrs": _attrs(obj)}
        f.visititems(visit)
    json.dump(tree, sys.stdout)


def wakes(npz, out):
    d = np.load(npz)
    with h5py.File(out, "w") as hf:
        for k in range(len(d["step"])):
            g = hf.create_group("step_" + str(int(d["step"][k])))
            g.attrs["step"] = int(d["step"][k])
            g.attrs["position"] = float(d["position"][k])
            g.attrs["mean_gamma"] = float(d["mean_gamma"][k])
            g.attrs["beam_energy"] = float(d["beam_energy"][k])
            g.attrs["element"] = str(d["element"][k])
            g.attrs["charge"] = float(d["charge"][k])
            shape = d["dE_dct"][k].shape
            g1 = g.create_group("longitudinal")
            g1.attrs["unit"] = "MeV/m"
            g1.create_dataset("x_grids", data=d["xmesh"][k].reshape(shape))
            g1.create_dataset("z_grids", data=d["zmesh"][k].reshape(shape))
            g1.create_dataset("dE_dct", data=d["dE_dct"][k])
            g2 = g.create_group("transverse")
            g2.attrs["unit"] = "MeV/m"
            g2.create_dataset("x_grids", data=d["xmesh"][k].reshape(shape))
            g2.create_dataset("z_grids", data=d["zmesh"][k].reshape(shape))
            g2.create_dataset("xkicks", data=d["x_kick"][k])


def _dict2hdf5(hf, dic, group=None):
    for key, item in dic.items():
        if not isinstance(item, dict):
            (group if group else hf).create_dataset(key, data=item)
        else:
            _dict2hdf5(hf, item, hf.create_group(key))


def statistics(npz, out):
    d = np.load(npz)
    st = {"twiss": {}}
    for k in d.files:
        if k.startswith("twiss/"):
            st["twiss"][k[6:]] = d[k]
    for k in ("slope", "sigma_x", "sigma_z", "sigma_energy", "mean_x", "mean_z", "mean_energy"):
        st[k] = d[k]
    with h5py.File(out, "w") as hf:
        hf.create_dataset(name="step_positions", data=d["step_positions"],
                          shape=d["step_positions"].shape)
        hf.create_dataset(name="coords", data=d["coords"])
        hf.create_dataset(name="n_vec", data=d["n_vec"])
        hf.create_dataset(name="tau_vec", data=d["tau_vec"])
        _dict2hdf5(hf, st)


if __name__ == "__main__":
    cmd = sys.argv[1]
    if cmd == "dump":
        dump(sys.argv[2])
    elif cmd == "wakes":
        wakes(sys.argv[2], sys.argv[3])
    elif cmd == "statistics":
        statistics(sys.argv[2], sys.argv[3])
    else:
        raise SystemExit("usage: h5_layout.py dump FILE | wakes NPZ OUT | statistics NPZ OUT")
