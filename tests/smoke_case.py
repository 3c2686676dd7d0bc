"""One small CSR step on cuda:0 through the C ABI, checked against the oracle.

Used by __graft_entry__.smoke(): deposit + density functions of a seeded
beam, then the wake on the reference's golden history, both compared with the
CPU oracle."""
from __future__ import annotations

import os

import numpy as np
import torch

from oracle import cpu_ref as O
from tests.gpu_util import DeviceHistory, DeviceLattice, rel_max, wake_params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def run_smoke(device="cuda:0"):
    from pydfcsr_amd import _lib, kernels
    from pydfcsr_amd.deposit import DF_tracker
    _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("smoke needs a HIP device")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(device)
    print("smoke: deposit + density functions (40x48 grid, 12000 particles)", flush=True)
    d = np.load(os.path.join(GOLDEN, "ref_get_df_corr.npz"))
    tr = DF_tracker(dict(xbins=40, zbins=48, filter_order=2, filter_window=7,
                         velocity_threhold=1000.0), device=device)
    tr.get_DF(t(d["x"]), t(d["z"]), t(d["px"]), 0.25)
    f = tr.current_fields()
    ref = O.get_DF(d["x"], d["z"], d["px"], 0.25, xbins=40, zbins=48, filter_order=2,
                   filter_window=7, velocity_threhold=1000.0)
    for k in ("density", "vx", "density_x", "density_z", "vx_x"):
        err = rel_max(f[k].cpu(), ref[k])
        assert err <= 1e-9, (k, err)
    print("smoke: wake on the golden history (3x4 mesh, 12x16 integration)", flush=True)
    h = np.load(os.path.join(GOLDEN, "ref_history.npz"))
    T = O.DFTracker(xbins=40, zbins=48, filter_order=2, filter_window=7,
                    velocity_threhold=1000.0, upper_limit=40)
    for k in range(len(h["times"])):
        x, z, px = h["inputs"][k]
        T.get_DF(x, z, px, float(h["times"][k]))
        T.append_DF()
        T.append_interpolant(float(h["fls"][k]), float(h["n_fl"]))
    T.build_interpolant()
    w = np.load(os.path.join(GOLDEN, "ref_wake.npz"))
    L = O.LatticeTables([dict(type="drift", L=0.1, nsep=1),
                         dict(type="dipole", L=0.3, angle=0.05, nsep=1),
                         dict(type="drift", L=0.4, nsep=2),
                         dict(type="dipole", L=0.3, angle=-0.05, nsep=1)])
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]), slope=w["slope"],
                mean_x=float(w["mean_x"]))
    xr, zr = w["xrange"], w["zrange"]
    P = wake_params(0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                    (xr[0], xr[-1], xr.size), (zr[0], zr[-1], zr.size))
    DH, DL = DeviceHistory(T, device), DeviceLattice(L, device)
    ws = _lib.Workspace(device)
    out = kernels.wake_mesh(DH.desc, DL.desc, P, _lib.DFCSR_PREC_FP64, device, ws)
    torch.cuda.synchronize()
    e1 = rel_max(out["dE_dct"].cpu().view(3, 4), w["dE_dct"])
    e2 = rel_max(out["x_kick"].cpu().view(3, 4), w["x_kick"])
    assert e1 <= 1e-10 and e2 <= 1e-10, (e1, e2)
    print(f"smoke: OK (wake rel err {e1:.2e}, {e2:.2e})", flush=True)
