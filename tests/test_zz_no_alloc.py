"""dfcsr_wake_mesh_noalloc (the wake entry point that allocates nothing)
against dfcsr_wake_mesh with a caller's workspace, through the C ABI.

Both run the same kernels; the no-allocation form keeps its partial sums and
observer data in the library's static device array instead of the
workspace.  Two calls in a row, the second with a smaller mesh and on the
other integration branch (the array keeps the first call's contents), must
each equal the workspace form bit for bit, in both precisions.  A mesh whose
workspace exceeds the array is rejected with DFCSR_ERANGE before anything
runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEV = "cuda"


DFCSR_ERANGE = -3   # include/dfcsr.h
LAT_ELEMENTS = [dict(type="drift", L=0.1, nsep=1), dict(type="dipole", L=0.3, angle=0.05, nsep=1),
                dict(type="drift", L=0.4, nsep=2), dict(type="dipole", L=0.3, angle=-0.05, nsep=1)]


def _history():
    """The golden history (tests/golden/ref_history.npz) rebuilt by the
    oracle, the golden wake's beam moments, and the lattice of both."""
    import os
    from oracle import cpu_ref as O
    gd = os.path.join(os.path.dirname(__file__), "golden")
    d = np.load(os.path.join(gd, "ref_history.npz"))
    T = O.DFTracker(xbins=40, zbins=48, filter_order=2, filter_window=7,
                    velocity_threhold=1000.0, upper_limit=40)
    for k in range(len(d["times"])):
        x, z, px = d["inputs"][k]
        T.get_DF(x, z, px, float(d["times"][k]))
        T.append_DF()
        T.append_interpolant(float(d["fls"][k]), float(d["n_fl"]))
    T.build_interpolant()
    w = np.load(os.path.join(gd, "ref_wake.npz"))
    return w, T, O.LatticeTables(LAT_ELEMENTS)


def test_noalloc_wake_equals_workspace_wake():
    import torch
    from pydfcsr_amd import _lib, kernels
    from tests.gpu_util import DeviceHistory, DeviceLattice, wake_params

    w, T, lat = _history()
    H = DeviceHistory(T, DEV, rings=("fp64", "fp32t", "cells", "cells32"))
    L = DeviceLattice(lat, DEV)
    lib = _lib.load()
    ws = _lib.Workspace(torch.device(DEV))
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]),
                slope=np.array(w["slope"]), mean_x=float(w["mean_x"]))
    # first call: 24 x 20 mesh, non-chirp branch; second: 5 x 7, chirp branch
    cases = [(24, 20, float(w["slope"][0])), (5, 7, 2.5)]
    for xb, zb, slope0 in cases:
        b = dict(beam, slope=np.array([slope0, beam["slope"][1]]))
        P = wake_params(0.3, b, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                        (-2e-4, 2e-4, xb), (-1.5e-4, 1.5e-4, zb))
        M = xb * zb
        for prec in (_lib.DFCSR_PREC_FP64, _lib.DFCSR_PREC_MIXED):
            ref = kernels.wake_mesh(H.desc, L.desc, P, prec, torch.device(DEV), ws)
            out = {k: torch.full((M,), np.nan, dtype=torch.float64, device=DEV)
                   for k in ("xmesh", "zmesh", "dE_dct", "x_kick")}
            rc = lib.dfcsr_wake_mesh_noalloc(ctypes.byref(H.desc), ctypes.byref(L.desc),
                                             ctypes.byref(P), prec, out["xmesh"].data_ptr(),
                                             out["zmesh"].data_ptr(), out["dE_dct"].data_ptr(),
                                             out["x_kick"].data_ptr(),
                                             _lib.stream_handle(torch.device(DEV)))
            assert rc == 0, _lib.load().dfcsr_last_error()
            torch.cuda.synchronize()
            for k in ("xmesh", "zmesh", "dE_dct", "x_kick"):
                a, r = out[k].cpu().numpy(), ref[k].cpu().numpy()
                assert np.isfinite(r).all() and np.array_equal(a, r), (xb, zb, prec, k)
            assert np.abs(ref["dE_dct"].cpu().numpy()).max() > 0


def test_noalloc_rejects_a_mesh_beyond_its_scratch():
    import torch
    from pydfcsr_amd import _lib
    from tests.gpu_util import DeviceHistory, DeviceLattice, wake_params

    w, T, lat = _history()
    H = DeviceHistory(T, DEV, rings=("fp64",))
    L = DeviceLattice(lat, DEV)
    lib = _lib.load()
    beam = dict(sigma_x=float(w["sigma_x"]), sigma_z=float(w["sigma_z"]),
                slope=np.array(w["slope"]), mean_x=float(w["mean_x"]))
    # 1024 x 1024 mesh points need far more partial sums than the 32 MiB scratch
    P = wake_params(0.3, beam, 0.05, 1.0, 12, 16, float(w["csr_scaling"]),
                    (-2e-4, 2e-4, 1024), (-1.5e-4, 1.5e-4, 1024))
    assert lib.dfcsr_wake_workspace_bytes(ctypes.byref(P)) > (32 << 20)
    dE = torch.zeros(1, dtype=torch.float64, device=DEV)
    rc = lib.dfcsr_wake_mesh_noalloc(ctypes.byref(H.desc), ctypes.byref(L.desc),
                                     ctypes.byref(P), _lib.DFCSR_PREC_FP64, None, None,
                                     dE.data_ptr(), dE.data_ptr(),
                                     _lib.stream_handle(torch.device(DEV)))
    assert rc == DFCSR_ERANGE
    assert b"scratch" in lib.dfcsr_last_error()
    torch.cuda.synchronize()
    assert dE.item() == 0.0
