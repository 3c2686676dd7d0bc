"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, rejects bad arguments before touching a device, and its
host-side Savitzky-Golay operator equals scipy's."""
import os
import re

import numpy as np
import pytest
from scipy.signal import savgol_coeffs, savgol_filter

from pydfcsr_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "dfcsr.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dfcsr_\w+)\s*\(", txt)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(_lib.EXPORTED) | {"dfcsr_savgol_operator"}
    assert lib.dfcsr_abi_version() == _lib.ABI_VERSION == 8


def test_argument_errors_without_device():
    lib = _lib.load()
    rc = lib.dfcsr_beam_moments(None, None, 10, None, 0, None, None)
    assert rc == -1 and b"NULL" in lib.dfcsr_last_error()
    rc = lib.dfcsr_df_process(1, 1, 1, 1, 1, 8, 0.0, 1.0, 8, 0.0, 1.0, 5.0, 9, 2, None, 1, 1 << 20,
                              None)
    assert rc == -1 and b"window_length" in lib.dfcsr_last_error()
    with pytest.raises(_lib.DfcsrError, match="workspace"):
        _lib.check(lib.dfcsr_df_process(1, 1, 1, 1, 1, 8, 0.0, 1.0, 8, 0.0, 1.0, 5.0, 5, 2, None,
                                        1, 8, None), "df_process")
    H = _lib.History(1, None, None, None, None, 4, 0, 1, 8, 8, 0, 0, 0, 1, 1, 1)
    rc = lib.dfcsr_csr_integrand(H, _lib.Lattice(), 0.0, 0.0, 0.0, 1, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"slice" in lib.dfcsr_last_error()
    # the integrand reads the float64 ring; a history needs at least one ring
    H = _lib.History(None, 2, 3, None, None, 4, 0, 2, 8, 8, 0, 0, 0, 1, 1, 1)
    rc = lib.dfcsr_csr_integrand(H, _lib.Lattice(), 0.0, 0.0, 0.0, 1, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"FP64 ring" in lib.dfcsr_last_error()
    H = _lib.History(None, None, None, None, None, 4, 0, 2, 8, 8, 0, 0, 0, 1, 1, 1)
    rc = lib.dfcsr_csr_integrand(H, _lib.Lattice(), 0.0, 0.0, 0.0, 1, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"without a ring" in lib.dfcsr_last_error()
    # the time differences of the float32 cell-major ring: both rings, a window
    # inside the ring; a window that starts past its newest slice has no work
    rc = lib.dfcsr_cells32_tdiff(1, None, 8, 8, 4, 0, 2, 0, None)
    assert rc == -1 and b"NULL ring" in lib.dfcsr_last_error()
    for cap, head, nt in ((4, 4, 2), (4, 0, 5), (0, 0, 0), (4, -1, 2)):
        rc = lib.dfcsr_cells32_tdiff(1, 2, 8, 8, cap, head, nt, 0, None)
        assert rc == -1 and b"cap" in lib.dfcsr_last_error(), (cap, head, nt)
    rc = lib.dfcsr_cells32_tdiff(1, 2, 0, 8, 4, 0, 2, 0, None)
    assert rc == -1 and b"grid" in lib.dfcsr_last_error()
    assert lib.dfcsr_cells32_tdiff(1, 2, 8, 8, 4, 0, 2, 2, None) == 0
    assert lib.dfcsr_cells32_tdiff(1, 2, 8, 8, 4, 0, 0, 0, None) == 0
    # the span table: a cell-major ring and the table, 16-bit z, a window in the ring
    rc = lib.dfcsr_history_zspan(None, None, 3, 8, 8, 4, 0, 2, 0, None)
    assert rc == -1 and b"NULL ring" in lib.dfcsr_last_error()
    rc = lib.dfcsr_history_zspan(1, None, None, 8, 8, 4, 0, 2, 0, None)
    assert rc == -1 and b"NULL ring" in lib.dfcsr_last_error()
    rc = lib.dfcsr_history_zspan(1, None, 3, 8, 70000, 4, 0, 2, 0, None)
    assert rc == -1 and b"65535" in lib.dfcsr_last_error()
    rc = lib.dfcsr_history_zspan(None, 2, 3, 8, 8, 4, 4, 2, 0, None)
    assert rc == -1 and b"cap" in lib.dfcsr_last_error()
    assert lib.dfcsr_history_zspan(1, None, 3, 8, 8, 4, 0, 2, 2, None) == 0


@pytest.mark.parametrize("window,order", [(5, 2), (7, 2), (9, 3), (5, 0), (9, 1), (21, 4)])
def test_savgol_operator_matches_scipy(window, order):
    from pydfcsr_amd.kernels import savgol_operator
    c, edge = savgol_operator(window, order)
    # scipy solves the fit by SVD in float64 (~1e-14 at window 21, order 4);
    # the operator here is solved in extended precision
    tol = 1e-14 if window < 21 else 5e-13
    assert np.allclose(c, savgol_coeffs(window, order), rtol=0, atol=tol)
    # the edge rows reproduce scipy's polynomial edge fit (mode='interp')
    rng = np.random.default_rng(window * 10 + order)
    x = rng.normal(size=3 * window)
    y = savgol_filter(x, window, order)
    h = window // 2
    mine_left = edge @ x[:window]
    mine_right = edge[::-1, ::-1] @ x[-window:]
    assert np.allclose(mine_left, y[:h], rtol=0, atol=1e-12)
    assert np.allclose(mine_right, y[-h:], rtol=0, atol=1e-12)
    mid = np.convolve(x, c[::-1], mode="valid")
    assert np.allclose(mid, y[h:-h], rtol=0, atol=100 * tol)


def test_savgol_operator_rejects_bad_window():
    lib = _lib.load()
    c = np.zeros(8)
    assert lib.dfcsr_savgol_operator(4, 2, c.ctypes.data, c.ctypes.data) == -1
    assert lib.dfcsr_savgol_operator(5, 5, c.ctypes.data, c.ctypes.data) == -1
    assert lib.dfcsr_savgol_operator(0, 0, c.ctypes.data, c.ctypes.data) == -1


def _params(slope0, mesh=64, nxz=200):
    return _lib.WakeParams(3.0, 1e-4, 1e-4, slope0, 0.0, 0.0, 0.7, 4.0, 1e-5, nxz, nxz, mesh, mesh,
                           -1e-4, 1e-4, -1e-4, 1e-4, None)


def far_plan(xbins, zbins, Nz):
    """Region 0 of a mesh, observer-major: (partial sums per mesh point,
    workgroups).  W lanes per observer block (64, or xbins rounded up to a
    power of two), S = 64 / W column blocks per wave, q workgroups of 4 waves
    per (z line, observer block): enough for 8192 waves, at most 32, no more
    column blocks than columns."""
    W = 1
    while W < xbins and W < 64:
        W *= 2
    S = 64 // W
    nob = -(-xbins // W)
    lines = zbins * nob
    q = -(-(8192 // 4) // lines)
    q = max(1, min(q, 32, max(1, -(-Nz // (4 * S)))))
    return 4 * S * q, zbins * nob * q


@pytest.mark.parametrize("slope0,near", [(-1.25, 3 * 4), (0.5, 2 * 4), (2.0, 3 * 4)])
def test_wake_workspace_follows_the_phase_plan(slope0, near):
    """Every later region has Nx rows in 4 chunks; 3 regions without chirp
    (|slope| <= 1), 4 with.  Region 0 of a mesh: one partial sum per column
    block (128 for a 64 x 64 mesh: 32 workgroups of 4 per z line); of
    explicit observers: 2 Nx rows in
    ceil(400/64) = 7 chunks.  Per mesh point: 2 partial sums per phase or
    column block + 48 doubles of observer data."""
    import ctypes
    lib = _lib.load()
    P = _params(slope0)
    M = 64 * 64
    far, _ = far_plan(64, 64, 200)
    assert far == 128
    assert lib.dfcsr_wake_workspace_bytes(ctypes.byref(P)) == M * (2 * (far + near) + 48) * 8 + 256
    P.observers = 8
    assert lib.dfcsr_wake_workspace_bytes(ctypes.byref(P)) == M * (2 * (7 + near) + 48) * 8 + 256


def test_wake_tasks_pack_host_logic():
    """Packing is host-only: counts workgroups and partial-sum bytes per task
    and rejects bad tasks before any device work (fake device pointers)."""
    import ctypes
    lib = _lib.load()
    k = 3
    hist = (_lib.History * k)(*[_lib.History(1, 2, 3, None, None, 8, 0, 5, 16, 16, 0, 0, 0, 1, 1, 1)
                                 for _ in range(k)])
    lat = (_lib.Lattice * k)(*[_lib.Lattice(3, 100, 1, 0.0, 0.01, 4, 5) for _ in range(k)])
    P = (_lib.WakeParams * k)(_params(-1.25, 8), _params(0.5, 16), _params(-1.25, 4))
    out = (_lib.c_dp * k)(*[8 + j for j in range(k)])
    nb = int(lib.dfcsr_wake_task_table_bytes(k))
    tbl = ctypes.create_string_buffer(nb)
    batch = _lib.WakeBatch()
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_MIXED, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == 0, lib.dfcsr_last_error()
    M = [64, 256, 16]
    plans = [far_plan(b, b, 200) for b in (8, 16, 4)]
    ph = [(plans[0][0], 12), (plans[1][0], 8), (plans[2][0], 12)]
    assert batch.tasks == k and batch.observers == sum(M)
    assert batch.blocks_far == sum(wg for _, wg in plans)
    assert batch.blocks_near == sum(m * b for m, (_, b) in zip(M, ph))
    assert batch.partial_bytes == sum(m * (2 * (a + b) + 48) * 8 for m, (a, b) in zip(M, ph))
    assert lib.dfcsr_wake_batched_workspace_bytes(k, P) == nb + batch.partial_bytes
    assert batch.precision == _lib.DFCSR_PREC_MIXED
    # a launch must use the precision the tasks were checked for
    rc = lib.dfcsr_wake_batched_launch(tbl, ctypes.byref(batch), _lib.DFCSR_PREC_FP64, 1,
                                       batch.partial_bytes, None)
    assert rc == -1 and b"packed for precision" in lib.dfcsr_last_error()
    # mixed precision needs one FP32 ring per task (either layout), FP64 the FP64 ring
    hist[1].data32 = None
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_MIXED, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == 0, lib.dfcsr_last_error()
    hist[1].data32t = None
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_MIXED, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == -1 and b"FP32" in lib.dfcsr_last_error()
    hist[1].data32 = 2
    hist[2].data = None
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_MIXED, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == 0, lib.dfcsr_last_error()
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_FP64, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == -1 and b"FP64 precision" in lib.dfcsr_last_error()
    hist[2].data = 1
    # the cell-major rings are optional
    hist[0].cells, hist[0].cells32 = 7, 8
    for prec in (_lib.DFCSR_PREC_FP64, _lib.DFCSR_PREC_MIXED):
        rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, prec, out, out, tbl, nb,
                                       ctypes.byref(batch))
        assert rc == 0, lib.dfcsr_last_error()
    # the wake addresses a cell within a plane with 32-bit byte offsets: a
    # plane of 11000 x 11000 float32 cells (2.4 GB) fits, float64 (4.8 GB) not
    hist[0].Xi = hist[0].Zi = 11000
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_MIXED, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == 0, lib.dfcsr_last_error()
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_FP64, out, out, tbl, nb,
                                   ctypes.byref(batch))
    assert rc == -1 and b"exceeds 4 GiB" in lib.dfcsr_last_error()
    hist[0].Xi = hist[0].Zi = 16
    rc = lib.dfcsr_wake_tasks_pack(k, hist, lat, P, _lib.DFCSR_PREC_FP64, out, out, tbl, 16,
                                   ctypes.byref(batch))
    assert rc != 0 and b"too small" in lib.dfcsr_last_error()


def test_bench_workload_lattices():
    """The three bench workloads: 400-step chicane, 100-step dipole, 1000-step large run."""
    import bench
    from pydfcsr_amd.lattice import Lattice
    for cfg, steps in ((bench.chicane_config(), 400), (bench.dipole_config(), 100),
                       (bench.large_config(), 1000)):
        lat = Lattice(cfg["input_lattice"])
        assert lat.total_steps - 1 == steps
        # whole steps per element (lattice.py:152-173); an element boundary
        # between two positions turns one step into the split partial steps
        # of the run loop (CSR.py:227-236)
        assert steps - 1 <= int(sum(lat.steps_per_element)) <= steps
    big = bench.large_config()
    assert big["input_beam"]["two_gaussian"]["ratio"] == (0.3, 0.7)
    assert big["particle_deposition"]["filter_window"] == 9
    assert big["particle_deposition"]["filter_order"] == 3
