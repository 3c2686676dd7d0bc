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
    assert lib.dfcsr_abi_version() == _lib.ABI_VERSION == 2


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
    H = _lib.History(1, None, 4, 0, 1, 8, 8, 0, 0, 0, 1, 1, 1)
    rc = lib.dfcsr_csr_integrand(H, _lib.Lattice(), 0.0, 0.0, 0.0, 1, 1, 1, 1, 1, 1, None)
    assert rc == -1 and b"slice" in lib.dfcsr_last_error()


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
