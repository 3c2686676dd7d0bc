"""2D Savitzky-Golay filter (SGolay_filter.sgolay2d, SGolay_filter.py:3-81) on
the CPU: the oracle against the reference's own outputs and the mathematics
that fixes them, the library's host-side kernels against exact rational
arithmetic, and its argument checks (no device needed)."""
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import cpu_ref as O

EINVAL, ERANGE = -1, -3   # DFCSR_EINVAL, DFCSR_ERANGE (include/dfcsr.h)
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_sgolay2d.npz")
MODES = (None, "col", "row", "both")


def _golden():
    d = np.load(GOLD)
    for i, (nx, nz, w, order) in enumerate(d["cases"]):
        want = {m: d[f"{m or 'none'}{i}"] for m in (None, "col", "row")}
        want["both"] = (d[f"bothcol{i}"], d[f"bothrow{i}"])
        yield d[f"z{i}"], int(w), int(order), want


def test_oracle_matches_the_reference_vectors():
    """tests/golden/ref_sgolay2d.npz, written by the reference's sgolay2d:
    every derivative mode, grids down to window_size // 2 + 1 a side."""
    n = 0
    for z, w, order, want in _golden():
        for m in MODES:
            got = O.sgolay2d(z, w, order, m)
            if m == "both":
                for g, r in zip(got, want[m]):
                    assert np.allclose(g, r, rtol=0, atol=1e-15)
            else:
                assert np.allclose(got, want[m], rtol=0, atol=1e-15), (w, order, m)
        n += 1
    assert n == 6


def test_notebook_ramp():
    """test/test.ipynb cell 23: the 'row' slope of [[1,2,3],[4,5,6],[7,8,9]]
    with window 3, order 1 is 1 everywhere (the 'col' slope is 3)."""
    z = np.arange(1.0, 10.0).reshape(3, 3)
    assert np.allclose(O.sgolay2d(z, 3, 1, "row"), 1.0, rtol=0, atol=1e-14)
    assert np.allclose(O.sgolay2d(z, 3, 1, "col"), 3.0, rtol=0, atol=1e-14)


@pytest.mark.parametrize("w,order", [(5, 1), (5, 2), (7, 3), (9, 2)])
def test_polynomials_of_the_fit_degree_are_reproduced(w, order):
    """Away from the padded border a least-squares fit of degree `order`
    reproduces any polynomial of that degree: the smoothed grid is the grid,
    'col' / 'row' are its exact slopes along the first / second axis per
    index step (which fixes the slope rows' sign and orientation)."""
    rng = np.random.default_rng(w * 10 + order)
    nx, nz, h = 23, 19, w // 2
    i, j = np.meshgrid(np.arange(nx, dtype=float), np.arange(nz, dtype=float), indexing="ij")
    coef = {t: rng.normal() for t in O.sgolay2d_fit_terms(order)}
    z = sum(c * i ** a * j ** b for (a, b), c in coef.items())
    dzi = sum(c * a * i ** (a - 1) * j ** b for (a, b), c in coef.items() if a)
    dzj = sum(c * b * i ** a * j ** (b - 1) for (a, b), c in coef.items() if b)
    inner = (slice(h, nx - h), slice(h, nz - h))
    scale = np.abs(z).max()
    assert np.allclose(O.sgolay2d(z, w, order)[inner], z[inner], rtol=0, atol=1e-10 * scale)
    col, row = O.sgolay2d(z, w, order, "both")
    assert np.allclose(col[inner], dzi[inner], rtol=0, atol=1e-10 * scale)
    assert np.allclose(row[inner], dzj[inner], rtol=0, atol=1e-10 * scale)


def test_padding_continues_a_rising_ramp():
    """For data rising along both axes every band and corner of the padding
    is the plane's own continuation (the mirrors about the edge values are
    then odd reflections), so the padded grid is one plane."""
    nx, nz, h = 7, 9, 3
    i, j = np.meshgrid(np.arange(-h, nx + h, dtype=float), np.arange(-h, nz + h, dtype=float),
                       indexing="ij")
    plane = 0.5 + 2.0 * i + 0.75 * j
    Z = O.sgolay2d_pad(plane[h:h + nx, h:h + nz], h)
    assert np.array_equal(Z, plane)
    # a falling edge is mirrored, not continued: the top band of a grid
    # falling along the first axis lies below its first row
    Z = O.sgolay2d_pad(-plane[h:h + nx, h:h + nz], h)
    assert np.all(Z[:h, h:h + nz] < Z[h, h:h + nz])


def _exact_pinv_rows(w, order, rows):
    h = w // 2
    terms = O.sgolay2d_fit_terms(order)
    pts = [(a - h, b - h) for a in range(w) for b in range(w)]
    A = [[Fraction(dx ** ex * dy ** ey) for ex, ey in terms] for dx, dy in pts]
    nt, P = len(terms), len(pts)
    G = [[sum(A[p][r] * A[p][c] for p in range(P)) for c in range(nt)] for r in range(nt)]
    X = [[A[p][r] for p in range(P)] for r in range(nt)]
    for col in range(nt):
        piv = next(r for r in range(col, nt) if G[r][col] != 0)
        G[col], G[piv], X[col], X[piv] = G[piv], G[col], X[piv], X[col]
        d = G[col][col]
        G[col] = [g / d for g in G[col]]
        X[col] = [x / d for x in X[col]]
        for r in range(nt):
            if r != col and G[r][col] != 0:
                f = G[r][col]
                G[r] = [a - f * b for a, b in zip(G[r], G[col])]
                X[r] = [a - f * b for a, b in zip(X[r], X[col])]
    return [np.array([float(x) for x in X[r]]).reshape(w, w) for r in rows]


@pytest.mark.parametrize("w,order", [(3, 1), (5, 2), (7, 3), (9, 8)])
def test_host_kernels_are_the_exact_pseudo_inverse_rows(w, order):
    """The library's kernels equal pinv(A) rows 0, -1, -2 computed in exact
    rational arithmetic ((A^T A)^-1 A^T, A of full column rank) to a few ulps,
    also at order 8 where numpy's pinv is off by 1e-11."""
    from pydfcsr_amd.kernels import sgolay2d_kernels
    r0, r1, r2 = _exact_pinv_rows(w, order, [0, 1, 2])
    k0 = sgolay2d_kernels(w, order, None)
    kc, kr = sgolay2d_kernels(w, order, "both")
    assert np.allclose(k0, r0, rtol=0, atol=4e-15 * np.abs(r0).max())
    assert np.allclose(kc, -r1, rtol=0, atol=4e-15 * np.abs(r1).max())
    assert np.allclose(kr, -r2, rtol=0, atol=4e-15 * np.abs(r2).max())
    assert np.array_equal(sgolay2d_kernels(w, order, "col"), kc)
    assert np.array_equal(sgolay2d_kernels(w, order, "row"), kr)


def test_host_kernels_match_numpy_pinv_at_low_order():
    from pydfcsr_amd.kernels import sgolay2d_kernels
    for w in range(3, 14, 2):
        h = w // 2
        off = np.arange(-h, h + 1, dtype=np.float64)
        dx, dy = np.meshgrid(off, off, indexing="ij")
        for order in range(0, min(w, 5)):
            A = np.stack([dx.ravel() ** a * dy.ravel() ** b
                          for a, b in O.sgolay2d_fit_terms(order)], axis=1)
            P = np.linalg.pinv(A)
            k0 = sgolay2d_kernels(w, order, None)
            assert np.allclose(k0.ravel(), P[0], rtol=0, atol=1e-13 * np.abs(P[0]).max())
            if order:
                kc, kr = sgolay2d_kernels(w, order, "both")
                assert np.allclose(kc.ravel(), -P[1], rtol=0, atol=1e-13 * np.abs(P[1]).max())
                assert np.allclose(kr.ravel(), -P[2], rtol=0, atol=1e-13 * np.abs(P[2]).max())


def test_argument_errors_without_device():
    from pydfcsr_amd import _lib
    lib = _lib.load()
    z = np.zeros((8, 8))
    o1, o2 = np.zeros_like(z), np.zeros_like(z)
    p, q, r = z.ctypes.data, o1.ctypes.data, o2.ctypes.data

    def call(nx, nz, w, order, deriv, out2=r):
        return lib.dfcsr_sgolay2d(p, nx, nz, w, order, deriv, q, out2, None)

    assert call(8, 8, 4, 1, 0) == EINVAL and b"odd" in lib.dfcsr_last_error()
    assert call(8, 8, 1, 0, 0) == EINVAL
    assert call(8, 8, 15, 1, 0) == ERANGE and b"exceeds 13" in lib.dfcsr_last_error()
    assert call(8, 8, 5, 5, 0) == EINVAL and b"order" in lib.dfcsr_last_error()
    assert call(8, 8, 5, 2, 4) == EINVAL and b"derivative" in lib.dfcsr_last_error()
    assert call(8, 8, 5, 0, 1) == EINVAL and b"order >= 1" in lib.dfcsr_last_error()
    assert call(8, 8, 5, 2, 3, None) == EINVAL and b"NULL" in lib.dfcsr_last_error()
    assert call(8, 2, 5, 2, 0) == EINVAL and b"smaller" in lib.dfcsr_last_error()
    assert lib.dfcsr_sgolay2d_kernels(4, 1, 0, q, r) == EINVAL
