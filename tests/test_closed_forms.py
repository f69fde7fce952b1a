"""The two closed forms the tile kernel relies on, evaluated on the CPU by the
library's host hooks and checked by brute force against the reference's
literal procedures:

* the Bresenham closed form (walk_window) against the literal loop of
  Standard::burn_line (rust/src/rasterization/burners.rs:53-84, restated in
  the oracle), on random segments and windows;
* the sort-free row coverage (crossing-count parity) against the sort /
  chunks_exact(2) / floor(x+0.5) span procedure of burn_polygon
  (burners.rs:302-316), including duplicates, +-inf and both NaN signs.
"""
import ctypes
import struct

import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib
from rusterize_amd._lib import Geometry

POINTER = ctypes.POINTER


def walk(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1, cap=1 << 16):
    xs = np.empty(cap, np.int64)
    ys = np.empty(cap, np.int64)
    n = _lib.lib().rz_debug_walk_cells(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1,
                                       xs.ctypes.data_as(POINTER(ctypes.c_int64)),
                                       ys.ctypes.data_as(POINTER(ctypes.c_int64)), cap)
    assert 0 <= n <= cap  # (-2: the index-stepped walk disagrees)
    return list(zip(xs[:n].tolist(), ys[:n].tolist()))


def reference_cells(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1):
    xs, ys = oracle.bresenham_cells(ix0, iy0, ix1, iy1)
    return [(x, y) for x, y in zip(xs.tolist(), ys.tolist()) if X0 <= x < X1 and Y0 <= y < Y1]


@pytest.mark.parametrize("seed", range(6))
def test_bresenham_closed_form_random(seed):
    rng = np.random.default_rng(seed)
    span = [8, 40, 300, 3000, 50, 200][seed]
    for _ in range(3000):
        ix0, iy0, ix1, iy1 = (int(v) for v in rng.integers(-span, span, 4))
        X0 = int(rng.integers(-span, span))
        Y0 = int(rng.integers(-span, span))
        X1 = X0 + int(rng.integers(0, 2 * span))
        Y1 = Y0 + int(rng.integers(0, 2 * span))
        assert walk(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1) == reference_cells(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1)


def test_bresenham_exhaustive_small():
    # every segment between cells of a 9x9 block, whole-plane window
    big = 1 << 20
    for ix0 in range(-4, 5):
        for iy0 in range(-4, 5):
            for ix1 in range(-4, 5):
                for iy1 in range(-4, 5):
                    assert walk(ix0, iy0, ix1, iy1, -big, big, -big, big) == \
                        reference_cells(ix0, iy0, ix1, iy1, -big, big, -big, big)


def test_bresenham_far_endpoints_128bit():
    # segments from far outside: the window bounds the work; the __int128
    # path is used beyond 2^30 and must agree with the literal loop where
    # the loop is affordable (checked on windows near the end points)
    rng = np.random.default_rng(7)
    for _ in range(200):
        ix0 = int(rng.integers(-(1 << 31) - 5000, -(1 << 31) + 5000))
        iy0 = int(rng.integers(-3000, 3000))
        ix1 = ix0 + int(rng.integers(1, 4000))
        iy1 = iy0 + int(rng.integers(-4000, 4000))
        X0 = ix0 + int(rng.integers(0, 100))
        Y0 = min(iy0, iy1) + int(rng.integers(0, 100))
        assert walk(ix0, iy0, ix1, iy1, X0, X0 + 64, Y0, Y0 + 16) == \
            reference_cells(ix0, iy0, ix1, iy1, X0, X0 + 64, Y0, Y0 + 16)



def test_bresenham_long_segments_wide_numerators():
    # segments of 1e5-3e5 major steps cut by 64 x 16 windows along them: the
    # window set-up's numerators (2n * j) pass 2^32, beyond its 32-bit
    # divisions, and must still give the literal loop's cells
    rng = np.random.default_rng(11)
    for _ in range(12):
        n = int(rng.integers(100_000, 300_000))
        m = int(rng.integers(0, n + 1))
        du, dw = (n, m) if rng.integers(2) else (m, n)
        ix0, iy0 = (int(v) for v in rng.integers(-1000, 1000, 2))
        ix1 = ix0 + du * (1 if rng.integers(2) else -1)
        iy1 = iy0 + dw * (1 if rng.integers(2) else -1)
        xs, ys = oracle.bresenham_cells(ix0, iy0, ix1, iy1)
        for c in rng.integers(0, len(xs), 8):
            X0 = int(xs[c]) - int(rng.integers(0, 64))
            Y0 = int(ys[c]) - int(rng.integers(0, 16))
            X1, Y1 = X0 + 64, Y0 + 16
            inside = (xs >= X0) & (xs < X1) & (ys >= Y0) & (ys < Y1)
            assert walk(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1) == \
                list(zip(xs[inside].tolist(), ys[inside].tolist()))

# ---------------------------------------------------------- row coverage

def total_key(x):
    b = struct.unpack("<q", struct.pack("<d", x))[0]
    return b ^ ((b >> 63) & 0x7FFFFFFFFFFFFFFF)


def spans_reference(xs, ncols):
    """burn_polygon's span procedure for one row (burners.rs:302-316)."""
    xs = sorted(xs, key=total_key)
    cov = np.zeros(ncols, dtype=np.uint8)

    def col(x):
        f = np.floor(x + 0.5)
        if f < 0.0:
            f = 0.0
        if f > ncols:
            f = float(ncols)
        return 0 if not f > 0 else int(f)

    for i in range(0, len(xs) - 1, 2):
        a, b = col(xs[i]), col(xs[i + 1])
        cov[a:b] = 1
    return cov


def coverage_hook(xs, ncols, C0, width):
    arr = np.ascontiguousarray(xs, dtype=np.float64)
    out = np.zeros(width, dtype=np.uint8)
    rc = _lib.lib().rz_debug_row_coverage(arr.ctypes.data_as(POINTER(ctypes.c_double)), len(arr), ncols, C0,
                                          width, out.ctypes.data_as(POINTER(ctypes.c_uint8)))
    assert rc == 0
    return out


NEG_NAN = struct.unpack("<d", struct.pack("<Q", 0xFFF8000000000000))[0]
POS_NAN = struct.unpack("<d", struct.pack("<Q", 0x7FF8000000000000))[0]


@pytest.mark.parametrize("width", [64, 128, 256])
@pytest.mark.parametrize("seed", range(4))
def test_row_coverage_matches_sorted_spans(width, seed):
    rng = np.random.default_rng(seed)
    ncols = int(rng.integers(1, 600))
    specials = [NEG_NAN, POS_NAN, float("inf"), -float("inf"), -0.0, 0.0, 0.5, float(ncols), ncols - 0.5]
    for _ in range(1500):
        n = int(rng.integers(0, 12))
        xs = list(rng.uniform(-50, ncols + 50, n))
        if rng.random() < 0.5 and xs:
            xs += [xs[int(rng.integers(0, len(xs)))]] * int(rng.integers(1, 3))  # duplicates
        if rng.random() < 0.5:
            xs += list(np.floor(rng.uniform(-5, ncols + 5, int(rng.integers(1, 4)))) + 0.5)  # ties
        if rng.random() < 0.3:
            xs += [specials[int(rng.integers(0, len(specials)))] for _ in range(int(rng.integers(1, 3)))]
        ref = spans_reference(xs, ncols)
        for C0 in range(0, ncols, width):
            got = coverage_hook(xs, ncols, C0, width)
            want = np.zeros(width, dtype=np.uint8)
            seg = ref[C0:C0 + width]
            want[:len(seg)] = seg
            np.testing.assert_array_equal(got[:len(seg)], want[:len(seg)], err_msg=f"xs={xs} C0={C0}")


# ------------------------------------------------ all-touched walker (device form)

def _at_cells(x0, y0, x1, y1, ncols, nrows, X0, X1, Y0, Y1, cap=1 << 16):
    xs = np.empty(cap, dtype=np.int64)
    ys = np.empty(cap, dtype=np.int64)
    n = _lib.lib().rz_debug_walk_all_touched(x0, y0, x1, y1, ncols, nrows, X0, X1, Y0, Y1,
                                              xs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                              ys.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), cap)
    assert n >= 0
    return list(zip(ys[:n].tolist(), xs[:n].tolist()))


def test_all_touched_walker_equals_the_oracle():
    # the tile kernel's walker (walk_all_touched) against the oracle's
    # restatement of burners.rs:115-246, cell for cell and in order, on whole
    # rasters and on windows of them
    rng = np.random.default_rng(23)
    ncols, nrows = 37, 29
    info = oracle.make_info(ncols, nrows, 0.0, float(ncols), 0.0, float(nrows), 1.0, 1.0)
    cases = []
    for _ in range(400):
        X0, X1 = rng.uniform(-8, ncols + 8, 2)
        Y0, Y1 = rng.uniform(-8, nrows + 8, 2)
        kind = rng.integers(0, 6)
        if kind == 1:
            X1 = X0 + rng.uniform(-0.009, 0.009)      # vertical branch
        elif kind == 2:
            Y1 = Y0 + rng.uniform(-0.009, 0.009)      # horizontal branch
        elif kind == 3:
            X0, Y0 = float(np.round(X0)), float(np.round(Y0))  # grid points
            X1, Y1 = float(np.round(X1)), float(np.round(Y1))
        elif kind == 4:
            X1 = X0 + (Y1 - Y0)                       # 45 degrees
        cases.append((float(X0), float(Y0), float(X1), float(Y1)))
    for X0, Y0, X1, Y1 in cases:
        g = Geometry.from_wkt([f"LINESTRING ({X0!r} {Y0!r}, {X1!r} {Y1!r})"])
        want = [tuple(c) for c in oracle.trace(g, info, "last", all_touched=True)[:, :2].tolist()]
        px0, py0, px1, py1 = X0, float(nrows) - Y0, X1, float(nrows) - Y1
        if not (min(px0, px1) < ncols and max(px0, px1) >= 0 and min(py0, py1) < nrows and max(py0, py1) >= 0):
            assert want == []  # extract_line's box test drops it (edges.rs:130)
            continue
        assert _at_cells(px0, py0, px1, py1, ncols, nrows, 0, ncols, 0, nrows) == want, (X0, Y0, X1, Y1)
        for _ in range(3):
            wx0, wy0 = int(rng.integers(0, ncols)), int(rng.integers(0, nrows))
            wx1, wy1 = wx0 + int(rng.integers(1, 20)), wy0 + int(rng.integers(1, 20))
            sub = [(r, c) for r, c in want if wx0 <= c < wx1 and wy0 <= r < wy1]
            assert _at_cells(px0, py0, px1, py1, ncols, nrows, wx0, wx1, wy0, wy1) == sub, (X0, Y0, X1, Y1)
