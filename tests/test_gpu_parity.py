"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the
reference's golden rasters.  Bit-exact for every function, dtype and input;
NaN payloads are compared as NaN (a generated NaN's payload is not defined
by the reference's source)."""
import os
import zlib

import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib, rusterize
from rusterize_amd._lib import Geometry
from tests.cases import EXPLODED_VALUES, EXPLODED_WKT, FIELD_RANGE, GEOMS, VALUES, sq

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def assert_same(gpu: np.ndarray, ref: np.ndarray, what: str = ""):
    assert gpu.shape == ref.shape, (gpu.shape, ref.shape)
    assert gpu.dtype == ref.dtype
    if gpu.dtype.kind == "f":
        gn, rn = np.isnan(gpu), np.isnan(ref)
        bad = (gn != rn) | (~gn & (gpu.view(f"u{gpu.itemsize}") != ref.view(f"u{ref.itemsize}")))
    else:
        bad = gpu != ref
    n = int(bad.sum())
    if n:
        idx = np.argwhere(bad)[:5]
        raise AssertionError(f"{what}: {n} pixels differ, e.g. {[(tuple(i), gpu[tuple(i)], ref[tuple(i)]) for i in idx]}")


def field_for(dtype, n, seed):
    lo, hi = FIELD_RANGE[dtype]
    rng = np.random.default_rng(seed)
    if dtype.startswith("float"):
        v = rng.integers(lo, hi, n).astype(dtype) + rng.random(n).astype(dtype)
    else:
        v = rng.integers(lo, hi, n).astype(dtype)
    return v


def run_both(geom, info, dtype, fun, field=None, scalar=1, layers=None, n_layers=1, bg=0):
    got = _lib.rasterize(geom, info, dtype, fun, field, scalar, layers, n_layers, bg)
    ref = oracle.rasterize(geom, oracle.copy_info(info), dtype, fun, field=field, scalar=scalar,
                           layer_of_geom=layers, n_layers=n_layers, background=bg)
    return got, ref


# ------------------------------------------------------------ golden rasters

def test_golden_standard_sum():
    # python/test/test_many.py:228-235 through the drop-in API
    r = rusterize(GEOMS, res=(1, 1), dtype="uint8", burn=VALUES, fun="sum", encoding="numpy").squeeze()
    assert_same(r, np.load(os.path.join(GOLDEN, "standard_output_sum.npy")), "standard_output_sum")


def test_golden_custom_shape():
    # python/test/test_many.py:329-346 (non-square pixels: per-line dedup)
    r = rusterize(EXPLODED_WKT, out_shape=(47, 319), dtype="uint8", burn=EXPLODED_VALUES, fun="sum",
                  encoding="numpy").squeeze()
    assert_same(r, np.load(os.path.join(GOLDEN, "standard_output_sum_custom_shape.npy")), "custom_shape")


# ---------------------------------------------------------------- KATs (R)

def test_r_known_answers():
    # R/rusterize/tests/testthat/test-geometry.R and test-boundary.R on a 4x4 grid
    kw = dict(out_shape=(4, 4), extent=(0, 0, 4, 4), encoding="numpy", background=0)
    m = rusterize(["LINESTRING (0 0, 4 4)"], burn=9, fun="last", **kw)[0]
    want = np.zeros((4, 4))
    for r, c in [(3, 1), (2, 2), (1, 3)]:
        want[r, c] = 9
    assert_same(m, want)
    mp = "MULTIPOLYGON (((0 0, 2 0, 2 2, 0 2, 0 0)), ((2 2, 4 2, 4 4, 2 4, 2 2)))"
    m = rusterize([mp], burn=9, fun="last", **kw)[0]
    want = np.zeros((4, 4))
    want[2:4, 0:2] = 9
    want[0:2, 2:4] = 9
    assert_same(m, want)
    two = [sq(0, 0, 4, 4), sq(0, 0, 4, 4)]
    for fun, val in [("sum", 12), ("min", 5), ("max", 7), ("first", 5), ("last", 7), ("any", 1)]:
        m = rusterize(two, burn=np.array([5.0, 7.0]), fun=fun, **kw)[0]
        assert_same(m, np.full((4, 4), float(val)), fun)
    m = rusterize(two, burn=np.array([1.0, 1.0]), fun="count", **kw)[0]
    assert_same(m, np.full((4, 4), 2.0))
    m = rusterize(["POINT (1.5 1.5)"], burn=9, fun="last", **kw)[0]
    assert m.sum() == 9 and (m > 0).sum() == 1
    assert rusterize([sq(0, 0, 4, 4)], out_shape=(4, 4), encoding="numpy").shape == (1, 4, 4)
    assert rusterize([sq(0, 0, 4, 4)], res=(1, 1), encoding="numpy").shape == (1, 5, 5)


# ------------------------------------------------- all functions x dtypes

MIXED_CACHE = {}


def mixed_geometry():
    if "g" not in MIXED_CACHE:
        stars = Geometry.synth(1, 11, 60, 300.0, 200.0, 3.0, 40.0)
        convex = Geometry.synth(2, 12, 80, 300.0, 200.0, 2.0, 60.0, 0.3, 0.4)
        lines = Geometry.synth(3, 13, 0, 300.0, 200.0, 40, 60, 20)
        wkts = [sq(-10, -10, 30, 20), "LINESTRING (-50 -50, 400 300)", "POINT (299.5 199.5)",
                "MULTIPOINT ((10.5 10.5), (10.5 10.5), (11.5 12.5))",
                "LINESTRING (10 10, 20 10, 20 20, 10 10)", "POLYGON EMPTY",
                "POLYGON ((100 50, 100 50.0000000000001, 150 50, 100 50))"]
        extra = Geometry.from_wkt(wkts)
        MIXED_CACHE["g"] = concat([stars, convex, lines, extra])
    return MIXED_CACHE["g"]


def concat(parts):
    gio, kinds, iro, rvo, xy, fl = [0], [], [0], [0], [], []
    for p in parts:
        base_i, base_r, base_v = len(kinds), len(rvo) - 1, rvo[-1]
        gio.extend((p.geom_item_offsets[1:] + base_i).tolist())
        kinds.extend(p.item_kind.tolist())
        iro.extend((p.item_ring_offsets[1:] + base_r).tolist())
        rvo.extend((p.ring_vertex_offsets[1:] + base_v).tolist())
        xy.append(p.xy)
        fl.extend(p.ring_flags.tolist())
    return Geometry(gio, kinds, iro, rvo, np.concatenate(xy), fl)


@pytest.mark.parametrize("dtype", _lib.DTYPES)
@pytest.mark.parametrize("fun", _lib.FUNS)
def test_mixed_all_functions(dtype, fun):
    g = mixed_geometry()
    info = _lib.raster_info(None, shape=(200, 300), extent=(0, 0, 300, 200))
    field = field_for(dtype, g.n_geoms, seed=zlib.crc32(f"{dtype}{fun}".encode()))
    # make some values equal to the background to exercise the `== bg` quirk
    field[::17] = 0
    for bg in ([0, 7] if not dtype.startswith("float") else [0, float("nan")]):
        got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
        assert_same(got, ref, f"{dtype} {fun} bg={bg}")


@pytest.mark.parametrize("binning", ["always", "never"])
def test_binning_modes_agree(binning, monkeypatch):
    monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    g = mixed_geometry()
    info = _lib.raster_info(None, shape=(200, 300), extent=(0, 0, 300, 200))
    for dtype, fun in [("float64", "sum"), ("uint8", "count"), ("int32", "first"), ("float32", "min")]:
        field = field_for(dtype, g.n_geoms, seed=5)
        got, ref = run_both(g, info, dtype, fun, field=field, bg=0)
        assert_same(got, ref, f"{binning} {dtype} {fun}")


def test_ragged_grids_and_offsets():
    g = mixed_geometry()
    for shape, ext in [((1, 1), (0, 0, 300, 200)), ((333, 517), (-20.5, -13.25, 310.0, 222.0)),
                       ((65, 129), (0, 0, 300, 200)), ((7, 3000), (0, 0, 300, 200))]:
        info = _lib.raster_info(None, shape=shape, extent=ext)
        for dtype, fun in [("float64", "sum"), ("uint8", "last"), ("int16", "max")]:
            field = field_for(dtype, g.n_geoms, seed=3)
            got, ref = run_both(g, info, dtype, fun, field=field, bg=0)
            assert_same(got, ref, f"{shape} {dtype} {fun}")


def test_non_square_pixels_dedup():
    g = mixed_geometry()
    info = _lib.raster_info(None, shape=(70, 300), extent=(0, 0, 300, 200))  # yres != xres
    for fun in ["sum", "count", "last"]:
        field = field_for("float64", g.n_geoms, seed=9)
        got, ref = run_both(g, info, "float64", fun, field=field, bg=float("nan"))
        assert_same(got, ref, fun)


def test_multiband_by():
    g = Geometry.synth(2, 21, 3000, 1000.0, 700.0, 2.0, 120.0, 0.1, 0.2)
    info = _lib.raster_info(None, shape=(700, 1000), extent=(0, 0, 1000, 700))
    layers = np.random.default_rng(1).integers(0, 5, g.n_geoms).astype(np.uint32)
    for dtype, fun in [("uint8", "first"), ("int32", "last"), ("float32", "max"), ("float64", "sum")]:
        field = field_for(dtype, g.n_geoms, seed=2)
        got, ref = run_both(g, info, dtype, fun, field=field, layers=layers, n_layers=5, bg=0)
        assert_same(got, ref, f"by {dtype} {fun}")


# ------------------------------------------------------ workload configs

def test_config1_star_polygons_f64_sum():
    g = Geometry.synth(1, 1, 1000, 2048.0, 2048.0, 5.0, 50.0)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 2048, 2048))
    field = np.empty(g.n_geoms)
    _lib.lib().rz_synth_field(9, 2, 101, g.n_geoms, 0.0, 1.0, field.ctypes.data)
    got, ref = run_both(g, info, "float64", "sum", field=field, bg=float("nan"))
    assert_same(got, ref, "config 1")


@pytest.mark.parametrize("dtype", ["uint8", "int32", "float32"])
def test_config2_reduced(dtype):
    # the convex-polygon workload at 1/100 of the polygons on a 4000^2 grid
    g = Geometry.synth(2, 2, 10000, 4000.0, 4000.0, 2.0, 200.0, 0.1, 0.2)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 4000, 4000))
    f32 = np.empty(g.n_geoms, dtype=np.int32)
    _lib.lib().rz_synth_field(6, 0, 102, g.n_geoms, 0.0, 1e6, f32.ctypes.data)
    field = f32.astype(dtype)  # numpy cast: wraps for uint8
    layers = np.random.default_rng(2).integers(0, 8, g.n_geoms).astype(np.uint32)
    for fun in ["count", "any", "first", "last", "min", "max"]:
        got, ref = run_both(g, info, dtype, fun, field=field, layers=layers, n_layers=8, bg=0)
        assert_same(got, ref, f"config2 {dtype} {fun}")


def test_config3_reduced():
    g = Geometry.synth(3, 3, 0, 2000.0, 2000.0, 5000, 20000, 500)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 2000, 2000))
    field = np.empty(g.n_geoms, dtype=np.float32)
    _lib.lib().rz_synth_field(8, 1, 103, g.n_geoms, 0.0, 1.0, field.ctypes.data)
    for fun in ["max", "sum"]:
        got, ref = run_both(g, info, "float32", fun, field=field, bg=float("nan"))
        assert_same(got, ref, f"config3 {fun}")


def zigzag(x0, y0, w, h, n, seed):
    rng = np.random.default_rng(seed)
    xs = x0 + rng.uniform(0, w, n)
    ys = y0 + rng.uniform(0, h, n)
    return "LINESTRING (" + ", ".join(f"{float(x)!r} {float(y)!r}" for x, y in zip(xs, ys)) + ")"


LINE_PATH_CACHE = {}


def line_path_geometry():
    """Geometry for the line / point tile kernel's paths: zigzag lines whose
    cells in one strip overflow the touched-pixel list (the dense apply),
    banded lines over many strips, long runs of one-vertex points with
    repeated cells, multi-points with repeated vertices, and polygons in
    between (each of them in geometry order)."""
    if "g" not in LINE_PATH_CACHE:
        rng = np.random.default_rng(21)
        wkts = []
        wkts.append(zigzag(3.0, 100.0, 60.0, 14.0, 600, 1))      # one strip, > 256 cells
        wkts.append(zigzag(70.0, 10.0, 120.0, 110.0, 400, 2))    # banded over many strips
        wkts += [f"POINT ({x:.3f} {y:.3f})" for x, y in rng.uniform(0, 64, (300, 2)) // 4 * 4 + 0.5]
        wkts.append(sq(20, 20, 50, 45))
        wkts += [f"POINT ({x:.3f} {y:.3f})" for x, y in rng.uniform(0, 64, (200, 2)) // 4 * 4 + 0.5]
        wkts.append("MULTIPOINT ((5.5 5.5), (5.5 5.5), (6.5 5.5), (5.5 5.5))")
        wkts.append(zigzag(0.0, 0.0, 250.0, 120.0, 3000, 3))    # banded, heavy overlap
        wkts.append(zigzag(100.0, 60.0, 30.0, 6.0, 800, 4))
        LINE_PATH_CACHE["g"] = Geometry.from_wkt(wkts)
    return LINE_PATH_CACHE["g"]


@pytest.mark.parametrize("binning", [None, "always", "never"])
def test_line_point_kernel_paths(binning, monkeypatch):
    if binning:
        monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    g = line_path_geometry()
    for dtype, fun, bg in [("float64", "sum", float("nan")), ("int32", "count", 0), ("uint8", "sum", 0),
                           ("float32", "first", float("nan")), ("int16", "last", 0), ("float64", "max", 0),
                           ("uint16", "min", 0), ("int64", "any", 0)]:
        field = field_for(dtype, g.n_geoms, seed=31)
        field[::7] = 0
        for xres in (1.0, 0.75):  # 0.75: non-square pixels, one write per pixel per line
            info = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 250, 125))
            got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
            assert_same(got, ref, f"line paths {binning} {dtype} {fun} xres={xres}")



GROUP_CACHE = {}


def line_group_geometry():
    """Runs of short line items that meet one strip together (walked as one
    group, each item writing its own marks or counters), overlapping so the
    item order decides first / last / sum, with one-vertex points between
    them (on a cell of a line of the group, repeated, or elsewhere); in
    between a line of more than 255 segments in one strip (walked alone
    under sum and count), two-vertex lines, multi-points with repeated
    cells, one-vertex points and a polygon, each ending a run."""
    if "g" not in GROUP_CACHE:
        rng = np.random.default_rng(41)
        wkts = []
        for run in range(12):
            x0, y0 = rng.uniform(0, 180, 2)
            for _ in range(int(rng.integers(2, 14))):
                z = zigzag(x0, y0, 20.0, 10.0, int(rng.integers(2, 9)), int(rng.integers(1 << 30)))
                wkts.append(z)
                # one-vertex points between the lines: on a vertex of the
                # line just drawn (its cell is written by a line of the
                # group), repeated, or elsewhere in the run's area
                r = rng.random()
                if r < 0.3:
                    vx, vy = z[len("LINESTRING ("):].split(",")[0].split()
                    wkts.append(f"POINT ({vx} {vy})")
                    if r < 0.1:
                        wkts.append(f"POINT ({vx} {vy})")
                elif r < 0.6:
                    px, py = rng.uniform(0, 20, 2)
                    wkts.append(f"POINT ({x0 + px} {y0 + py / 2})")
            if run == 3:
                wkts.append(zigzag(x0, y0 + 2.0, 60.0, 12.0, 400, 7))   # > 255 segments in one strip
            if run == 5:
                wkts.append(f"MULTIPOINT (({x0 + 1.5} {y0 + 1.5}), ({x0 + 1.5} {y0 + 1.5}), ({x0 + 2.5} {y0 + 1.5}))")
            if run == 7:
                wkts.append(f"POINT ({x0 + 3.5} {y0 + 2.5})")
            if run == 9:
                wkts.append(sq(x0, y0, x0 + 12, y0 + 7))
            if run == 10:
                wkts.append(f"LINESTRING ({x0} {y0}, {x0 + 40.5} {y0 + 3.25})")
        GROUP_CACHE["g"] = Geometry.from_wkt(wkts)
    return GROUP_CACHE["g"]


@pytest.mark.parametrize("binning", [None, "always", "never"])
def test_line_item_groups(binning, monkeypatch):
    if binning:
        monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    g = line_group_geometry()
    for dtype, fun, bg in [("float32", "sum", float("nan")), ("float64", "sum", 0), ("int32", "count", 0),
                           ("uint8", "sum", 0), ("uint16", "count", 0), ("float32", "first", float("nan")),
                           ("int32", "last", 0), ("float64", "max", float("nan")), ("uint8", "min", 0),
                           ("int16", "any", 0)]:
        field = field_for(dtype, g.n_geoms, seed=43)
        for xres in (1.0, 0.75):  # 0.75: non-square pixels, one write per pixel per line
            info = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 220, 210))
            got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
            assert_same(got, ref, f"line groups {binning} {dtype} {fun} xres={xres}")
            if xres == 1.0 and fun in ("sum", "last"):
                got = _lib.rasterize(g, info, dtype, fun, field, None, None, 1, bg, all_touched=True)
                ref = oracle.rasterize(g, oracle.copy_info(info), dtype, fun, field=field, background=bg,
                                       all_touched=True)
                assert_same(got, ref, f"line groups all_touched {binning} {dtype} {fun}")

def sum_error_bound(g, info, dtype, field):
    """8 n eps sum|v| per pixel: the rounding of n additions in any order,
    from the oracle's count of writes and its sum of |v|."""
    eps = float(np.finfo(np.dtype(dtype)).eps)
    n = oracle.rasterize(g, oracle.copy_info(info), "float64", "count", scalar=1, background=0)
    s = oracle.rasterize(g, oracle.copy_info(info), "float64", "sum", field=np.abs(field).astype(np.float64),
                         background=0)
    return 8.0 * n * eps * s


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_unordered_sum_mode_within_its_bound(dtype):
    # RZ_SUM_UNORDERED on the lines / points / collections workload (C3
    # reduced): every pixel within 8 n eps sum|v| of the reference's ordered
    # fold (f64: below 1e-12 relative wherever the sum does not cancel), the
    # same written / unwritten pixels; the ordered mode stays bit-exact
    g = Geometry.synth(3, 3, 0, 2000.0, 2000.0, 5000, 20000, 500)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 2000, 2000))
    field = np.random.default_rng(17).standard_normal(g.n_geoms).astype(dtype)
    ref = oracle.rasterize(g, oracle.copy_info(info), dtype, "sum", field=field, background=float("nan"))
    ordered = _lib.rasterize(g, info, dtype, "sum", field, None, None, 1, float("nan"), sum_mode=0)
    assert_same(ordered, ref, f"ordered sum {dtype}")
    got = _lib.rasterize(g, info, dtype, "sum", field, None, None, 1, float("nan"), sum_mode=1)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    err = np.abs(got[ok].astype(np.float64) - ref[ok].astype(np.float64))
    bound = sum_error_bound(g, info, dtype, field)[ok]
    assert (err <= bound).all(), float((err - bound).max())
    if dtype == "float64":
        # where the sum keeps a tenth of sum|v| (no heavy cancellation): 1e-12 relative
        absum = sum_error_bound(g, info, dtype, field)[ok] / (8.0 * np.finfo(np.float64).eps)
        n = oracle.rasterize(g, oracle.copy_info(info), "float64", "count", scalar=1, background=0)[ok]
        keep = np.abs(ref[ok]) >= 0.1 * absum / np.maximum(n, 1)
        assert (err[keep] <= 1e-12 * np.abs(ref[ok][keep])).all()


def test_far_line_segments():
    # a segment whose floored end point lies beyond 2^30 pixels is walked in
    # 128 bits from its world coordinates; its line is never banded
    wkts = ["LINESTRING (-1100000000.25 40.5, 240.5 44.5, 100.5 100.5)", "POINT (3.5 80.5)",
            zigzag(0.0, 0.0, 250.0, 120.0, 200, 5)]
    g = Geometry.from_wkt(wkts)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 250, 125))
    field = field_for("float64", g.n_geoms, seed=41)
    got, ref = run_both(g, info, "float64", "sum", field=field, bg=float("nan"))
    assert_same(got, ref, "far segment")


@pytest.mark.parametrize("binning", [None, "always", "never"])
def test_long_rings_edge_bands(binning, monkeypatch):
    # many-vertex rings spanning many tile rows take the banded-edge path
    # (edges binned by 64-row tile row); overlapping stars, holes and parts
    if binning:
        monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    stars = Geometry.synth(4, 51, 150, 1500.0, 1000.0, 20.0, 450.0, 65.0, 3000.0)
    holed = Geometry.synth(2, 52, 40, 1500.0, 1000.0, 300.0, 900.0, 0.5, 1.0)
    g = concat([stars, holed])
    info = _lib.raster_info(None, shape=(1000, 1500), extent=(0, 0, 1500, 1000))
    for dtype, fun, bg in [("float64", "sum", float("nan")), ("int32", "last", 0), ("uint8", "count", 0),
                           ("float32", "min", float("nan"))]:
        field = field_for(dtype, g.n_geoms, seed=7)
        got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
        assert_same(got, ref, f"bands {binning} {dtype} {fun}")


def test_heavy_tiles_precomputed_coverage():
    # thousands of overlapping polygons on a small raster: every tile holds
    # more than 32 items, so polygon coverage comes from the parallel pre-pass
    dense = Geometry.synth(4, 61, 3000, 320.0, 300.0, 2.0, 40.0, 4.0, 600.0)
    lines = Geometry.synth(3, 62, 0, 320.0, 300.0, 200, 300, 20)
    g = concat([dense, lines])
    info = _lib.raster_info(None, shape=(300, 320), extent=(0, 0, 320, 300))
    for dtype, fun, bg in [("float64", "sum", float("nan")), ("int32", "last", 0), ("uint8", "count", 0),
                           ("float32", "max", float("nan")), ("int16", "first", 0), ("uint32", "min", 0)]:
        field = field_for(dtype, g.n_geoms, seed=8)
        got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
        assert_same(got, ref, f"heavy {dtype} {fun}")


# ------------------------------------------------ polygon coverage layouts

def polygon_rings(rings_per_geom):
    """A Geometry of one polygon item per entry; each entry is a list of rings
    (arrays of (x, y) vertices, closed by the caller)."""
    gio, kinds, iro, rvo, xy = [0], [], [0], [0], []
    for rings in rings_per_geom:
        for r in rings:
            r = np.asarray(r, dtype=np.float64).reshape(-1, 2)
            xy.append(r)
            rvo.append(rvo[-1] + len(r))
        kinds.append(0)
        iro.append(iro[-1] + len(rings))
        gio.append(len(kinds))
    return Geometry(gio, kinds, iro, rvo, np.concatenate(xy))


def degenerate_geometry():
    """Rings with NaN (of either sign), infinite and huge coordinates: NaN
    crossings (sorted first or dropped), rows with an odd crossing count,
    skipped segments and edges clamped to the raster, next to ordinary rings
    they overlap."""
    nan, inf = float("nan"), float("inf")
    star = [(150 + (60 if k % 2 == 0 else 25) * np.cos(k * np.pi / 40),
             100 + (60 if k % 2 == 0 else 25) * np.sin(k * np.pi / 40)) for k in range(80)]
    star.append(star[0])
    return polygon_rings([
        [[(10, 10), (80, 15), (nan, 60), (20, 90), (10, 10)]],           # NaN vertex
        [[(100, 5), (140, 5), (inf, 40), (100, 40), (100, 5)]],          # +inf vertex
        [[(200, 20), (260, 30), (-inf, 70), (200, 20)]],                 # -inf vertex
        [[(30, 120), (1e300, 130), (60, 190), (30, 120)]],               # huge x
        [[(250, -1e300), (290, 150), (240, 150), (250, -1e300)]],        # huge y
        [[(120, 150), (125, nan), (180, 190), (120, 150)]],              # NaN y
        [[(40, 40), (inf, 45), (inf, 80), (45, 85), (40, 40)]],          # inf - inf slope
        [[(160, 100), (-nan, 130), (190, 170), (160, 100)]],             # negative NaN x
        [star],
        [[(5, 5), (295, 5), (295, 195), (5, 195), (5, 5)],               # frame with a hole
         [(100, 60), (200, 60), (200, 140), (100, 140), (100, 60)]],
    ])


def test_degenerate_coordinates():
    g = degenerate_geometry()
    info = _lib.raster_info(None, shape=(200, 300), extent=(0, 0, 300, 200))
    for dtype, fun, bg in [("float64", "sum", float("nan")), ("int32", "count", 0), ("uint8", "last", 0)]:
        field = field_for(dtype, g.n_geoms, seed=9)
        got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
        assert_same(got, ref, f"degenerate {dtype} {fun}")


@pytest.mark.parametrize("wide,layout", [("never", None), ("always", "rows"), ("always", "edges"),
                                         (None, None)])
def test_coverage_paths_agree(wide, layout, monkeypatch):
    # every polygon through one coverage path: in-tile (never wide), or the
    # wide pre-pass in one layout (row-major or edge-major chunks), on
    # ordinary, degenerate, banded and heavy-tile inputs, in both item-list
    # forms
    for var, val in (("RUSTERIZE_WIDE", wide), ("RUSTERIZE_COVERAGE", layout)):
        if val:
            monkeypatch.setenv(var, val)
    stars = Geometry.synth(4, 51, 60, 1500.0, 1000.0, 20.0, 450.0, 65.0, 3000.0)
    dense = Geometry.synth(4, 61, 1500, 320.0, 300.0, 2.0, 40.0, 4.0, 600.0)
    cases = [(mixed_geometry(), (200, 300)), (degenerate_geometry(), (200, 300)),
             (stars, (1000, 1500)), (dense, (300, 320))]
    for g, (h, w) in cases:
        info = _lib.raster_info(None, shape=(h, w), extent=(0, 0, w, h))
        for binning in [None, "always", "never"]:
            if binning:
                monkeypatch.setenv("RUSTERIZE_BINNING", binning)
            else:
                monkeypatch.delenv("RUSTERIZE_BINNING", raising=False)
            for dtype, fun, bg in [("float64", "sum", float("nan")), ("uint8", "count", 0),
                                   ("int32", "first", 0)]:
                field = field_for(dtype, g.n_geoms, seed=11)
                got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
                assert_same(got, ref, f"{wide} {layout} {binning} {g.n_geoms} geoms {dtype} {fun}")


@pytest.mark.parametrize("binning", [None, "always"])
def test_wide_masks_past_their_cap(binning, monkeypatch):
    # 300 polygons over a 512 x 512 raster, every one on the wide pre-pass:
    # their row masks (about 1.2M words) exceed the cap (the larger of 2 bytes
    # per output pixel and 2^20 words), so the items past it compute their
    # coverage per tile; the raster is still the oracle's (a single call
    # sizes the masks for the cap without reading totals back; binned, from
    # the totals)
    monkeypatch.setenv("RUSTERIZE_WIDE", "always")
    if binning:
        monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    rng = np.random.default_rng(12)
    wkts = []
    for _ in range(300):
        x0, y0 = rng.uniform(-20, 40, 2)
        x1, y1 = rng.uniform(470, 530, 2)
        wkts.append(f"POLYGON (({x0} {y0}, {x1} {y0 + rng.uniform(-10, 10)}, {x1} {y1}, "
                    f"{x0 + rng.uniform(-10, 10)} {y1}, {x0} {y0}))")
    g = Geometry.from_wkt(wkts)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 512, 512))
    for dtype, fun, bg in (("float64", "sum", float("nan")), ("uint8", "count", 0)):
        field = field_for(dtype, g.n_geoms, seed=19)
        got, ref = run_both(g, info, dtype, fun, field=field, bg=bg)
        assert_same(got, ref, f"wide past cap {binning} {dtype} {fun}")


def test_item_box_merged_across_waves():
    # a polygon whose left and right edges lie outside the raster, its
    # vertices split between two waves of the vertex pass (items before it
    # shift it across a 64-vertex boundary): each wave's part of its box has
    # an empty column range, and only merged do they span the raster
    big = "POLYGON ((-12.4 0.8, 517.2 0.7, 517.2 497.3, -6.7 497.3, -12.4 0.8))"
    for shift in (0, 12, 24, 25, 26, 38):  # 25, 38: the right edge in one wave, the left edge in the next
        wkts = [sq(10 + k, 10, 11 + k, 11) for k in range(shift)] + [big]
        g = Geometry.from_wkt(wkts)
        info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 512, 512))
        field = np.arange(1, g.n_geoms + 1, dtype=np.float64)
        got, ref = run_both(g, info, "float64", "count", field=field, bg=0)
        assert_same(got, ref, f"shift {shift}")
