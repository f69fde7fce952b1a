"""GPU parity of all_touched=True (AllTouched::burn_line, rust/src/rasterization/
burners.rs:94-247; polygons burn their outline too, burn_geometry.rs:76-166,
219-238; sum and count dedup through the PixelCache, prelude.rs:116-118,
pixel_cache.rs, writers.rs:14-60) against the oracle, bit for bit: every
pixel function, several dtypes, square and non-square pixels, both
item-list forms and both coverage paths, the reference's own all-touched
cases (python/test/test_many.py:237-262, 299-327, 348-374, 407-440, which
compare with GDAL: absent here, so the oracle stands in), and polygons whose
fill reaches past their outline's PixelCache window, where a fill pixel's
wrapped cache index lands on an outline cell and the pixel is skipped."""
import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib, rusterize
from rusterize_amd._lib import Geometry
from tests.cases import EXPLODED_VALUES, EXPLODED_WKT
from tests.test_gpu_parity import assert_same, concat, field_for, line_path_geometry, mixed_geometry

pytestmark = pytest.mark.gpu


def run_at(geom, info, dtype, fun, field, bg=0, layers=None, n_layers=1):
    got = _lib.rasterize(geom, info, dtype, fun, field, None, layers, n_layers, bg, all_touched=True)
    ref = oracle.rasterize(geom, oracle.copy_info(info), dtype, fun, field=field, layer_of_geom=layers,
                           n_layers=n_layers, background=bg, all_touched=True)
    return got, ref


def at_geometry():
    return concat([mixed_geometry(), line_path_geometry()])


@pytest.mark.parametrize("fun", _lib.FUNS)
def test_all_touched_every_function(fun):
    g = at_geometry()
    for xres in (1.0, 0.75):  # 0.75: non-square pixels
        info = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 300, 200))
        for dtype, bg in (("float64", float("nan")), ("int32", 0), ("uint8", 3), ("float32", 0)):
            field = field_for(dtype, g.n_geoms, seed=41)
            field[::9] = 0
            got, ref = run_at(g, info, dtype, fun, field, bg)
            assert_same(got, ref, f"all_touched {fun} {dtype} xres={xres}")


@pytest.mark.parametrize("binning", ["always", "never"])
@pytest.mark.parametrize("wide", ["always", "never"])
def test_all_touched_item_lists_and_coverage_paths(binning, wide, monkeypatch):
    monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    monkeypatch.setenv("RUSTERIZE_WIDE", wide)
    g = at_geometry()
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 300, 200))
    layers = np.random.default_rng(6).integers(0, 3, g.n_geoms).astype(np.uint32)
    for dtype, fun in (("float64", "sum"), ("uint16", "count"), ("int16", "last"), ("float32", "min")):
        field = field_for(dtype, g.n_geoms, seed=42)
        got, ref = run_at(g, info, dtype, fun, field, 0, layers, 3)
        assert_same(got, ref, f"all_touched {binning} {wide} {dtype} {fun}")


def test_reference_all_touched_cases():
    # the exploded test layer at res 0.5, on a custom extent, a custom shape,
    # and both, fun sum in uint8 (python/test/test_many.py:237-262, 299-327,
    # 348-374, 407-440)
    g = Geometry.from_wkt(EXPLODED_WKT)
    field = np.asarray(EXPLODED_VALUES, dtype=np.uint8)
    extent = (-349.0, -507.0, 1.0, 0.0)
    cases = [dict(res=(0.5, 0.5)), dict(res=(1, 1), extent=extent), dict(out_shape=(47, 319)),
             dict(out_shape=(47, 319), extent=extent)]
    for kw in cases:
        r = rusterize(EXPLODED_WKT, dtype="uint8", burn=EXPLODED_VALUES, fun="sum", encoding="numpy",
                      all_touched=True, **kw)
        res = kw.get("res")
        info = _lib.raster_info(g, resolution=(float(res[0]), float(res[1])) if res else None,
                                shape=kw.get("out_shape"), extent=kw.get("extent"))
        ref = oracle.rasterize(g, oracle.copy_info(info), "uint8", "sum", field=field, background=0,
                               all_touched=True)
        assert_same(r, ref, f"reference case {kw}")
        std = rusterize(EXPLODED_WKT, dtype="uint8", burn=EXPLODED_VALUES, fun="sum", encoding="numpy", **kw)
        assert int((r > 0).sum()) >= int((std > 0).sum())  # all-touched burns at least as many pixels


# Polygons with one outline segment inside the raster (0..100 x 0..100, 1 x 1
# pixels) and every other edge outside it, so the PixelCache window is that
# segment's box while the fill covers most of the raster: fill pixels outside
# the window wrap onto cache bits (pixel_cache.rs:38-59).
ALIAS_WKT = [
    # pixel corner (-5, 10)-(10, -5) top left, fill to the lower right
    "POLYGON ((-5 90, 10 105, 200 105, 200 -100, -5 -100, -5 90))",
    # fractional, negative origin (trunc and floor of xmin differ)
    "POLYGON ((-5.5 89.7, 10.7 105.2, 200 105.2, 200 -100, -5.5 -100, -5.5 89.7))",
    # bottom right corner, fill to the upper left
    "POLYGON ((95 -5, 105 5, 105 200, -100 200, -100 -5, 95 -5))",
    # a diagonal through the middle with far-away vertices
    "POLYGON ((40 60, 60 40, 300 40, 300 -300, -300 -300, -300 60, 40 60))",
]


def test_all_touched_cache_aliasing():
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 100, 100))
    union = np.zeros((1, 100, 100), dtype=np.int32)
    for k, wkt in enumerate(ALIAS_WKT):
        sub = Geometry.from_wkt([wkt])
        for dtype, fun in (("float64", "sum"), ("int32", "count"), ("uint8", "sum"), ("int32", "last")):
            got, ref = run_at(sub, info, dtype, fun, np.array([3], dtype=dtype))
            assert_same(got, ref, f"alias polygon {k} {dtype} {fun}")
        union += oracle.rasterize(sub, oracle.copy_info(info), "int32", "any", field=np.ones(1, np.int32),
                                  background=0, all_touched=True)
    g = Geometry.from_wkt(ALIAS_WKT)
    field = np.ones(g.n_geoms, dtype=np.int32)
    got, ref = run_at(g, info, "int32", "count", field)
    assert_same(got, ref, "alias polygons together")
    # the case is meaningful: some fill pixels are skipped (count != union)
    assert (ref != union).any()


def test_all_touched_star_polygons_at_full_size():
    # the 1,000-star workload on 2048 x 2048, float64 sum and last
    g = Geometry.synth(1, 1, 1000, 2048.0, 2048.0, 5.0, 50.0)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 2048, 2048))
    field = np.random.default_rng(5).random(g.n_geoms)
    for fun in ("sum", "last"):
        got, ref = run_at(g, info, "float64", fun, field, float("nan"))
        assert_same(got, ref, f"stars all_touched {fun}")


def test_all_touched_sum_in_a_captured_graph():
    # all-touched sum / count in a captured batch: the PixelCache words of
    # the polygons whose fill leaves their cache window get a fixed capacity
    # at capture; every replay equals the oracle
    from rusterize_amd.device import BatchGraph, Context, DeviceBuffer, DeviceGeometry, Job, batch_rasterize, grid
    ctx = Context(0)
    ga, gm = Geometry.from_wkt(ALIAS_WKT), at_geometry()
    da, dm = DeviceGeometry(ctx, ga), DeviceGeometry(ctx, gm)
    fa = np.arange(1, ga.n_geoms + 1, dtype=np.int32)
    fm = field_for("float64", gm.n_geoms, seed=44)
    bufs = [DeviceBuffer.from_array(ctx, fa), DeviceBuffer.from_array(ctx, fm),
            DeviceBuffer(ctx, 100 * 100 * 4), DeviceBuffer(ctx, 300 * 200 * 8)]
    jobs = [Job(da, grid(100, 100, 0, 0), "int32", "count", bufs[2], field=bufs[0], all_touched=True),
            Job(dm, grid(300, 200, 0, 0), "float64", "sum", bufs[3], field=bufs[1], background=float("nan"),
                all_touched=True)]
    refs = [oracle.rasterize(ga, oracle.copy_info(jobs[0].info), "int32", "count", field=fa, background=0,
                             all_touched=True),
            oracle.rasterize(gm, oracle.copy_info(jobs[1].info), "float64", "sum", field=fm,
                             background=float("nan"), all_touched=True)]
    g = None
    try:
        g = BatchGraph(jobs, headroom=1.0)
        for rep in range(3):
            for j in jobs:
                j.out.upload(np.full(j.out_bytes, 0x5A, np.uint8))
            g.launch()
            g.synchronize()
            for k, j in enumerate(jobs):
                assert_same(j.result(), refs[k], f"replay {rep} job {k}")
            batch_rasterize(ctx, jobs[::-1])  # eager work between replays
            ctx.synchronize()
    finally:
        if g is not None:
            g.close()
        for b in bufs:
            b.free()
        da.free()
        dm.free()
        ctx.close()
