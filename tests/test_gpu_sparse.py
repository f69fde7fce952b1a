"""GPU parity of encoding="sparse" (SparseArrayWriter, rust/src/encoding/
writers.rs:85-142; SparseArray, arrays.rs:61-209): the device's triplets
equal the oracle's write trace in order, row, column and value, band after
band; the documented triplets of python/docs/python.md:106-136; and the
dense raster equals the materialised sparse array (build_array) for every
pixel function."""
import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib, rusterize
from rusterize_amd._lib import Geometry
from rusterize_amd.sparse import rasterize_sparse
from tests.cases import GEOMS, VALUES
from tests.test_gpu_parity import assert_same, concat, field_for, line_path_geometry, mixed_geometry

pytestmark = pytest.mark.gpu


def expected_triplets(geom, info, field, layers, n_layers, fun="last", all_touched=False):
    """The oracle's writes in order, regrouped band after band (a stable
    sort by band keeps geometry order inside each band)."""
    tr = oracle.trace(geom, oracle.copy_info(info), fun, all_touched=all_touched)
    g = tr[:, 2].astype(np.int64)
    band = layers[g] if layers is not None else np.zeros(len(g), dtype=np.uint32)
    order = np.argsort(band, kind="stable")
    tr = tr[order]
    counts = np.bincount(band.astype(np.int64), minlength=n_layers if layers is not None else 1)
    return tr[:, 0], tr[:, 1], field[tr[:, 2].astype(np.int64)], counts


def check_sparse(geom, info, dtype, field, layers=None, n_layers=1, what="", fun="last", all_touched=False):
    sp = rasterize_sparse(geom, info, dtype, fun, field, None, layers, n_layers, 0,
                          [f"b{i}" for i in range(n_layers if layers is not None else 1)], all_touched=all_touched)
    rows, cols, vals = sp.triplets()
    er, ec, ev, ecounts = expected_triplets(geom, info, field, layers, n_layers, fun, all_touched)
    assert sp.counts().tolist() == ecounts.tolist(), what
    np.testing.assert_array_equal(rows, er, err_msg=what)
    np.testing.assert_array_equal(cols, ec, err_msg=what)
    assert_same(vals, ev.astype(vals.dtype), what)
    return sp


def test_documented_triplets():
    # python/docs/python.md:106-136 through the drop-in API
    sp = rusterize(GEOMS, res=(1, 1), burn=VALUES, fun="sum", encoding="sparse")
    assert sp.shape() == (1, 131, 361)
    assert sp.extent() == (-180.5, -70.5, 180.5, 60.5)
    df = sp.to_frame()
    assert len(df) == 29363 and list(df.columns) == ["row", "col", "values"]
    assert df[["row", "col"]].values[:5].tolist() == [[6, 40], [6, 41], [6, 42], [7, 39], [7, 40]]
    assert df[["row", "col"]].values[-5:].tolist() == [[39, 286], [39, 287], [39, 288], [39, 289], [39, 290]]
    assert df["values"].values[:5].tolist() == [1.0] * 5 and df["values"].values[-5:].tolist() == [5.0] * 5
    # materialised, the sparse array is the dense raster
    dense = rusterize(GEOMS, res=(1, 1), burn=VALUES, fun="sum", encoding="numpy")
    assert_same(sp.to_numpy(), dense, "documented sum")
    g = Geometry.from_wkt(GEOMS)
    info = _lib.raster_info(g, resolution=(1.0, 1.0))
    check_sparse(g, info, "float64", np.asarray(VALUES, dtype=np.float64), what="documented triplets")


@pytest.mark.parametrize("xres", [1.0, 0.75])
def test_triplets_match_the_write_trace(xres):
    # polygons (holes, parts, slivers), lines (far, closed, banded, zigzag),
    # points, multi-points; 0.75: non-square pixels, one write per pixel per line
    g = concat([mixed_geometry(), line_path_geometry()])
    info = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 300, 200))
    for dtype in ("float64", "int32", "uint8"):
        field = field_for(dtype, g.n_geoms, seed=71)
        check_sparse(g, info, dtype, field, what=f"trace {dtype} xres={xres}")
    layers = np.random.default_rng(3).integers(0, 5, g.n_geoms).astype(np.uint32)
    field = field_for("float32", g.n_geoms, seed=72)
    check_sparse(g, info, "float32", field, layers, 5, what=f"trace 5 bands xres={xres}")


@pytest.mark.parametrize("fun", _lib.FUNS)
def test_dense_equals_materialised_sparse(fun):
    g = concat([mixed_geometry(), line_path_geometry()])
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 300, 200))
    layers = np.random.default_rng(4).integers(0, 3, g.n_geoms).astype(np.uint32)
    for dtype, bg in (("float64", float("nan")), ("int16", 0), ("uint8", 7)):
        field = field_for(dtype, g.n_geoms, seed=73)
        field[::11] = 0
        sp = rasterize_sparse(g, info, dtype, fun, field, None, layers, 3, bg, ["a", "b", "c"])
        dense = _lib.rasterize(g, info, dtype, fun, field, None, layers, 3, bg)
        assert_same(sp.to_numpy(), dense, f"{fun} {dtype}")
        ref = oracle.rasterize(g, oracle.copy_info(info), dtype, fun, field=field, layer_of_geom=layers, n_layers=3,
                               background=bg)
        assert_same(sp.to_numpy(), ref, f"{fun} {dtype} vs oracle")


def test_config1_sparse_at_full_size():
    # the 1,000-star workload: every triplet against the oracle's trace
    g = Geometry.synth(1, 1, 1000, 2048.0, 2048.0, 5.0, 50.0)
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 2048, 2048))
    field = np.random.default_rng(5).random(g.n_geoms)
    sp = check_sparse(g, info, "float64", field, what="stars")
    dense = _lib.rasterize(g, info, "float64", "sum", field, None, None, 1, float("nan"))
    sp._fun, sp._bg = "sum", float("nan")
    assert_same(sp.to_numpy(), dense, "stars sum")


@pytest.mark.parametrize("fun", ["last", "sum", "count"])
@pytest.mark.parametrize("xres", [1.0, 0.75])
def test_all_touched_triplets_match_the_write_trace(fun, xres):
    # all-touched (burners.rs:94-247): lines through the GDAL walker, polygons
    # their outline then their rows; sum / count write an outline pixel once
    # and skip fill pixels the outline's PixelCache contains, also through
    # wrapped cache indices (the alias polygons)
    from tests.test_gpu_all_touched import ALIAS_WKT
    g = concat([mixed_geometry(), line_path_geometry()])
    info = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 300, 200))
    for dtype in ("float64", "int32"):
        field = field_for(dtype, g.n_geoms, seed=81)
        check_sparse(g, info, dtype, field, what=f"all-touched {fun} {dtype} xres={xres}", fun=fun,
                     all_touched=True)
    layers = np.random.default_rng(8).integers(0, 3, g.n_geoms).astype(np.uint32)
    field = field_for("float32", g.n_geoms, seed=82)
    check_sparse(g, info, "float32", field, layers, 3, what=f"all-touched 3 bands {fun}", fun=fun, all_touched=True)
    ga = Geometry.from_wkt(ALIAS_WKT)
    info_a = _lib.raster_info(None, resolution=(xres, 1.0), extent=(0, 0, 100, 100))
    check_sparse(ga, info_a, "int32", np.arange(1, ga.n_geoms + 1, dtype=np.int32), what=f"alias {fun}", fun=fun,
                 all_touched=True)


def test_all_touched_dense_equals_materialised_sparse():
    g = concat([mixed_geometry(), line_path_geometry()])
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, 300, 200))
    for fun in _lib.FUNS:
        field = field_for("float64", g.n_geoms, seed=83)
        sp = rasterize_sparse(g, info, "float64", fun, field, None, None, 1, float("nan"), ["a"], all_touched=True)
        dense = _lib.rasterize(g, info, "float64", fun, field, None, None, 1, float("nan"), all_touched=True)
        assert_same(sp.to_numpy(), dense, f"all-touched {fun}")
