"""CPU tests of the C ABI: the library loads and exports every declared entry
point; the grid builder, the pixel-function / dtype parsers, the WKT / WKB
ingestion and the seeded generators behave like the reference (checked
against the oracle and hand-built WKB); device entry points fail loudly
without a GPU."""
import ctypes
import re
import struct
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib
from rusterize_amd._lib import Geometry
from tests.cases import GEOMS, sq

HEADER = Path(__file__).resolve().parent.parent / "include" / "rusterize_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(rz_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) > 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert set(names) <= set(_lib.EXPORTED), set(names) - set(_lib.EXPORTED)


def test_version_and_enums():
    L = _lib.lib()
    assert b"gfx950" in L.rz_version()
    for i, name in enumerate(_lib.FUNS):
        assert _lib.fun_code(name) == i
    with pytest.raises(ValueError, match="Unknown pixel function"):
        _lib.fun_code("bogus")
    out = ctypes.c_int()
    assert L.rz_dtype_parse(b"weird", ctypes.byref(out)) == _lib.RZ_TYPE_ERROR
    for i, name in enumerate(_lib.DTYPES):
        assert L.rz_dtype_parse(name.encode(), ctypes.byref(out)) == 0 and out.value == i
        assert L.rz_dtype_size(i) == np.dtype(name).itemsize


GRID_CASES = [
    dict(shape=(4, 4)), dict(resolution=(1, 1)), dict(resolution=(0.3, 0.7)), dict(shape=(47, 319)),
    dict(shape=(10, 20), extent=(0, 0, 4, 4)), dict(resolution=(0.5, 2.0), extent=(-3.3, -1.1, 7.7, 9.9)),
    dict(resolution=(0.5, 0.5), tap=True), dict(resolution=(0.7, 0.3), tap=True, extent=(-3.3, -1.1, 7.7, 9.9)),
    dict(resolution=(3.0, 3.0), extent=(-1e6, 2e5, 1e6 + 0.1, 7e5)),
]


@pytest.mark.parametrize("case", GRID_CASES)
def test_grid_builder_matches_oracle(case):
    g = Geometry.from_wkt(GEOMS + [sq(-1.25, 3.5, 7.75, 9.0)])
    a = _lib.raster_info(g, **case)
    b = oracle.raster_info(g, **case)
    assert (a.ncols, a.nrows, a.xmin, a.xmax, a.ymin, a.ymax, a.xres, a.yres) == b.as_tuple()


def test_grid_builder_errors():
    g = Geometry.from_wkt([sq(0, 0, 4, 4)])
    for kw, msg in [
        (dict(), "Must set at least one of `shape` or `resolution`"),
        (dict(shape=(4, 4), resolution=(1, 1)), "Shape and resolution are mutually exclusive"),
        (dict(shape=(0, 4)), "Shape values must be > 0."),
        (dict(resolution=(-1, 1)), "Resolution values must be > 0."),
        (dict(resolution=(1, 1), extent=(0, 0, 0, 0)), "Unspecified extent"),
    ]:
        with pytest.raises(RuntimeError, match=re.escape(msg)):
            _lib.raster_info(g, **kw)
        with pytest.raises(oracle.OracleError, match=re.escape(msg)):
            oracle.raster_info(g, **kw)


def test_bbox_ignores_holes():
    # geo's Polygon bounding rect uses the exterior ring only
    poly = "POLYGON ((0 0, 4 0, 4 4, 0 4, 0 0), (1 1, 9 1, 9 2, 1 1))"
    g = Geometry.from_wkt([poly])
    a = _lib.raster_info(g, shape=(4, 4))
    assert (a.xmin, a.xmax) == (0.0, 4.0)


# ------------------------------------------------------------- ingestion

def wkb_point(x, y, le=True):
    e = "<" if le else ">"
    return struct.pack(e + "BIdd", 1 if le else 0, 1, x, y)


def wkb_line(pts, t=2):
    return struct.pack("<BII", 1, t, len(pts)) + b"".join(struct.pack("<dd", *p) for p in pts)


def wkb_polygon(rings):
    out = struct.pack("<BII", 1, 3, len(rings))
    for r in rings:
        out += struct.pack("<I", len(r)) + b"".join(struct.pack("<dd", *p) for p in r)
    return out


def wkb_multi(t, members):
    return struct.pack("<BII", 1, t, len(members)) + b"".join(members)


def same_geometry(a: Geometry, b: Geometry):
    for f in ["geom_item_offsets", "item_kind", "item_ring_offsets", "ring_vertex_offsets", "xy", "ring_flags"]:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)


def test_wkt_and_wkb_agree():
    sq_ring = [(0, 0), (4, 0), (4, 4), (0, 4), (0, 0)]
    hole = [(1, 1), (2, 1), (2, 2), (1, 1)]
    wkb = [
        wkb_point(1.5, 2.5),
        wkb_point(1.5, 2.5, le=False),
        wkb_line([(0, 0), (3, 4), (5, 1)]),
        wkb_polygon([sq_ring, hole]),
        wkb_multi(4, [wkb_point(1, 1), wkb_point(2, 2)]),
        wkb_multi(5, [wkb_line([(0, 0), (1, 1)]), wkb_line([(2, 2), (3, 3)])]),
        wkb_multi(6, [wkb_polygon([sq_ring]), wkb_polygon([[(5, 5), (6, 5), (6, 6), (5, 5)]])]),
        wkb_multi(7, [wkb_point(1, 1), wkb_multi(7, [wkb_polygon([sq_ring])]), wkb_line([(0, 0), (9, 9)])]),
    ]
    wkt = [
        "POINT (1.5 2.5)", "POINT(1.5 2.5)", "LINESTRING (0 0, 3 4, 5 1)",
        "POLYGON ((0 0, 4 0, 4 4, 0 4, 0 0), (1 1, 2 1, 2 2, 1 1))",
        "MULTIPOINT ((1 1), (2 2))", "MULTILINESTRING ((0 0, 1 1), (2 2, 3 3))",
        "MULTIPOLYGON (((0 0, 4 0, 4 4, 0 4, 0 0)), ((5 5, 6 5, 6 6, 5 5)))",
        "GEOMETRYCOLLECTION (POINT (1 1), GEOMETRYCOLLECTION (POLYGON ((0 0, 4 0, 4 4, 0 4, 0 0))), LINESTRING (0 0, 9 9))",
    ]
    same_geometry(Geometry.from_wkb(wkb), Geometry.from_wkt(wkt))


def test_wkt_forms():
    g = Geometry.from_wkt(["multipoint (1 2, 3 4)", "POINT Z (1 2 3)", "LINESTRING M (0 0 5, 1 1 6)",
                           "POLYGON ((0 0, 1 0, 1 1))", "POINT EMPTY", "GEOMETRYCOLLECTION EMPTY",
                           "POLYGON ZM ((0 0 1 2, 1 0 1 2, 1 1 1 2, 0 0 1 2))"])
    assert g.n_geoms == 7
    # an open polygon ring is closed like geo-types Polygon::new does
    r = g.item_ring_offsets[g.geom_item_offsets[3]]
    v0, v1 = g.ring_vertex_offsets[r], g.ring_vertex_offsets[r + 1]
    assert v1 - v0 == 4 and tuple(g.xy[v0]) == tuple(g.xy[v1 - 1])
    # empty geometries own no items but keep their index (field alignment)
    assert g.geom_item_offsets[4] == g.geom_item_offsets[5] == g.geom_item_offsets[6]


def test_wkb_empty_point_is_skipped():
    nan = float("nan")
    g = Geometry.from_wkb([wkb_point(1, 1), wkb_point(nan, nan), wkb_point(2, 2)])
    assert g.n_geoms == 2  # parse_geometry.rs:113-116 filters it out


def test_parse_errors():
    with pytest.raises(ValueError, match="Cannot parse geometry"):
        Geometry.from_wkt(["POLYGON ((0 0, 1 1"])
    with pytest.raises(ValueError, match="Cannot parse geometry"):
        Geometry.from_wkt(["CIRCLE (0 0, 1)"])
    with pytest.raises(ValueError, match="Cannot parse geometry"):
        Geometry.from_wkb([b"\x01\x01\x00"])


def test_wkb_counts_beyond_the_blob_are_rejected():
    # a ring / member / vertex count larger than the remaining bytes is a
    # parse error before anything is allocated for it
    for blob in (b"\x01" + struct.pack("<II", 3, 0xFFFFFFFF),                    # polygon rings
                 b"\x01" + struct.pack("<II", 2, 0xFFFFFFFF),                    # line vertices
                 b"\x01" + struct.pack("<II", 4, 0x7FFFFFFF),                    # multipoint members
                 b"\x01" + struct.pack("<II", 6, 0x10000000),                    # multipolygon members
                 b"\x01" + struct.pack("<II", 7, 0xFFFFFFFF),                    # collection members
                 b"\x01" + struct.pack("<III", 3, 1, 0x40000000)):               # ring vertices
        with pytest.raises(ValueError, match="Cannot parse geometry"):
            Geometry.from_wkb([blob])
    # truncated after a valid header
    with pytest.raises(ValueError, match="truncated"):
        Geometry.from_wkb([wkb_point(1, 2)[:-3]])


def test_nesting_depth_is_capped():
    deep = "GEOMETRYCOLLECTION (" * 100 + "POINT (1 1)" + ")" * 100
    with pytest.raises(ValueError, match="nested deeper"):
        Geometry.from_wkt([deep])
    ok = "GEOMETRYCOLLECTION (" * 10 + "POINT (1 1)" + ")" * 10
    assert Geometry.from_wkt([ok]).n_items == 1
    blob = wkb_point(1, 1)
    for _ in range(100):
        blob = b"\x01" + struct.pack("<II", 7, 1) + blob
    with pytest.raises(ValueError, match="nested deeper"):
        Geometry.from_wkb([blob])


def test_rings_without_items_are_rejected():
    # n_items == 0 with rings / vertices: the per-vertex maps would index an
    # item that does not exist, so both the host copy and the device upload
    # refuse the view before any launch
    g = _lib.RzGeometry()
    gio = np.zeros(2, np.int64)
    rvo = np.array([0, 3], np.int64)
    xy = np.zeros(6, np.float64)
    g.n_geoms, g.n_items, g.n_rings, g.n_vertices = 1, 0, 1, 3
    g.geom_item_offsets = gio.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    g.item_kind = None
    g.item_ring_offsets = None
    g.ring_vertex_offsets = rvo.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    g.xy = xy.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    g.ring_flags = None
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.rz_geometry_from_arrays(ctypes.byref(g), ctypes.byref(h)) == _lib.RZ_VALUE_ERROR
    assert L.rz_device_geometry_upload(None, ctypes.byref(g), ctypes.byref(h)) == _lib.RZ_VALUE_ERROR
    assert "without items" in _lib.last_error()


def test_synth_is_deterministic():
    a = Geometry.synth(2, 5, 200, 100.0, 100.0, 2.0, 50.0, 0.1, 0.2)
    b = Geometry.synth(2, 5, 200, 100.0, 100.0, 2.0, 50.0, 0.1, 0.2)
    same_geometry(a, b)
    assert a.n_geoms == 200 and (a.ring_flags == 1).sum() > 0
    c = Geometry.synth(3, 6, 0, 100.0, 100.0, 30, 40, 10)
    assert c.n_geoms == 80
    kinds = set(c.item_kind.tolist())
    assert kinds == {0, 1, 2}
    f = np.empty(1000, np.uint8)
    _lib.lib().rz_synth_field(0, 0, 3, 1000, 0.0, 1e6, f.ctypes.data)
    f32 = np.empty(1000, np.int32)
    _lib.lib().rz_synth_field(6, 0, 3, 1000, 0.0, 1e6, f32.ctypes.data)
    np.testing.assert_array_equal(f, f32.astype(np.uint8))  # numpy wrap semantics


def device_available() -> bool:
    """True when a HIP device is visible (counting devices does not
    initialise the HIP runtime)."""
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return Path("/dev/kfd").exists()


def test_device_entry_points_fail_loudly_without_gpu():
    if device_available():
        pytest.skip("a GPU is present; the GPU tests cover the device path")
    g = Geometry.from_wkt([sq(0, 0, 4, 4)])
    info = _lib.raster_info(g, shape=(4, 4))
    with pytest.raises(_lib.RusterizeNativeError, match="HIP|device"):
        _lib.rasterize(g, info, "float64", "last", None, 1, None, 1, 0.0)


def test_job_stream_generator():
    # the seeded job stream: spec sizes match the generated geometry, rings are
    # closed stars around their centre, and the distribution bounds hold
    for k in (0, 15, 20, 21, 31, 100):
        sp = _lib.job_spec(4, k)
        assert 256 <= sp["side"] <= 8192 and 10 <= sp["n_poly"] <= 50000 and sp["fun"] in _lib.FUNS
        if sp["n_vertices"] > 300000:
            continue
        g = Geometry.synth_job(4, k)
        assert g.n_geoms == sp["n_poly"] and g.n_vertices == sp["n_vertices"]
        rvo = g.ring_vertex_offsets
        n = np.diff(rvo) - 1
        assert n.min() >= 4 and n.max() <= 4096
        first, last = g.xy[rvo[:-1]], g.xy[rvo[1:] - 1]
        assert np.array_equal(first, last)  # closed rings
        f = _lib.job_field(4, k)
        assert f.shape == (sp["n_poly"],) and f.min() >= 0 and f.max() < 1
    # determinism across calls
    g = Geometry.synth_job(4, 21)
    assert np.array_equal(Geometry.synth_job(4, 21).xy, g.xy)
