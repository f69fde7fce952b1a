"""ctypes binding of librusterize_hip.so (the C ABI in include/rusterize_hip.h).

The shared library is built in-tree by ``rusterize_amd.build``; importing this
module never falls back to anything else: a missing library raises at load,
and a missing GPU raises RuntimeError from the device calls.
"""
from __future__ import annotations

import ctypes
import os
import threading
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("RUSTERIZE_LIBRARY", Path(__file__).resolve().parent / "librusterize_hip.so"))

DTYPES = ("uint8", "uint16", "uint32", "uint64", "int8", "int16", "int32", "int64", "float32", "float64")
FUNS = ("sum", "first", "last", "min", "max", "count", "any")
ITEM_POLYGONS, ITEM_LINES, ITEM_POINTS = 0, 1, 2

RZ_OK, RZ_VALUE_ERROR, RZ_RUNTIME_ERROR, RZ_TYPE_ERROR, RZ_DEVICE_ERROR, RZ_PARSE_ERROR, RZ_CAPACITY_ERROR = range(7)


class RzGeometry(ctypes.Structure):
    _fields_ = [
        ("n_geoms", c_int64), ("n_items", c_int64), ("n_rings", c_int64), ("n_vertices", c_int64),
        ("geom_item_offsets", POINTER(c_int64)), ("item_kind", POINTER(c_int32)),
        ("item_ring_offsets", POINTER(c_int64)), ("ring_vertex_offsets", POINTER(c_int64)),
        ("xy", POINTER(c_double)), ("ring_flags", POINTER(c_uint8)),
    ]


class RzRasterInfo(ctypes.Structure):
    _fields_ = [
        ("ncols", c_uint64), ("nrows", c_uint64), ("xmin", c_double), ("xmax", c_double),
        ("ymin", c_double), ("ymax", c_double), ("xres", c_double), ("yres", c_double), ("epsg", c_int32),
    ]


class RzJob(ctypes.Structure):
    """rz_job: one job of rz_batch_rasterize (device pointers, raw value bits)."""
    _fields_ = [
        ("geometry", c_void_p), ("info", RzRasterInfo), ("dtype", c_int), ("fun", c_int),
        ("field_dev", c_void_p), ("scalar_bits", c_uint64), ("layer_dev", c_void_p), ("n_layers", c_uint32),
        ("background_bits", c_uint64), ("out_dev", c_void_p), ("sum_mode", c_int), ("all_touched", c_int),
    ]


class RzServerStats(ctypes.Structure):
    _fields_ = [(name, c_uint64) for name in ("queued", "in_flight_batches", "batches", "completed",
                                              "memory_used", "memory_peak", "memory_budget",
                                              "alloc_us_sum", "run_us_sum", "wait_us_sum",
                                              "alloc_us_max", "run_us_max", "wait_us_max")]


class RusterizeNativeError(RuntimeError):
    """A failing call into librusterize_hip; `status` is the rz_status code."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


_lib = None
_lock = threading.Lock()

# (name, restype, argtypes) of every exported entry point in include/rusterize_hip.h
_SIGNATURES = [
    ("rz_last_error", c_char_p, []),
    ("rz_version", c_char_p, []),
    ("rz_raster_info_build", c_int, [POINTER(RzGeometry), c_int, c_uint64, c_uint64, c_int, POINTER(c_double),
                                     c_int, c_double, c_double, c_int, c_int32, POINTER(RzRasterInfo)]),
    ("rz_pixel_function_parse", c_int, [c_char_p, POINTER(c_int)]),
    ("rz_dtype_parse", c_int, [c_char_p, POINTER(c_int)]),
    ("rz_dtype_size", c_size_t, [c_int]),
    ("rz_geometry_from_wkt", c_int, [POINTER(c_char_p), POINTER(c_int64), c_int64, POINTER(c_void_p)]),
    ("rz_geometry_from_wkb", c_int, [POINTER(c_char_p), POINTER(c_int64), c_int64, POINTER(c_void_p)]),
    ("rz_geometry_from_arrays", c_int, [POINTER(RzGeometry), POINTER(c_void_p)]),
    ("rz_geometry_view", None, [c_void_p, POINTER(RzGeometry)]),
    ("rz_geometry_free", None, [c_void_p]),
    ("rz_synth_geometry", c_int, [c_int, c_uint64, c_int64, c_double, c_double, c_double, c_double, c_double,
                                  c_double, POINTER(c_void_p)]),
    ("rz_synth_field", c_int, [c_int, c_int, c_uint64, c_int64, c_double, c_double, c_void_p]),
    ("rz_synth_job_spec", c_int, [c_uint64, c_int64, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                  POINTER(c_int64)]),
    ("rz_synth_job_geometry", c_int, [c_uint64, c_int64, POINTER(c_void_p)]),
    ("rz_synth_job_field", c_int, [c_uint64, c_int64, c_void_p]),
    ("rz_rasterize", c_int, [POINTER(RzGeometry), POINTER(RzRasterInfo), c_int, c_int, c_void_p, c_void_p,
                             POINTER(c_uint32), c_uint32, c_void_p, c_int, c_int, c_void_p]),
    ("rz_context_create", c_int, [c_int, POINTER(c_void_p)]),
    ("rz_context_destroy", c_int, [c_void_p]),
    ("rz_context_synchronize", c_int, [c_void_p]),
    ("rz_context_workspace_bytes", c_size_t, [c_void_p]),
    ("rz_device_malloc", c_int, [c_void_p, c_size_t, POINTER(c_void_p)]),
    ("rz_device_free", c_int, [c_void_p, c_void_p]),
    ("rz_memcpy_h2d", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("rz_memcpy_d2h", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("rz_memcpy_d2d", c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    ("rz_device_geometry_upload", c_int, [c_void_p, POINTER(RzGeometry), POINTER(c_void_p)]),
    ("rz_device_geometry_free", c_int, [c_void_p]),
    ("rz_device_geometry_count", c_int64, [c_void_p]),
    ("rz_device_geometry_set_xy", c_int, [c_void_p, c_void_p, c_int64]),
    ("rz_device_synth_job", c_int, [c_void_p, c_uint64, c_int64, POINTER(c_void_p)]),
    ("rz_rasterize_device", c_int, [c_void_p, c_void_p, POINTER(RzRasterInfo), c_int, c_int, c_void_p, c_void_p,
                                    c_void_p, c_uint32, c_void_p, c_int, c_int, c_void_p]),
    ("rz_rasterize_device_noalloc", c_int, [c_void_p, c_void_p, POINTER(RzRasterInfo), c_int, c_int, c_void_p,
                                            c_void_p, c_void_p, c_uint32, c_void_p, c_int, c_int, c_void_p]),
    ("rz_rasterize_sparse", c_int, [POINTER(RzGeometry), POINTER(RzRasterInfo), c_int, c_int, c_void_p, c_void_p,
                                    POINTER(c_uint32), c_uint32, c_int, POINTER(c_void_p)]),
    ("rz_sparse_counts", c_int, [c_void_p, POINTER(c_uint64)]),
    ("rz_sparse_copy", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("rz_sparse_to_dense", c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    ("rz_sparse_free", c_int, [c_void_p]),
    ("rz_batch_rasterize", c_int, [c_void_p, POINTER(RzJob), c_int64]),
    ("rz_batch_graph_create", c_int, [c_int, POINTER(RzJob), c_int64, c_double, POINTER(c_void_p)]),
    ("rz_batch_graph_launch", c_int, [c_void_p]),
    ("rz_batch_graph_synchronize", c_int, [c_void_p]),
    ("rz_batch_graph_destroy", c_int, [c_void_p]),
    ("rz_autotune", c_int, [c_void_p, POINTER(RzJob), c_int64, c_int, POINTER(c_int64)]),
    ("rz_autotune_table", c_int64, [POINTER(c_uint32), POINTER(c_uint32), POINTER(c_uint32), c_int64]),
    ("rz_autotune_set", c_int, [c_uint32, c_uint32, c_uint32]),
    ("rz_server_create", c_int, [c_int, c_uint64, c_int, c_int, POINTER(c_void_p)]),
    ("rz_server_destroy", c_int, [c_void_p]),
    ("rz_server_submit", c_int, [c_void_p, POINTER(RzJob), POINTER(c_uint64)]),
    ("rz_server_wait", c_int, [c_void_p, c_uint64, c_int, POINTER(c_void_p), POINTER(c_double)]),
    ("rz_server_release", c_int, [c_void_p, c_uint64]),
    ("rz_server_get_stats", c_int, [c_void_p, POINTER(RzServerStats)]),
    ("rz_context_set_profiling", c_int, [c_void_p, c_int]),
    ("rz_context_stage_times", c_int, [c_void_p, POINTER(c_double), c_int]),
    ("rz_context_reset_stage_times", None, [c_void_p]),
    ("rz_stage_name", c_char_p, [c_int]),
    ("rz_context_last_counters", c_int, [c_void_p, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64)]),
    ("rz_debug_walk_cells", c_int64, [c_int64] * 8 + [POINTER(c_int64), POINTER(c_int64), c_int64]),
    ("rz_debug_row_coverage", c_int, [POINTER(c_double), c_int64, c_int64, c_int32, c_int32, POINTER(c_uint8)]),
    ("rz_debug_walk_all_touched", c_int64, [c_double] * 4 + [c_int64] * 6 + [POINTER(c_int64), POINTER(c_int64),
                                                                            c_int64]),
]

EXPORTED = tuple(name for name, _, _ in _SIGNATURES)


def lib():
    """Loads librusterize_hip.so (once) and declares every signature."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(f"{LIB_PATH} is not built; run `python -m rusterize_amd.build`")
            handle = ctypes.CDLL(str(LIB_PATH))
            for name, res, args in _SIGNATURES:
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def last_error() -> str:
    return (lib().rz_last_error() or b"").decode("utf-8", "replace")


def check(status: int) -> None:
    """Raises the Python exception the reference raises for a status code."""
    if status == RZ_OK:
        return
    msg = last_error()
    if status == RZ_VALUE_ERROR:
        raise ValueError(msg)
    if status == RZ_TYPE_ERROR:
        raise TypeError(msg)
    if status == RZ_PARSE_ERROR:
        raise ValueError(msg)
    raise RusterizeNativeError(status, msg)


def np_dtype(name: str) -> np.dtype:
    if name not in DTYPES:
        raise TypeError("Invalid dtype or encoding provided.")
    return np.dtype(name)


def dtype_code(name: str) -> int:
    return DTYPES.index(np_dtype(name).name)


def fun_code(name: str) -> int:
    out = c_int()
    check(lib().rz_pixel_function_parse(name.encode(), ctypes.byref(out)))
    return out.value


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(POINTER(ctype))


class Geometry:
    """A flattened geometry set (the rz_geometry model of the C ABI).

    Geometry g owns items geom_item_offsets[g]:geom_item_offsets[g+1]; an item
    is a polygon set (all rings of a Polygon / MultiPolygon), a line set or a
    point set; rings index vertices in xy (n_vertices x 2, world units).
    """

    def __init__(self, geom_item_offsets, item_kind, item_ring_offsets, ring_vertex_offsets, xy, ring_flags=None):
        self.geom_item_offsets = np.ascontiguousarray(geom_item_offsets, dtype=np.int64)
        self.item_kind = np.ascontiguousarray(item_kind, dtype=np.int32)
        self.item_ring_offsets = np.ascontiguousarray(item_ring_offsets, dtype=np.int64)
        self.ring_vertex_offsets = np.ascontiguousarray(ring_vertex_offsets, dtype=np.int64)
        self.xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        n_rings = len(self.ring_vertex_offsets) - 1
        if ring_flags is None:
            ring_flags = np.zeros(n_rings, dtype=np.uint8)
        self.ring_flags = np.ascontiguousarray(ring_flags, dtype=np.uint8)

    # ---- constructors
    @classmethod
    def _from_native(cls, handle: c_void_p) -> "Geometry":
        view = RzGeometry()
        lib().rz_geometry_view(handle, ctypes.byref(view))
        try:
            def arr(p, n, dt):
                if n == 0:
                    return np.zeros(0, dtype=dt)
                return np.ctypeslib.as_array(p, shape=(n,)).copy()
            return cls(
                arr(view.geom_item_offsets, view.n_geoms + 1, np.int64),
                arr(view.item_kind, view.n_items, np.int32),
                arr(view.item_ring_offsets, view.n_items + 1, np.int64),
                arr(view.ring_vertex_offsets, view.n_rings + 1, np.int64),
                arr(view.xy, 2 * view.n_vertices, np.float64),
                arr(view.ring_flags, view.n_rings, np.uint8),
            )
        finally:
            lib().rz_geometry_free(handle)

    @classmethod
    def from_wkt(cls, texts) -> "Geometry":
        enc = [t.encode("utf-8") if isinstance(t, str) else bytes(t) for t in texts]
        n = len(enc)
        ptrs = (c_char_p * max(n, 1))(*enc)
        lens = np.array([len(e) for e in enc], dtype=np.int64)
        h = c_void_p()
        check(lib().rz_geometry_from_wkt(ptrs, _ptr(lens, c_int64), n, ctypes.byref(h)))
        return cls._from_native(h)

    @classmethod
    def from_wkb(cls, blobs) -> "Geometry":
        enc = [bytes(b) for b in blobs]
        n = len(enc)
        ptrs = (c_char_p * max(n, 1))(*enc)
        lens = np.array([len(e) for e in enc], dtype=np.int64)
        h = c_void_p()
        check(lib().rz_geometry_from_wkb(ptrs, _ptr(lens, c_int64), n, ctypes.byref(h)))
        return cls._from_native(h)

    @classmethod
    def synth(cls, kind: int, seed: int, count: int, grid_w: float, grid_h: float,
              p0: float = 0.0, p1: float = 0.0, p2: float = 0.0, p3: float = 0.0) -> "Geometry":
        h = c_void_p()
        check(lib().rz_synth_geometry(kind, seed, count, grid_w, grid_h, p0, p1, p2, p3, ctypes.byref(h)))
        return cls._from_native(h)

    @classmethod
    def synth_job(cls, seed: int, k: int) -> "Geometry":
        """Job k of the seeded batched-workload job stream (rz_synth_job_geometry)."""
        h = c_void_p()
        check(lib().rz_synth_job_geometry(seed, k, ctypes.byref(h)))
        return cls._from_native(h)

    # ---- views
    @property
    def n_geoms(self) -> int:
        return len(self.geom_item_offsets) - 1

    @property
    def n_items(self) -> int:
        return len(self.item_kind)

    @property
    def n_rings(self) -> int:
        return len(self.ring_vertex_offsets) - 1

    @property
    def n_vertices(self) -> int:
        return len(self.xy)

    def __len__(self) -> int:
        return self.n_geoms

    def as_struct(self) -> RzGeometry:
        g = RzGeometry()
        g.n_geoms = self.n_geoms
        g.n_items = self.n_items
        g.n_rings = self.n_rings
        g.n_vertices = self.n_vertices
        g.geom_item_offsets = _ptr(self.geom_item_offsets, c_int64)
        g.item_kind = _ptr(self.item_kind, c_int32)
        g.item_ring_offsets = _ptr(self.item_ring_offsets, c_int64)
        g.ring_vertex_offsets = _ptr(self.ring_vertex_offsets, c_int64)
        g.xy = _ptr(self.xy, c_double)
        g.ring_flags = _ptr(self.ring_flags, c_uint8)
        g._keep = self  # keep the arrays alive with the struct
        return g

    def subset(self, idx) -> "Geometry":
        """The geometries at the given indices, in that order."""
        idx = np.asarray(idx, dtype=np.int64)
        gio, iro, rvo = self.geom_item_offsets, self.item_ring_offsets, self.ring_vertex_offsets
        kinds, ir, rv, xy, fl = [], [0], [0], [], []
        go = [0]
        for g in idx:
            for it in range(gio[g], gio[g + 1]):
                kinds.append(self.item_kind[it])
                for r in range(iro[it], iro[it + 1]):
                    v0, v1 = rvo[r], rvo[r + 1]
                    xy.append(self.xy[v0:v1])
                    fl.append(self.ring_flags[r])
                    rv.append(rv[-1] + (v1 - v0))
                ir.append(len(rv) - 1)
            go.append(len(kinds))
        xy_arr = np.concatenate(xy) if xy else np.zeros((0, 2))
        return Geometry(go, kinds, ir, rv, xy_arr, fl)


def raster_info(geom: Geometry | None, shape=None, extent=None, resolution=None, tap=False, epsg=None) -> RzRasterInfo:
    """RawRasterInfo::build (python/src/geo/raster.rs:16-42) through the C ABI."""
    out = RzRasterInfo()
    ext = (c_double * 4)(*([float(e) for e in extent] if extent is not None else [0, 0, 0, 0]))
    gs = geom.as_struct() if geom is not None else None
    nrows, ncols = (int(shape[0]), int(shape[1])) if shape is not None else (0, 0)
    xres, yres = (float(resolution[0]), float(resolution[1])) if resolution is not None else (0.0, 0.0)
    st = lib().rz_raster_info_build(ctypes.byref(gs) if gs is not None else None, shape is not None, nrows, ncols,
                                    extent is not None, ext, resolution is not None, xres, yres, bool(tap),
                                    -1 if epsg is None else int(epsg), ctypes.byref(out))
    if st != RZ_OK:
        # the binding maps builder errors to RuntimeError (python/src/rusterize.rs:146-148)
        raise RuntimeError(last_error())
    return out


def rasterize(geom: Geometry, info: RzRasterInfo, dtype: str, fun: str, field: np.ndarray | None,
              scalar, layer_of_geom: np.ndarray | None, n_layers: int, background, all_touched: bool = False,
              sum_mode: int = 0) -> np.ndarray:
    """rz_rasterize: host buffers in, a (bands, nrows, ncols) numpy array out."""
    dt = np_dtype(dtype)
    code = dtype_code(dtype)
    fcode = fun_code(fun)
    gs = geom.as_struct()
    bands = n_layers if layer_of_geom is not None else 1
    out = np.empty((bands, int(info.nrows), int(info.ncols)), dtype=dt)
    bg = np.array([background], dtype=dt)
    sc = np.array([scalar if scalar is not None else 0], dtype=dt)
    fld = None
    if field is not None:
        fld = np.ascontiguousarray(field, dtype=dt)
        if len(fld) != geom.n_geoms:
            raise ValueError("Geometry and field lengths must match")
    lay = None
    if layer_of_geom is not None:
        lay = np.ascontiguousarray(layer_of_geom, dtype=np.uint32)
        if len(lay) != geom.n_geoms:
            raise ValueError("Geometry and by lengths must match")
    st = lib().rz_rasterize(ctypes.byref(gs), ctypes.byref(info), code, fcode,
                            fld.ctypes.data if fld is not None else None, sc.ctypes.data,
                            _ptr(lay, c_uint32) if lay is not None else None, bands, bg.ctypes.data,
                            int(bool(all_touched)), int(sum_mode), out.ctypes.data)
    check(st)
    return out


def job_spec(seed: int, k: int) -> dict:
    """Size of job k of the seeded job stream: side, n_poly, fun, n_vertices."""
    side, n_poly, fun, nv = c_int32(), c_int32(), c_int32(), c_int64()
    check(lib().rz_synth_job_spec(seed, k, ctypes.byref(side), ctypes.byref(n_poly), ctypes.byref(fun),
                                  ctypes.byref(nv)))
    return {"side": side.value, "n_poly": n_poly.value, "fun": FUNS[fun.value], "n_vertices": nv.value}


def job_field(seed: int, k: int) -> np.ndarray:
    """Field values (float64) of job k of the seeded job stream."""
    n = job_spec(seed, k)["n_poly"]
    out = np.empty(n, dtype=np.float64)
    check(lib().rz_synth_job_field(seed, k, out.ctypes.data))
    return out
