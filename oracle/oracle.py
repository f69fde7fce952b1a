"""ctypes wrapper of the CPU oracle (oracle/rz_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline,
never as a product code path.  The oracle restates the reference's CPU
algorithm (see the header of rz_oracle.c for the file-by-file mapping).

Geometry inputs use the flattened layout of include/rusterize_hip.h: any
object with the arrays geom_item_offsets, item_kind, item_ring_offsets,
ring_vertex_offsets, xy and (optional) ring_flags.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_uint8, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"
DTYPES = ("uint8", "uint16", "uint32", "uint64", "int8", "int16", "int32", "int64", "float32", "float64")
FUNS = ("sum", "first", "last", "min", "max", "count", "any")


class _Geom(ctypes.Structure):
    _fields_ = [
        ("n_geoms", c_int64), ("n_items", c_int64), ("n_rings", c_int64), ("n_vertices", c_int64),
        ("geom_item_offsets", POINTER(c_int64)), ("item_kind", POINTER(c_int32)),
        ("item_ring_offsets", POINTER(c_int64)), ("ring_vertex_offsets", POINTER(c_int64)),
        ("xy", POINTER(c_double)), ("ring_flags", POINTER(c_uint8)),
    ]


class RasterInfo(ctypes.Structure):
    _fields_ = [
        ("ncols", c_uint64), ("nrows", c_uint64), ("xmin", c_double), ("xmax", c_double),
        ("ymin", c_double), ("ymax", c_double), ("xres", c_double), ("yres", c_double), ("epsg", c_int32),
    ]

    def as_tuple(self):
        return (self.ncols, self.nrows, self.xmin, self.xmax, self.ymin, self.ymax, self.xres, self.yres)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        h = ctypes.CDLL(str(LIB))
        h.oracle_last_error.restype = c_char_p
        h.oracle_raster_info_build.restype = c_int
        h.oracle_raster_info_build.argtypes = [POINTER(_Geom), c_int, c_uint64, c_uint64, c_int, POINTER(c_double),
                                               c_int, c_double, c_double, c_int, c_int32, POINTER(RasterInfo)]
        h.oracle_rasterize.restype = c_int
        h.oracle_rasterize.argtypes = [POINTER(_Geom), POINTER(RasterInfo), c_int, c_int, c_void_p, c_void_p,
                                       POINTER(c_uint32), c_uint32, c_void_p, c_int, c_uint64, c_uint64, c_int,
                                       c_void_p, POINTER(c_uint64)]
        h.oracle_trace.restype = c_int64
        h.oracle_trace.argtypes = [POINTER(_Geom), POINTER(RasterInfo), c_int, c_int, POINTER(c_uint64), c_int64]
        h.oracle_bresenham_cells.restype = c_int64
        h.oracle_bresenham_cells.argtypes = [c_int64] * 4 + [POINTER(c_int64), POINTER(c_int64), c_int64]
        _lib = h
    return _lib


def _struct(g):
    s = _Geom()
    arrays = dict(
        gio=np.ascontiguousarray(g.geom_item_offsets, dtype=np.int64),
        ik=np.ascontiguousarray(g.item_kind, dtype=np.int32),
        iro=np.ascontiguousarray(g.item_ring_offsets, dtype=np.int64),
        rvo=np.ascontiguousarray(g.ring_vertex_offsets, dtype=np.int64),
        xy=np.ascontiguousarray(g.xy, dtype=np.float64).reshape(-1),
    )
    rf = getattr(g, "ring_flags", None)
    arrays["rf"] = np.ascontiguousarray(rf if rf is not None else np.zeros(len(arrays["rvo"]) - 1), dtype=np.uint8)
    s.n_geoms = len(arrays["gio"]) - 1
    s.n_items = len(arrays["ik"])
    s.n_rings = len(arrays["rvo"]) - 1
    s.n_vertices = len(arrays["xy"]) // 2
    s.geom_item_offsets = arrays["gio"].ctypes.data_as(POINTER(c_int64))
    s.item_kind = arrays["ik"].ctypes.data_as(POINTER(c_int32))
    s.item_ring_offsets = arrays["iro"].ctypes.data_as(POINTER(c_int64))
    s.ring_vertex_offsets = arrays["rvo"].ctypes.data_as(POINTER(c_int64))
    s.xy = arrays["xy"].ctypes.data_as(POINTER(c_double))
    s.ring_flags = arrays["rf"].ctypes.data_as(POINTER(c_uint8))
    s._keep = arrays
    return s


class OracleError(Exception):
    pass


def raster_info(geom=None, shape=None, extent=None, resolution=None, tap=False, epsg=None) -> RasterInfo:
    """RasterInfoBuilder (rust/src/geo/raster.rs:50-156)."""
    out = RasterInfo()
    ext = (c_double * 4)(*([float(e) for e in extent] if extent is not None else [0, 0, 0, 0]))
    gs = _struct(geom) if geom is not None else None
    nr, nc = (int(shape[0]), int(shape[1])) if shape is not None else (0, 0)
    xr, yr = (float(resolution[0]), float(resolution[1])) if resolution is not None else (0.0, 0.0)
    st = lib().oracle_raster_info_build(ctypes.byref(gs) if gs is not None else None, shape is not None, nr, nc,
                                        extent is not None, ext, resolution is not None, xr, yr, bool(tap),
                                        -1 if epsg is None else int(epsg), ctypes.byref(out))
    if st != 0:
        raise OracleError(lib().oracle_last_error().decode())
    return out


def make_info(ncols, nrows, xmin, xmax, ymin, ymax, xres, yres) -> RasterInfo:
    ri = RasterInfo()
    ri.ncols, ri.nrows = ncols, nrows
    ri.xmin, ri.xmax, ri.ymin, ri.ymax = xmin, xmax, ymin, ymax
    ri.xres, ri.yres = xres, yres
    ri.epsg = -1
    return ri


def copy_info(info) -> RasterInfo:
    return make_info(info.ncols, info.nrows, info.xmin, info.xmax, info.ymin, info.ymax, info.xres, info.yres)


def rasterize(geom, info, dtype: str, fun: str, field=None, scalar=1, layer_of_geom=None, n_layers=1,
              background=0, all_touched=False, rows=None, threads=None, return_burns=False):
    """DenseArray::build (rust/src/rasterize.rs:71-116) on the CPU.

    rows=(r0, r1) restricts the output to that row window (same computation,
    writes outside the window dropped).  Returns (bands, r1-r0, ncols)."""
    info = copy_info(info)
    dt = np.dtype(dtype)
    gs = _struct(geom)
    r0, r1 = (0, int(info.nrows)) if rows is None else (int(rows[0]), min(int(rows[1]), int(info.nrows)))
    bands = n_layers if layer_of_geom is not None else 1
    out = np.empty((bands, max(r1 - r0, 0), int(info.ncols)), dtype=dt)
    bg = np.array([background], dtype=dt)
    sc = np.array([scalar if scalar is not None else 0], dtype=dt)
    fld = np.ascontiguousarray(field, dtype=dt) if field is not None else None
    lay = np.ascontiguousarray(layer_of_geom, dtype=np.uint32) if layer_of_geom is not None else None
    burns = c_uint64()
    nthreads = threads if threads is not None else min(bands, os.cpu_count() or 1)
    st = lib().oracle_rasterize(ctypes.byref(gs), ctypes.byref(info), DTYPES.index(dt.name), FUNS.index(fun),
                                fld.ctypes.data if fld is not None else None, sc.ctypes.data,
                                lay.ctypes.data_as(POINTER(c_uint32)) if lay is not None else None, bands,
                                bg.ctypes.data, int(bool(all_touched)), r0, r1, nthreads, out.ctypes.data,
                                ctypes.byref(burns))
    if st != 0:
        raise OracleError(lib().oracle_last_error().decode())
    return (out, burns.value) if return_burns else out


def trace(geom, info, fun: str = "last", all_touched=False) -> np.ndarray:
    """Every pixel write in reference order as rows of (row, col, geometry)."""
    info = copy_info(info)
    gs = _struct(geom)
    n = lib().oracle_trace(ctypes.byref(gs), ctypes.byref(info), FUNS.index(fun), int(bool(all_touched)), None, 0)
    cells = np.empty((max(n, 1), 3), dtype=np.uint64)
    lib().oracle_trace(ctypes.byref(gs), ctypes.byref(info), FUNS.index(fun), int(bool(all_touched)),
                       cells.ctypes.data_as(POINTER(c_uint64)), n)
    return cells[:n]


def bresenham_cells(ix0, iy0, ix1, iy1, cap=1 << 20):
    """The literal loop of Standard::burn_line (burners.rs:53-84)."""
    xs = np.empty(cap, dtype=np.int64)
    ys = np.empty(cap, dtype=np.int64)
    n = lib().oracle_bresenham_cells(ix0, iy0, ix1, iy1, xs.ctypes.data_as(POINTER(c_int64)),
                                     ys.ctypes.data_as(POINTER(c_int64)), cap)
    if n < 0:
        raise OracleError("segment longer than cap")
    return xs[:n], ys[:n]
