"""encoding="sparse": the mirror of the reference's SparseArray
(rust/src/encoding/arrays.rs:61-209) and its Python class PySparseArray
(python/src/encoding/pyarray.rs:48-140).

The triplets are built on the GPU by rz_rasterize_sparse and stay there
until the object is dropped; to_numpy() folds them back into a dense raster
on the GPU (rz_sparse_to_dense, SparseArray::build_array), to_frame() copies
them out.  polars is not available here, so to_frame() returns a pandas
DataFrame with the reference's columns (``band`` for multi-band arrays,
``row``, ``col``, ``values``) and dtypes.
"""
from __future__ import annotations

import ctypes
from ctypes import c_uint32, c_uint64, c_void_p

import numpy as np

from . import _lib


class SparseArray:
    def __init__(self, handle: c_void_p, info, band_names, dtype: str, fun: str, background):
        self._h = handle
        self._info = info
        self.band_names = list(band_names)
        self.dtype = dtype
        self._fun = fun
        self._bg = background

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _lib.lib().rz_sparse_free(h)
            self._h = None

    # ---- spatial information (arrays.rs:143-171)
    def shape(self) -> tuple[int, int, int]:
        return (len(self.band_names), int(self._info.nrows), int(self._info.ncols))

    def extent(self) -> tuple[float, float, float, float]:
        i = self._info
        return (i.xmin, i.ymin, i.xmax, i.ymax)

    def resolution(self) -> tuple[float, float]:
        return (self._info.xres, self._info.yres)

    def epsg(self):
        return None if self._info.epsg < 0 else int(self._info.epsg)

    def size_hint(self) -> str:
        """Estimated size of the materialised array (pyarray.rs:78-90)."""
        b, r, c = self.shape()
        n = np.dtype(self.dtype).itemsize * b * r * c
        if n < 1000:
            return f"{n} bytes"
        if n < 1_000_000:
            return f"{n / 1000.0:.2f} KB"
        if n < 1_000_000_000:
            return f"{n / 1_000_000.0:.2f} MB"
        return f"{n / 1_000_000_000.0:.2f} GB"

    def __repr__(self) -> str:
        epsg = self.epsg()
        return (f"SparseArray:\n- Shape: {self.shape()}\n- Extent: {self.extent()}\n- Resolution: "
                f"{self.resolution()}\n- EPSG: {'None' if epsg is None else epsg}\n- Estimated size: "
                f"{self.size_hint()}")

    # ---- data
    def counts(self) -> np.ndarray:
        """Triplets of each band (SparseArray offsets)."""
        out = np.zeros(len(self.band_names), dtype=np.uint64)
        _lib.check(_lib.lib().rz_sparse_counts(self._h, out.ctypes.data_as(ctypes.POINTER(c_uint64))))
        return out

    def triplets(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """Every (row, col, value) in write order, band after band."""
        n = int(self.counts().sum())
        rows = np.empty(n, dtype=np.uint64)
        cols = np.empty(n, dtype=np.uint64)
        vals = np.empty(n, dtype=np.dtype(self.dtype))
        _lib.check(_lib.lib().rz_sparse_copy(self._h, rows.ctypes.data, cols.ctypes.data, vals.ctypes.data))
        return rows, cols, vals

    def to_numpy(self) -> np.ndarray:
        """SparseArray::build_array (arrays.rs:102-141), folded on the GPU."""
        dt = np.dtype(self.dtype)
        out = np.empty(self.shape(), dtype=dt)
        bg = np.array([self._bg], dtype=dt)
        _lib.check(_lib.lib().rz_sparse_to_dense(self._h, _lib.fun_code(self._fun), bg.ctypes.data,
                                                 out.ctypes.data))
        return out

    def to_xarray(self):
        import xarray  # noqa: F401  (absent here: the reference's ModuleNotFoundError)
        raise NotImplementedError("to_xarray() output is not implemented")

    def to_frame(self):
        """SparseArray::to_frame (arrays.rs:177-207) as a pandas DataFrame."""
        import pandas as pd

        rows, cols, vals = self.triplets()
        data = {}
        counts = self.counts()
        if len(counts) > 1:
            data["band"] = np.repeat(np.arange(1, len(counts) + 1, dtype=np.uint64), counts.astype(np.int64))
        data["row"] = rows
        data["col"] = cols
        data["values"] = vals
        return pd.DataFrame(data)


def rasterize_sparse(geom, info, dtype: str, fun: str, field, scalar, layer_of_geom, n_layers: int,
                     background, band_names, all_touched: bool = False) -> SparseArray:
    """rz_rasterize_sparse: the call's triplets on the GPU."""
    dt = _lib.np_dtype(dtype)
    gs = geom.as_struct()
    sc = np.array([scalar if scalar is not None else 0], dtype=dt)
    fld = np.ascontiguousarray(field, dtype=dt) if field is not None else None
    lay = np.ascontiguousarray(layer_of_geom, dtype=np.uint32) if layer_of_geom is not None else None
    h = c_void_p()
    st = _lib.lib().rz_rasterize_sparse(ctypes.byref(gs), ctypes.byref(info), _lib.dtype_code(dtype),
                                        _lib.fun_code(fun), fld.ctypes.data if fld is not None else None,
                                        sc.ctypes.data,
                                        lay.ctypes.data_as(ctypes.POINTER(c_uint32)) if lay is not None else None,
                                        n_layers if lay is not None else 1, int(bool(all_touched)),
                                        ctypes.byref(h))
    _lib.check(st)
    return SparseArray(h, info, band_names, dtype, fun, background)
