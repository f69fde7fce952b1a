"""rusterize_amd -- MI355X (gfx950) rasteriser behind rusterize's Python API.

``rusterize()`` keeps the signature, argument checks, coercions and error
messages of the reference's ``rusterize.rusterize``
(python/python/rusterize/__init__.py:83-356) and ``_rusterize``
(python/src/rusterize.rs:45-184); the burn itself runs in hand-written HIP
kernels through the C ABI of librusterize_hip.so (include/rusterize_hip.h).
There is no CPU fallback: without the built library or a gfx950 GPU the call
raises.
"""
from __future__ import annotations

import importlib.util
from types import NoneType

import numpy as np

from . import _lib
from ._lib import Geometry, RusterizeNativeError

__all__ = ["rusterize", "_rusterize", "Geometry", "GeoFrame", "PanicException", "RusterizeNativeError"]
__version__ = "0.1.0"


class PanicException(BaseException):
    """Stands for pyo3_runtime.PanicException: the reference panics (via
    ``unimplemented!``) on an unsupported dtype or encoding
    (python/src/rusterize.rs:28)."""


class GeoFrame:
    """A minimal table of geometries plus attribute columns.

    It stands in for the polars / geopandas frames the reference accepts
    (neither library is available here): ``GeoFrame({"geometry": [...wkt or
    wkb...], "value": [...], "grp": [...]})``.  ``field`` and ``by`` name
    columns of it exactly as they name DataFrame columns in the reference.
    """

    def __init__(self, columns: dict):
        if "geometry" not in columns:
            raise ValueError("If `polars.DataFrame`, a 'geometry' column is expected.")
        self.columns = dict(columns)
        n = len(self.columns["geometry"])
        for k, v in self.columns.items():
            if len(v) != n:
                raise ValueError(f"column {k!r} has {len(v)} rows, expected {n}")

    def __len__(self) -> int:
        return len(self.columns["geometry"])

    def is_empty(self) -> bool:
        return len(self) == 0


def _xarray_available() -> bool:
    return importlib.util.find_spec("xarray") is not None and importlib.util.find_spec("rioxarray") is not None


def _check_for_geopandas(data) -> bool:
    return importlib.util.find_spec("geopandas") is not None and type(data).__module__.startswith("geopandas")


# ------------------------------------------------------------- geometry --

def _parse_geometry(obj) -> Geometry:
    """ParsedGeometry::extract, python/src/geo/parse_geometry.rs:36-75."""
    if isinstance(obj, Geometry):
        if obj.n_geoms == 0:
            raise ValueError("No geometries found.")
        return obj
    if hasattr(obj, "geom_type"):
        return _parse_wkb_seq(_shapely_to_wkb(obj))
    if isinstance(obj, list) or type(obj).__name__ == "ndarray":
        if len(obj) == 0:
            raise ValueError("No geometries found.")
        first = obj[0]
        if isinstance(first, (bytes, bytearray, memoryview)):
            return _parse_wkb_seq(obj)
        if isinstance(first, str):
            return _parse_wkt_seq(obj)
        if hasattr(first, "geom_type"):
            return _parse_wkb_seq(_shapely_to_wkb(obj))
        raise ValueError("Sequence must contain geometries as shapely Geometry, bytes (WKB), or string (WKT).")
    raise TypeError("Unsupported geometry input type.")


def _shapely_to_wkb(obj):
    try:
        import shapely
    except ImportError as e:  # pragma: no cover - shapely is not installed here
        raise ModuleNotFoundError("shapely is required for shapely geometries") from e
    if not shapely.__version__.startswith("2"):
        raise ValueError("Shapely version 2 required")
    return list(shapely.to_wkb(obj, output_dimension=2, include_srid=False, flavor="iso"))


def _bail_if_empty(g: Geometry) -> Geometry:
    if g.n_geoms == 0:
        raise ValueError("Could not parse geometry. Only WKT or WKB formats are supported.")
    return g


def _parse_wkt_seq(seq) -> Geometry:
    return _bail_if_empty(Geometry.from_wkt([str(s) for s in seq]))


def _parse_wkb_seq(seq) -> Geometry:
    return _bail_if_empty(Geometry.from_wkb([bytes(b) for b in seq]))


# ----------------------------------------------------------- coercions --

_INT_RANGES = {
    "uint8": (0, 2**8 - 1), "uint16": (0, 2**16 - 1), "uint32": (0, 2**32 - 1), "uint64": (0, 2**64 - 1),
    "int8": (-(2**7), 2**7 - 1), "int16": (-(2**15), 2**15 - 1), "int32": (-(2**31), 2**31 - 1),
    "int64": (-(2**63), 2**63 - 1),
}


def _extract_scalar(obj, dtype: str):
    """PyO3 `obj.extract::<N>()`: returns the value or raises TypeError /
    OverflowError the way the extraction fails."""
    if dtype in _INT_RANGES:
        if isinstance(obj, (float, np.floating)) or not hasattr(obj, "__index__"):
            raise TypeError(f"'{type(obj).__name__}' object cannot be interpreted as an integer")
        v = int(obj.__index__())
        lo, hi = _INT_RANGES[dtype]
        if not lo <= v <= hi:
            raise OverflowError("out of range integral type conversion attempted")
        return v
    if not hasattr(obj, "__float__") and not hasattr(obj, "__index__"):
        raise TypeError(f"must be real number, not {type(obj).__name__}")
    return float(obj)


def _background(pybackground, dtype: str):
    """`pybackground.and_then(|b| b.extract().ok()).unwrap_or_default()`
    (python/src/rusterize.rs:50-53): anything that does not extract as N --
    NaN or a float for an integer dtype, None -- becomes 0."""
    if pybackground is None:
        return 0
    try:
        return _extract_scalar(pybackground, dtype)
    except (TypeError, OverflowError, ValueError):
        return 0


def _field_source(pyburn, dtype: str, n_geoms: int):
    """FieldSource from `pyburn` (python/src/rusterize.rs:79-88): a scalar if
    it extracts as N, else a 1-D numpy array of exactly dtype N."""
    if pyburn is None:
        return None, 1
    try:
        return None, _extract_scalar(pyburn, dtype)
    except (TypeError, OverflowError, ValueError):
        pass
    if not (isinstance(pyburn, np.ndarray) and pyburn.ndim == 1 and pyburn.dtype == np.dtype(dtype)):
        raise TypeError(f"'{type(pyburn).__name__}' object cannot be converted to 'PyArray<T, D>'")
    if len(pyburn) != n_geoms:
        raise RuntimeError("Geometry and field lengths must match")
    return np.ascontiguousarray(pyburn), None


def _group_keys(by_values):
    """group_keys (rust/src/rasterize.rs:198-205): bands are the distinct keys
    in byte-lexicographic order (UTF-8 byte order == code point order)."""
    keys = [str(k) for k in by_values]
    names = sorted(set(keys), key=lambda s: s.encode("utf-8"))
    index = {k: i for i, k in enumerate(names)}
    return np.array([index[k] for k in keys], dtype=np.uint32), names


# ------------------------------------------------------------ binding --

def _rusterize(geometry, raw_raster_info: dict, pypixel_fn: str, pydf=None, pyfield=None, pyby=None, pyburn=None,
               pybackground=None, pytouched: bool = False, pyencoding: str = "xarray", pydtype: str = "float64"):
    """Mirror of the PyO3 `_rusterize` (python/src/rusterize.rs:127-184)."""
    geoms = _parse_geometry(geometry)
    info = _lib.raster_info(geoms, shape=raw_raster_info.get("shape"), extent=raw_raster_info.get("extent"),
                            resolution=raw_raster_info.get("resolution"), tap=bool(raw_raster_info.get("tap", False)),
                            epsg=raw_raster_info.get("epsg"))
    if pypixel_fn not in _lib.FUNS:
        raise ValueError("Unknown pixel function")
    if pydtype not in _lib.DTYPES or pyencoding not in ("xarray", "numpy", "sparse"):
        raise PanicException("not implemented: Invalid dtype or encoding provided.")

    background = _background(pybackground, pydtype)
    field, scalar = None, None
    if pydf is not None and pyfield is not None:
        col = np.asarray(pydf.columns[pyfield])
        field = col.astype(np.dtype(pydtype))  # numpy cast: integers wrap
    else:
        field, scalar = _field_source(pyburn, pydtype, geoms.n_geoms)
    layer, names, n_layers = None, ["band_1"], 1
    if pydf is not None and pyby is not None:
        col = pydf.columns[pyby]
        if any(v is None for v in col):
            raise RuntimeError("Found nulls in `by` column. Consider droppping them.")
        layer, names = _group_keys(col)
        n_layers = len(names)
        if len(layer) != geoms.n_geoms:
            raise RuntimeError("Geometry and by lengths must match")
    if field is not None and len(field) != geoms.n_geoms:
        raise RuntimeError("Geometry and field lengths must match")
    if pyencoding == "xarray":
        # the reference builds a DataArray from the raster; xarray / rioxarray
        # are not available here, so this fails before any device work
        import xarray  # noqa: F401
        raise NotImplementedError("encoding='xarray' output is not implemented")

    if pyencoding == "sparse":
        from .sparse import rasterize_sparse

        return rasterize_sparse(geoms, info, pydtype, pypixel_fn, field, scalar, layer, n_layers, background,
                                names, all_touched=bool(pytouched))
    return _lib.rasterize(geoms, info, pydtype, pypixel_fn, field, scalar, layer, n_layers, background,
                          all_touched=bool(pytouched))


def rusterize(
    data,
    like=None,
    res: tuple | list | None = None,
    out_shape: tuple | list | None = None,
    extent: tuple | list | None = None,
    field: str | None = None,
    by: str | None = None,
    burn: int | float | np.ndarray | None = None,
    fun: str = "last",
    background: int | float | None = np.nan,
    encoding: str = "xarray",
    all_touched: bool = False,
    tap: bool = False,
    dtype: str = "float64",
):
    """Rasterise geometries (python/python/rusterize/__init__.py:83-356).

    ``data`` is a list / numpy array of WKT strings or WKB bytes, a
    ``Geometry`` (already flattened) or a ``GeoFrame`` (geometries plus
    columns for ``field`` / ``by``).  Output with ``encoding="numpy"`` is a
    (bands, rows, cols) array; the other arguments mean what they mean in the
    reference.
    """
    if isinstance(data, (list, np.ndarray, Geometry)):
        data_type = "raw"
    elif _check_for_geopandas(data):  # pragma: no cover - geopandas is not installed here
        if data.empty:
            raise ValueError("Input data is empty.")
        data_type = "geopandas"
    elif isinstance(data, GeoFrame):
        if data.is_empty():
            raise ValueError("Input data is empty.")
        data_type = "frame"
    else:
        raise TypeError(
            "`data` must be either geopandas.GeoDataFrame, geopandas.GeoSeries, polars.DataFrame, list, or numpy.ndarray"
        )

    if not isinstance(res, (tuple, list, NoneType)):
        raise TypeError("`resolution` must be a tuple or list of (xres, yres).")
    if not isinstance(out_shape, (tuple, list, NoneType)):
        raise TypeError("`out_shape` must be a tuple or list of (nrows, ncols).")
    if not isinstance(extent, (tuple, list, NoneType)):
        raise TypeError("`extent` must be a tuple or list of (xmin, ymin, xmax, ymax).")
    if not isinstance(field, (str, NoneType)):
        raise TypeError("`field` must be a string column name.")
    if not isinstance(by, (str, NoneType)):
        raise TypeError("`by` must be a string column name.")
    if not isinstance(burn, (int, float, np.ndarray, NoneType)):
        raise TypeError("`burn` must be an integer, float, or a numpy.ndarray.")
    if not isinstance(fun, str):
        raise TypeError("`pixel_fn` must be one of sum, first, last, min, max, count, or any.")
    if not isinstance(background, (int, float, NoneType)):
        raise TypeError("`background` must be integer, float, or None.")
    if not isinstance(encoding, str):
        raise TypeError("`encoding` must be one of 'xarray', 'numpy', or 'sparse'.")
    if not isinstance(all_touched, bool):
        raise TypeError("`all_touched` must be a boolean.")
    if not isinstance(tap, bool):
        raise TypeError("`tap` must be a boolean.")
    if not isinstance(dtype, str):
        raise TypeError(
            "`dtype` must be a one of 'uint8', 'uint16', 'uint32', 'uint64', 'int8', 'int16', 'int32', 'int64', 'float32', 'float64'"
        )

    if encoding not in ["xarray", "numpy", "sparse"]:
        raise ValueError("`encoding` must be one of `xarray`, 'numpy', or `sparse`.")
    if encoding == "xarray" and not _xarray_available():
        raise ModuleNotFoundError(
            "`xarray` and `rioxarray` must be installed if encoding is `xarray`. Install with `pip install xarray rioxarray`."
        )
    if field and burn is not None:
        raise ValueError("Only one of `field` or `burn` can be specified.")
    if isinstance(burn, np.ndarray) and burn.size != len(data):
        raise ValueError("If `burn` is a `numpy.ndarray`, it must have the same length as `data`.")

    _bounds = None
    _res = None
    _shape = None
    if like is not None:
        if not _xarray_available():
            raise TypeError("`like` must be a xarray.DataArray or xarray.Dataset")
        if any((res, out_shape, extent)):  # pragma: no cover - needs xarray
            raise ValueError("`like` is mutually exclusive with `res`, `out_shape`, and `extent`.")
        if not hasattr(like, "rio"):  # pragma: no cover
            raise AttributeError("The `like` object must have a 'rio' accessor.")
        try:  # pragma: no cover
            _shape = like.squeeze().shape
            _bounds = like.rio.bounds()
        except Exception as e:  # pragma: no cover
            raise AttributeError("No spatial dimension found for like object") from e
    else:
        if not res and not out_shape and not extent:
            raise ValueError("One of `res`, `out_shape`, or `extent` must be provided.")
        if res and out_shape:
            raise ValueError("`res` and `out_shape` are mutually exclusive; provide only one.")
        if extent:
            if not res and not out_shape:
                raise ValueError("Must also specify `res` or `out_shape` with extent.")
            if len(extent) != 4 or all(e == 0 for e in extent):
                raise ValueError("`extent` must be a tuple or list of (xmin, ymin, xmax, ymax).")
            _bounds = extent
        if res:
            if len(res) != 2 or any(r <= 0 for r in res) or any(not isinstance(r, (int, float)) for r in res):
                raise ValueError("`res` must be 2 positive numbers.")
            _res = res
        if out_shape:
            if len(out_shape) != 2 or any(s <= 0 for s in out_shape) or any(not isinstance(s, int) for s in out_shape):
                raise ValueError("`out_shape` must be 2 positive integers.")
            _shape = out_shape

    cols = list(set([col for col in (field, by) if col and col != "geometry"]))
    df = None
    epsg = None
    if data_type == "frame":
        if cols:
            missing = [c for c in cols if c not in data.columns]
            if missing:
                raise KeyError("Column not found in polars DataFrame.")
            df = data
        geometries = list(data.columns["geometry"])
    else:
        geometries = data

    if isinstance(burn, np.ndarray) and burn.dtype != dtype:
        burn = np.ascontiguousarray(burn, dtype=dtype)

    raw_raster_info = {"shape": _shape, "extent": _bounds, "resolution": _res, "tap": tap, "epsg": epsg}
    return _rusterize(geometries, raw_raster_info, fun, df, field, by, burn, background, all_touched, encoding, dtype)
