"""Device-resident runtime of librusterize_hip from Python: contexts, device
buffers, resident geometry sets, single calls and batches.

This is the layer a serving or batching host uses: geometry sets are
uploaded once and rasterised any number of times with new fields and grids;
outputs stay on the device until downloaded.  Every call goes through the C
ABI of include/rusterize_hip.h; there is no CPU fallback (a missing GPU is a
RusterizeNativeError from rz_context_create).
"""
from __future__ import annotations

from ctypes import byref, c_double, c_int64, c_uint32, c_uint64, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import Geometry, RzJob, RzRasterInfo, check


def value_bits(value, dtype: str) -> int:
    """Raw little-endian bits of `value` cast to `dtype` (the C ABI's value form)."""
    b = np.array([value]).astype(_lib.np_dtype(dtype)).tobytes()
    return int.from_bytes(b.ljust(8, b"\0"), "little")


class Context:
    """rz_context: one HIP stream and a grow-only workspace on `device`."""

    def __init__(self, device: int = 0):
        self.L = _lib.lib()
        self.ptr = c_void_p()
        check(self.L.rz_context_create(int(device), byref(self.ptr)))

    def close(self) -> None:
        if self.ptr:
            self.L.rz_context_destroy(self.ptr)
            self.ptr = c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def synchronize(self) -> None:
        check(self.L.rz_context_synchronize(self.ptr))

    def set_profiling(self, level: int) -> None:
        check(self.L.rz_context_set_profiling(self.ptr, int(level)))

    def stage_times(self) -> dict:
        t = (c_double * 8)()
        n = self.L.rz_context_stage_times(self.ptr, t, 8)
        return {self.L.rz_stage_name(i).decode(): t[i] for i in range(n)}

    def reset_stage_times(self) -> None:
        self.L.rz_context_reset_stage_times(self.ptr)

    def last_counters(self) -> tuple[int, int, int]:
        b, p, t = c_uint64(), c_uint64(), c_uint64()
        check(self.L.rz_context_last_counters(self.ptr, byref(b), byref(p), byref(t)))
        return b.value, p.value, t.value

    def workspace_bytes(self) -> int:
        return int(self.L.rz_context_workspace_bytes(self.ptr))


class DeviceBuffer:
    """Device memory owned by a context (rz_device_malloc / rz_device_free)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx, self.nbytes = ctx, int(nbytes)
        self.ptr = c_void_p()
        check(ctx.L.rz_device_malloc(ctx.ptr, max(self.nbytes, 1), byref(self.ptr)))

    @classmethod
    def from_array(cls, ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(ctx, a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        check(self.ctx.L.rz_memcpy_h2d(self.ctx.ptr, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested array larger than the device buffer")
        check(self.ctx.L.rz_memcpy_d2h(self.ctx.ptr, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            self.ctx.L.rz_device_free(self.ctx.ptr, self.ptr)
            self.ptr = c_void_p()


class DeviceGeometry:
    """A geometry set resident on the device (rz_device_geometry_upload)."""

    def __init__(self, ctx: Context, geom: Geometry):
        self.ctx = ctx
        self.ptr = c_void_p()
        self.n_geoms = geom.n_geoms
        self.n_items, self.n_vertices = geom.n_items, geom.n_vertices
        s = geom.as_struct()
        check(ctx.L.rz_device_geometry_upload(ctx.ptr, byref(s), byref(self.ptr)))

    @classmethod
    def synth_job(cls, ctx: Context, seed: int, k: int) -> "DeviceGeometry":
        """Job k of the seeded job stream, generated on the device (rz_device_synth_job)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.ptr = c_void_p()
        spec = _lib.job_spec(seed, k)
        self.n_geoms = self.n_items = spec["n_poly"]
        self.n_vertices = spec["n_vertices"]
        check(ctx.L.rz_device_synth_job(ctx.ptr, seed, k, byref(self.ptr)))
        return self

    def set_xy(self, xy: np.ndarray) -> None:
        """New coordinates for the same topology (rz_device_geometry_set_xy)."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        if xy.size != 2 * self.n_vertices:
            raise ValueError("set_xy needs one (x, y) pair per resident vertex")
        check(self.ctx.L.rz_device_geometry_set_xy(self.ptr, xy.ctypes.data, self.n_vertices))

    def free(self) -> None:
        if self.ptr:
            self.ctx.L.rz_device_geometry_free(self.ptr)
            self.ptr = c_void_p()


@dataclass
class Job:
    """One rasterise job: the arguments of rz_rasterize_device."""
    geometry: DeviceGeometry
    info: RzRasterInfo
    dtype: str
    fun: str
    out: DeviceBuffer | None  # None for server requests (the server allocates it)
    field: DeviceBuffer | None = None
    scalar: float = 1
    layers: DeviceBuffer | None = None
    n_layers: int = 1
    background: float = 0
    sum_mode: int = 0  # rz_sum_mode: 0 ordered (exact), 1 unordered (float sums of lines and points)
    all_touched: bool = False

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.n_layers if self.layers is not None else 1, int(self.info.nrows), int(self.info.ncols))

    @property
    def out_bytes(self) -> int:
        b, r, c = self.shape
        return b * r * c * _lib.np_dtype(self.dtype).itemsize

    def as_struct(self) -> RzJob:
        j = RzJob()
        j.geometry = self.geometry.ptr
        j.info = self.info
        j.dtype = _lib.dtype_code(self.dtype)
        j.fun = _lib.fun_code(self.fun)
        j.field_dev = self.field.ptr if self.field is not None else None
        j.scalar_bits = value_bits(self.scalar, self.dtype)
        j.layer_dev = self.layers.ptr if self.layers is not None else None
        j.n_layers = self.n_layers if self.layers is not None else 1
        j.background_bits = value_bits(self.background, self.dtype)
        j.out_dev = self.out.ptr if self.out is not None else None
        j.sum_mode = self.sum_mode
        j.all_touched = int(bool(self.all_touched))
        return j

    def result(self) -> np.ndarray:
        return self.out.download(self.shape, _lib.np_dtype(self.dtype))


def rasterize_device(ctx: Context, job: Job) -> None:
    """One call (rz_rasterize_device), asynchronous on the context's stream."""
    dt = _lib.np_dtype(job.dtype)
    scalar = np.array([job.scalar]).astype(dt)
    bg = np.array([job.background]).astype(dt)
    check(ctx.L.rz_rasterize_device(
        ctx.ptr, job.geometry.ptr, byref(job.info), _lib.dtype_code(job.dtype), _lib.fun_code(job.fun),
        job.field.ptr if job.field is not None else None, None if job.field is not None else scalar.ctypes.data,
        job.layers.ptr if job.layers is not None else None, job.n_layers if job.layers is not None else 1,
        bg.ctypes.data, int(bool(job.all_touched)), job.sum_mode, job.out.ptr))


def batch_rasterize(ctx: Context, jobs: list[Job]) -> None:
    """All jobs as one pipeline (rz_batch_rasterize), asynchronous."""
    arr = (RzJob * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = j.as_struct()
    check(ctx.L.rz_batch_rasterize(ctx.ptr, arr, len(jobs)))


class BatchGraph:
    """rz_batch_graph: a batch captured in a HIP graph (own context and
    stream); launch() replays it, synchronize() waits and checks it."""

    def __init__(self, jobs: list[Job], device: int = 0, headroom: float = 1.0):
        self.L = _lib.lib()
        self.ptr = c_void_p()
        self.jobs = list(jobs)  # keeps the buffers alive
        arr = (RzJob * max(len(jobs), 1))()
        for i, j in enumerate(jobs):
            arr[i] = j.as_struct()
        check(self.L.rz_batch_graph_create(int(device), arr, len(jobs), float(headroom), byref(self.ptr)))

    def launch(self) -> None:
        check(self.L.rz_batch_graph_launch(self.ptr))

    def synchronize(self) -> None:
        check(self.L.rz_batch_graph_synchronize(self.ptr))

    def close(self) -> None:
        if self.ptr:
            self.L.rz_batch_graph_destroy(self.ptr)
            self.ptr = c_void_p()


class Server:
    """rz_server: admission by a device-memory budget and continuous batching
    over n_streams contexts.  Requests are Jobs whose `out` is None; the
    server allocates every output on the device until `release`."""

    def __init__(self, device: int = 0, budget_bytes: int = 200 * 10**9, n_streams: int = 2,
                 max_batch: int = 64):
        self.L = _lib.lib()
        self.ptr = c_void_p()
        check(self.L.rz_server_create(int(device), int(budget_bytes), int(n_streams), int(max_batch),
                                      byref(self.ptr)))
        self._jobs = {}  # ticket -> Job (keeps its buffers alive until released)

    def submit(self, job: Job) -> int:
        t = c_uint64()
        s = job.as_struct()
        check(self.L.rz_server_submit(self.ptr, byref(s), byref(t)))
        self._jobs[t.value] = job
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> tuple[int, float]:
        """(device address of the output, latency in ms from submission to completion)."""
        p, lat = c_void_p(), c_double()
        check(self.L.rz_server_wait(self.ptr, ticket, int(timeout_ms), byref(p), byref(lat)))
        return p.value, lat.value

    def result(self, ctx: Context, ticket: int) -> np.ndarray:
        """Waits for a request and downloads its output through `ctx`."""
        p, _ = self.wait(ticket)
        job = self._jobs[ticket]
        out = np.empty(job.shape, dtype=_lib.np_dtype(job.dtype))
        check(ctx.L.rz_memcpy_d2h(ctx.ptr, out.ctypes.data, c_void_p(p), out.nbytes))
        return out

    def release(self, ticket: int) -> None:
        check(self.L.rz_server_release(self.ptr, ticket))
        self._jobs.pop(ticket, None)

    def stats(self) -> dict:
        st = _lib.RzServerStats()
        check(self.L.rz_server_get_stats(self.ptr, byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def close(self) -> None:
        if self.ptr:
            self.L.rz_server_destroy(self.ptr)
            self.ptr = c_void_p()
            self._jobs.clear()


def autotune(ctx: Context, jobs: list[Job], reps: int = 2) -> int:
    """Times every launch-knob candidate on each job (rz_autotune) and stores
    the fastest per size class; returns the number of new classes."""
    arr = (RzJob * max(len(jobs), 1))()
    for i, j in enumerate(jobs):
        arr[i] = j.as_struct()
    n = c_int64()
    check(ctx.L.rz_autotune(ctx.ptr, arr, len(jobs), int(reps), byref(n)))
    return n.value


def autotune_table() -> dict:
    """The process-wide table: size class -> (tiles per grab, heavy-tile items)."""
    L = _lib.lib()
    n = L.rz_autotune_table(None, None, None, 0)
    cls, ch, hv = (c_uint32 * max(n, 1))(), (c_uint32 * max(n, 1))(), (c_uint32 * max(n, 1))()
    n = min(n, L.rz_autotune_table(cls, ch, hv, n))
    return {int(cls[i]): (int(ch[i]), int(hv[i])) for i in range(n)}


def autotune_clear() -> None:
    L = _lib.lib()
    for c in autotune_table():
        check(L.rz_autotune_set(c, 0, 0))


def grid(ncols: int, nrows: int, xmin: float = 0.0, ymin: float = 0.0, res: float = 1.0) -> RzRasterInfo:
    """An explicit grid: ncols x nrows pixels of `res` from (xmin, ymin)."""
    return _lib.raster_info(None, resolution=(res, res), extent=(xmin, ymin, xmin + ncols * res, ymin + nrows * res))


__all__ = ["autotune", "autotune_clear", "autotune_table", "BatchGraph", "Context", "DeviceBuffer", "DeviceGeometry", "Job", "Server", "rasterize_device", "batch_rasterize",
           "grid", "value_bits"]
