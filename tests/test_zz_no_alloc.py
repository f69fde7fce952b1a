"""The allocation-free entry point (rz_rasterize_device_noalloc) against
rz_rasterize_device, through the C ABI: two calls in a row on one device, the
second with a smaller input, each bit-identical to the allocating call and to
the oracle; oversized inputs fail with a capacity error."""
from ctypes import byref

import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib
from rusterize_amd._lib import Geometry
from rusterize_amd.device import Context, DeviceBuffer, DeviceGeometry, Job, grid, rasterize_device
from tests.test_gpu_parity import assert_same, field_for, mixed_geometry

pytestmark = pytest.mark.gpu


def noalloc(ctx, job):
    dt = _lib.np_dtype(job.dtype)
    bg = np.array([job.background]).astype(dt)
    scalar = np.array([job.scalar]).astype(dt)
    return ctx.L.rz_rasterize_device_noalloc(
        ctx.ptr, job.geometry.ptr, byref(job.info), _lib.dtype_code(job.dtype), _lib.fun_code(job.fun),
        job.field.ptr if job.field is not None else None, None if job.field is not None else scalar.ctypes.data,
        job.layers.ptr if job.layers is not None else None, job.n_layers if job.layers is not None else 1,
        bg.ctypes.data, 0, 0, job.out.ptr)


# (geometry, ncols, nrows, xmin, ymin, dtype, fun, bands, background): the
# second case is smaller in every respect (vertices, items, tiles)
CASES = [
    (lambda: Geometry.synth(3, 41, 0, 900.0, 600.0, 200, 500, 40), 900, 600, 0, 0, "float32", "sum", 2,
     float("nan")),
    (mixed_geometry, 300, 200, 0, 0, "uint8", "count", 1, 0),
]


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_noalloc_equals_allocating_call_twice_in_a_row(ctx):
    for k, (make, nc, nr, x0, y0, dt, fun, bands, bg) in enumerate(CASES):
        host = make()
        print(f"noalloc case {k}: {host.n_geoms} geometries, {nc}x{nr}, {dt} {fun}", flush=True)
        dg = DeviceGeometry(ctx, host)
        field = field_for(dt, host.n_geoms, seed=50 + k)
        layers = (np.arange(host.n_geoms) % bands).astype(np.uint32) if bands > 1 else None
        info = grid(nc, nr, x0, y0)
        d_field = DeviceBuffer.from_array(ctx, field)
        d_layers = DeviceBuffer.from_array(ctx, layers) if layers is not None else None
        nbytes = bands * nr * nc * np.dtype(dt).itemsize
        ref_job = Job(dg, info, dt, fun, DeviceBuffer(ctx, nbytes), field=d_field, layers=d_layers,
                      n_layers=bands, background=bg)
        na_job = Job(dg, info, dt, fun, DeviceBuffer(ctx, nbytes), field=d_field, layers=d_layers,
                     n_layers=bands, background=bg)
        try:
            rasterize_device(ctx, ref_job)
            ctx.synchronize()
            st = noalloc(ctx, na_job)
            assert st == 0, _lib.last_error()
            got, want = na_job.result(), ref_job.result()
            assert_same(got, want, f"case {k}: allocation-free vs allocating call")
            oref = oracle.rasterize(host, oracle.copy_info(info), dt, fun, field=field, layer_of_geom=layers,
                                    n_layers=bands, background=bg)
            assert_same(got, oref, f"case {k}: allocation-free vs oracle")
        finally:
            for b in (ref_job.out, na_job.out, d_field) + ((d_layers,) if d_layers is not None else ()):
                b.free()
            dg.free()


def test_noalloc_rejects_a_workspace_beyond_its_array(ctx):
    # about 3.6M vertices need more than the 64 MiB array
    host = Geometry.synth(2, 42, 400000, 4000.0, 4000.0, 2.0, 20.0)
    dg = DeviceGeometry(ctx, host)
    out = DeviceBuffer(ctx, 64 * 64 * 8)
    try:
        job = Job(dg, grid(64, 64, 0, 0), "float64", "last", out, scalar=1, background=0)
        st = noalloc(ctx, job)
        assert st == _lib.RZ_CAPACITY_ERROR
        assert "allocation-free" in _lib.last_error()
    finally:
        out.free()
        dg.free()
