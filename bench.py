"""Benchmark of the rasterisation hot path on one MI355X.

Workload (the large single-raster case): 1,000,000 random convex polygons
with 4-12 vertices and diameters log-uniform in [2, 200] px, 10 % of them
MultiPolygons of 2-4 parts and 20 % with 1-3 holes at half scale, on a
40,000 x 40,000 grid (res 1), an int32 field uniform on [0, 1e6) and a
`by` column of 8 classes (8 output bands).  One step rasterises it with
each of count, any, first, last, min and max into uint8, int32 and float32
outputs (18 calls), with geometry, fields and output resident in HBM.

Prints one JSON line (see the keys below).  Timing: W untimed warm-up steps,
then exactly K steps bracketed by device synchronisation; the tile kernel's
time comes from HIP events recorded on the context's own stream.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from ctypes import byref, c_double, c_uint64, c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from rusterize_amd import _lib  # noqa: E402
from rusterize_amd._lib import Geometry, check  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
FUNS = ["count", "any", "first", "last", "min", "max"]
DTYPES = ["uint8", "int32", "float32"]
W = H = 40000
N_POLY = 1_000_000
N_BANDS = 8


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def make_inputs(n_poly: int):
    g = Geometry.synth(2, 2, n_poly, float(W), float(H), 2.0, 200.0, 0.1, 0.2)
    f32 = np.empty(g.n_geoms, np.int32)
    check(_lib.lib().rz_synth_field(6, 0, 202, g.n_geoms, 0.0, 1e6, f32.ctypes.data))
    lay = np.empty(g.n_geoms, np.uint32)
    check(_lib.lib().rz_synth_field(2, 0, 203, g.n_geoms, 0.0, float(N_BANDS), lay.ctypes.data))
    return g, f32, lay


def b_alg(g: Geometry, esz: int) -> int:
    """Algorithmic bytes of one call: 16 B per input vertex, the field and
    band index per geometry, and the raster written once."""
    return 16 * g.n_vertices + esz * g.n_geoms + 4 * g.n_geoms + esz * N_BANDS * W * H


def b_alg_tile(g: Geometry, esz: int) -> int:
    """Algorithmic bytes of the tile kernel: the raster written once, every
    polygon edge record (32 B) and field value read once."""
    return esz * N_BANDS * W * H + 32 * g.n_vertices + esz * g.n_geoms


def cpu_baseline(g, f32, lay, seconds_cap: float):
    """The CPU restatement timed on a bounded sample: the top 5,000-row band
    (one eighth of the grid, exactly the same computation for those rows)
    for fun=last into int32, one thread per output band like the reference's
    rayon loop."""
    from oracle import oracle

    info = oracle.make_info(W, 5000, 0.0, float(W), float(H - 5000), float(H), 1.0, 1.0)
    threads = min(N_BANDS, os.cpu_count() or 1)
    t0 = time.perf_counter()
    _, burns = oracle.rasterize(g, info, "int32", "last", field=f32, layer_of_geom=lay, n_layers=N_BANDS,
                                background=0, threads=threads, return_burns=True)
    dt = time.perf_counter() - t0
    return {
        "value": burns / dt,
        "unit": "pixel-burns/s",
        "cores": threads,
        "kind": "port",
        "sample": f"rows 0-5000 of 40000 (8 bands x 5000 x 40000 int32), fun=last, all 1M polygons; "
                  f"{burns} burns in {dt:.2f} s incl. background fill; CPU {cpu_model()}, nproc={os.cpu_count()}",
    }


def copy_calibration(nbytes: int) -> int:
    """One device-to-device copy of nbytes (after a warm-up copy): reading
    and writing nbytes each is what a traffic counter should report for it."""
    L = _lib.lib()
    ctx = c_void_p()
    check(L.rz_context_create(0, byref(ctx)))
    src, dst = c_void_p(), c_void_p()
    check(L.rz_device_malloc(ctx, nbytes, byref(src)))
    check(L.rz_device_malloc(ctx, nbytes, byref(dst)))
    for _ in range(2):
        check(L.rz_memcpy_d2d(ctx, dst, src, nbytes))
        check(L.rz_context_synchronize(ctx))
    L.rz_device_free(ctx, src)
    L.rz_device_free(ctx, dst)
    L.rz_context_destroy(ctx)
    print(json.dumps({"copy_bytes": nbytes}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--polygons", type=int, default=N_POLY)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only", default=None,
                    help="DTYPE:FUN -- one warm-up and one timed call of that combination only "
                         "(for counter collection)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run only a device-to-device copy of this many bytes (the plain copy "
                         "kernel that calibrates HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_calibration(args.copy_bytes)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the rasteriser shards by output rows with no exchange "
              "(see DESIGN.md); multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    L = _lib.lib()
    t0 = time.perf_counter()
    g, f32, lay = make_inputs(args.polygons)
    log(f"inputs: {g.n_geoms} geometries, {g.n_rings} rings, {g.n_vertices} vertices "
        f"({time.perf_counter() - t0:.1f} s)")

    ctx = c_void_p()
    check(L.rz_context_create(0, byref(ctx)))
    dg = c_void_p()
    gs = g.as_struct()
    check(L.rz_device_geometry_upload(ctx, byref(gs), byref(dg)))
    d_field = {}
    for dt in DTYPES:
        host = f32.astype(dt)  # numpy cast: uint8 wraps
        p = c_void_p()
        check(L.rz_device_malloc(ctx, host.nbytes, byref(p)))
        check(L.rz_memcpy_h2d(ctx, p, host.ctypes.data, host.nbytes))
        d_field[dt] = p
    d_lay = c_void_p()
    check(L.rz_device_malloc(ctx, lay.nbytes, byref(d_lay)))
    check(L.rz_memcpy_h2d(ctx, d_lay, lay.ctypes.data, lay.nbytes))
    out_bytes = 4 * N_BANDS * W * H
    d_out = c_void_p()
    check(L.rz_device_malloc(ctx, out_bytes, byref(d_out)))
    info = _lib.raster_info(None, resolution=(1.0, 1.0), extent=(0, 0, W, H))
    bg = {dt: np.zeros(1, dtype=dt) for dt in DTYPES}

    def call(dt, fun):
        check(L.rz_rasterize_device(ctx, dg, byref(info), _lib.dtype_code(dt), _lib.fun_code(fun), d_field[dt],
                                    None, d_lay, N_BANDS, bg[dt].ctypes.data, 0, 0, d_out))

    def step():
        for dt in DTYPES:
            for fun in FUNS:
                call(dt, fun)

    if args.only:
        dt, fun = args.only.split(":")
        call(dt, fun)  # warm-up
        check(L.rz_context_synchronize(ctx))
        check(L.rz_context_set_profiling(ctx, 1))
        L.rz_context_reset_stage_times(ctx)
        call(dt, fun)
        check(L.rz_context_synchronize(ctx))
        times = (c_double * 8)()
        n = L.rz_context_stage_times(ctx, times, 8)
        print(json.dumps({"only": args.only, "stages_ms": {L.rz_stage_name(i).decode(): times[i] for i in range(n)},
                          "b_alg_tile": b_alg_tile(g, np.dtype(dt).itemsize)}), flush=True)
        return 0

    # HIP events around every stage of every call, recorded on the context's
    # stream and resolved after the timed region (no added synchronisation)
    check(L.rz_context_set_profiling(ctx, 1))
    for _ in range(args.warmup):
        step()
    check(L.rz_context_synchronize(ctx))
    L.rz_context_reset_stage_times(ctx)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        log(f"step {k + 1}/{args.steps} queued")
    check(L.rz_context_synchronize(ctx))
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    times = (c_double * 8)()
    n = L.rz_context_stage_times(ctx, times, 8)
    stages = {L.rz_stage_name(i).decode(): times[i] / args.steps for i in range(n)}  # ms per step
    tile_ms = stages["tile"]  # summed over the 18 tile launches of a step

    # per-dtype breakdown (untimed): 6 calls each
    by_dtype = {}
    for dt in DTYPES:
        L.rz_context_reset_stage_times(ctx)
        for fun in FUNS:
            call(dt, fun)
        n = L.rz_context_stage_times(ctx, times, 8)
        st = {L.rz_stage_name(i).decode(): times[i] / len(FUNS) for i in range(n)}
        byt = b_alg_tile(g, np.dtype(dt).itemsize)
        by_dtype[dt] = {"tile_ms": round(st["tile"], 4), "GB/s": round(byt / (st["tile"] * 1e-3) / 1e9, 1),
                        "frac": round(byt / (st["tile"] * 1e-3) / HBM_PEAK, 4),
                        "stages_ms": {k: round(v, 4) for k, v in st.items()}}
    # exact burn counts (profiling level 2 adds counters; untimed)
    check(L.rz_context_set_profiling(ctx, 2))
    burns_call = {}
    for dt in DTYPES:
        call(dt, "max")
        burns, pairs, tiles = c_uint64(), c_uint64(), c_uint64()
        check(L.rz_context_last_counters(ctx, byref(burns), byref(pairs), byref(tiles)))
        burns_call[dt] = (burns.value, pairs.value, tiles.value)
        by_dtype[dt]["pairs"] = pairs.value
        by_dtype[dt]["tiles"] = tiles.value
    check(L.rz_context_set_profiling(ctx, 0))
    # every fun writes the same pixels, so a step burns 6x each dtype's count
    burns_per_step = sum(len(FUNS) * burns_call[dt][0] for dt in DTYPES)
    tile_bytes = len(FUNS) * sum(b_alg_tile(g, np.dtype(dt).itemsize) for dt in DTYPES)
    call_bytes = len(FUNS) * sum(b_alg(g, np.dtype(dt).itemsize) for dt in DTYPES)
    call_ms = sum(stages.values())
    achieved = tile_bytes / (tile_ms * 1e-3)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(g, f32, lay, 30.0)

    for p in list(d_field.values()) + [d_lay, d_out]:
        L.rz_device_free(ctx, p)
    L.rz_device_geometry_free(dg)
    L.rz_context_destroy(ctx)

    result = {
        "metric": "pixel-burns/s + HBM GB/s vs roofline; batched/served vs plain p50/p99",
        "value": burns_per_step / (ms_per_step * 1e-3),
        "unit": "pixel-burns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 crossings; u8/i32/f32 folds",
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {
            "workload": f"{g.n_geoms} convex polygons (10% multi, 20% holed) on {W}x{H}, by=8 bands, "
                        f"funs {'/'.join(FUNS)} x dtypes {'/'.join(DTYPES)} per step",
            "polygons": g.n_geoms, "vertices": g.n_vertices, "grid": [H, W], "bands": N_BANDS,
            "calls_per_step": len(FUNS) * len(DTYPES), "burns_per_step": burns_per_step,
            "stages_ms_per_step": {k: round(v, 3) for k, v in stages.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved / 1e9, 1),
            "peak": HBM_PEAK / 1e9,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK, 4),
            "traffic": None,
            "kernel": "k_tile (18 launches per step; bytes-weighted)",
            "by_dtype": by_dtype,
            "call_GB/s": round(call_bytes / (call_ms * 1e-3) / 1e9, 1),
            "call_frac": round(call_bytes / (call_ms * 1e-3) / HBM_PEAK, 4),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
