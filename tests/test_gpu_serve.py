"""GPU tests of the serving loop (rz_server): requests submitted while
others run are batched with them, and every result equals the same request
run alone (rz_rasterize_device) bit for bit; admission respects the memory
budget."""
import threading
import time

import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib
from rusterize_amd._lib import Geometry
from rusterize_amd.device import (Context, DeviceBuffer, DeviceGeometry, Job, Server, grid,
                                  rasterize_device)
from tests.test_gpu_parity import assert_same, field_for, mixed_geometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def request_mix(ctx, n, seed):
    """n requests over a few resident geometry sets; later requests reuse a
    set with a new extent, field and function (the cached-geometry case)."""
    sets = [Geometry.synth(1, 71, 300, 900.0, 700.0, 3.0, 60.0),
            Geometry.synth(2, 72, 1500, 900.0, 700.0, 2.0, 100.0, 0.1, 0.3),
            Geometry.synth(3, 73, 0, 900.0, 700.0, 200, 500, 40),
            Geometry.synth(4, 74, 400, 900.0, 700.0, 2.0, 40.0, 4.0, 600.0),
            mixed_geometry()]
    dev = [DeviceGeometry(ctx, g) for g in sets]
    rng = np.random.default_rng(seed)
    reqs, bufs = [], []
    for k in range(n):
        gi = int(rng.integers(0, len(sets)))
        g = sets[gi]
        dt = ["float64", "float32", "int32", "uint8", "int16"][int(rng.integers(0, 5))]
        fun = _lib.FUNS[int(rng.integers(0, 7))]
        nc, nr = int(rng.integers(16, 900)), int(rng.integers(16, 700))
        x0, y0 = float(rng.integers(-40, 300)), float(rng.integers(-40, 300))
        field = field_for(dt, g.n_geoms, seed=1000 + k)
        d_field = DeviceBuffer.from_array(ctx, field)
        bufs.append(d_field)
        bg = float("nan") if dt.startswith("float") else 0
        reqs.append((Job(dev[gi], grid(nc, nr, x0, y0), dt, fun, None, field=d_field, background=bg),
                     g, field))
    return reqs, dev, bufs


def alone(ctx, job):
    out = DeviceBuffer(ctx, job.out_bytes)
    j = Job(job.geometry, job.info, job.dtype, job.fun, out, field=job.field, background=job.background)
    rasterize_device(ctx, j)
    ctx.synchronize()
    r = j.result()
    out.free()
    return r


def test_served_results_equal_alone(ctx):
    reqs, dev, bufs = request_mix(ctx, 40, seed=3)
    srv = Server(0, budget_bytes=4 * 10**9, n_streams=2, max_batch=8)
    try:
        tickets = []

        def producer():  # a few spaced arrivals, then a burst that queues behind them
            for k, (job, _, _) in enumerate(reqs):
                tickets.append(srv.submit(job))
                if k < 8:
                    time.sleep(0.0005)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        for k, ((job, g, field), t) in enumerate(zip(reqs, tickets)):
            got = srv.result(ctx, t)
            assert_same(got, alone(ctx, job), f"request {k} served vs alone")
            if k % 8 == 0:  # and a sample against the oracle
                ref = oracle.rasterize(g, oracle.copy_info(job.info), job.dtype, job.fun, field=field,
                                       background=job.background)
                assert_same(got, ref, f"request {k} served vs oracle")
            srv.release(t)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["queued"] == 0
        assert st["batches"] < len(reqs)  # requests were batched together
        assert st["memory_peak"] <= st["memory_budget"]
    finally:
        srv.close()
        for b in bufs:
            b.free()
        for d in dev:
            d.free()


def test_admission_waits_for_memory_and_rejects_oversized(ctx):
    g = Geometry.synth(1, 81, 200, 1000.0, 1000.0, 3.0, 40.0)
    dg = DeviceGeometry(ctx, g)
    field = DeviceBuffer.from_array(ctx, field_for("float64", g.n_geoms, seed=1))
    # each output is 8 MB; the budget holds only a few at once
    srv = Server(0, budget_bytes=40 * 10**6, n_streams=2, max_batch=4)
    try:
        job = Job(dg, grid(1000, 1000), "float64", "sum", None, field=field, background=float("nan"))
        ref = alone(ctx, job)
        tickets = [srv.submit(job) for _ in range(12)]
        for t in tickets:  # completing requests and releasing them lets the rest in
            assert_same(srv.result(ctx, t), ref, "budgeted request")
            srv.release(t)
        assert srv.stats()["memory_peak"] <= 40 * 10**6
        big = Job(dg, grid(4000, 4000), "float64", "sum", None, field=field, background=float("nan"))
        t = srv.submit(big)  # 128 MB output: can never fit
        with pytest.raises(_lib.RusterizeNativeError, match="budget"):
            srv.wait(t, timeout_ms=20000)
    finally:
        srv.close()
        field.free()
        dg.free()


def test_served_stream_requests_match_oracle(ctx):
    # the served workload's request stream at test size: Poisson arrivals
    # (500 requests/s) drawn from the seeded batched-job stream (jobs of at most
    # 2,048 px and 150,000 vertices, generated on the device), 30 % of them
    # reusing a cached geometry set with a new window, field and function;
    # every response against the oracle
    seed = 4
    picks = []
    for k in range(600):
        sp = _lib.job_spec(seed, k)
        if sp["side"] <= 2048 and sp["n_vertices"] <= 150000:
            picks.append((k, sp))
        if len(picks) == 45:
            break
    rng = np.random.default_rng(11)
    cache, reqs, bufs, devs = [], [], [], []
    srv = Server(0, budget_bytes=8 * 10**9, n_streams=2, max_batch=16)
    try:
        t, nxt = 0.0, 0
        for _ in range(64):
            t += rng.exponential(1.0 / 500.0)
            if cache and (rng.random() < 0.3 or nxt == len(picks)):
                k, sp, dg, host = cache[int(rng.integers(0, len(cache)))]
                w = max(16, int(sp["side"] * rng.uniform(0.3, 1.0)))
                x0, y0 = float(rng.integers(-64, sp["side"] - w + 64)), float(rng.integers(-64, sp["side"] - w + 64))
                field = rng.random(host.n_geoms)
                info, fun = grid(w, w, x0, y0), _lib.FUNS[int(rng.integers(0, 7))]
            else:
                k, sp = picks[nxt]
                nxt += 1
                dg = DeviceGeometry.synth_job(ctx, seed, k)
                devs.append(dg)
                host = Geometry.synth_job(seed, k)
                cache.append((k, sp, dg, host))
                field = _lib.job_field(seed, k)
                info, fun = grid(sp["side"], sp["side"]), sp["fun"]
            d_field = DeviceBuffer.from_array(ctx, field)
            bufs.append(d_field)
            reqs.append((t, Job(dg, info, "float64", fun, None, field=d_field, background=float("nan")), host, field, k))
        ctx.synchronize()
        tickets = [None] * len(reqs)
        t0 = time.perf_counter()

        def producer():
            for i, (ta, job, _, _, _) in enumerate(reqs):
                d = t0 + ta - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(job)

        th = threading.Thread(target=producer)
        th.start()
        th.join()
        for i, (ta, job, host, field, k) in enumerate(reqs):
            got = srv.result(ctx, tickets[i])
            ref = oracle.rasterize(host, oracle.copy_info(job.info), "float64", job.fun, field=field,
                                   background=float("nan"))
            assert_same(got, ref, f"served request {i} (stream job {k}, {job.fun}, {job.shape})")
            srv.release(tickets[i])
        # (at 500 requests/s most requests find the server idle and run as a
        # batch of one; batching under load is covered above)
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["batches"] <= len(reqs)
    finally:
        srv.close()
        for b in bufs:
            b.free()
        for d in devs:
            d.free()
