"""GPU parity of batches (rz_batch_rasterize): every job of a batch of uneven
extents, dtypes, pixel functions and geometry sets equals the oracle and its
own single call bit for bit, whatever the batch composition or order."""
import numpy as np
import pytest

from oracle import oracle
from rusterize_amd import _lib
from rusterize_amd._lib import Geometry
from rusterize_amd.device import Context, DeviceBuffer, DeviceGeometry, Job, batch_rasterize, grid, rasterize_device
from tests.test_gpu_parity import assert_same, field_for, mixed_geometry

pytestmark = pytest.mark.gpu


def geometry_sets():
    return [
        Geometry.synth(1, 31, 400, 1000.0, 700.0, 3.0, 60.0),                    # stars
        Geometry.synth(2, 32, 2500, 1000.0, 700.0, 2.0, 120.0, 0.1, 0.3),       # convex, multi, holes
        Geometry.synth(3, 33, 0, 1000.0, 700.0, 300, 800, 60),                  # lines, points, collections
        mixed_geometry(),                                                        # all kinds + edge cases
        Geometry.from_wkt(["POLYGON EMPTY", "POINT (3.5 2.5)"]),
        Geometry.synth(2, 34, 20, 1000.0, 700.0, 50.0, 400.0, 0.0, 0.5),        # few large polygons
    ]


# (geometry set, ncols, nrows, xmin, ymin, dtype, fun, bands, background)
SPECS = [
    (0, 1000, 700, 0, 0, "float64", "sum", 1, float("nan")),
    (0, 257, 129, 100.5, 300.25, "uint8", "count", 1, 0),
    (1, 1000, 700, 0, 0, "int32", "first", 4, 0),
    (1, 333, 517, -20, -13, "float32", "max", 1, float("nan")),
    (1, 64, 64, 500, 500, "uint16", "last", 1, 7),
    (2, 1000, 700, 0, 0, "float32", "sum", 1, float("nan")),
    (2, 800, 300, 50, 200, "int16", "min", 2, 0),
    (3, 300, 200, 0, 0, "uint8", "any", 1, 0),
    (3, 300, 200, 0, 0, "float64", "last", 3, 0),
    (3, 129, 65, 10, 20, "int64", "sum", 1, 0),
    (4, 8, 6, 0, 0, "float64", "sum", 1, 0),
    (5, 1000, 700, 0, 0, "uint32", "min", 1, 0),
    (5, 2048, 1100, -500, -200, "int8", "count", 1, 0),
    (0, 1, 1, 10, 10, "float32", "first", 1, 0),
]


class Batch:
    def __init__(self, ctx, specs, seed=0):
        self.ctx = ctx
        self.host_geoms = geometry_sets()
        self.dev_geoms = [DeviceGeometry(ctx, g) for g in self.host_geoms]
        self.jobs, self.refs = [], []
        rng = np.random.default_rng(seed)
        self.buffers = []
        for k, (gi, nc, nr, x0, y0, dt, fun, bands, bg) in enumerate(specs):
            g = self.host_geoms[gi]
            field = field_for(dt, g.n_geoms, seed=100 + k)
            if g.n_geoms:
                field[:: 13] = 0
            layers = rng.integers(0, bands, g.n_geoms).astype(np.uint32) if bands > 1 else None
            info = grid(nc, nr, x0, y0)
            d_field = DeviceBuffer.from_array(ctx, field)
            d_layers = DeviceBuffer.from_array(ctx, layers) if layers is not None else None
            job = Job(self.dev_geoms[gi], info, dt, fun, DeviceBuffer(ctx, bands * nr * nc * np.dtype(dt).itemsize),
                      field=d_field, layers=d_layers, n_layers=bands, background=bg)
            self.jobs.append(job)
            self.buffers += [d_field, job.out] + ([d_layers] if d_layers is not None else [])
            self.refs.append(oracle.rasterize(g, oracle.copy_info(info), dt, fun, field=field, layer_of_geom=layers,
                                              n_layers=bands, background=bg))

    def free(self):
        for b in self.buffers:
            b.free()
        for g in self.dev_geoms:
            g.free()


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("binning", [None, "always", "never"])
def test_batch_matches_oracle_and_single_calls(ctx, binning, monkeypatch):
    if binning:
        monkeypatch.setenv("RUSTERIZE_BINNING", binning)
    b = Batch(ctx, SPECS)
    try:
        batch_rasterize(ctx, b.jobs)
        ctx.synchronize()
        batched = [j.result() for j in b.jobs]
        for k, (got, ref) in enumerate(zip(batched, b.refs)):
            assert_same(got, ref, f"job {k} {SPECS[k]} vs oracle")
        for k, j in enumerate(b.jobs):
            rasterize_device(ctx, j)
            ctx.synchronize()
            assert_same(j.result(), batched[k], f"job {k} {SPECS[k]} single vs batch")
    finally:
        b.free()


def test_batch_composition_and_order_do_not_matter(ctx):
    b = Batch(ctx, SPECS, seed=1)
    try:
        order = list(range(len(SPECS)))[::-1]
        batch_rasterize(ctx, [b.jobs[k] for k in order])
        ctx.synchronize()
        for k in order:
            assert_same(b.jobs[k].result(), b.refs[k], f"reversed job {k}")
        for part in (order[::2], order[1::2], [3], [10, 0]):
            for k in part:  # clobber the outputs first
                DeviceBuffer.upload(b.jobs[k].out, np.full(b.jobs[k].out_bytes, 0xA5, np.uint8))
            batch_rasterize(ctx, [b.jobs[k] for k in part])
            ctx.synchronize()
            for k in part:
                assert_same(b.jobs[k].result(), b.refs[k], f"subset {part} job {k}")
    finally:
        b.free()


def test_batch_of_many_small_jobs(ctx):
    # 300 jobs over one resident geometry set, each a different window / function
    g = Geometry.synth(1, 41, 300, 1000.0, 700.0, 3.0, 40.0)
    dg = DeviceGeometry(ctx, g)
    rng = np.random.default_rng(7)
    field = field_for("float64", g.n_geoms, seed=8)
    d_field = DeviceBuffer.from_array(ctx, field)
    jobs, refs, bufs = [], [], []
    try:
        for k in range(300):
            nc, nr = int(rng.integers(1, 300)), int(rng.integers(1, 300))
            x0, y0 = float(rng.integers(-50, 900)), float(rng.integers(-50, 600))
            fun = _lib.FUNS[k % 7]
            info = grid(nc, nr, x0, y0)
            out = DeviceBuffer(ctx, nr * nc * 8)
            bufs.append(out)
            jobs.append(Job(dg, info, "float64", fun, out, field=d_field, background=float("nan")))
            refs.append(oracle.rasterize(g, oracle.copy_info(info), "float64", fun, field=field,
                                         background=float("nan")))
        batch_rasterize(ctx, jobs)
        ctx.synchronize()
        for k, (j, ref) in enumerate(zip(jobs, refs)):
            assert_same(j.result(), ref, f"small job {k}")
    finally:
        for b in bufs:
            b.free()
        d_field.free()
        dg.free()


def test_batch_rejects_bad_jobs(ctx):
    g = Geometry.from_wkt(["POINT (1 1)"])
    dg = DeviceGeometry(ctx, g)
    out = DeviceBuffer(ctx, 64)
    try:
        ok = Job(dg, grid(4, 4), "uint8", "sum", out)
        bad = Job(dg, grid(4, 4), "uint8", "sum", out)
        s = bad.as_struct()
        s.fun = 9
        arr = (_lib.RzJob * 2)(ok.as_struct(), s)
        with pytest.raises(ValueError, match="Unknown pixel function"):
            _lib.check(ctx.L.rz_batch_rasterize(ctx.ptr, arr, 2))
        batch_rasterize(ctx, [])  # an empty batch is a no-op
    finally:
        out.free()
        dg.free()


def test_job_stream_device_generator_matches_oracle(ctx):
    # jobs of the seeded 4,096-job stream: generated on the device, batched,
    # and checked against the oracle on the host generator's geometry
    seed, picks = 4, []
    for k in range(400):
        sp = _lib.job_spec(seed, k)
        if sp["n_vertices"] <= 150000 and sp["side"] <= 2048:
            picks.append((k, sp))
        if len(picks) == 6:
            break
    geoms, jobs, refs, bufs = [], [], [], []
    try:
        for k, sp in picks:
            dg = DeviceGeometry.synth_job(ctx, seed, k)
            geoms.append(dg)
            field = _lib.job_field(seed, k)
            d_field = DeviceBuffer.from_array(ctx, field)
            out = DeviceBuffer(ctx, sp["side"] ** 2 * 8)
            bufs += [d_field, out]
            info = grid(sp["side"], sp["side"])
            jobs.append(Job(dg, info, "float64", sp["fun"], out, field=d_field, background=float("nan")))
            refs.append(oracle.rasterize(Geometry.synth_job(seed, k), oracle.copy_info(info), "float64", sp["fun"],
                                         field=field, background=float("nan")))
        batch_rasterize(ctx, jobs)
        ctx.synchronize()
        for (k, sp), j, ref in zip(picks, jobs, refs):
            assert_same(j.result(), ref, f"stream job {k} {sp}")
    finally:
        for b in bufs:
            b.free()
        for g in geoms:
            g.free()


@pytest.mark.parametrize("wide", [None, "always"])
def test_captured_batch_graph_replays(ctx, wide, monkeypatch):
    # a captured batch replayed several times, with eager batches between the
    # replays and new field values before one of them; every replay equals the
    # eager result of the same inputs (also with every polygon item taking the
    # wide coverage pre-pass inside the graph)
    from rusterize_amd.device import BatchGraph
    if wide:
        monkeypatch.setenv("RUSTERIZE_WIDE", wide)
    b = Batch(ctx, SPECS, seed=2)
    g = None
    try:
        batch_rasterize(ctx, b.jobs)
        ctx.synchronize()
        eager = [j.result() for j in b.jobs]
        g = BatchGraph(b.jobs, headroom=1.0)
        for rep in range(3):
            for j in b.jobs:  # clobber the outputs
                j.out.upload(np.full(j.out_bytes, 0x5A, np.uint8))
            g.launch()
            g.synchronize()
            for k, j in enumerate(b.jobs):
                assert_same(j.result(), eager[k], f"replay {rep} job {k}")
            batch_rasterize(ctx, b.jobs[::-1])  # eager work between replays
            ctx.synchronize()
        # refill one job's field: the next replay reads the new values
        k = 2
        f2 = field_for(b.jobs[k].dtype, b.host_geoms[SPECS[k][0]].n_geoms, seed=4242)
        b.jobs[k].field.upload(f2)
        g.launch()
        g.synchronize()
        ref = oracle.rasterize(b.host_geoms[SPECS[k][0]], oracle.copy_info(b.jobs[k].info), b.jobs[k].dtype,
                               b.jobs[k].fun, field=f2, layer_of_geom=None if SPECS[k][7] == 1 else None,
                               n_layers=1, background=b.jobs[k].background) if SPECS[k][7] == 1 else None
        batch_rasterize(ctx, [b.jobs[k]])
        ctx.synchronize()
        eager_new = b.jobs[k].result()
        g.launch()
        g.synchronize()
        assert_same(b.jobs[k].result(), eager_new, "replay after a field refill")
        if ref is not None:
            assert_same(eager_new, ref, "refilled field vs oracle")
    finally:
        if g is not None:
            g.close()
        b.free()


def test_autotuner_fills_table_and_keeps_results(ctx):
    from rusterize_amd.device import autotune, autotune_clear, autotune_table
    b = Batch(ctx, SPECS[:6], seed=5)
    try:
        autotune_clear()
        n = autotune(ctx, b.jobs[:4], reps=1)
        table = autotune_table()
        assert n >= 1 and len(table) == n
        for cls, (chunk, heavy) in table.items():
            assert chunk in (1, 2, 4, 8) and heavy in (16, 32, 64, 1 << 30)
        batch_rasterize(ctx, b.jobs)  # tuned launches give the same rasters
        ctx.synchronize()
        for k, j in enumerate(b.jobs):
            assert_same(j.result(), b.refs[k], f"tuned job {k}")
    finally:
        autotune_clear()
        b.free()


def test_large_stream_jobs_batched_and_replayed(ctx):
    # jobs of the seeded 4,096-job stream at full size: 7,652 px with rings of
    # up to 4,047 vertices (job 41), 31,047 polygons on 6,952 px (job 761),
    # 30,570 polygons on 681 px (job 267) and 8,009 px (job 334); batched,
    # then replayed from a captured graph, every output against the oracle
    # on the host generator's geometry
    from rusterize_amd.device import BatchGraph
    seed, picks = 4, (41, 267, 334, 761)
    geoms, jobs, refs, bufs = [], [], [], []
    g = None
    try:
        for k in picks:
            sp = _lib.job_spec(seed, k)
            dg = DeviceGeometry.synth_job(ctx, seed, k)
            geoms.append(dg)
            field = _lib.job_field(seed, k)
            d_field = DeviceBuffer.from_array(ctx, field)
            out = DeviceBuffer(ctx, sp["side"] ** 2 * 8)
            bufs += [d_field, out]
            info = grid(sp["side"], sp["side"])
            jobs.append(Job(dg, info, "float64", sp["fun"], out, field=d_field, background=float("nan")))
            print(f"stream job {k}: {sp}", flush=True)
            refs.append(oracle.rasterize(Geometry.synth_job(seed, k), oracle.copy_info(info), "float64", sp["fun"],
                                         field=field, background=float("nan")))
        batch_rasterize(ctx, jobs)
        ctx.synchronize()
        for k, j, ref in zip(picks, jobs, refs):
            assert_same(j.result(), ref, f"stream job {k} batched")
        g = BatchGraph(jobs, headroom=1.0)
        for rep in range(2):
            for j in jobs:
                j.out.upload(np.full(j.out_bytes, 0x3C, np.uint8))
            g.launch()
            g.synchronize()
            for k, j, ref in zip(picks, jobs, refs):
                assert_same(j.result(), ref, f"stream job {k} graph replay {rep}")
    finally:
        if g is not None:
            g.close()
        for b in bufs:
            b.free()
        for d in geoms:
            d.free()


def test_out_of_range_layer_fails_the_batch_and_spares_other_jobs(ctx):
    # a band index >= n_layers burns nothing and fails the call; the other
    # jobs of the batch still equal the oracle
    b = Batch(ctx, SPECS, seed=6)
    try:
        k_bad = 2  # 4 bands
        g = b.host_geoms[SPECS[k_bad][0]]
        layers = np.zeros(g.n_geoms, np.uint32)
        layers[5] = 4
        b.jobs[k_bad].layers.upload(layers)
        batch_rasterize(ctx, b.jobs)
        with pytest.raises(ValueError, match="layer index out of range"):
            ctx.synchronize()
        for k, j in enumerate(b.jobs):
            if k != k_bad:
                assert_same(j.result(), b.refs[k], f"job {k} next to the failing one")
        # the failing geometry alone is dropped: the rest of that job burns as
        # the oracle does without it
        keep = np.ones(g.n_geoms, bool)
        keep[5] = False
        field = field_for(SPECS[k_bad][5], g.n_geoms, seed=100 + k_bad)
        field[::13] = 0
        ref = oracle.rasterize(g.subset(np.flatnonzero(keep)), oracle.copy_info(b.jobs[k_bad].info), SPECS[k_bad][5],
                               SPECS[k_bad][6], field=field[keep], layer_of_geom=layers[keep], n_layers=4,
                               background=SPECS[k_bad][8])
        assert_same(b.jobs[k_bad].result(), ref, "failing job without its bad geometry")
        ctx.synchronize()  # the error was reported once
    finally:
        b.free()


def _star(cx, cy, rx, ry, n=200):
    """A closed star ring of n + 1 vertices in world coordinates (the grid
    below maps pixel row y to world y = 256 - y)."""
    k = np.arange(n)
    r = np.where(k % 2 == 0, 1.0, 0.45)
    ring = np.stack([cx + rx * r * np.cos(2 * np.pi * k / n), 256.0 - (cy + ry * r * np.sin(2 * np.pi * k / n))], 1)
    return np.vstack([ring, ring[:1]])


def _grown_sets(b_cy, b_ry):
    """Forty long rings (151 vertices) at rows b_cy +- b_ry, then one ring of
    10,001 vertices over rows 68..124; the same topology whatever b_cy, b_ry."""
    from tests.test_gpu_parity import polygon_rings
    rings = [[_star(12 + 12 * k, b_cy, 10, b_ry, 150)] for k in range(40)]
    rings.append([_star(256, 96, 200, 28, 10000)])
    return polygon_rings(rings)


@pytest.mark.parametrize("wide", [None, "always"])
def test_graph_replay_past_band_and_mask_capacities(ctx, wide, monkeypatch):
    # a batch captured while forty long rings were flat (one 16-row strip:
    # no edge bands, one strip of wide masks) and one large ring was banded,
    # replayed after the forty grew to two strips inside the same tiles:
    # their band entries outgrow the captured capacity, so the rings past it
    # (the large one among them) fall back to their whole edge list, and with
    # every polygon on the wide pre-pass the masks outgrow theirs too, so
    # the items past it take per-tile coverage; the replay still equals the
    # oracle.  Growing into more tiles outgrows the pairs and fails the
    # replay with a capacity error.
    from rusterize_amd.device import BatchGraph
    if wide:
        monkeypatch.setenv("RUSTERIZE_WIDE", wide)
    monkeypatch.setenv("RUSTERIZE_BINNING", "always")  # 41 items would be scanned, with no pairs
    flat, tall, taller = _grown_sets(8.0, 3.0), _grown_sets(16.0, 12.0), _grown_sets(100.0, 95.0)
    info = grid(512, 256, 0, 0)
    dg = DeviceGeometry(ctx, flat)
    field = field_for("float64", flat.n_geoms, seed=77)
    d_field = DeviceBuffer.from_array(ctx, field)
    out = DeviceBuffer(ctx, 512 * 256 * 8)
    jobs = [Job(dg, info, "float64", fun, out, field=d_field, background=float("nan")) for fun in ("sum",)]
    g = None
    try:
        g = BatchGraph(jobs, headroom=1.0)
        g.launch()
        g.synchronize()
        ref = oracle.rasterize(flat, oracle.copy_info(info), "float64", "sum", field=field, background=float("nan"))
        assert_same(jobs[0].result(), ref, "captured (flat)")
        dg.set_xy(tall.xy)
        for rep in range(2):
            g.launch()
            g.synchronize()
            ref = oracle.rasterize(tall, oracle.copy_info(info), "float64", "sum", field=field,
                                   background=float("nan"))
            assert_same(jobs[0].result(), ref, f"replay {rep} past the band / mask capacities")
            rasterize_device(ctx, jobs[0])  # eager work between replays
            ctx.synchronize()
            assert_same(jobs[0].result(), ref, "eager on the grown rings")
        dg.set_xy(taller.xy)
        g.launch()
        with pytest.raises(Exception, match="capture it again"):
            g.synchronize()
    finally:
        if g is not None:
            g.close()
        d_field.free()
        out.free()
        dg.free()
