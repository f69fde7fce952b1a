/*
 * rusterize_hip.h -- C ABI of the MI355X (gfx950) rasteriser.
 *
 * This is the drop-in boundary for rusterize's rasterisation path.  Every
 * entry point below names the reference interface it replaces (paths are
 * relative to the rusterize source tree).  All calls are plain C: pointers,
 * sizes and enums, no C++ or torch types, no exceptions across the boundary.
 *
 * Ownership: input buffers are borrowed for the duration of a call (the
 * reference borrows `&[Geometry]` and `ArrayView1`, rust/src/rasterize.rs:54-62);
 * outputs are caller-allocated.  Errors are returned as rz_status and the
 * message of the last failing call on the calling thread is available from
 * rz_last_error(); messages reuse the reference's wording
 * (rust/src/error.rs:3-12, rust/src/geo/raster.rs:50-156,
 * rust/src/rasterize.rs:207-229, rust/src/rasterization/pixel_functions.rs:18-33).
 *
 * Threading: calls on distinct rz_context objects may run concurrently;
 * one context serialises its work on its own HIP stream.
 */
#ifndef RUSTERIZE_HIP_H
#define RUSTERIZE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- enums -- */

typedef enum rz_status {
  RZ_OK = 0,
  RZ_VALUE_ERROR = 1,   /* RusterizeError::ValueError  (rust/src/error.rs:8) */
  RZ_RUNTIME_ERROR = 2, /* RusterizeError::RuntimeError (rust/src/error.rs:6) */
  RZ_TYPE_ERROR = 3,    /* unsupported dtype / encoding (python/src/rusterize.rs:28) */
  RZ_DEVICE_ERROR = 4,  /* HIP runtime failure or no usable gfx950 device */
  RZ_PARSE_ERROR = 5,   /* WKT / WKB decoding failure (python/src/geo/parse_geometry.rs:77-88) */
  RZ_CAPACITY_ERROR = 6 /* request exceeds a stated limit (memory budget, index width) */
} rz_status;

/* Output element type; order follows python/src/rusterize.rs:171-182. */
typedef enum rz_dtype {
  RZ_U8 = 0, RZ_U16 = 1, RZ_U32 = 2, RZ_U64 = 3,
  RZ_I8 = 4, RZ_I16 = 5, RZ_I32 = 6, RZ_I64 = 7,
  RZ_F32 = 8, RZ_F64 = 9
} rz_dtype;

/* Pixel function; rust/src/rasterization/pixel_functions.rs:8-16. */
typedef enum rz_fun {
  RZ_SUM = 0, RZ_FIRST = 1, RZ_LAST = 2, RZ_MIN = 3, RZ_MAX = 4, RZ_COUNT = 5, RZ_ANY = 6
} rz_fun;

/* How `sum` is reduced on the device.  ORDERED folds every pixel's writes
 * in geometry order and is bit-identical to the reference
 * (pixel_functions.rs:56-65).  UNORDERED adds the writes of lines and
 * points with LDS atomics as they land, in no fixed order, with no per-line
 * pass; a pixel is the sum of its writes' values (the background where it
 * has none).  For float outputs that differs from the reference by rounding
 * only: |got - ref| <= 8 n eps sum|v| for n writes of values v (eps of the
 * output type), provided no value is NaN and no partial sum equals the
 * background (the reference restarts a pixel there).  Integer outputs and
 * polygon-only launches always take the ordered fold. */
typedef enum rz_sum_mode { RZ_SUM_ORDERED = 0, RZ_SUM_UNORDERED = 1 } rz_sum_mode;

/* Item kinds of the flattened geometry model (see rz_geometry). */
typedef enum rz_item_kind {
  RZ_ITEM_POLYGONS = 0, /* all rings of a Polygon / MultiPolygon: one even-odd edge list */
  RZ_ITEM_LINES = 1,    /* all lines of a LineString / MultiLineString / Line           */
  RZ_ITEM_POINTS = 2    /* a Point or the points of a MultiPoint                        */
} rz_item_kind;

/* ------------------------------------------------------------ raster info -- */

/* rust/src/geo/raster.rs:9-20 (RasterInfo).  epsg < 0 means "None". */
typedef struct rz_raster_info {
  uint64_t ncols;
  uint64_t nrows;
  double xmin, xmax, ymin, ymax;
  double xres, yres;
  int32_t epsg;
} rz_raster_info;

/* --------------------------------------------------------- geometry model -- */

/*
 * Flattened, GeoArrow-like geometry set.  Geometry g (one field value, one
 * `by` key) owns items [geom_item_offsets[g], geom_item_offsets[g+1]); item i
 * has kind item_kind[i] and rings [item_ring_offsets[i], item_ring_offsets[i+1]);
 * ring r has vertices [ring_vertex_offsets[r], ring_vertex_offsets[r+1]) in
 * xy (interleaved x, y world coordinates).
 *
 * The mapping from geo-types follows the burn dispatch of
 * rust/src/rasterization/burn_geometry.rs:24-210:
 *   Point            -> one POINTS item with one vertex            (31-36)
 *   MultiPoint       -> one POINTS item, one ring of n vertices    (37-44)
 *   Polygon          -> one POLYGONS item: exterior + holes        (76-118)
 *   MultiPolygon     -> one POLYGONS item: every ring of every part (120-166)
 *   LineString/Line  -> one LINES item with one ring               (56-59, 168-187)
 *   MultiLineString  -> one LINES item: one ring per line          (189-210)
 *   GeometryCollection -> its members' items, in order, recursively (64-74)
 * Polygon rings are closed (first vertex repeated) as geo-types'
 * Polygon::new does; line rings are kept as given.
 */
typedef struct rz_geometry {
  int64_t n_geoms;
  int64_t n_items;
  int64_t n_rings;
  int64_t n_vertices;
  const int64_t* geom_item_offsets;   /* [n_geoms + 1] */
  const int32_t* item_kind;           /* [n_items]     */
  const int64_t* item_ring_offsets;   /* [n_items + 1] */
  const int64_t* ring_vertex_offsets; /* [n_rings + 1] */
  const double* xy;                   /* [2 * n_vertices] */
  const uint8_t* ring_flags;          /* [n_rings] or NULL; bit 0 = polygon interior ring.
                                         Only extent inference reads it: geo's BoundingRect
                                         of a Polygon uses the exterior ring alone. */
} rz_geometry;

/* Owned geometry buffers produced by the parsers / generators below. */
typedef struct rz_geometry_buffers rz_geometry_buffers;

/* ------------------------------------------------------------------ errors -- */

/* Message of the last failing call on this thread ("" if none). */
const char* rz_last_error(void);

/* Library version string. */
const char* rz_version(void);

/* ------------------------------------------------------------ host helpers -- */

/* Replaces RasterInfoBuilder::{build, build_with, finalize}
 * (rust/src/geo/raster.rs:50-156) as driven by RawRasterInfo::build
 * (python/src/geo/raster.rs:16-42).  When has_extent is 0 the extent is
 * inferred from `geoms` (bounding rectangle of every vertex). */
rz_status rz_raster_info_build(const rz_geometry* geoms, int has_shape, uint64_t nrows,
                               uint64_t ncols, int has_extent, const double* extent4,
                               int has_res, double xres, double yres, int tap, int32_t epsg,
                               rz_raster_info* out);

/* Replaces `impl FromStr for PixelFunction`
 * (rust/src/rasterization/pixel_functions.rs:18-33). */
rz_status rz_pixel_function_parse(const char* name, rz_fun* out);

/* Replaces the dtype match of python/src/rusterize.rs:166-183. */
rz_status rz_dtype_parse(const char* name, rz_dtype* out);
size_t rz_dtype_size(rz_dtype dtype);

/* Geometry ingestion; replaces parse_sequence_wkt / parse_sequence_wkb and
 * try_parse_*_to_geometry (python/src/geo/parse_geometry.rs:77-133).
 * `texts` / `blobs` are n pointers with explicit byte lengths. WKB geometries
 * that geo cannot represent (empty points) are skipped like the reference's
 * filter at parse_geometry.rs:113-116. */
rz_status rz_geometry_from_wkt(const char* const* texts, const int64_t* lengths, int64_t n,
                               rz_geometry_buffers** out);
rz_status rz_geometry_from_wkb(const uint8_t* const* blobs, const int64_t* lengths, int64_t n,
                               rz_geometry_buffers** out);
/* Build from caller-supplied flat arrays (copied). */
rz_status rz_geometry_from_arrays(const rz_geometry* src, rz_geometry_buffers** out);
void rz_geometry_view(const rz_geometry_buffers* bufs, rz_geometry* view);
void rz_geometry_free(rz_geometry_buffers* bufs);

/* Seeded synthetic workloads (SplitMix64) used by tests and benchmarks.
 * kind: 1 = star polygons, 2 = convex polygons with parts/holes,
 *       3 = random-walk lines + points + collections, 4 = mixed job.
 * The four knobs p0..p3 are interpreted per kind (documented at the
 * generators in rusterize_amd/csrc/rz_host.cpp). */
rz_status rz_synth_geometry(int kind, uint64_t seed, int64_t count, double grid_w, double grid_h,
                            double p0, double p1, double p2, double p3,
                            rz_geometry_buffers** out);
/* Fill n values of `dtype` from a seeded distribution:
 * dist 0 = integer-valued uniform [lo, hi) cast with numpy's wrap,
 * dist 1 = standard normal rounded to the dtype,
 * dist 2 = real uniform [lo, hi) rounded to the dtype. */
rz_status rz_synth_field(rz_dtype dtype, int dist, uint64_t seed, int64_t n, double lo, double hi,
                         void* out);

/* No reference counterpart: the seeded job stream of the batched and served
 * workloads (job k: square extent 256..8192 px, 10..50,000 star polygons of
 * 4..4,096 vertices, all log-uniform; a pixel function; float64 field
 * U[0, 1)), documented in rusterize_amd/csrc/rz_synth_job.h.  The host and
 * device generators produce identical coordinates.
 * rz_synth_job_spec: extent side, polygon count, pixel function (rz_fun) and
 *   vertex count of job k (any pointer may be NULL).
 * rz_synth_job_geometry: job k's geometry on the host.
 * rz_synth_job_field: job k's n_poly field values.
 * rz_device_synth_job: job k's geometry generated directly on the device
 *   (declared with the device runtime below). */
rz_status rz_synth_job_spec(uint64_t seed, int64_t k, int32_t* side, int32_t* n_poly, int32_t* fun,
                            int64_t* n_vertices);
rz_status rz_synth_job_geometry(uint64_t seed, int64_t k, rz_geometry_buffers** out);
rz_status rz_synth_job_field(uint64_t seed, int64_t k, double* out);

/* --------------------------------------------------------- rasterisation -- */

/*
 * Host-in / host-out rasterisation: replaces
 *   `geoms.rasterize::<DenseArray<N>>(ctx)` (rust/src/rasterize.rs:54-116) and
 *   `rusterize_py_impl` for dense encodings (python/src/rusterize.rs:45-125).
 * field: n_geoms values of `dtype` (FieldSource::Array) or NULL, in which case
 *   *scalar is used for every geometry (FieldSource::Scalar).
 * layer_of_geom: NULL for one band, else the band index of every geometry
 *   (the BTreeMap grouping of rust/src/rasterize.rs:198-205 done by the caller:
 *   bands in byte-lexicographic key order); n_layers bands are produced.
 * out: n_layers * nrows * ncols elements of `dtype`, C order (band, row, col).
 */
rz_status rz_rasterize(const rz_geometry* geoms, const rz_raster_info* info, rz_dtype dtype,
                       rz_fun fun, const void* field, const void* scalar,
                       const uint32_t* layer_of_geom, uint32_t n_layers, const void* background,
                       int all_touched, rz_sum_mode sum_mode, void* out);

/* ------------------------------------------------- device-resident runtime -- */

typedef struct rz_context rz_context;
typedef struct rz_device_geometry rz_device_geometry;

/* A context owns one HIP stream (default priority) and a grow-only workspace
 * on `device`. */
rz_status rz_context_create(int device, rz_context** out);
rz_status rz_context_destroy(rz_context* ctx);
rz_status rz_context_synchronize(rz_context* ctx);
/* Workspace bytes currently held by the context. */
size_t rz_context_workspace_bytes(const rz_context* ctx);

rz_status rz_device_malloc(rz_context* ctx, size_t bytes, void** out);
rz_status rz_device_free(rz_context* ctx, void* ptr);
rz_status rz_memcpy_h2d(rz_context* ctx, void* dst, const void* src, size_t bytes);
rz_status rz_memcpy_d2h(rz_context* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy on the context's stream (asynchronous); also the
 * plain copy kernel used to calibrate HBM traffic counters. */
rz_status rz_memcpy_d2d(rz_context* ctx, void* dst, const void* src, size_t bytes);

/* Uploads a geometry set once; it stays resident until freed and can be
 * rasterised any number of times with different fields and grids. */
rz_status rz_device_geometry_upload(rz_context* ctx, const rz_geometry* host,
                                    rz_device_geometry** out);
rz_status rz_device_geometry_free(rz_device_geometry* dg);
int64_t rz_device_geometry_count(const rz_device_geometry* dg);
/* Replaces the coordinates (n_vertices x,y pairs, the set's own count) of a
 * resident set in place, keeping its topology; returns once the copy is
 * done.  Captured batches over the set read the new coordinates on their
 * next replay.  (No counterpart in the reference, whose geometries are
 * parsed per call.) */
rz_status rz_device_geometry_set_xy(rz_device_geometry* dg, const double* xy, int64_t n_vertices);
/* Job k of the seeded job stream (see rz_synth_job_spec), generated on the
 * device into a resident geometry set. */
rz_status rz_device_synth_job(rz_context* ctx, uint64_t seed, int64_t k, rz_device_geometry** out);

/* Same semantics as rz_rasterize with every array on the device:
 * field_dev (or NULL + scalar_host), layer_dev (or NULL), out_dev.
 * Asynchronous on the context's stream. */
rz_status rz_rasterize_device(rz_context* ctx, const rz_device_geometry* dg,
                              const rz_raster_info* info, rz_dtype dtype, rz_fun fun,
                              const void* field_dev, const void* scalar_host,
                              const uint32_t* layer_dev, uint32_t n_layers,
                              const void* background_host, int all_touched,
                              rz_sum_mode sum_mode, void* out_dev);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as rz_rasterize_device, with no
 * allocation of any kind: the workspace is a fixed 64 MiB array in the
 * library's device code (one per device, held until the process ends),
 * whose address rz_context_create resolves once per device.  Every item box
 * is tested by every tile strip (no binning), so it suits small and medium
 * geometry sets; a set whose workspace exceeds 64 MiB (about 49 B per vertex
 * plus 88 B per item) fails with RZ_CAPACITY_ERROR, and all_touched with
 * RZ_VALUE_ERROR.  The call is synchronous: it returns once out_dev is
 * written.  Calls on one device share the array and are serialised by the
 * library; ctx must come from rz_context_create. */
rz_status rz_rasterize_device_noalloc(rz_context* ctx, const rz_device_geometry* dg,
                                      const rz_raster_info* info, rz_dtype dtype, rz_fun fun,
                                      const void* field_dev, const void* scalar_host,
                                      const uint32_t* layer_dev, uint32_t n_layers,
                                      const void* background_host, int all_touched,
                                      rz_sum_mode sum_mode, void* out_dev);

/* ---------------------------------------------------------- sparse output -- */

/* encoding="sparse": replaces SparseArrayWriter (rust/src/encoding/writers.rs:
 * 85-142) and SparseArray (rust/src/encoding/arrays.rs:61-209).  Every write
 * of the call becomes a (row, col, value) triplet, in the reference's order:
 * bands in order; in a band its geometries in order; a polygon's rows top to
 * bottom and each row's columns left to right; a line's kept segments in
 * order and each segment's cells in Bresenham order, the end cell last;
 * points in order.  value is the geometry's field value (or the scalar).
 * With all_touched, a line's segments are walked by the all-touched walker,
 * and a polygon writes its outline that way before its rows; fun decides,
 * as in the reference's context (prelude.rs:116-118), whether sum / count
 * write each outline pixel once and skip fill pixels the outline's
 * PixelCache contains (LineWriter / FillWriter, writers.rs:14-60).
 * The triplets stay on the device until rz_sparse_free. */
typedef struct rz_sparse rz_sparse;
rz_status rz_rasterize_sparse(const rz_geometry* geoms, const rz_raster_info* info, rz_dtype dtype,
                              rz_fun fun, const void* field, const void* scalar, const uint32_t* layer_of_geom,
                              uint32_t n_layers, int all_touched, rz_sparse** out);
/* Triplets of each band (n_layers values; SparseArray offsets). */
rz_status rz_sparse_counts(const rz_sparse* sparse, uint64_t* per_band);
/* Every triplet, band after band, into host arrays of the total count
 * (rows and cols as u64, values of the call's dtype); any may be NULL. */
rz_status rz_sparse_copy(const rz_sparse* sparse, uint64_t* rows, uint64_t* cols, void* values);
/* SparseArray::build_array (arrays.rs:102-141): each band's triplets folded
 * in order with pixel function fun, from the background, into out
 * (host, [n_layers][nrows][ncols] of the call's dtype). */
rz_status rz_sparse_to_dense(const rz_sparse* sparse, rz_fun fun, const void* background, void* out);
rz_status rz_sparse_free(rz_sparse* sparse);

/* ------------------------------------------------------------- batching -- */

/* No reference counterpart (the reference rasterises one call at a time):
 * one job of a batch, i.e. the arguments of one rz_rasterize_device call.
 * Values are the raw little-endian bits of the job's dtype. */
typedef struct rz_job {
  const rz_device_geometry* geometry;
  rz_raster_info info;
  rz_dtype dtype;
  rz_fun fun;
  const void* field_dev;     /* [n_geoms] values of dtype, or NULL: scalar_bits */
  uint64_t scalar_bits;
  const uint32_t* layer_dev; /* [n_geoms] band of every geometry, or NULL: one band */
  uint32_t n_layers;
  uint64_t background_bits;
  void* out_dev;             /* [n_layers][nrows][ncols] of dtype, disjoint per job */
  rz_sum_mode sum_mode;      /* fun = sum with a float dtype: ordered (exact) or unordered */
  int all_touched;           /* AllTouched strategy (burners.rs:94-247) */
} rz_job;

/* Rasterises n_jobs independent jobs (any mix of sizes, dtypes and pixel
 * functions) as one pipeline: one edge/descriptor/binning pass over all of
 * them and one persistent work-queue tile launch per (dtype, function)
 * group.  Every output equals the job's own rz_rasterize_device output.
 * Asynchronous on the context's stream. */
rz_status rz_batch_rasterize(rz_context* ctx, const rz_job* jobs, int64_t n_jobs);

/* No reference counterpart: a batch captured in a HIP graph on a context of
 * its own (own stream and workspace).  One eager pass sizes the bins (the
 * captured capacities are its totals x headroom); the captured pipeline then
 * never returns to the host.  Each launch rasterises the same jobs again:
 * field and band buffers may be refilled between launches (their values are
 * read at launch time); geometry, grids, pixel functions and output pointers
 * are fixed.  A launch whose bins outgrow the capacities fails in
 * rz_batch_graph_synchronize (RZ_CAPACITY_ERROR): capture again. */
typedef struct rz_batch_graph rz_batch_graph;
rz_status rz_batch_graph_create(int device, const rz_job* jobs, int64_t n_jobs, double headroom,
                                rz_batch_graph** out);
rz_status rz_batch_graph_launch(rz_batch_graph* graph);
rz_status rz_batch_graph_synchronize(rz_batch_graph* graph);
rz_status rz_batch_graph_destroy(rz_batch_graph* graph);

/* ------------------------------------------------------------- autotuner -- */

/* No reference counterpart: per size class (tile width, tiles of a launch,
 * items per tile) the tiles a work-queue grab takes and the item count from
 * which a tile's polygon coverage is precomputed in parallel.  rz_autotune
 * times every candidate on each job (reps timed runs each) and stores the
 * fastest for that job's class in a process-wide table that every later call
 * and batch uses; run it while no other rasterisation runs in the process.
 * rz_autotune_table lists the table (returns the entry count, fills up to
 * cap); rz_autotune_set adds or (0, 0) removes an entry. */
rz_status rz_autotune(rz_context* ctx, const rz_job* jobs, int64_t n_jobs, int reps,
                      int64_t* n_classes);
int64_t rz_autotune_table(uint32_t* size_class, uint32_t* chunk, uint32_t* heavy_min, int64_t cap);
rz_status rz_autotune_set(uint32_t size_class, uint32_t chunk, uint32_t heavy_min);

/* --------------------------------------------------------------- serving -- */

/* No reference counterpart: a serving loop for streams of rasterise
 * requests (continuous batching).  A worker thread owns n_streams contexts;
 * whenever one is idle it admits queued requests in arrival order while their
 * outputs and estimated workspace fit memory_budget bytes (at most max_batch
 * per batch) and runs them as one rz_batch_rasterize.  Batches on different
 * contexts overlap on the device.  Outputs are allocated by the server and
 * stay on the device until released.  Every output equals the request's own
 * rz_rasterize_device output. */
typedef struct rz_server rz_server;
typedef struct rz_server_stats {
  uint64_t queued, in_flight_batches, batches, completed;
  uint64_t memory_used, memory_peak, memory_budget;
  /* host time of a batch's phases, microseconds, summed over batches and
   * the largest single one: output allocation and eviction, run_batch
   * (launches and its host round trips), the wait for the device */
  uint64_t alloc_us_sum, run_us_sum, wait_us_sum;
  uint64_t alloc_us_max, run_us_max, wait_us_max;
} rz_server_stats;

rz_status rz_server_create(int device, uint64_t memory_budget, int n_streams, int max_batch,
                           rz_server** out);
/* Waits for every submitted request, then frees everything. */
rz_status rz_server_destroy(rz_server* server);
/* Queues a request (job->out_dev is ignored); the geometry, field and band
 * map must stay valid until the request completes. */
rz_status rz_server_submit(rz_server* server, const rz_job* job, uint64_t* ticket);
/* Waits for a request (timeout_ms < 0: no limit) and gives its device output
 * and its latency from submission to completion. */
rz_status rz_server_wait(rz_server* server, uint64_t ticket, int timeout_ms, void** out_dev,
                         double* latency_ms);
/* Returns a completed request's output to the budget. */
rz_status rz_server_release(rz_server* server, uint64_t ticket);
rz_status rz_server_get_stats(rz_server* server, rz_server_stats* out);

/* Profiling level: 0 off; 1 records HIP events around every stage of every
 * rz_rasterize_device call on the context's stream (no added
 * synchronisation); 2 also counts exact pixel-burns; 3 records only the two
 * events around the tile launches (the other stages read 0).
 * rz_context_stage_times waits for the recorded calls and writes up to n
 * per-stage sums (milliseconds since the last reset) in the order given by
 * rz_stage_name(i); returns the number of stages. */
rz_status rz_context_set_profiling(rz_context* ctx, int level);
int rz_context_stage_times(rz_context* ctx, double* ms, int n);
void rz_context_reset_stage_times(rz_context* ctx);
const char* rz_stage_name(int i);

/* Counters of the last call: pixel-burns (writes reaching the dense writer,
 * rust/src/encoding/writers.rs:68-71), tile/item pairs and tiles.  Burns are
 * counted at profiling level 2 only. */
rz_status rz_context_last_counters(rz_context* ctx, uint64_t* burns, uint64_t* pairs,
                                   uint64_t* tiles);

/* ------------------------------------------------------------ test hooks -- */

/* No reference counterpart: host evaluations of the closed forms the tile
 * kernel uses, so CPU tests can check them against the literal reference
 * loops by brute force.
 * rz_debug_walk_cells: the cells Standard::burn_line
 * (rust/src/rasterization/burners.rs:53-84) visits between integer cells
 * (ix0,iy0) -> (ix1,iy1), end cell excluded, restricted to [X0,X1) x [Y0,Y1);
 * returns the count and stores up to cap cells (-2: the kernel's
 * index-stepped form of the walk gave other cells).
 * rz_debug_row_coverage: which of the columns [C0, C0+width) the spans of
 * burn_polygon (burners.rs:302-316) cover for the crossings xs[0..n) of one
 * row; width is 64, 128 or 256.
 * rz_debug_walk_all_touched: the cells AllTouchedBase::burn_line
 * (burners.rs:115-246) writes for one edge in pixel coordinates on an
 * ncols x nrows raster, in order, restricted to [X0,X1) x [Y0,Y1); returns
 * the count (or -1 past the walk's step limit) and stores up to cap cells. */
int64_t rz_debug_walk_cells(int64_t ix0, int64_t iy0, int64_t ix1, int64_t iy1, int64_t X0,
                            int64_t X1, int64_t Y0, int64_t Y1, int64_t* xs, int64_t* ys,
                            int64_t cap);
int rz_debug_row_coverage(const double* xs, int64_t n, int64_t ncols, int32_t C0, int32_t width,
                          uint8_t* out);
int64_t rz_debug_walk_all_touched(double x0, double y0, double x1, double y1, int64_t ncols, int64_t nrows,
                                  int64_t X0, int64_t X1, int64_t Y0, int64_t Y1, int64_t* xs, int64_t* ys,
                                  int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* RUSTERIZE_HIP_H */
