/*
 * rz_oracle.c -- CPU restatement of rusterize's rasterisation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only
 * as the checker / CPU baseline.  The product (rusterize_amd) never links,
 * loads or calls it.
 *
 * It restates, function by function, the reference Rust sources (paths are
 * relative to the rusterize tree; the Rust toolchain is absent here so the
 * reference itself cannot be built):
 *   rust/src/geo/raster.rs            -> oracle_raster_info_build
 *   rust/src/geo/edges.rs             -> poly_edge_new, intersect_at, extract_*
 *   rust/src/rasterization/burners.rs -> burn_line_standard, burn_line_all_touched,
 *                                        burn_point, burn_polygon
 *   rust/src/rasterization/burn_geometry.rs -> burn_item / burn_geometry
 *   rust/src/rasterization/pixel_cache.rs   -> pixel_cache_*
 *   rust/src/rasterization/pixel_functions.rs -> the PX_* folds
 *   rust/src/encoding/writers.rs      -> dense / line / fill / trace writers
 *   rust/src/rasterize.rs             -> oracle_rasterize (process, per-band threads)
 * with Rust's language semantics made explicit: saturating float->int `as`,
 * f64::total_cmp, wrapping integer arithmetic (release profile), and no FP
 * contraction (build with -ffp-contract=off; see oracle/Makefile).
 *
 * What pins each part (tests/test_oracle.py, tests/test_closed_forms.py):
 *   raster info      golden rasters' grids, R test-boundary.R known answers
 *   polygon path     golden standard_output_sum (polygon with a hole), exact
 *                    rational point-in-polygon on random polygons, rectangle
 *                    closed form, R known answers
 *   standard lines   both golden rasters, brute-force Bresenham vs the
 *                    literal loop (exhaustive small cases, random, 128-bit)
 *   points           golden rasters, R and Rust known answers
 *   pixel cache      golden standard_output_sum_custom_shape (xres != yres)
 *   pixel functions  golden rasters (sum), R known answers, documented
 *                    triplets of python/docs/python.md:106-136
 *   all-touched      (tests/test_all_touched_oracle.py; the reference tests
 *                    it only against GDAL, which is absent) the R known
 *                    answer, closed forms of the axis-aligned branches and
 *                    of slope-1 diagonals, and exact rational geometry on
 *                    generic segments and triangles: a segment burns its
 *                    supercover, a polygon every cell whose interior it
 *                    meets; count with dedup is 1 on the touched set; a
 *                    planted walker slip fails the supercover check.
 *                    Cases within the walker's tolerances (segments through
 *                    grid corners, near-axis segments) are pinned only by
 *                    the closed forms and the R answer.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/rusterize_hip.h"

#define ORACLE_OK 0
#define ORACLE_VALUE_ERROR 1
#define ORACLE_RUNTIME_ERROR 2

static __thread const char* g_err = "";
const char* oracle_last_error(void) { return g_err; }

/* ------------------------------------------------ Rust scalar semantics -- */

/* `v as usize` for f64: NaN -> 0, saturating at 0 and u64::MAX. */
static inline uint64_t sat_usize(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551616.0) return UINT64_MAX;
  return (uint64_t)v;
}
/* `v as isize` for f64: NaN -> 0, saturating. */
static inline int64_t sat_isize(double v) {
  if (v != v) return 0;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  return (int64_t)v;
}
/* wrapping i64 arithmetic (Rust release builds) */
static inline int64_t wadd(int64_t a, int64_t b) { return (int64_t)((uint64_t)a + (uint64_t)b); }
static inline int64_t wsub(int64_t a, int64_t b) { return (int64_t)((uint64_t)a - (uint64_t)b); }
static inline int64_t wabs(int64_t a) { return a < 0 ? (int64_t)(0 - (uint64_t)a) : a; }
static inline int64_t wneg(int64_t a) { return (int64_t)(0 - (uint64_t)a); }
/* f64::total_cmp key */
static inline int64_t total_key(double x) {
  int64_t b;
  memcpy(&b, &x, 8);
  b ^= (int64_t)(((uint64_t)(b >> 63)) >> 1);
  return b;
}
/* f64::min / f64::max: a NaN operand yields the other operand */
static inline double rmin(double a, double b) { return fmin(a, b); }
static inline double rmax(double a, double b) { return fmax(a, b); }
/* f64::clamp (NaN stays NaN) */
static inline double rclamp(double v, double lo, double hi) {
  if (v < lo) v = lo;
  if (v > hi) v = hi;
  return v;
}

/* -------------------------------------------------------- growable vec -- */

typedef struct { void* p; size_t n, cap, esz; } vec_t;
static void vec_init(vec_t* v, size_t esz) { v->p = NULL; v->n = v->cap = 0; v->esz = esz; }
static void* vec_push(vec_t* v) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 64;
    v->p = realloc(v->p, v->cap * v->esz);
  }
  return (char*)v->p + (v->n++) * v->esz;
}
static void vec_free(vec_t* v) { free(v->p); v->p = NULL; v->n = v->cap = 0; }

/* --------------------------------------------------------------- edges -- */

/* rust/src/geo/edges.rs:5-14 */
typedef struct { uint64_t x, y; } point_edge;
/* rust/src/geo/edges.rs:17-24 */
typedef struct { uint64_t ystart, yend; double x0, y0, dxdy, x_at_yline; } poly_edge;
/* rust/src/geo/edges.rs:59-65 */
typedef struct { double x0, y0, x1, y1; int is_closed; } line_edge;

/* PolyEdge::new, rust/src/geo/edges.rs:27-46 */
static poly_edge poly_edge_new(double x0, double y0, double x1, double y1) {
  double xt, yt, xb, yb;
  if (y0 < y1) { xt = x0; yt = y0; xb = x1; yb = y1; }
  else { xt = x1; yt = y1; xb = x0; yb = y0; }
  poly_edge e;
  e.ystart = sat_usize(ceil(yt - 0.5));
  e.yend = sat_usize(ceil(yb - 0.5));
  e.dxdy = (xb - xt) / (yb - yt);
  e.x0 = xt;
  e.y0 = yt;
  e.x_at_yline = INFINITY;
  return e;
}

/* PolyEdge::intersect_at, rust/src/geo/edges.rs:50-55 */
static inline double intersect_at(const poly_edge* e, uint64_t yline) {
  double center_y = (double)yline + 0.5;
  return e->x0 + (center_y - e->y0) * e->dxdy;
}

/* extract_point, rust/src/geo/edges.rs:79-88 */
static void extract_point(vec_t* edges, double px, double py, const rz_raster_info* ri) {
  double x = (px - ri->xmin) / ri->xres;
  double y = (ri->ymax - py) / ri->yres;
  if (x >= 0.0 && x < (double)ri->ncols && y >= 0.0 && y < (double)ri->nrows) {
    point_edge* pe = (point_edge*)vec_push(edges);
    pe->x = sat_usize(x);
    pe->y = sat_usize(y);
  }
}

/* extract_ring, rust/src/geo/edges.rs:90-110 */
static void extract_ring(vec_t* edges, const double* xy, int64_t n, const rz_raster_info* ri) {
  double rows = (double)ri->nrows;
  for (int64_t i = 0; i + 1 < n; ++i) {
    double x0 = (xy[2 * i] - ri->xmin) / ri->xres;
    double y0 = (ri->ymax - xy[2 * i + 1]) / ri->yres;
    double x1 = (xy[2 * i + 2] - ri->xmin) / ri->xres;
    double y1 = (ri->ymax - xy[2 * i + 3]) / ri->yres;
    if (fabs(y0 - y1) >= 2.220446049250313e-16) {
      double min_y = rmin(y0, y1);
      double max_y = rmax(y0, y1);
      if (min_y < rows && max_y >= 0.0) {
        *(poly_edge*)vec_push(edges) = poly_edge_new(x0, y0, x1, y1);
      }
    }
  }
}

/* geo-types LineString::is_closed: first coordinate == last coordinate
 * (an empty line string counts as closed). */
static int line_is_closed(const double* xy, int64_t n) {
  if (n == 0) return 1;
  return xy[0] == xy[2 * (n - 1)] && xy[1] == xy[2 * (n - 1) + 1];
}

/* extract_line, rust/src/geo/edges.rs:112-134 */
static void extract_line(vec_t* edges, const double* xy, int64_t n, const rz_raster_info* ri) {
  double rows = (double)ri->nrows, cols = (double)ri->ncols;
  int is_closed = line_is_closed(xy, n);
  for (int64_t i = 0; i + 1 < n; ++i) {
    double x0 = (xy[2 * i] - ri->xmin) / ri->xres;
    double y0 = (ri->ymax - xy[2 * i + 1]) / ri->yres;
    double x1 = (xy[2 * i + 2] - ri->xmin) / ri->xres;
    double y1 = (ri->ymax - xy[2 * i + 3]) / ri->yres;
    double min_x = rmin(x0, x1), max_x = rmax(x0, x1);
    double min_y = rmin(y0, y1), max_y = rmax(y0, y1);
    if (min_x < cols && max_x >= 0.0 && min_y < rows && max_y >= 0.0) {
      line_edge* le = (line_edge*)vec_push(edges);
      le->x0 = x0; le->y0 = y0; le->x1 = x1; le->y1 = y1; le->is_closed = is_closed;
    }
  }
}

/* --------------------------------------------------------- pixel cache -- */

/* PixelCache, rust/src/rasterization/pixel_cache.rs:7-59 */
typedef struct { uint64_t* bits; uint64_t nbits; uint64_t width; int64_t xmin, ymin; } pixel_cache;

static void pixel_cache_new(pixel_cache* c, const line_edge* e, size_t n) {
  double xmin = 1.7976931348623157e308, ymin = 1.7976931348623157e308;
  double xmax = -1.7976931348623157e308, ymax = -1.7976931348623157e308;
  for (size_t i = 0; i < n; ++i) {
    xmin = rmin(rmin(xmin, e[i].x0), e[i].x1);
    ymin = rmin(rmin(ymin, e[i].y0), e[i].y1);
    xmax = rmax(rmax(xmax, e[i].x0), e[i].x1);
    ymax = rmax(rmax(ymax, e[i].y0), e[i].y1);
  }
  uint64_t width = sat_usize(floor(xmax) - floor(xmin)) + 1;
  uint64_t length = sat_usize(floor(ymax) - floor(ymin)) + 1;
  c->width = width;
  c->nbits = width * length;
  c->bits = (uint64_t*)calloc(c->nbits / 64 + 1, 8);
  c->xmin = sat_isize(xmin);
  c->ymin = sat_isize(ymin);
}
static inline uint64_t pixel_cache_index(const pixel_cache* c, uint64_t x, uint64_t y) {
  uint64_t lx = (uint64_t)wsub((int64_t)x, c->xmin);
  uint64_t ly = (uint64_t)wsub((int64_t)y, c->ymin);
  return ly * c->width + lx;
}
static int pixel_cache_insert(pixel_cache* c, uint64_t x, uint64_t y) {
  uint64_t i = pixel_cache_index(c, x, y);
  if (i >= c->nbits) abort(); /* FixedBitSet::insert panics out of range */
  if (c->bits[i >> 6] & (1ull << (i & 63))) return 0;
  c->bits[i >> 6] |= 1ull << (i & 63);
  return 1;
}
static int pixel_cache_contains(const pixel_cache* c, uint64_t x, uint64_t y) {
  uint64_t i = pixel_cache_index(c, x, y);
  if (i >= c->nbits) return 0; /* FixedBitSet::contains is false out of range */
  return (c->bits[i >> 6] >> (i & 63)) & 1;
}
static void pixel_cache_free(pixel_cache* c) { free(c->bits); c->bits = NULL; }

/* ------------------------------------------------------------- writers -- */

/* PixelWriter trait, rust/src/encoding/writers.rs:10-12 */
typedef struct writer writer;
struct writer {
  void (*write)(writer* self, uint64_t y, uint64_t x);
};

typedef void (*pxfn_t)(void* band, uint64_t idx, const void* value, const void* bg);

/* DenseArrayWriter (writers.rs:62-78) restricted to the row window
 * [row0, row1) of the band; writes outside the window are dropped (the
 * window is the same computation observed on fewer rows). */
typedef struct {
  writer base;
  char* band;
  uint64_t ncols, row0, row1;
  size_t esz;
  pxfn_t fn;
  const void* value;
  const void* bg;
  uint64_t burns;
} dense_writer;
static void dense_write(writer* w, uint64_t y, uint64_t x) {
  dense_writer* d = (dense_writer*)w;
  if (y < d->row0 || y >= d->row1) return;
  d->burns++;
  d->fn(d->band, (y - d->row0) * d->ncols + x, d->value, d->bg);
}

/* Collector: records every write (the test writer of
 * rust/src/rasterization/burn_geometry.rs:251-260, and the write stream of
 * SparseArrayWriter, writers.rs:85-99). */
typedef struct {
  writer base;
  vec_t cells; /* (y, x, geometry index) as 3 x u64 */
  uint64_t geom;
} trace_writer;
static void trace_write(writer* w, uint64_t y, uint64_t x) {
  trace_writer* t = (trace_writer*)w;
  uint64_t* c = (uint64_t*)vec_push(&t->cells);
  c[0] = y; c[1] = x; c[2] = t->geom;
}

/* LineWriter, writers.rs:14-36 */
typedef struct { writer base; writer* inner; pixel_cache* cache; } line_writer;
static void line_write(writer* w, uint64_t y, uint64_t x) {
  line_writer* l = (line_writer*)w;
  if (pixel_cache_insert(l->cache, x, y)) l->inner->write(l->inner, y, x);
}
/* FillWriter, writers.rs:38-60 */
typedef struct { writer base; writer* inner; pixel_cache* cache; } fill_writer;
static void fill_write(writer* w, uint64_t y, uint64_t x) {
  fill_writer* f = (fill_writer*)w;
  if (!pixel_cache_contains(f->cache, x, y)) f->inner->write(f->inner, y, x);
}

/* ------------------------------------------------------------- burners -- */

/* Standard::burn_line, rust/src/rasterization/burners.rs:35-92 */
static void burn_line_standard(const line_edge* edges, size_t n, const rz_raster_info* ri,
                               writer* w) {
  if (n == 0) return;
  int64_t nrows = (int64_t)ri->nrows, ncols = (int64_t)ri->ncols;
  size_t last_idx = n - 1;
  for (size_t idx = 0; idx < n; ++idx) {
    const line_edge* e = &edges[idx];
    int64_t ix0 = sat_isize(floor(e->x0));
    int64_t ix1 = sat_isize(floor(e->x1));
    int64_t iy0 = sat_isize(floor(e->y0));
    int64_t iy1 = sat_isize(floor(e->y1));
    int64_t dx = wabs(wsub(ix1, ix0));
    int64_t dy = wneg(wabs(wsub(iy1, iy0)));
    int64_t sx = ix0 < ix1 ? 1 : -1;
    int64_t sy = iy0 < iy1 ? 1 : -1;
    int64_t err = wadd(dx, dy);
    while (ix0 != ix1 || iy0 != iy1) {
      if (ix0 >= 0 && ix0 < ncols && iy0 >= 0 && iy0 < nrows) w->write(w, (uint64_t)iy0, (uint64_t)ix0);
      int64_t e2 = wadd(err, err);
      if (e2 >= dy) { err = wadd(err, dy); ix0 = wadd(ix0, sx); }
      if (e2 <= dx) { err = wadd(err, dx); iy0 = wadd(iy0, sy); }
    }
    if (idx == last_idx && !e->is_closed && ix0 >= 0 && ix0 < ncols && iy0 >= 0 && iy0 < nrows)
      w->write(w, (uint64_t)iy0, (uint64_t)ix0);
  }
}

/* AllTouchedBase::burn_line, rust/src/rasterization/burners.rs:94-247
 * (GDAL llrasterize all-touched walker). */
#define EPSILON_INTERSECT 1e-4
#define TOLERANCE 1e-9
static void burn_line_all_touched(const line_edge* edges, size_t n, const rz_raster_info* ri,
                                  writer* w) {
  if (n == 0) return;
  int64_t nrows = (int64_t)ri->nrows, ncols = (int64_t)ri->ncols;
  double nrows_f = (double)ri->nrows, ncols_f = (double)ri->ncols;
  for (size_t k = 0; k < n; ++k) {
    double dfx = edges[k].x0, dfy = edges[k].y0, dfxe = edges[k].x1, dfye = edges[k].y1, t;
    if (dfx > dfxe) {
      t = dfx; dfx = dfxe; dfxe = t;
      t = dfy; dfy = dfye; dfye = t;
    }
    if (fabs(dfx - dfxe) < 0.01) { /* vertical */
      if (dfye < dfy) { t = dfy; dfy = dfye; dfye = t; }
      int64_t ix = sat_isize(floor(dfxe));
      int64_t iy = sat_isize(floor(dfy));
      int64_t iy_end = sat_isize(floor(dfye - EPSILON_INTERSECT));
      if (ix < 0 || ix >= ncols) continue;
      if (iy < 0) iy = 0;
      if (iy_end > nrows - 1) iy_end = nrows - 1;
      for (int64_t y = iy; y <= iy_end; ++y) w->write(w, (uint64_t)y, (uint64_t)ix);
      continue;
    }
    if (fabs(dfy - dfye) < 0.01) { /* horizontal */
      if (dfxe < dfx) { t = dfx; dfx = dfxe; dfxe = t; }
      int64_t ix = sat_isize(floor(dfx));
      int64_t iy = sat_isize(floor(dfy));
      int64_t ix_end = sat_isize(floor(dfxe - EPSILON_INTERSECT));
      if (iy < 0 || iy >= nrows) continue;
      if (ix < 0) ix = 0;
      if (ix_end > ncols - 1) ix_end = ncols - 1;
      for (int64_t x = ix; x <= ix_end; ++x) w->write(w, (uint64_t)iy, (uint64_t)x);
      continue;
    }
    double slope = (dfye - dfy) / (dfxe - dfx);
    double inv_slope = 1.0 / slope;
    if (dfx < 0.0) { dfy += (0.0 - dfx) * slope; dfx = 0.0; }
    if (dfxe > ncols_f) { dfye += (ncols_f - dfxe) * slope; dfxe = ncols_f; }
    if (dfy < 0.0) { dfx += (0.0 - dfy) * inv_slope; dfy = 0.0; }
    else if (dfy > nrows_f) { dfx += (nrows_f - dfy) * inv_slope; dfy = nrows_f; }
    if (dfye < 0.0) { dfxe += (0.0 - dfye) * inv_slope; }
    else if (dfye > nrows_f) { dfxe += (nrows_f - dfye) * inv_slope; }
    dfx = rclamp(dfx, 0.0, ncols_f);
    dfxe = rclamp(dfxe, 0.0, ncols_f);
    while (dfx >= 0.0 && dfx < dfxe) {
      int64_t ix = sat_isize(floor(dfx));
      int64_t iy = sat_isize(floor(dfy));
      if (ix >= 0 && ix < ncols && iy >= 0 && iy < nrows) w->write(w, (uint64_t)iy, (uint64_t)ix);
      double sx = floor(dfx + 1.0) - dfx;
      double sy = sx * slope;
      if (sat_isize(floor(dfy + sy)) == iy) {
        dfx += sx; dfy += sy;
      } else if (slope < 0.0) {
        sy = (double)iy - dfy;
        if (sy > -TOLERANCE) sy = -TOLERANCE;
        sx = sy / slope;
        dfx += sx; dfy += sy;
      } else {
        sy = (double)wadd(iy, 1) - dfy;
        if (sy < TOLERANCE) sy = TOLERANCE;
        sx = sy / slope;
        dfx += sx; dfy += sy;
      }
    }
  }
}

/* burn_point, rust/src/rasterization/burners.rs:249-258 */
static void burn_point(const point_edge* pts, size_t n, writer* w) {
  for (size_t i = 0; i < n; ++i) w->write(w, pts[i].y, pts[i].x);
}

static int cmp_ystart(const void* a, const void* b) {
  uint64_t ya = ((const poly_edge*)a)->ystart, yb = ((const poly_edge*)b)->ystart;
  return ya < yb ? -1 : ya > yb;
}
static int cmp_xline(const void* a, const void* b) {
  int64_t ka = total_key(((const poly_edge*)a)->x_at_yline);
  int64_t kb = total_key(((const poly_edge*)b)->x_at_yline);
  return ka < kb ? -1 : ka > kb;
}

/* burn_polygon, rust/src/rasterization/burners.rs:260-320 */
static void burn_polygon(vec_t* polyedges, const rz_raster_info* ri, writer* w) {
  size_t n = polyedges->n;
  if (n == 0) return;
  poly_edge* pe = (poly_edge*)polyedges->p;
  qsort(pe, n, sizeof(poly_edge), cmp_ystart);
  uint64_t yline = pe[0].ystart;
  poly_edge* active = (poly_edge*)malloc(n * sizeof(poly_edge));
  size_t nact = 0, next = 0;
  double ncols = (double)ri->ncols;
  while (yline < ri->nrows && (nact > 0 || next < n)) {
    /* transfer current edges to active edges (partition_point + drain) */
    while (next < n && pe[next].ystart <= yline) active[nact++] = pe[next++];
    /* remove finished edges (retain keeps order) */
    size_t k = 0;
    for (size_t i = 0; i < nact; ++i)
      if (active[i].yend > yline) active[k++] = active[i];
    nact = k;
    if (nact == 0) { yline += 1; continue; }
    for (size_t i = 0; i < nact; ++i) active[i].x_at_yline = intersect_at(&active[i], yline);
    qsort(active, nact, sizeof(poly_edge), cmp_xline);
    for (size_t i = 0; i + 1 < nact; i += 2) { /* chunks_exact(2) */
      double x1 = active[i].x_at_yline, x2 = active[i + 1].x_at_yline;
      uint64_t xstart = sat_usize(rclamp(floor(x1 + 0.5), 0.0, ncols));
      uint64_t xend = sat_usize(rclamp(floor(x2 + 0.5), 0.0, ncols));
      for (uint64_t xp = xstart; xp < xend; ++xp) w->write(w, yline, xp);
    }
    yline += 1;
  }
  free(active);
}

/* --------------------------------------------------- geometry dispatch -- */

typedef struct {
  const rz_geometry* g;
  const rz_raster_info* ri;
  int all_touched;
  int dedup; /* RasterizeContext::requires_dedup, rust/src/prelude.rs:116-118 */
} burn_ctx;

static void burn_lines_with(const burn_ctx* c, vec_t* lines, writer* w) {
  if (c->all_touched) burn_line_all_touched((line_edge*)lines->p, lines->n, c->ri, w);
  else burn_line_standard((line_edge*)lines->p, lines->n, c->ri, w);
}

/* One item of the flattened model; see the mapping in rusterize_hip.h which
 * follows rust/src/rasterization/burn_geometry.rs:24-243. */
static void burn_item(const burn_ctx* c, int64_t item, writer* w) {
  const rz_geometry* g = c->g;
  const rz_raster_info* ri = c->ri;
  int64_t r0 = g->item_ring_offsets[item], r1 = g->item_ring_offsets[item + 1];
  int kind = g->item_kind[item];
  if (kind == RZ_ITEM_POINTS) { /* burn_geometry.rs:31-44 */
    vec_t pts; vec_init(&pts, sizeof(point_edge));
    for (int64_t r = r0; r < r1; ++r)
      for (int64_t v = g->ring_vertex_offsets[r]; v < g->ring_vertex_offsets[r + 1]; ++v)
        extract_point(&pts, g->xy[2 * v], g->xy[2 * v + 1], ri);
    burn_point((point_edge*)pts.p, pts.n, w);
    vec_free(&pts);
  } else if (kind == RZ_ITEM_POLYGONS) { /* burn_geometry.rs:76-166 */
    vec_t edges; vec_init(&edges, sizeof(poly_edge));
    for (int64_t r = r0; r < r1; ++r) {
      int64_t v0 = g->ring_vertex_offsets[r], v1 = g->ring_vertex_offsets[r + 1];
      extract_ring(&edges, g->xy + 2 * v0, v1 - v0, ri);
    }
    if (c->all_touched) {
      vec_t lines; vec_init(&lines, sizeof(line_edge));
      for (int64_t r = r0; r < r1; ++r) {
        int64_t v0 = g->ring_vertex_offsets[r], v1 = g->ring_vertex_offsets[r + 1];
        extract_line(&lines, g->xy + 2 * v0, v1 - v0, ri);
      }
      if (c->dedup) { /* handle_polygon, burn_geometry.rs:225-234 */
        pixel_cache cache;
        pixel_cache_new(&cache, (line_edge*)lines.p, lines.n);
        line_writer lw = {{line_write}, w, &cache};
        burn_lines_with(c, &lines, &lw.base);
        fill_writer fw = {{fill_write}, w, &cache};
        burn_polygon(&edges, ri, &fw.base);
        pixel_cache_free(&cache);
      } else { /* burn_geometry.rs:235-238 */
        burn_lines_with(c, &lines, w);
        burn_polygon(&edges, ri, w);
      }
      vec_free(&lines);
    } else {
      burn_polygon(&edges, ri, w);
    }
    vec_free(&edges);
  } else { /* LINES: burn_geometry.rs:168-210 */
    vec_t lines; vec_init(&lines, sizeof(line_edge));
    for (int64_t r = r0; r < r1; ++r) {
      int64_t v0 = g->ring_vertex_offsets[r], v1 = g->ring_vertex_offsets[r + 1];
      extract_line(&lines, g->xy + 2 * v0, v1 - v0, ri);
    }
    if (ri->xres != ri->yres || c->dedup) {
      pixel_cache cache;
      pixel_cache_new(&cache, (line_edge*)lines.p, lines.n);
      line_writer lw = {{line_write}, w, &cache};
      burn_lines_with(c, &lines, &lw.base);
      pixel_cache_free(&cache);
    } else {
      burn_lines_with(c, &lines, w);
    }
    vec_free(&lines);
  }
}

static void burn_geometry(const burn_ctx* c, int64_t gi, writer* w) {
  for (int64_t it = c->g->geom_item_offsets[gi]; it < c->g->geom_item_offsets[gi + 1]; ++it)
    burn_item(c, it, w);
}

/* ---------------------------------------------------- pixel functions -- */

/* rust/src/rasterization/pixel_functions.rs:55-123 with NaNAware
 * (rust/src/prelude.rs:13-40): is_nan is false for integers.  Integer `+=`
 * wraps (release profile), done in the unsigned type U. */
#define ISNAN_F(v) ((v) != (v))
#define ISNAN_I(v) (0)
#define DEFINE_PX(T, U, NAME, ISNAN)                                                        \
  static void px_sum_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {          \
    T* p = (T*)a + i; T v = *(const T*)vp, bg = *(const T*)bp;                            \
    if (*p == bg || ISNAN(*p) || ISNAN(v)) *p = v; else *p = (T)((U)*p + (U)v);             \
  }                                                                                         \
  static void px_first_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {        \
    T* p = (T*)a + i; T v = *(const T*)vp, bg = *(const T*)bp;                            \
    if (*p == bg || ISNAN(*p)) *p = v;                                                      \
  }                                                                                         \
  static void px_last_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {         \
    (void)bp; ((T*)a)[i] = *(const T*)vp;                                                   \
  }                                                                                         \
  static void px_min_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {          \
    T* p = (T*)a + i; T v = *(const T*)vp, bg = *(const T*)bp;                            \
    if (*p == bg || ISNAN(*p) || *p > v) *p = v;                                            \
  }                                                                                         \
  static void px_max_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {          \
    T* p = (T*)a + i; T v = *(const T*)vp, bg = *(const T*)bp;                            \
    if (*p == bg || ISNAN(*p) || *p < v) *p = v;                                            \
  }                                                                                         \
  static void px_count_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {        \
    (void)vp; T* p = (T*)a + i; T bg = *(const T*)bp;                                      \
    if (*p == bg || ISNAN(*p)) *p = (T)1; else *p = (T)((U)*p + (U)1);                      \
  }                                                                                         \
  static void px_any_##NAME(void* a, uint64_t i, const void* vp, const void* bp) {          \
    (void)vp; (void)bp; ((T*)a)[i] = (T)1;                                                  \
  }                                                                                         \
  static const pxfn_t px_table_##NAME[7] = {px_sum_##NAME, px_first_##NAME, px_last_##NAME, \
                                            px_min_##NAME, px_max_##NAME, px_count_##NAME, \
                                            px_any_##NAME};

DEFINE_PX(uint8_t, uint8_t, u8, ISNAN_I)
DEFINE_PX(uint16_t, uint16_t, u16, ISNAN_I)
DEFINE_PX(uint32_t, uint32_t, u32, ISNAN_I)
DEFINE_PX(uint64_t, uint64_t, u64, ISNAN_I)
DEFINE_PX(int8_t, uint8_t, i8, ISNAN_I)
DEFINE_PX(int16_t, uint16_t, i16, ISNAN_I)
DEFINE_PX(int32_t, uint32_t, i32, ISNAN_I)
DEFINE_PX(int64_t, uint64_t, i64, ISNAN_I)
DEFINE_PX(float, float, f32, ISNAN_F)
DEFINE_PX(double, double, f64, ISNAN_F)

static const pxfn_t* px_tables[10] = {px_table_u8, px_table_u16, px_table_u32, px_table_u64,
                                      px_table_i8, px_table_i16, px_table_i32, px_table_i64,
                                      px_table_f32, px_table_f64};
static const size_t dtype_sizes[10] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8};

/* ---------------------------------------------------------- raster info -- */

/* RasterInfoBuilder::{build, build_with, finalize}, rust/src/geo/raster.rs:50-156 */
int oracle_raster_info_build(const rz_geometry* g, int has_shape, uint64_t nrows, uint64_t ncols,
                             int has_extent, const double* extent, int has_res, double xres,
                             double yres, int tap, int32_t epsg, rz_raster_info* out) {
  double xmin, ymin, xmax, ymax;
  int inferred;
  if (has_extent) { /* build(), raster.rs:50-64 */
    int all_zero = 1;
    for (int i = 0; i < 4; ++i) all_zero &= total_key(extent[i]) == total_key(0.0);
    if (all_zero) { g_err = "Unspecified extent (all zeros)."; return ORACLE_VALUE_ERROR; }
    xmin = extent[0]; ymin = extent[1]; xmax = extent[2]; ymax = extent[3];
    inferred = 0;
  } else { /* build_with(), raster.rs:67-93: fold of geo BoundingRect; a
              Polygon's bounding rect covers its exterior ring only */
    int have = 0;
    double bx0 = 0, by0 = 0, bx1 = 0, by1 = 0;
    for (int64_t r = 0; g && r < g->n_rings; ++r) {
      if (g->ring_flags && (g->ring_flags[r] & 1)) continue;
      for (int64_t v = g->ring_vertex_offsets[r]; v < g->ring_vertex_offsets[r + 1]; ++v) {
        double x = g->xy[2 * v], y = g->xy[2 * v + 1];
        if (!have) { bx0 = bx1 = x; by0 = by1 = y; have = 1; }
        else { bx0 = rmin(bx0, x); by0 = rmin(by0, y); bx1 = rmax(bx1, x); by1 = rmax(by1, y); }
      }
    }
    if (!have) { g_err = "Cannot infer bounding box from geometry."; return ORACLE_RUNTIME_ERROR; }
    xmin = bx0; ymin = by0; xmax = bx1; ymax = by1;
    inferred = 1;
  }
  /* finalize, raster.rs:95-156 */
  if (!has_shape && !has_res) { g_err = "Must set at least one of `shape` or `resolution`"; return ORACLE_VALUE_ERROR; }
  if (has_shape && has_res) { g_err = "Shape and resolution are mutually exclusive; provide only one"; return ORACLE_VALUE_ERROR; }
  if (!has_shape) { nrows = 0; ncols = 0; }
  if (!has_res) { xres = 0.0; yres = 0.0; }
  if (has_shape && (nrows == 0 || ncols == 0)) { g_err = "Shape values must be > 0."; return ORACLE_VALUE_ERROR; }
  if (has_res && (xres <= 0.0 || yres <= 0.0)) { g_err = "Resolution values must be > 0."; return ORACLE_VALUE_ERROR; }
  if (inferred && !tap && has_res) {
    xmin -= xres / 2.0; xmax += xres / 2.0; ymin -= yres / 2.0; ymax += yres / 2.0;
  }
  if (!has_res) {
    xres = (xmax - xmin) / (double)ncols;
    yres = (ymax - ymin) / (double)nrows;
  } else if (tap) {
    xmin = floor(xmin / xres) * xres;
    xmax = ceil(xmax / xres) * xres;
    ymin = floor(ymin / yres) * yres;
    ymax = ceil(ymax / yres) * yres;
  }
  if (!has_shape) {
    nrows = sat_usize(0.5 + (ymax - ymin) / yres);
    ncols = sat_usize(0.5 + (xmax - xmin) / xres);
  }
  out->ncols = ncols; out->nrows = nrows;
  out->xmin = xmin; out->xmax = xmax; out->ymin = ymin; out->ymax = ymax;
  out->xres = xres; out->yres = yres; out->epsg = epsg;
  return ORACLE_OK;
}

/* ------------------------------------------------------------- process -- */

typedef struct {
  const burn_ctx* c;
  const int64_t* idx; /* geometry indices of this band, ascending */
  int64_t n_idx;
  dense_writer dw;
  const void* field;  /* NULL -> scalar */
  const void* scalar;
  size_t esz;
} band_job;

/* process, rust/src/rasterize.rs:159-196: sequential loop over the band's
 * geometry indices in ascending order. */
static void* run_band(void* arg) {
  band_job* j = (band_job*)arg;
  for (int64_t k = 0; k < j->n_idx; ++k) {
    int64_t gi = j->idx[k];
    j->dw.value = j->field ? (const void*)((const char*)j->field + (size_t)gi * j->esz) : j->scalar;
    burn_geometry(j->c, gi, &j->dw.base);
  }
  return NULL;
}

/*
 * DenseArray::build, rust/src/rasterize.rs:71-116.
 * out: n_layers x (row1-row0) x ncols elements, pre-filled here with the
 * background (RasterInfo::build_raster, raster.rs:23-28).  Bands run on up
 * to n_threads threads, one band per thread like the rayon loop at
 * rasterize.rs:89-101 (single-threaded inside a band).
 */
int oracle_rasterize(const rz_geometry* g, const rz_raster_info* ri, int dtype, int fun,
                     const void* field, const void* scalar, const uint32_t* layer_of_geom,
                     uint32_t n_layers, const void* bg, int all_touched, uint64_t row0,
                     uint64_t row1, int n_threads, void* out, uint64_t* burns_out) {
  if (dtype < 0 || dtype > 9 || fun < 0 || fun > 6) { g_err = "bad dtype/fun"; return ORACLE_VALUE_ERROR; }
  if (!layer_of_geom) n_layers = 1;
  if (row1 > ri->nrows) row1 = ri->nrows;
  if (row0 > row1) row0 = row1;
  size_t esz = dtype_sizes[dtype];
  uint64_t band_elems = (row1 - row0) * ri->ncols;
  /* Array3::from_elem: fill by doubling copies of the first element */
  uint64_t total_bytes = band_elems * n_layers * esz;
  if (total_bytes) {
    memcpy(out, bg, esz);
    for (uint64_t done = esz; done < total_bytes;) {
      uint64_t n = done < total_bytes - done ? done : total_bytes - done;
      memcpy((char*)out + done, out, n);
      done += n;
    }
  }

  burn_ctx c = {g, ri, all_touched, all_touched && (fun == RZ_SUM || fun == RZ_COUNT)};
  /* group_keys, rasterize.rs:198-205 (the caller supplies sorted band ids) */
  int64_t* counts = (int64_t*)calloc(n_layers + 1, sizeof(int64_t));
  int64_t* idx = (int64_t*)malloc((size_t)(g->n_geoms > 0 ? g->n_geoms : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < g->n_geoms; ++i) {
    uint32_t l = layer_of_geom ? layer_of_geom[i] : 0;
    if (l >= n_layers) { free(counts); free(idx); g_err = "layer out of range"; return ORACLE_VALUE_ERROR; }
    counts[l + 1]++;
  }
  for (uint32_t l = 0; l < n_layers; ++l) counts[l + 1] += counts[l];
  int64_t* cursor = (int64_t*)malloc((n_layers + 1) * sizeof(int64_t));
  memcpy(cursor, counts, (n_layers + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < g->n_geoms; ++i) idx[cursor[layer_of_geom ? layer_of_geom[i] : 0]++] = i;

  band_job* jobs = (band_job*)calloc(n_layers, sizeof(band_job));
  for (uint32_t l = 0; l < n_layers; ++l) {
    band_job* j = &jobs[l];
    j->c = &c; j->idx = idx + counts[l]; j->n_idx = counts[l + 1] - counts[l];
    j->dw.base.write = dense_write;
    j->dw.band = (char*)out + (size_t)l * band_elems * esz;
    j->dw.ncols = ri->ncols; j->dw.row0 = row0; j->dw.row1 = row1;
    j->dw.esz = esz; j->dw.fn = px_tables[dtype][fun]; j->dw.bg = bg;
    j->field = field; j->scalar = scalar; j->esz = esz;
  }
  if (n_threads < 1) n_threads = 1;
  for (uint32_t base = 0; base < n_layers; base += (uint32_t)n_threads) {
    uint32_t nb = n_layers - base < (uint32_t)n_threads ? n_layers - base : (uint32_t)n_threads;
    if (nb == 1) { run_band(&jobs[base]); continue; }
    pthread_t* th = (pthread_t*)malloc(nb * sizeof(pthread_t));
    for (uint32_t k = 0; k < nb; ++k) pthread_create(&th[k], NULL, run_band, &jobs[base + k]);
    for (uint32_t k = 0; k < nb; ++k) pthread_join(th[k], NULL);
    free(th);
  }
  uint64_t burns = 0;
  for (uint32_t l = 0; l < n_layers; ++l) burns += jobs[l].dw.burns;
  if (burns_out) *burns_out = burns;
  free(jobs); free(cursor); free(idx); free(counts);
  return ORACLE_OK;
}

/*
 * Write trace of one band holding every geometry (the Collector of
 * burn_geometry.rs:251-260; the triplet stream of SparseArrayWriter,
 * writers.rs:93-99).  Returns the number of writes; fills up to cap
 * (y, x, geometry) triplets into cells (3 * cap u64) when cells != NULL.
 */
int64_t oracle_trace(const rz_geometry* g, const rz_raster_info* ri, int fun, int all_touched,
                     uint64_t* cells, int64_t cap) {
  burn_ctx c = {g, ri, all_touched, all_touched && (fun == RZ_SUM || fun == RZ_COUNT)};
  trace_writer tw;
  tw.base.write = trace_write;
  vec_init(&tw.cells, 3 * sizeof(uint64_t));
  for (int64_t i = 0; i < g->n_geoms; ++i) {
    tw.geom = (uint64_t)i;
    burn_geometry(&c, i, &tw.base);
  }
  int64_t n = (int64_t)tw.cells.n;
  if (cells) memcpy(cells, tw.cells.p, (size_t)(n < cap ? n : cap) * 3 * sizeof(uint64_t));
  vec_free(&tw.cells);
  return n;
}

/* The literal Bresenham loop of Standard::burn_line (burners.rs:53-84) on
 * one segment in integer cells, for brute-force checks of closed forms:
 * writes the visited cells (excluding the end cell) to xs/ys, returns the
 * count (or -1 if more than cap). */
int64_t oracle_bresenham_cells(int64_t ix0, int64_t iy0, int64_t ix1, int64_t iy1, int64_t* xs,
                               int64_t* ys, int64_t cap) {
  int64_t dx = wabs(wsub(ix1, ix0));
  int64_t dy = wneg(wabs(wsub(iy1, iy0)));
  int64_t sx = ix0 < ix1 ? 1 : -1;
  int64_t sy = iy0 < iy1 ? 1 : -1;
  int64_t err = wadd(dx, dy), n = 0;
  while (ix0 != ix1 || iy0 != iy1) {
    if (n >= cap) return -1;
    xs[n] = ix0; ys[n] = iy0; ++n;
    int64_t e2 = wadd(err, err);
    if (e2 >= dy) { err = wadd(err, dy); ix0 = wadd(ix0, sx); }
    if (e2 <= dx) { err = wadd(err, dx); iy0 = wadd(iy0, sy); }
  }
  return n;
}
