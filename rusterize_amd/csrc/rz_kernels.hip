// rz_kernels.hip -- gfx950 kernels: device-side edge construction, item
// descriptors and binning.  The tile kernel templates are in rz_tile_t.h;
// the pipeline is described in rz_kernels.h.
#include "rz_device.h"

namespace rzk {

// ------------------------------------------------------------- item maps --

__global__ void k_item_geom(const int64_t* geom_item_off, int64_t n_geoms, int32_t* item_geom) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_geoms) return;
  for (int64_t i = geom_item_off[g]; i < geom_item_off[g + 1]; ++i) item_geom[i] = (int32_t)g;
}

__global__ void k_ring_item(const int64_t* item_ring_off, const int64_t* ring_voff, int64_t n_items,
                            int32_t* ring_item, int64_t* item_vbeg) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_items) return;
  int64_t r0 = item_ring_off[i], r1 = item_ring_off[i + 1];
  for (int64_t r = r0; r < r1; ++r) ring_item[r] = (int32_t)i;
  item_vbeg[i] = ring_voff[r0];
  if (i == n_items - 1) item_vbeg[n_items] = ring_voff[r1];
}

// The per-vertex record every rasterise call reads (built once per resident
// geometry set, so k_prep needs one load per vertex instead of the chain
// vertex -> ring -> item -> kind): see the VI_* layout.
__global__ void k_vert_info(const int64_t* ring_voff, const int32_t* ring_item, const int32_t* item_kind,
                            const double* xy, int64_t n_rings, uint64_t* vert_info) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rings) return;
  const int64_t rbeg = ring_voff[r], rend = ring_voff[r + 1];
  if (rend <= rbeg) return;
  const int32_t item = ring_item[r];
  // geo-types LineString::is_closed: first coordinate == last coordinate
  const bool closed = xy[2 * rbeg] == xy[2 * (rend - 1)] && xy[2 * rbeg + 1] == xy[2 * (rend - 1) + 1];
  const uint64_t base = (uint64_t)(uint32_t)item | ((uint64_t)(item_kind[item] & 3) << VI_KIND_SHIFT) |
                        (closed ? VI_CLOSED : 0ull);
  for (int64_t v = rbeg; v < rend; ++v)
    vert_info[v] = base | (v + 1 < rend ? VI_HAS_NEXT : 0ull) | (v == rbeg ? VI_RING_START : 0ull);
}

// Start of a call: every item's box and statistic reset, and the words the
// call counts into (status and burn counters, tile queues, the zero tails of
// the per-item scans) cleared, in one launch.
__global__ void k_init(InitArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n_items) {
    a.box[i] = ItemBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    a.last[i] = 0ull;
    if (a.win) {  // (f64::MAX, f64::MAX, f64::MIN, f64::MIN): the folds' start, pixel_cache.rs:18-29
      const uint64_t hi = f64_key(1.7976931348623157e308), lo = f64_key(-1.7976931348623157e308);
#pragma unroll
      for (int k = 0; k < 8; k += 4) {
        a.win[8 * i + k + 0] = hi;
        a.win[8 * i + k + 1] = hi;
        a.win[8 * i + k + 2] = lo;
        a.win[8 * i + k + 3] = lo;
      }
    }
  }
  uint64_t t = (uint64_t)i;
#pragma unroll
  for (int r = 0; r < InitArgs::kRanges; ++r) {
    if (t < a.n_zero[r]) {
      a.zero[r][t] = 0u;
      break;
    }
    t -= a.n_zero[r];
  }
}

// Clears n words (a kernel rather than a memset node, so captured batches
// replay it reliably).
__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Fills keys[n .. cap) with a key above every tile (graph-captured batches
// sort a fixed capacity; the filler sorts to the end and is never read).
__global__ void k_fill_keys(uint32_t* keys, const uint64_t* n_pairs, uint64_t cap) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cap && i >= *n_pairs) keys[i] = 0xFFFFFFFFu;
}

// ------------------------------------------------------------------ prep --

// Segmented (by item) min/max reduction across the wave, then one set of
// atomics per item segment.  Items own contiguous vertex ranges, so lanes of
// one item are contiguous.  `stat` is item_last's contribution: polygon
// lanes add the rows their edge crosses (saturating: a ranking statistic
// only), line lanes take the maximum of (kept segment + 1).  An item has one
// kind, so every segment combines with one of the two.
__device__ inline void reduce_box(int64_t item, ItemBox b, uint64_t stat, bool line, ItemBox* out,
                                  unsigned long long* last_out, bool active) {
  const int lane = threadIdx.x & 63;
  int64_t it = active ? item : -1;
  // polygon row sums saturate below 2^31: 32-bit shuffles unless a line
  // vertex (a 64-bit vertex index) is in the wave
  const bool any_line = __ballot(line) != 0;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int64_t oit = __shfl_up(it, off, 64);
    int32_t r0 = __shfl_up(b.r0, off, 64), r1 = __shfl_up(b.r1, off, 64);
    int32_t c0 = __shfl_up(b.c0, off, 64), c1 = __shfl_up(b.c1, off, 64);
    const uint64_t ostat = any_line ? (uint64_t)__shfl_up(stat, off, 64)
                                    : (uint64_t)(uint32_t)__shfl_up((int)(uint32_t)stat, off, 64);
    if (lane >= off && oit == it) {
      b.r0 = min(b.r0, r0);
      b.r1 = max(b.r1, r1);
      b.c0 = min(b.c0, c0);
      b.c1 = max(b.c1, c1);
      stat = line ? max(stat, ostat) : min(stat + ostat, (uint64_t)0x7FFFFFFFu);
    }
  }
  int64_t nit = __shfl_down(it, 1, 64);
  bool tail = (lane == 63) || (nit != it);
  // A segment's box is merged whenever some edge set it, even when its own
  // column range is empty: a polygon edge left or right of the raster
  // clamps to an empty range, and the item's other edges, possibly in
  // another wave, widen the merged box to the span between them.
  if (active && tail && b.r0 < b.r1) {
    atomicMin(&out[item].r0, b.r0);
    atomicMax(&out[item].r1, b.r1);
    atomicMin(&out[item].c0, b.c0);
    atomicMax(&out[item].c1, b.c1);
  }
  if (active && tail && stat) {
    if (line) atomicMax(&last_out[item], (unsigned long long)stat);
    else atomicAdd(&last_out[item], (unsigned long long)stat);
  }
}

// One thread per vertex of the batch: world -> pixel (edges.rs:94-97),
// polygon edge records (extract_ring + PolyEdge::new, edges.rs:27-46,
// 90-110), line / point pixel coordinates (extract_line / extract_point,
// edges.rs:79-88, 112-134) and each item's pixel bounding box.
__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = v < a.n_vertices;
  int64_t item = -1;
  ItemBox box{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  uint64_t stat = 0;  // polygon edges: rows [ystart, yend) crossed; lines: kept segment + 1
  bool line = false;
  bool odd = false;   // polygon edges: some row may end with an odd crossing count
  bool far = false;   // line segments: a floored coordinate beyond +-2^30
  auto vertex = [&](const RZ_CONST JobDev& J) {
    const int64_t lv = v - J.v_base;
    // world coordinates of this vertex and of the next one in the job, all
    // loaded alongside the vertex record (none waits on it)
    // (the next index is clamped to the job, so no load sits behind a branch;
    // the last vertex of a job has no successor in its ring)
    const double* xy = J.xy;
    const int64_t ln = min(lv + 1, J.nv - 1);
    const uint64_t vi = J.vert_info[lv];
    const double X0 = xy[2 * lv], Y0 = xy[2 * lv + 1];
    const double X1 = xy[2 * ln], Y1 = xy[2 * ln + 1];
    item = J.i_base + (int64_t)(uint32_t)vi;
    const int32_t kind = (int32_t)((vi >> VI_KIND_SHIFT) & 3u);
    const bool has_next = (vi & VI_HAS_NEXT) != 0;
    const bool closed = (vi & VI_CLOSED) != 0;
    const double rows = (double)J.nrows, cols = (double)J.ncols;
    // world -> pixel; a NaN coordinate stays that NaN (x86 semantics, x86_nan)
    auto px_x = [&](double X) { return isnan(X) ? X : (X - J.xmin) / J.xres; };
    auto px_y = [&](double Y) { return isnan(Y) ? Y : (J.ymax - Y) / J.yres; };
    const double x0 = px_x(X0);
    const double y0 = px_y(Y0);
    if (kind == 0) {  // polygon rings: extract_ring + PolyEdge::new
      PolyEdge e{0.0, 0.0, 0.0, 0, 0};
      // a ring left open by the caller (its last coordinate is not its
      // first) can leave rows with an odd crossing count
      if ((vi & VI_RING_START) && !closed) odd = true;
      if (has_next) {
        const double x1 = px_x(X1);
        const double y1 = px_y(Y1);
        if (!(fabs(y0 - y1) >= 2.220446049250313e-16)) {
          // a skipped segment whose end points straddle a pixel-centre row
          // (or are not numbers) leaves rows with an odd crossing count
          if (isnan(y0) || isnan(y1) || ceil(fmin(y0, y1) - 0.5) != ceil(fmax(y0, y1) - 0.5)) odd = true;
        } else {  // edges.rs:100
          const double min_y = fmin(y0, y1), max_y = fmax(y0, y1);
          if (min_y < rows && max_y >= 0.0) {  // edges.rs:105
            double xt, yt, xb, yb;
            if (y0 < y1) { xt = x0; yt = y0; xb = x1; yb = y1; }
            else { xt = x1; yt = y1; xb = x0; yb = y0; }
            e.ystart = sat_usize_cap(ceil(yt - 0.5), J.nrows);
            e.yend = sat_usize_cap(ceil(yb - 0.5), J.nrows);
            // (yb - yt is neither NaN nor zero here)
            const double dx = x86_nan(xb - xt, xb, xt);
            e.dxdy = x86_nan(dx / (yb - yt), dx, 0.0);
            e.x0 = xt;
            e.y0 = yt;
            // With a NaN x0 or slope, every crossing of the edge is NaN with
            // that NaN's sign (x0's when both are NaN).  The reference sorts a
            // positive NaN crossing last, where it closes an empty span or is
            // the odd one out: as if it were absent, so the edge is emptied.
            // A negative one sorts first and spans from column 0, like -inf:
            // the edge is pinned there.
            if (isnan(xt) || isnan(e.dxdy)) {
              if (signbit(isnan(xt) ? xt : e.dxdy)) {
                e.x0 = -__builtin_inf();
                e.y0 = 0.0;
                e.dxdy = 0.0;
              } else {
                if (e.ystart < e.yend) odd = true;  // its rows lose one crossing each
                e.ystart = e.yend = 0;
              }
            }
            if (e.ystart < e.yend) {
              box.r0 = e.ystart;
              box.r1 = e.yend;
              stat = (uint64_t)(e.yend - e.ystart);
              // Crossings of this edge lie between its end points up to a few
              // ulps, so their span columns are within one pixel of these.
              const double ax = fabs(xt), bx = fabs(xb);
              if (isnan(xt) || isnan(xb) || ax > 1.0e12 || bx > 1.0e12) {
                box.c0 = 0;
                box.c1 = J.ncols;
              } else {
                const double lo = floor(fmin(xt, xb) + 0.5) - 1.0;
                const double hi = floor(fmax(xt, xb) + 0.5) + 1.0;
                box.c0 = (int32_t)fmin(fmax(lo, 0.0), cols);
                box.c1 = (int32_t)fmin(fmax(hi, 0.0), cols);
              }
            }
          }
        }
      }
      a.edges[v] = e;
      const bool at_cache = a.item_win && (J.fun == 0 || J.fun == 5);
      if (J.all_touched && at_cache) {
        // the box of every vertex, which bounds the fill (a NaN coordinate
        // may pin a crossing anywhere: the box becomes infinite)
        unsigned long long* w = a.item_win + 8 * item + 4;
        const double inf = __builtin_inf();
        atomicMin(&w[0], f64_key(isnan(x0) ? -inf : x0));
        atomicMax(&w[2], f64_key(isnan(x0) ? inf : x0));
        atomicMin(&w[1], f64_key(isnan(y0) ? -inf : y0));
        atomicMax(&w[3], f64_key(isnan(y0) ? inf : y0));
      }
      if (J.all_touched && has_next) {
        // the outline segment the all-touched polygon also burns as a line
        // (burn_geometry.rs:89-106: extract_line of every ring, edges.rs:130)
        const double x1 = px_x(X1);
        const double y1 = px_y(Y1);
        const double mnx = fmin(x0, x1), mxx = fmax(x0, x1);
        const double mny = fmin(y0, y1), mxy = fmax(y0, y1);
        if (mnx < cols && mxx >= 0.0 && mny < rows && mxy >= 0.0) {
          // its cells lie within the floors of its end points (one pixel of
          // margin for the walker's 1e-9 steps), merged into the item's box
          const int32_t c0 = (int32_t)max((int64_t)0, min(sat_isize(floor(mnx)) - 1, (int64_t)J.ncols));
          const int32_t c1 = (int32_t)max((int64_t)0, min(sat_isize(floor(mxx)) + 2, (int64_t)J.ncols));
          const int32_t r0 = (int32_t)max((int64_t)0, min(sat_isize(floor(mny)) - 1, (int64_t)J.nrows));
          const int32_t r1 = (int32_t)max((int64_t)0, min(sat_isize(floor(mxy)) + 2, (int64_t)J.nrows));
          if (c0 < c1 && r0 < r1) {
            box.r0 = min(box.r0, r0);
            box.r1 = max(box.r1, r1);
            box.c0 = min(box.c0, c0);
            box.c1 = max(box.c1, c1);
          }
          // the PixelCache box: f64::min / max over both end points (a NaN
          // leaves the fold unchanged)
          if (at_cache) {
            unsigned long long* w = a.item_win + 8 * item;
            if (!isnan(x0)) { atomicMin(&w[0], f64_key(x0)); atomicMax(&w[2], f64_key(x0)); }
            if (!isnan(x1)) { atomicMin(&w[0], f64_key(x1)); atomicMax(&w[2], f64_key(x1)); }
            if (!isnan(y0)) { atomicMin(&w[1], f64_key(y0)); atomicMax(&w[3], f64_key(y0)); }
            if (!isnan(y1)) { atomicMin(&w[1], f64_key(y1)); atomicMax(&w[3], f64_key(y1)); }
          }
        }
      }
    } else if (kind == 1) {  // line rings: extract_line
      line = true;
      uint4 rec = make_uint4(kLsNone, 0u, 0u, 0u);
      if (has_next) {
        const double x1 = px_x(X1);
        const double y1 = px_y(Y1);
        const double mnx = fmin(x0, x1), mxx = fmax(x0, x1);
        const double mny = fmin(y0, y1), mxy = fmax(y0, y1);
        if (mnx < cols && mxx >= 0.0 && mny < rows && mxy >= 0.0) {  // edges.rs:130
          stat = (uint64_t)(v + 1);
          // Bresenham visits cells between the floored end points only.
          const int64_t fx0 = sat_isize(floor(mnx)), fx1 = sat_isize(floor(mxx));
          const int64_t fy0 = sat_isize(floor(mny)), fy1 = sat_isize(floor(mxy));
          // (all-touched walks get one pixel of margin for their 1e-9 steps)
          const int64_t m = J.all_touched ? 1 : 0;
          box.c0 = (int32_t)max((int64_t)0, min(fx0 - m, (int64_t)J.ncols));
          box.c1 = (int32_t)max((int64_t)0, min(fx1 + 1 + m, (int64_t)J.ncols));
          box.r0 = (int32_t)max((int64_t)0, min(fy0 - m, (int64_t)J.nrows));
          box.r1 = (int32_t)max((int64_t)0, min(fy1 + 1 + m, (int64_t)J.nrows));
          // the segment's end cells (burners.rs:54-57), biased into 31 bits
          const int64_t ix0 = sat_isize(floor(x0)), iy0 = sat_isize(floor(y0));
          const int64_t ix1 = sat_isize(floor(x1)), iy1 = sat_isize(floor(y1));
          if (fx0 > -kLsBias && fx1 < kLsBias && fy0 > -kLsBias && fy1 < kLsBias) {
            rec = make_uint4((uint32_t)(ix0 + kLsBias), (uint32_t)(iy0 + kLsBias), (uint32_t)(ix1 + kLsBias),
                             (uint32_t)(iy1 + kLsBias) | (closed ? kLsFlag : 0u));
          } else {
            rec = make_uint4(kLsFar, 0u, 0u, closed ? kLsFlag : 0u);
            far = true;
          }
        }
      }
      a.lseg[v] = rec;
    } else {  // points: extract_point
      uint4 rec = make_uint4(kLsNone, 0u, 0u, 0u);
      if (x0 >= 0.0 && x0 < cols && y0 >= 0.0 && y0 < rows) {  // edges.rs:85
        const int32_t cx = (int32_t)x0, cy = (int32_t)y0;
        box = ItemBox{cy, cy + 1, cx, cx + 1};
        rec = make_uint4((uint32_t)cx, (uint32_t)cy, 0u, 0u);
      }
      a.lseg[v] = rec;
    }
  };
  // A wave inside one job (the common case) reads the job record with
  // scalar loads, all at once; a wave across a job boundary per lane.
  const int64_t v0 = __builtin_amdgcn_readfirstlane(v);
  const int j0 = find_job(as_const(a.job_vend), a.n_jobs, v0);
  const int64_t v_end = min(v0 + 64, a.n_vertices);  // the wave's vertices: [v0, v_end)
  if (v_end <= as_const(a.job_vend)[j0]) {
    if (active) vertex(as_const(a.jobs)[j0]);
  } else if (active) {
    vertex(as_const(a.jobs)[find_job(a.job_vend, a.n_jobs, v)]);
  }
  reduce_box(item, box, stat, line, a.item_box, a.item_last, active);
  if (odd) atomicOr(&a.item_last[item], kItemOddBit);
  if (far) atomicOr(&a.item_last[item], kItemFarBit);
}

// ------------------------------------------------------- item descriptors --

__device__ inline bool tile_rect(int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t tw,
                                 int32_t& tx0, int32_t& tx1, int32_t& ty0, int32_t& ty1) {
  if (!(r0 < r1 && c0 < c1)) return false;
  tx0 = c0 / tw;
  tx1 = (c1 - 1) / tw;
  ty0 = r0 / TILE_H;
  ty1 = (r1 - 1) / TILE_H;
  return true;
}

// One thread per item: resolve the field value and band of the owning
// geometry (rust/src/rasterize.rs:169-179) into the item's descriptor and,
// for binned jobs, count the tiles its box covers.
__global__ void k_item_desc(DescArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const int j = find_job_wave(a.job_iend, a.n_jobs, i);
  const RZ_CONST JobDev& J = as_const(a.jobs)[j];
  const int64_t li = i - J.i_base;
  ItemBox b = a.item_box[i];
  const int32_t g = J.item_geom[li];
  uint32_t layer = J.layer_of_geom ? J.layer_of_geom[g] : 0u;
  if (layer >= J.n_layers) {
    // an out-of-range band index burns nothing (its tiles would be another
    // band's or another job's) and fails the call
    atomicOr(a.status, 16u);
    layer = 0u;
    b = ItemBox{0, 0, 0, 0};
  }
  ItemDesc d;
  d.r0 = b.r0;
  d.r1 = b.r1;
  d.c0 = b.c0;
  d.c1 = b.c1;
  d.vbeg = J.v_base + J.item_vbeg[li];
  d.nv = (int32_t)(J.item_vbeg[li + 1] - J.item_vbeg[li]);
  d.kind_layer = (uint32_t)J.item_kind[li] | (layer << 2);
  uint64_t v = J.scalar_bits;
  if (J.field) {
    const uint8_t* f = (const uint8_t*)J.field + (size_t)g * (size_t)J.esz;
    v = 0;
    for (int k = 0; k < J.esz; ++k) v |= (uint64_t)f[k] << (8 * k);
  }
  d.value = v;
  const int32_t kind = J.item_kind[li];
  const uint64_t S_raw = a.item_last[i];
  const bool tall = d.r0 < d.r1 && (d.r1 - 1) / kBandRows > d.r0 / kBandRows;
  const bool banded = a.band_n && tall &&
                      ((kind == 0 && d.nv > kBandMinEdges) ||
                       (kind == 1 && d.nv > kLineBandMinSegs && !(S_raw & kItemFarBit) && !J.all_touched));
  d.last = kind == 1 ? (S_raw & ~kItemFarBit) : 0ull;
  if (kind == 0 && J.all_touched) d.last = (uint64_t)i;  // (its outline box is found by item)
  if (a.at_words) {
    // sum / count dedup (prelude.rs:116-118): the fill skips pixels in the
    // outline's PixelCache; a fill pixel outside the cache box may alias a
    // cache bit (unravel_index wraps), so such an item's cache is kept in
    // global memory, where the tile kernel looks those pixels up
    uint32_t words = 0;
    CacheWin cw;
    if (kind == 0 && J.all_touched && (J.fun == 0 || J.fun == 5) && d.r0 < d.r1 && d.c0 < d.c1 &&
        item_cache(a.item_win, i, cw)) {
      // fill pixels lie in the floors of the vertex box (crossings are
      // between edge end points; spans start at floor(x + 0.5) and end
      // before it), inside the raster; the window holds columns
      // [x0, x0 + width) and rows [y0, y0 + length)
      const unsigned long long* w = a.item_win + 8 * i + 4;
      auto lo = [](double v, int64_t n) { return v < 0.0 ? (int64_t)0 : (v >= (double)n ? n : (int64_t)v); };
      const int64_t fc0 = lo(floor(f64_from_key(w[0])), J.ncols), fc1 = lo(floor(f64_from_key(w[2])), J.ncols - 1);
      const int64_t fr0 = lo(floor(f64_from_key(w[1])), J.nrows), fr1 = lo(floor(f64_from_key(w[3])), J.nrows - 1);
      const uint64_t length = cw.width ? cw.cap / cw.width : 0;
      const bool inside = fc0 > fc1 || fr0 > fr1 ||
                          (fc0 >= cw.x0 && (uint64_t)fc1 - (uint64_t)cw.x0 < cw.width && fr0 >= cw.y0 &&
                           (uint64_t)fr1 - (uint64_t)cw.y0 < length);
      if (!inside) {
        if (cw.cap > ((uint64_t)1 << 34)) atomicOr(a.status, 32u);  // more than 2 GiB of cache bits
        else words = (uint32_t)((cw.cap + 63) / 64);
      }
    }
    a.at_words[i] = words;
  }
  d.job = (uint32_t)j;
  d.band = 0;
  d.flags = 0;
  d.wide = 0;
  if (a.scan_box && J.scan_compact) a.scan_box[i] = scan_box_of(d.r0, d.r1, d.c0, d.c1, J.tile_w);
  if (kind == 0 && (S_raw & kItemOddBit)) d.flags |= kItemOddRows;
  if (a.at_words && a.at_words[i]) d.flags |= kItemAtCache;
  a.desc[i] = d;
  if (a.wide_units) {
    // Work of one strip, from the rows the edges cross (S) over the item's
    // height (H): an edge crossing s rows meets about (s + 15) / H of the
    // item's strips, so a strip sees n ~ (S + 15 nv) / H edges.  In steps of
    // 64 crossing tests: in a tile, row-major coverage takes n / 4 steps,
    // repeated in each of the tw tiles of the item's width.  The wide
    // pre-pass does the strip once, in n / 4 steps or, edge-major, in the
    // longest crossing run of each chunk of 64 loaded edges (all nv, or the n
    // of the strip's band; taken as 2 S / nv + 1 rows, at most 16), plus
    // about 16 steps of set-up and mask stores, and the tiles then read one
    // mask word each (about half a step).  Wide when that is at least a
    // quarter cheaper.
    uint32_t units = 0, words = 0;
    if (kind == 0 && d.nv > 0 && d.r0 < d.r1 && d.c0 < d.c1 && a.wide_mode != 1) {
      const float S = (float)(S_raw & ~kItemOddBit), nv = (float)d.nv, H = (float)(d.r1 - d.r0);
      const float n = fminf(nv, (S + 15.0f * nv) / H);
      const float chunks = ceilf((banded ? n : nv) / 64.0f);
      const float run = fminf(16.0f, 2.0f * S / nv + 1.0f);
      const float tw = (float)((d.c1 - 1) / J.tile_w - d.c0 / J.tile_w + 1);
      const float in_tile = tw * fmaxf(1.0f, n / 4.0f);
      const float wide = fminf(n / 4.0f, chunks * run) + 16.0f + 0.5f * tw;
      if (a.wide_mode == 2 || in_tile > 1.25f * wide) {
        const int32_t ns = wide_strip_count(d.r0, d.r1), nw = wide_row_words(d.c0, d.c1);
        units = (uint32_t)ns * (uint32_t)((nw + kWideWords - 1) / kWideWords);
        words = (uint32_t)ns * 16u * (uint32_t)nw;
      }
    }
    a.wide_units[i] = units;
    a.wide_words[i] = words;
  }
  if (a.band_n)
    a.band_n[i] = banded ? (uint32_t)((d.r1 - 1) / kBandRows - d.r0 / kBandRows + 1) : 0u;
  if (a.counts) {
    int32_t tx0, tx1, ty0, ty1;
    uint32_t c = 0;
    if (!J.scan && tile_rect(d.r0, d.r1, d.c0, d.c1, J.tile_w, tx0, tx1, ty0, ty1))
      c = (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1);
    a.counts[i] = c;
  }
}

// ------------------------------------------------------------ edge bands --

// Items whose band entries would pass the capacity stay unbanded (their
// strips then scan the whole edge list: slower, same result).
// Assigned items are also appended to the list the band passes walk (any
// order: every item owns its entries), one atomic per block (one per wave
// contended measurably on millions of items).
constexpr int kBandAssignBlock = 1024;
__global__ __launch_bounds__(kBandAssignBlock) void k_band_assign(ItemDesc* desc, const uint32_t* band_n,
                                                                  const uint64_t* band_base, int64_t n_items,
                                                                  uint64_t cap_entries, uint32_t* items,
                                                                  uint32_t* n_list) {
  __shared__ uint32_t s_n[kBandAssignBlock / 64], s_base;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n_items && band_n[i] && band_base[i] + band_n[i] <= cap_entries;
  if (ok) desc[i].band = (uint32_t)band_base[i] + 1u;
  const uint64_t m = __ballot(ok);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_n[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t n = 0;
    for (int w = 0; w < kBandAssignBlock / 64; ++w) n += s_n[w];
    s_base = n ? atomicAdd(n_list, n) : 0u;
  }
  __syncthreads();
  uint32_t base = s_base;
  for (int w = 0; w < wave; ++w) base += s_n[w];
  if (ok) items[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
}

// Wide items whose units and mask words fit the capacities (every item
// below them in the scans does too) take the wide pre-pass: flag, mask base
// and one unit-list entry per (strip, window) unit.
__global__ void k_wide_assign(WideAssignArgs a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const uint32_t nu = a.wide_units[i];
  if (!nu) return;
  const uint64_t ub = a.unit_base[i], wb = a.word_base[i];
  if (ub + nu > a.unit_cap || a.word_base[i + 1] > a.word_cap) return;  // stays per tile
  a.desc[i].flags |= kItemWide;
  a.desc[i].wide = (uint32_t)(wb >> 4);  // every item's words are a multiple of 16
  for (uint32_t k = 0; k < nu; ++k) a.unit_item[ub + k] = (uint32_t)i;
}

// ------------------------------------------------- all-touched PixelCache --

// The PixelCache of each kItemAtCache item (burn_geometry.rs:225-234 with
// pixel_cache.rs:38-59): every cell its kept outline segments walk sets the
// bit at the cell's cache index (cells inside the raster; an index past the
// capacity, which the reference rejects, sets nothing).  One wave per item.
__global__ __launch_bounds__(256) void k_at_cache(AtCacheArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t i = __builtin_amdgcn_readfirstlane(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= a.n_items) return;
  const RZ_CONST ItemDesc& d = as_const(a.desc)[i];
  if (!(d.flags & kItemAtCache)) return;
  if (a.at_base[i + 1] > a.cap_words) {  // a captured batch's cache capacity is exceeded
    if (lane == 0) atomicOr(a.status, 8u);
    return;
  }
  const RZ_CONST JobDev& J = as_const(a.jobs)[d.job];
  CacheWin cw;
  if (!item_cache(a.item_win, i, cw)) return;
  unsigned long long* bits = a.at_bits + a.at_base[i];
  bool ok = true;
  for (int64_t e = lane; e < d.nv; e += 64) {
    double x0, y0, x1, y1;
    if (!kept_segment(J, d.vbeg + e - J.v_base, x0, y0, x1, y1)) continue;
    ok &= walk_all_touched(x0, y0, x1, y1, J.ncols, J.nrows, INT64_MIN, INT64_MAX, INT64_MIN, INT64_MAX,
                           [&](int64_t x, int64_t y) {
                             const uint64_t idx = cache_index(cw, x, y);
                             if (idx < cw.cap) atomicOr(&bits[idx >> 6], 1ull << (idx & 63));
                           });
  }
  if (!ok) atomicOr(a.status, 64u);
}

// Count pass, one wave per banded item (of the list k_band_assign built): band_cnt of each of its strips,
// from a difference array in LDS (+1 at an edge's first strip, -1 after its
// last) and a running sum, so an edge costs two LDS atomics however many
// strips it meets.  Plain stores: every item owns its band entries.
__global__ __launch_bounds__(256) void k_band_count(BandArgs a) {
  __shared__ uint32_t s_cnt[4][kBandLds + 1];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t* cnt = s_cnt[wave];
  const uint32_t n_list = *a.n_list;
  for (uint32_t li = blockIdx.x * 4u + (uint32_t)wave; li < n_list; li += gridDim.x * 4u) {
    const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.items[li]);  // wave-uniform
    const uint32_t ns = as_const(a.band_n)[i];
    const RZ_CONST ItemDesc& d = as_const(a.desc)[i];
    const int32_t sr0 = d.r0 / kBandRows, sr1 = (d.r1 - 1) / kBandRows;
    const PolyEdge* eb = a.edges + d.vbeg;
    for (uint32_t s0 = 0; s0 < ns; s0 += kBandLds) {
      const uint32_t L = min((uint32_t)kBandLds, ns - s0);
      for (uint32_t k = lane; k <= L; k += 64) cnt[k] = 0u;
      wave_sync();
      // each lane's next edge is loaded while the current one is counted;
      // its rows [ystart, yend), or for a line segment those of its cells
      const bool line = (d.kind_layer & 3u) == 1u;  // wave-uniform
      const uint4* sb = a.lseg + d.vbeg;
      auto rows_of = [&](int32_t e) -> int2 {
        if (!line) return *reinterpret_cast<const int2*>(&eb[e].ystart);
        const uint4 r = sb[e];
        if (r.x >= kLsFar) return make_int2(0, 0);
        const int32_t y0 = (int32_t)((int64_t)r.y - kLsBias), y1 = (int32_t)((int64_t)(r.w & ~kLsFlag) - kLsBias);
        return make_int2(max(min(y0, y1), 0), max(y0, y1) + 1);
      };
      int2 ny = make_int2(0, 0);
      if (lane < d.nv) ny = rows_of(lane);
      for (int32_t e = lane; e < d.nv; e += 64) {
        const int2 yy = ny;
        if (e + 64 < d.nv) ny = rows_of(e + 64);
        if (!(yy.x < yy.y)) continue;
        const int32_t b0 = max(max(yy.x / kBandRows, sr0) - sr0, (int32_t)s0);
        const int32_t b1 = min(min((yy.y - 1) / kBandRows, sr1) - sr0, (int32_t)(s0 + L - 1));
        if (b0 > b1) continue;
        atomicAdd(&cnt[b0 - s0], 1u);
        atomicSub(&cnt[b1 - s0 + 1], 1u);
      }
      wave_sync();
      uint32_t run = 0;
      for (uint32_t k0 = 0; k0 < L; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint32_t x = k < L ? cnt[k] : 0u;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
          if (lane >= off) x += y;
        }
        if (k < L) a.band_cnt[d.band - 1u + s0 + k] = run + x;
        run += (uint32_t)__shfl((int)x, 63, 64);
      }
      wave_sync();
    }
  }
}

// Scatter pass, one wave per banded item: each edge is copied into every
// strip it meets, at a slot taken from the strip's LDS cursor.
__global__ __launch_bounds__(256) void k_band_scatter(BandArgs a) {
  __shared__ uint32_t s_pos[4][kBandLds];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t* pos = s_pos[wave];
  const uint32_t n_list = *a.n_list;
  for (uint32_t li = blockIdx.x * 4u + (uint32_t)wave; li < n_list; li += gridDim.x * 4u) {
    const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.items[li]);  // wave-uniform
    const uint32_t ns = as_const(a.band_n)[i];
    const RZ_CONST ItemDesc& d = as_const(a.desc)[i];
    const int32_t sr0 = d.r0 / kBandRows, sr1 = (d.r1 - 1) / kBandRows;
    const PolyEdge* eb = a.edges + d.vbeg;
    const uint64_t base = as_const(a.band_off)[d.band - 1u];
    for (uint32_t s0 = 0; s0 < ns; s0 += kBandLds) {
      const uint32_t L = min((uint32_t)kBandLds, ns - s0);
      for (uint32_t k = lane; k < L; k += 64) pos[k] = (uint32_t)(a.band_off[d.band - 1u + s0 + k] - base);
      wave_sync();
      if ((d.kind_layer & 3u) == 1u) {  // a line item (wave-uniform): segment records
        const uint4* sb = a.lseg + d.vbeg;
        uint4 nr = make_uint4(kLsNone, 0u, 0u, 0u);
        if (lane < d.nv) nr = sb[lane];
        for (int32_t e = lane; e < d.nv; e += 64) {
          uint4 r = nr;
          if (e + 64 < d.nv) nr = sb[e + 64];
          if (r.x >= kLsFar) continue;
          const int32_t y0 = (int32_t)((int64_t)r.y - kLsBias);
          const int32_t y1 = (int32_t)((int64_t)(r.w & ~kLsFlag) - kLsBias);
          const int32_t ys = max(min(y0, y1), 0), ye = max(y0, y1) + 1;
          // the band record's flag: this is the last kept edge of an open
          // line, whose end cell is written (burners.rs:86-89)
          const bool end = (uint64_t)(d.vbeg + e + 1) == d.last && !(r.w & kLsFlag);
          r.w = (r.w & ~kLsFlag) | (end ? kLsFlag : 0u);
          const int32_t b0 = max(max(ys / kBandRows, sr0) - sr0, (int32_t)s0);
          const int32_t b1 = min(min((ye - 1) / kBandRows, sr1) - sr0, (int32_t)(s0 + L - 1));
          for (int32_t b = b0; b <= b1; ++b) {
            const uint64_t slot = base + atomicAdd(&pos[b - s0], 1u);
            if (slot < a.edge_cap) a.band_segs[slot] = r;
            else atomicOr(a.status, 8u);
          }
        }
        wave_sync();
        continue;
      }
      // the record as two 16-byte vectors (no struct copy through scratch);
      // each lane's next edge is loaded while the current one is copied
      uint4 nlo = make_uint4(0u, 0u, 0u, 0u), nhi = nlo;
      if (lane < d.nv) {
        nlo = reinterpret_cast<const uint4*>(eb + lane)[0];
        nhi = reinterpret_cast<const uint4*>(eb + lane)[1];
      }
      for (int32_t e = lane; e < d.nv; e += 64) {
        const uint4 lo = nlo, hi = nhi;  // hi.z = ystart, hi.w = yend
        if (e + 64 < d.nv) {
          nlo = reinterpret_cast<const uint4*>(eb + e + 64)[0];
          nhi = reinterpret_cast<const uint4*>(eb + e + 64)[1];
        }
        const int32_t ys = (int32_t)hi.z, ye = (int32_t)hi.w;
        if (!(ys < ye)) continue;
        const int32_t b0 = max(max(ys / kBandRows, sr0) - sr0, (int32_t)s0);
        const int32_t b1 = min(min((ye - 1) / kBandRows, sr1) - sr0, (int32_t)(s0 + L - 1));
        for (int32_t b = b0; b <= b1; ++b) {
          const uint64_t slot = base + atomicAdd(&pos[b - s0], 1u);
          if (slot < a.edge_cap) {
            uint4* dst = reinterpret_cast<uint4*>(a.band_edges + slot);
            dst[0] = lo;
            dst[1] = hi;
          } else {
            atomicOr(a.status, 8u);
          }
        }
      }
      wave_sync();
    }
  }
}

// --------------------------------------------------------------- binning --

// Pairs (tile key, item) in item order; a stable sort by key then keeps
// geometry order inside every tile.
__global__ void k_bin_emit(BinArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  const ItemDesc d = a.desc[i];
  const JobDev& J = a.jobs[d.job];
  if (J.scan) return;
  int32_t tx0, tx1, ty0, ty1;
  if (!tile_rect(d.r0, d.r1, d.c0, d.c1, J.tile_w, tx0, tx1, ty0, ty1)) return;
  uint64_t o = a.offsets[i];
  if (a.offsets[i + 1] > a.cap) {  // capacity of a captured batch exceeded
    // the one item whose range crosses the capacity fills its slots below it
    // with keys past every tile, so every slot the fixed-capacity sort reads
    // is written in this pass (no stale key reaches the later kernels)
    atomicOr(a.status, 4u);
    for (; o < a.cap; ++o) {
      a.keys[o] = 0xFFFFFFFFu;
      a.vals[o] = 0xFFFFFFFFu;
    }
    return;
  }
  const uint32_t base = J.tile_base + (d.kind_layer >> 2) * J.tiles_per_layer;
  for (int32_t ty = ty0; ty <= ty1; ++ty)
    for (int32_t tx = tx0; tx <= tx1; ++tx) {
      a.keys[o] = base + (uint32_t)ty * (uint32_t)J.ntx + (uint32_t)tx;
      a.vals[o] = (uint32_t)i;
      ++o;
    }
}

// Descriptors in pair order, so a tile's items are one contiguous run.
// n_pairs is the device total; cap bounds it (a captured batch sorts a
// fixed capacity).
__global__ void k_gather_desc(const ItemDesc* desc, int64_t n_items, const uint32_t* pair_item,
                              const uint64_t* n_pairs, uint64_t cap, ItemDesc* pair_desc) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= min(*n_pairs, cap)) return;
  const uint32_t it = pair_item[i];
  if (it < (uint64_t)n_items) pair_desc[i] = desc[it];
}

// tile_off[t] = first pair index whose key >= t, for t in [0, n_tiles]; all
// lists are empty when the pairs overflowed their capacity (status bit 2).
__global__ void k_tile_offsets(const uint32_t* keys, const uint64_t* n_pairs, uint64_t cap,
                               const uint32_t* status, uint32_t n_tiles, uint32_t* tile_off) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  if (*status & 4u) {
    tile_off[t] = 0;
    return;
  }
  uint32_t lo = 0, hi = (uint32_t)min(*n_pairs, cap);
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < (uint32_t)t) lo = mid + 1;
    else hi = mid;
  }
  tile_off[t] = lo;
}

}  // namespace rzk
