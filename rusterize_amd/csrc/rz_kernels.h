// rz_kernels.h -- gfx950 kernels of the rasteriser.
//
// Pipeline of one rasterise call or batch of calls (all on one HIP stream):
//   k_prep        one thread per vertex: world->pixel transform, polygon edge
//                 records (rust/src/geo/edges.rs:27-46, 90-110), line / point
//                 pixel coordinates (edges.rs:79-88, 112-134) and a per-item
//                 pixel bounding box (segmented wave reduction + atomics).
//   k_item_desc   one thread per item: its descriptor (field value, band)
//                 and tile count.
//   scan / k_bin_emit / stable radix sort / k_tile_offsets / k_gather_desc
//                 sort-middle binning of items into 64-row tiles, keeping
//                 geometry order inside every tile.
//   k_tile        persistent work-queue kernel over 16-row tile strips: the
//                 background fill of RasterInfo::build_raster
//                 (rust/src/geo/raster.rs:23-28) and every burn of every item
//                 of the strip, folded in geometry order in registers, then
//                 written to HBM exactly once.
//
// Exactness: all f64 arithmetic is compiled with -ffp-contract=off, divisions
// and floor/ceil are IEEE exact, so every crossing and every pixel-centre
// rounding equals the reference's CPU value bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace rzk {

constexpr int TILE_H = 64;       // rows per tile (4 waves x 16 rows)
constexpr int WAVE_ROWS = 16;    // rows per wave strip
constexpr int NUM_XCD = 8;       // MI355X: 8 XCDs, workgroups dispatched round-robin

// Tile width in pixels for an element size: rows of 256 bytes (512 for
// 16-bit and 64-bit elements), 16 KB or more per tile.
__host__ __device__ constexpr int tile_w(int esz) { return esz <= 2 ? 256 : 64; }

// Edge record of one polygon segment (PolyEdge, rust/src/geo/edges.rs:17-24).
// ystart / yend are the reference's usize values clamped to nrows, which is
// exact for the row test `ystart <= r < yend` with r < nrows.  Invalid
// (skipped) segments have ystart == yend == 0.
struct __align__(32) PolyEdge {
  double x0, y0, dxdy;
  int32_t ystart, yend;
};

// Segment record of one line vertex (extract_line, rust/src/geo/edges.rs:
// 112-134), 16 bytes, built by k_prep: the floored pixel coordinates of the
// segment's end points (the cells Bresenham starts and stops at,
// burners.rs:54-57), each biased by kLsBias so it fits 31 bits:
//   x = ix0 + kLsBias, y = iy0 + kLsBias, z = ix1 + kLsBias,
//   w = (iy1 + kLsBias) | kLsFlag when flagged.
// In the per-vertex array the flag says the segment's ring is closed
// (geo-types LineString::is_closed); in an edge band it says the segment's
// end cell is written (the last kept edge of an open line, burners.rs:86-89).
// x == kLsNone: the vertex starts no kept segment (last vertex of its ring, or
// dropped by extract_line's bounding-box test); x == kLsFar: a kept segment
// with a floored coordinate beyond +-2^30 (recomputed from the world
// coordinates, in 64 or 128 bits).  A point vertex (extract_point,
// edges.rs:79-88) holds its cell, x = column, y = row, or x == kLsNone
// outside the raster.
constexpr uint32_t kLsNone = 0xFFFFFFFFu, kLsFar = 0xFFFFFFFEu, kLsFlag = 0x80000000u;
constexpr int64_t kLsBias = (int64_t)1 << 30;

// Per-vertex record of a resident geometry set: bits 0-31 the local item,
// 32-33 its kind, then whether the vertex has a next one in its ring, whether
// the ring is closed (first coordinate == last), and whether the vertex
// starts its ring.
constexpr int VI_KIND_SHIFT = 32;
constexpr uint64_t VI_HAS_NEXT = 1ull << 34, VI_CLOSED = 1ull << 35, VI_RING_START = 1ull << 36;

// Per-item pixel bounding box: rows [r0, r1), cols [c0, c1).
struct __align__(16) ItemBox {
  int32_t r0, r1, c0, c1;
};

// One rasterise job of a batch (a single call is a batch of one).  Every
// per-call array is indexed in the batch's global vertex / item / tile
// spaces; a job owns the ranges [v_base, v_base + nv), [i_base, i_base + ni)
// and [tile_base, tile_base + n_tiles).  The geometry pointers are the
// job's resident geometry set, indexed locally.
struct JobDev {
  const double* xy;
  const uint64_t* vert_info;  // per local vertex: item, kind and ring flags (VI_*)
  const int64_t* ring_voff;
  const int32_t* ring_item;
  const int32_t* item_kind;
  const int32_t* item_geom;
  const int64_t* item_vbeg;  // [ni + 1] local vertex range of every item
  int64_t v_base, i_base, nv, ni;
  double xmin, ymax, xres, yres;
  int32_t nrows, ncols, ntx, nty;
  int32_t esz, tile_w;       // element size of the output dtype, tile width in pixels
  uint32_t tiles_per_layer, tile_base, n_tiles;
  int32_t fun;               // pixel function (rz_fun)
  int32_t dedup_lines;       // xres != yres (burn_geometry.rs:179, 202)
  int32_t scan;              // 1: no bins, the tile kernel tests every item of the job
  int32_t sum_unordered;     // fun = sum, float output: lines and points add with LDS atomics
  int32_t all_touched;       // AllTouched strategy: lines walked cell by cell, polygons with their outline
  int32_t scan_compact;      // scan job of one band whose strip and tile indices fit 16 bits: its strips
                             // test the 8-byte item boxes of scan_box
  const void* field;         // [n_geoms] of the batch dtype, or null
  uint64_t scalar_bits, bg_bits;
  const uint32_t* layer_of_geom;  // may be null
  uint32_t n_layers;         // bands of the output; a layer index >= n_layers is an error
  void* out;                 // [layers][nrows][ncols]
};

// Read-only tables of a launch (job table, descriptors, tile offsets) are
// read through the constant address space: they do not change while a
// kernel runs, so the compiler may use scalar loads for uniform indices even
// though the kernel also stores to global memory.
#define RZ_CONST __attribute__((address_space(4)))
// LDS (group segment) pointers, for the direct global -> LDS loads
#define RZ_LDS __attribute__((address_space(3)))
template <typename T>
__device__ inline const RZ_CONST T* as_const(const T* p) {
  return (const RZ_CONST T*)p;
}

// Index of the job whose exclusive prefix end first exceeds x.
template <typename P>
__device__ inline int find_job(P ends, int n, int64_t x) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ends[mid] > x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// find_job for one index per lane of a wave whose indices ascend with the
// lane: one (scalar) search for the wave's first active lane, and a search of
// its own only in a lane past that job's end (rare: jobs are long).
__device__ inline int find_job_wave(const int64_t* ends, int n, int64_t x) {
  const int64_t x0 = __builtin_amdgcn_readfirstlane(x);
  const int j0 = find_job(as_const(ends), n, x0);  // uniform: scalar loads
  return x < as_const(ends)[j0] ? j0 : find_job(ends, n, x);
}

struct PrepArgs {
  const JobDev* jobs;
  const int64_t* job_vend;  // [n_jobs] prefix ends of the vertex ranges
  int n_jobs;
  int64_t n_vertices;       // batch total
  PolyEdge* edges;          // [n_vertices]
  uint4* lseg;              // [n_vertices] segment / cell records of line and point vertices
  ItemBox* item_box;        // [n_items]
  unsigned long long* item_last;  // [n_items] lines: last kept segment + 1 (global; 0: none);
                                  // polygons: the sum over edges of the rows each crosses
  unsigned long long* item_win;   // [8 n_items] all-touched sum / count polygons (or null): the f64
                                  // box of their kept outline segments (min x, min y, max x, max y as
                                  // ordered keys), then that of all their vertices (a NaN widens it
                                  // to infinity)
};

// Order-preserving u64 key of an f64 (for atomic min / max): the sign bit
// flipped for positives, every bit for negatives.
__host__ __device__ inline uint64_t f64_key(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ inline double f64_from_key(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
  double v;
  __builtin_memcpy(&v, &b, 8);
  return v;
}

// The PixelCache of an all-touched polygon (rust/src/rasterization/
// pixel_cache.rs:16-36) in integers: origin (xmin as isize, ymin as isize),
// width and length in pixels, capacity width * length (usize arithmetic,
// wrapping), from the f64 box of its kept outline segments.
struct CacheWin {
  int64_t x0, y0;
  uint64_t width, cap;
};
__host__ __device__ inline uint64_t as_usize(double v) {
  if (!(v > 0.0)) return 0;
  if (v >= 18446744073709551616.0) return ~0ull;
  return (uint64_t)v;
}
__host__ __device__ inline CacheWin cache_window(double xmin, double ymin, double xmax, double ymax) {
  CacheWin w;
  w.width = as_usize(floor(xmax) - floor(xmin)) + 1u;
  const uint64_t length = as_usize(floor(ymax) - floor(ymin)) + 1u;
  auto isz = [](double v) -> int64_t {
    if (v != v) return 0;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    if (v >= 9223372036854775808.0) return INT64_MAX;
    return (int64_t)v;
  };
  w.x0 = isz(xmin);
  w.y0 = isz(ymin);
  w.cap = w.width * length;
  return w;
}
// unravel_index (pixel_cache.rs:39-44), wrapping
__host__ __device__ inline uint64_t cache_index(const CacheWin& w, int64_t x, int64_t y) {
  return ((uint64_t)y - (uint64_t)w.y0) * w.width + ((uint64_t)x - (uint64_t)w.x0);
}
// the cache window of item i from its outline box keys (item_win); false when
// no kept segment gave both axes a finite extent (an empty cache: it contains
// no pixel of the raster)
__host__ __device__ inline bool item_cache(const unsigned long long* win, int64_t i, CacheWin& cw) {
  const unsigned long long* w = win + 8 * i;
  const double xmin = f64_from_key(w[0]), ymin = f64_from_key(w[1]);
  const double xmax = f64_from_key(w[2]), ymax = f64_from_key(w[3]);
  if (!(xmin <= xmax && ymin <= ymax)) return false;
  cw = cache_window(xmin, ymin, xmax, ymax);
  return true;
}

// Per-call item descriptor (64 B, one scalar load): everything the tile
// kernel needs about an item, with the owning geometry's field value and
// output band already resolved.  vbeg / last are global vertex indices.
struct __align__(64) ItemDesc {
  int32_t r0, r1, c0, c1;   // pixel box: rows [r0, r1), cols [c0, c1)
  int64_t vbeg;             // first vertex (edge / segment / point index)
  int32_t nv;               // vertex count
  uint32_t kind_layer;      // item kind | output band << 2
  uint64_t value;           // field value bits (dtype of the call)
  uint64_t last;            // lines: last kept segment + 1 (0: none; kItemFarBit masked off)
  uint32_t job;
  uint32_t band;            // banded polygon or line item: first band entry + 1 (0: not banded)
  uint32_t flags;           // kItem* bits
  uint32_t wide;            // kItemWide: first word of the item's row masks / 16
};

// ItemDesc::flags.  kItemWide: the polygon's coverage is precomputed once
// per 16-row strip over its whole column range (k_wide_cover) instead of in
// every output tile it meets: for items several tiles wide with many edges
// per strip, and for items of many short edges (which the pre-pass visits
// edge by edge).
constexpr uint32_t kItemWide = 1u;
// kItemOddRows: some row of the polygon may have an odd crossing count (an
// edge emptied for a positive-NaN slope or origin, or a segment skipped as
// horizontal that straddles a pixel-centre row).  Only then does a row's
// coverage need its rightmost crossing: with an even count the prefix
// parity is already zero from there on.  (Rings are closed, so their
// crossing counts are even otherwise.)
constexpr uint32_t kItemOddRows = 2u;
// kItemAtCache: an all-touched sum / count polygon whose fill may reach
// pixels outside its PixelCache window; those pixels look the cache up in
// global memory (at_bits), where k_at_cache built it.
constexpr uint32_t kItemAtCache = 4u;
// k_prep marks such items in bit 63 of their item_last word.
constexpr unsigned long long kItemOddBit = 1ull << 63;
// ... and line items with a kLsFar segment in bit 62 (they are never banded).
constexpr unsigned long long kItemFarBit = 1ull << 62;

// Row masks of a kItemWide item: strip k (rows from (r0 / 16 + k) * 16) holds
// 16 rows of nw = (c1 - 1) / 64 - c0 / 64 + 1 words each (word w covers
// columns (c0 / 64 + w) * 64 .. + 63), at wide_masks[wide * 16 + (16 k + row)
// * nw + w].  A unit of k_wide_cover is one strip of one item and at most
// kWideWords words of it (a window).
constexpr int kWideWords = 32;

__host__ __device__ inline int32_t wide_row_words(int32_t c0, int32_t c1) { return ((c1 - 1) >> 6) - (c0 >> 6) + 1; }
__host__ __device__ inline int32_t wide_strip_count(int32_t r0, int32_t r1) { return (r1 - 1) / 16 - r0 / 16 + 1; }

// Edge bands: a polygon item with more than kBandMinEdges edges that spans
// two or more 16-row strips gets its edges binned by strip (an edge goes to
// every strip its rows [ystart, yend) meet), so a strip visits only its own
// edges instead of the whole ring set.  Band entry
// band - 1 + (Rw / kBandRows - r0 / kBandRows) of such an item holds the
// edges of the strip starting at row Rw in
// band_edges[band_off[e] .. band_off[e + 1]) (any order: the crossing
// toggles and the maximum commute).
// Line items with more than kLineBandMinSegs segments over two or more
// strips (and no kLsFar segment) are banded the same way: a segment goes to
// every strip the rows of its cells [min(iy0, iy1), max(iy0, iy1)] meet, as
// a 16-byte segment record in band_segs[band_off[e] ..) whose flag marks the
// end cell (any order: a line's writes all carry one value, so only their
// number per pixel matters).
constexpr int kBandMinEdges = 128;
constexpr int kLineBandMinSegs = 32;
constexpr int kBandRows = WAVE_ROWS;

struct InitArgs {
  ItemBox* box;             // [n_items]
  unsigned long long* last; // [n_items]
  unsigned long long* win;  // [8 n_items] or null: reset to two empty f64 boxes
  int64_t n_items;
  static constexpr int kRanges = 7;
  uint32_t* zero[kRanges];  // word ranges cleared (n_zero[r] words from zero[r])
  uint64_t n_zero[kRanges];
  // Adds a range to clear; false when every slot is taken.
  __host__ bool clear(void* p, uint64_t n) {
    for (int r = 0; r < kRanges; ++r)
      if (n_zero[r] == 0) {
        zero[r] = (uint32_t*)p;
        n_zero[r] = n;
        return true;
      }
    return n == 0;
  }
  __host__ uint64_t words() const {
    uint64_t w = 0;
    for (int r = 0; r < kRanges; ++r) w += n_zero[r];
    return w;
  }
};

struct DescArgs {
  const JobDev* jobs;
  const int64_t* job_iend;  // [n_jobs] prefix ends of the item ranges
  int n_jobs;
  int64_t n_items;
  const ItemBox* item_box;
  const unsigned long long* item_last;  // lines: last kept segment + 1; polygons: rows their edges cross
  ItemDesc* desc;           // [n_items]
  uint32_t* counts;         // [n_items] tiles per item, or null (no binned job)
  uint32_t* band_n;         // [n_items] tile rows of a banded polygon item (0: not banded)
  int wide_mode;            // wide pre-pass: 0 chosen per item, 1 never, 2 every polygon item
  uint32_t* wide_units;     // [n_items] k_wide_cover units of a wide item (0: not wide), or null
  uint32_t* wide_words;     // [n_items] row-mask words of a wide item
  uint32_t* status;         // bit 4 set when a geometry's layer index is >= the job's n_layers
  const unsigned long long* item_win;  // all-touched polygons' outline boxes (null: none)
  uint2* scan_box;          // [n_items] boxes of items of scan_compact jobs (or null), see ScanBox
  uint32_t* at_words;       // [n_items] words of the PixelCache an all-touched dedup polygon needs
                            // in global memory (its fill reaches beyond the cache box), or null
};

// k_at_cache: the PixelCache bits of the kItemAtCache items, one wave per
// item, lanes over its outline segments.
struct AtCacheArgs {
  const JobDev* jobs;
  const ItemDesc* desc;     // [n_items]
  int64_t n_items;
  const unsigned long long* item_win;
  const uint64_t* at_base;  // [n_items + 1] exclusive scan of at_words
  unsigned long long* at_bits;
  uint64_t cap_words;       // words of at_bits (an item past them fails the call: bit 8)
  uint32_t* status;         // bit 64: a walk did not end
};

// Wide items within the capacities get their flag, mask base and unit list.
struct WideAssignArgs {
  ItemDesc* desc;
  int64_t n_items;
  const uint32_t* wide_units;
  const uint64_t* unit_base;  // [n_items + 1] exclusive scan of wide_units
  const uint64_t* word_base;  // [n_items + 1] exclusive scan of wide_words
  uint64_t unit_cap, word_cap;
  uint32_t* unit_item;        // [unit_cap] item of every unit
};

struct WideCoverArgs {
  const JobDev* jobs;
  const ItemDesc* desc;       // [n_items], wide fields assigned
  int64_t n_items;
  const uint32_t* unit_item;  // [units]
  const uint64_t* unit_base;  // [n_items + 1]
  const uint64_t* n_units;    // unit total (device side), clamped to unit_cap
  uint64_t unit_cap;
  const PolyEdge* edges;
  const uint64_t* band_off;
  const PolyEdge* band_edges;
  uint64_t* wide_masks;
  int cov_mode;               // layout per chunk of edges: 0 the cheaper, 1 row-major, 2 edge-major
  // heavy tiles (null heavy_cnt: none): their items' masks also go to the
  // heavy slots (see TileArgs), at the item's place in the tile's pair list
  const uint32_t* heavy_cnt;  // [n_tiles] items of a heavy tile, 0 otherwise
  const uint64_t* heavy_base;
  const uint64_t* pair_off;   // [n_items + 1] an item's first pair in emission order
  uint64_t pair_cap;          // pairs emitted at most (an item past it has none)
  const uint32_t* wide_j;     // [pair_cap] by emission index: a wide item's place in a heavy tile
  uint64_t* masks;
  uint32_t slot_words;
};

// Per-item edge-band build: one wave per item (grid-stride), counters of
// the item's strips in wave-private LDS, at most kBandLds strips per pass
// (taller items take several passes over their edges).
constexpr int kBandLds = 1024;
struct BandArgs {
  const ItemDesc* desc;     // [n_items], band fields assigned
  const uint32_t* band_n;   // [n_items] strips of a banded item
  const uint32_t* items;    // [*n_list] the banded items with assigned entries
  const uint32_t* n_list;
  const PolyEdge* edges;    // [n_vertices]
  const uint4* lseg;        // [n_vertices] line segment records
  uint32_t* band_cnt;       // [n_bands] edges per band entry (count pass)
  const uint64_t* band_off; // [n_bands + 1] exclusive scan of band_cnt (scatter pass)
  PolyEdge* band_edges;     // [edge_cap] polygon entries (scatter pass), or null
  uint4* band_segs;         // [edge_cap] line entries (scatter pass), or null
  uint64_t edge_cap;
  uint32_t* status;         // bit 3 set when the banded edges exceed edge_cap
};

struct BinArgs {
  const JobDev* jobs;
  const ItemDesc* desc;
  int64_t n_items;
  const uint64_t* offsets;  // [n_items + 1] exclusive prefix of the tile counts
  uint64_t cap;             // capacity of keys / vals
  uint32_t* keys;           // [cap]
  uint32_t* vals;           // [cap]
  uint32_t* status;         // bit 2 set when the pairs exceed cap
};

// Work queue of the tile kernel: one counter per (XCD, strip) pair, each
// counting the tile blocks handed out from that XCD's eighth of the launch.
constexpr int kQueueStride = 32;  // u32 words between counters (one 128-B line each)
constexpr int kQueueWords = NUM_XCD * 4 * kQueueStride;

// An item box in units of 16-row strips and tiles, both ends inclusive, for
// the box tests of scan jobs: x = first strip | last strip << 16, y = first
// tile column | last tile column << 16 (an empty box: first strip 0xFFFF,
// last 0).  A strip meets the item iff its strip and tile-column indices lie
// in both ranges, which is the pixel-box test (rows [r0, r1), columns
// [c0, c1)) in those units.
__host__ __device__ inline uint2 scan_box_of(int32_t r0, int32_t r1, int32_t c0, int32_t c1, int32_t tw) {
  if (!(r0 < r1 && c0 < c1)) return make_uint2(0xFFFFu, 0u);
  return make_uint2((uint32_t)(r0 / WAVE_ROWS) | ((uint32_t)((r1 - 1) / WAVE_ROWS) << 16),
                    (uint32_t)(c0 / tw) | ((uint32_t)((c1 - 1) / tw) << 16));
}
__host__ __device__ inline bool scan_box_meets(uint2 b, uint32_t strip, uint32_t tx) {
  return (b.x & 0xFFFFu) <= strip && strip <= (b.x >> 16) && (b.y & 0xFFFFu) <= tx && tx <= (b.y >> 16);
}

struct TileArgs {
  const JobDev* jobs;
  const int64_t* job_tend;  // [n_jobs] prefix ends of the tile ranges
  int n_jobs;
  uint32_t t_lo, t_hi;      // tiles of this launch (one pixel-function group)
  uint32_t chunk;           // tiles per queue grab (>= 1)
  uint32_t* queue;          // [kQueueWords], zero at launch
  // binned jobs: pairs sorted by tile, tile_off[t]..tile_off[t+1]
  const uint32_t* tile_off;  // [n_tiles + 1], or null when no job is binned
  const ItemDesc* pair_desc;  // descriptors in pair order
  const uint2* scan_box;      // [n_items] compact boxes of scan_compact jobs' items
  const uint64_t* band_off;   // edge bands of banded polygon and line items
  const PolyEdge* band_edges;
  const uint4* band_segs;
  // heavy tiles: heavy_base[t] .. heavy_base[t + 1] index the tile's items in
  // the pre-pass masks (null: no heavy tiles); slot_words u64 per strip mask,
  // item j of tile t, strip s at slot heavy_base[t] * 4 + s * (items of t) + j
  const uint64_t* heavy_base;
  const uint64_t* masks;
  uint32_t slot_words;
  uint32_t heavy_min;       // items from which a tile is heavy
  const uint64_t* wide_masks;  // row masks of kItemWide items (null: none)
  // heavy tiles of polygon-only groups go to k_tile_heavy (heavy_split = 1):
  // the heavy tiles in tile order, each tile's rank in that list, and the
  // launch's work counter
  int heavy_split;
  const uint32_t* heavy_list;
  const uint32_t* heavy_rank;  // [n_tiles + 1]
  uint32_t* heavy_ctr;
  // scan jobs: every item of the job is tested against the tile
  const ItemDesc* desc;
  const PolyEdge* edges;
  const uint4* lseg;        // line / point vertex records (kLs*)
  // all-touched polygons with sum / count: outline boxes (PixelCache
  // windows) and, for those whose fill reaches beyond them, the PixelCache
  // bits in global memory (at_bits + at_base[item] words)
  const unsigned long long* item_win;
  const uint64_t* at_base;
  const unsigned long long* at_bits;
  uint64_t at_cap;          // words of at_bits
  unsigned long long* burns;  // [kBurnSlots] counters, or null
  uint32_t* status;         // error flags
};

constexpr int kBurnSlots = 1024;

// Items per tile from which a tile's polygon coverage is precomputed by the
// k_cover pre-pass (see rz_tile.hip).
constexpr uint32_t kHeavyItems = 32;

struct CoverArgs {
  const JobDev* jobs;
  const int64_t* job_tend;
  int n_jobs;
  const uint32_t* pair_tile;    // [n_pairs] tile of every pair (sorted keys)
  const ItemDesc* pair_desc;    // [n_pairs]
  const uint64_t* heavy_base;   // [n_tiles + 1]
  const uint32_t* cover_list;   // [*cover_count] pair indices of heavy polygon pairs
  const uint32_t* cover_count;
  uint32_t n_tiles;
  const PolyEdge* edges;
  const uint64_t* band_off;
  const PolyEdge* band_edges;
  const uint32_t* tile_off;     // [n_tiles + 1]
  uint64_t* masks;              // [4 x heavy pairs][slot_words], strip-major per tile
  uint32_t slot_words;
};

// ------------------------------------------------ polygon row coverage --
//
// Pixel c of a row is inside a reference span [a_2k, a_2k+1) (the sorted,
// paired crossings of rust/src/rasterization/burners.rs:302-316) iff the
// number m of crossings whose span column a_i <= c is odd and c < max_i a_i:
// sorted columns are non-decreasing, so m odd puts c in [a_{m-1}, a_m), and
// m == n (c >= max a) is the dropped odd tail of chunks_exact(2).  Crossings
// that f64::total_cmp sorts after +inf (positive NaN) have column 0 but sit at
// the end of the sorted list, where they only pair into empty spans; they are
// therefore ignored.  This needs no sort: per crossing one toggle bit at its
// column, then an inclusive prefix-XOR along the row.

// Span column of a crossing: `(x + 0.5).floor().clamp(0.0, ncols) as usize`
// (burners.rs:310-311); NaN -> 0.
__host__ __device__ inline int32_t span_col(double x, double ncols_f) {
  double f = floor(x + 0.5);
  if (f < 0.0) f = 0.0;
  if (f > ncols_f) f = ncols_f;
  return (f > 0.0) ? (int32_t)f : 0;
}

__host__ __device__ inline bool is_pos_nan(double x) { return x != x && !signbit(x); }

// Adds one crossing column to the toggle state of a TW-wide row window
// starting at C0: `lp` is the parity of crossings left of the window.
__host__ __device__ inline void toggle_crossing(int32_t col, int32_t C0, int32_t TW,
                                                uint64_t* tog, uint32_t& lp, int32_t& amax) {
  amax = col > amax ? col : amax;
  const int32_t rel = col - C0;
  if (rel <= 0) lp ^= 1u;
  else if (rel < TW) tog[rel >> 6] ^= 1ull << (rel & 63);
}

// Branch-free form of toggle_crossing for NWORD-word windows, applied only
// where `on` (inactive edges of a predicated loop contribute nothing).
template <int NWORD>
__host__ __device__ inline void toggle_crossing_if(bool on, int32_t col, int32_t C0,
                                                   uint64_t* tog, uint32_t& lp, int32_t& amax) {
  amax = (on && col > amax) ? col : amax;
  const int32_t rel = col - C0;
  lp ^= (on && rel <= 0) ? 1u : 0u;
  const bool in = on && rel > 0 && rel < NWORD * 64;
  const uint64_t bit = 1ull << (rel & 63);
  for (int w = 0; w < NWORD; ++w) tog[w] ^= (in && (rel >> 6) == w) ? bit : 0ull;
}

// Inclusive prefix-XOR of the bits of one 64-bit word.
__host__ __device__ inline uint64_t prefix_xor64(uint64_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  x ^= x << 32;
  return x;
}

// Coverage bits of columns [C0 + 64*w, C0 + 64*w + 64) of the window from
// the toggles of the whole window (NWORD words): bit = odd prefix parity and
// column < amax.
template <int NWORD>
__host__ __device__ inline uint64_t row_coverage_word(const uint64_t* tog, uint32_t lp,
                                                      int32_t amax, int32_t C0, int w) {
  uint32_t carry = lp;
  uint64_t word = 0;
  for (int k = 0; k < NWORD; ++k) {
    const uint64_t x = prefix_xor64(tog[k]) ^ (carry ? ~0ull : 0ull);
    if (k == w) word = x;
    carry ^= (uint32_t)__builtin_popcountll(tog[k]) & 1u;
  }
  const int32_t lim = amax - (C0 + w * 64);
  return word & (lim <= 0 ? 0ull : (lim >= 64 ? ~0ull : ((1ull << lim) - 1ull)));
}

// floor(a / b) for 0 <= a, 0 < b.  When both fit 52 bits the f64 quotient
// is used: a and b convert exactly and the correctly rounded quotient is
// within 1/2 of a / b (which is below 2^52), so its truncation is the floor
// or one off, and one integer check settles it.  The 64-bit integer
// division it replaces is a long instruction sequence on the GPU.
template <typename I>
__host__ __device__ inline I div_floor_pos(I a, I b) {
  if (a < ((I)1 << 52) && b < ((I)1 << 52)) {
    I q = (I)(int64_t)((double)(int64_t)a / (double)(int64_t)b);
    const I r = a - q * b;
    if (r < 0) --q;
    else if (r >= b) ++q;
    return q;
  }
  return a / b;
}

// Cells of Standard::burn_line for one segment between integer cells
// (rust/src/rasterization/burners.rs:53-84), excluding the end cell, that
// fall in the window [X0,X1) x [Y0,Y1), in loop order.  Closed form of the
// loop: with dx = |ix1-ix0|, D = |iy1-iy0|, an x-major segment (dx >= D)
// visits on iteration k in [0, dx)
//     (ix0 + sx*k, iy0 + sy*floor((2*D*k + dx) / (2*dx)))
// and a y-major one (D > dx) visits on iteration k in [0, D)
//     (ix0 + sx*floor((2*dx*k + D) / (2*D)), iy0 + sy*k).
// (err after i x-steps and j y-steps is dx*(1+j) - D*(1+i); the x-major loop
// steps x every iteration and y exactly when dx*(2j+1) <= 2*D*(k+1), which is
// the recurrence of that floor; symmetrically for y-major.)  The window
// bounds the k range, so a long segment costs only its cells in the tile.
// I must hold 2*max(dx,D)^2: int64_t for |coords| < 2^30, __int128 beyond.
template <typename I, typename F>
__host__ __device__ inline void walk_window(int64_t ix0, int64_t iy0, int64_t ix1, int64_t iy1,
                                            int64_t X0, int64_t X1, int64_t Y0, int64_t Y1,
                                            F&& emit) {
  const int64_t dxs = ix1 - ix0, dys = iy1 - iy0;
  const I dx = (I)(dxs < 0 ? -dxs : dxs), D = (I)(dys < 0 ? -dys : dys);
  const int64_t sx = ix0 < ix1 ? 1 : -1, sy = iy0 < iy1 ? 1 : -1;
  const bool xmaj = dx >= D;
  const I n = xmaj ? dx : D;  // iterations (major steps)
  const I m = xmaj ? D : dx;  // minor steps in total
  if (n == 0) return;
  const int64_t u0 = xmaj ? ix0 : iy0, w0 = xmaj ? iy0 : ix0;
  const int64_t su = xmaj ? sx : sy, sw = xmaj ? sy : sx;
  const int64_t U0 = xmaj ? X0 : Y0, U1 = xmaj ? X1 : Y1;
  const int64_t W0 = xmaj ? Y0 : X0, W1 = xmaj ? Y1 : X1;
  I klo, khi;
  {
    int64_t lo, hi;  // k in [lo, hi] from the major-axis window
    if (su > 0) { lo = U0 - u0; hi = U1 - 1 - u0; }
    else { lo = u0 - (U1 - 1); hi = u0 - U0; }
    if (hi < 0 || (I)lo >= n || lo > hi) return;
    klo = (I)(lo < 0 ? 0 : lo);
    khi = (I)hi >= n ? n - 1 : (I)hi;
  }
  int64_t jlo, jhi;  // minor index window
  if (sw > 0) { jlo = W0 - w0; jhi = W1 - 1 - w0; }
  else { jlo = w0 - (W1 - 1); jhi = w0 - W0; }
  if (jhi < 0 || (I)jlo > m || jlo > jhi) return;
  const I two_n = (I)2 * n, two_m = (I)2 * m;
  if (m == 0) {
    if (jlo > 0) return;
  } else {
    if (jlo > 0) {  // j(k) >= jlo  <=>  k >= ceil((2n*jlo - n) / 2m)
      const I num = two_n * (I)jlo - n;
      const I kmin = div_floor_pos<I>(num + two_m - 1, two_m);
      if (kmin > klo) klo = kmin;
    }
    if ((I)jhi < m) {  // j(k) <= jhi  <=>  k <= ceil((2n*(jhi+1) - n) / 2m) - 1
      const I num = two_n * (I)(jhi + 1) - n;
      const I kmax = div_floor_pos<I>(num + two_m - 1, two_m) - 1;
      if (kmax < khi) khi = kmax;
    }
  }
  if (klo > khi) return;
  // j and its remainder at klo, then stepped: each iteration adds 2m to the
  // numerator and at most one 2n (m <= n).
  I numer = two_m * klo + n;
  I j = div_floor_pos<I>(numer, two_n);
  I rem = numer - j * two_n;
  for (I k = klo; k <= khi; ++k) {
    const int64_t u = u0 + su * (int64_t)k, w = w0 + sw * (int64_t)j;
    if (xmaj) emit(u, w);
    else emit(w, u);
    rem += two_m;
    if (rem >= two_n) {
      rem -= two_n;
      ++j;
    }
  }
}

// floor(a / b) for a >= 0, b > 0: a 32-bit division when both fit (short
// segments), else div_floor_pos.
__host__ __device__ inline int64_t div_pos32(int64_t a, int64_t b) {
  if ((((uint64_t)a | (uint64_t)b) >> 32) == 0) return (int64_t)((uint32_t)a / (uint32_t)b);
  return div_floor_pos<int64_t>(a, b);
}

// The set-up of walk_window32: the first cell in the window as (u, w) along
// the major / minor axes, their steps, the remainder state and the number
// of cells; false when no cell falls in the window.
struct Walk32 {
  int32_t u, w, su, sw, cnt;
  uint32_t rem, tm, gap;
  bool xmaj;
};

__host__ __device__ inline bool walk32_setup(int32_t ix0, int32_t iy0, int32_t ix1, int32_t iy1, int32_t X0,
                                             int32_t X1, int32_t Y0, int32_t Y1, Walk32& s) {
  const int32_t dxs = ix1 - ix0, dys = iy1 - iy0;
  const int64_t dx = dxs < 0 ? -(int64_t)dxs : dxs, D = dys < 0 ? -(int64_t)dys : dys;
  const int32_t sx = ix0 < ix1 ? 1 : -1, sy = iy0 < iy1 ? 1 : -1;
  const bool xmaj = dx >= D;
  const int64_t n = xmaj ? dx : D;  // iterations (major steps)
  const int64_t m = xmaj ? D : dx;  // minor steps in total
  if (n == 0) return false;
  const int32_t u0 = xmaj ? ix0 : iy0, w0 = xmaj ? iy0 : ix0;
  const int32_t su = xmaj ? sx : sy, sw = xmaj ? sy : sx;
  const int64_t U0 = xmaj ? X0 : Y0, U1 = xmaj ? X1 : Y1;
  const int64_t W0 = xmaj ? Y0 : X0, W1 = xmaj ? Y1 : X1;
  int64_t klo, khi;
  {
    int64_t lo, hi;  // k in [lo, hi] from the major-axis window
    if (su > 0) { lo = U0 - u0; hi = U1 - 1 - u0; }
    else { lo = u0 - (U1 - 1); hi = u0 - U0; }
    if (hi < 0 || lo >= n || lo > hi) return false;
    klo = lo < 0 ? 0 : lo;
    khi = hi >= n ? n - 1 : hi;
  }
  int64_t jlo, jhi;  // minor index window
  if (sw > 0) { jlo = W0 - w0; jhi = W1 - 1 - w0; }
  else { jlo = w0 - (W1 - 1); jhi = w0 - W0; }
  if (jhi < 0 || jlo > m || jlo > jhi) return false;
  const int64_t two_n = 2 * n, two_m = 2 * m;
  if (m == 0) {
    if (jlo > 0) return false;
  } else {
    if (jlo > 0) {
      const int64_t kmin = div_pos32(two_n * jlo - n + two_m - 1, two_m);
      if (kmin > klo) klo = kmin;
    }
    if (jhi < m) {
      const int64_t kmax = div_pos32(two_n * (jhi + 1) - n + two_m - 1, two_m) - 1;
      if (kmax < khi) khi = kmax;
    }
  }
  if (klo > khi) return false;
  // (from the first cell: j = floor(n / 2n) = 0, remainder n)
  int64_t j = 0;
  uint32_t rem = (uint32_t)n;
  if (klo > 0) {
    const int64_t numer = two_m * klo + n;
    j = div_pos32(numer, two_n);
    rem = (uint32_t)(numer - j * two_n);
  }
  s.u = u0 + su * (int32_t)klo;
  s.w = w0 + sw * (int32_t)j;
  s.su = su;
  s.sw = sw;
  s.cnt = (int32_t)(khi - klo + 1);
  s.rem = rem;
  s.tm = (uint32_t)two_m;
  s.gap = (uint32_t)(two_n - two_m);
  s.xmaj = xmaj;
  return true;
}

// walk_window for end cells within +-2^30 and a window inside the int32
// range (every tile strip): the same cells in the same order, the k range
// set up in 64 bits and the loop stepped in 32 (its positions stay inside
// the window; the remainder test rem + 2m >= 2n is rem >= 2n - 2m, so the
// remainder, below 2n < 2^32, never overflows).
template <typename F>
__host__ __device__ inline void walk_window32(int32_t ix0, int32_t iy0, int32_t ix1, int32_t iy1,
                                              int32_t X0, int32_t X1, int32_t Y0, int32_t Y1, F&& emit) {
  Walk32 s;
  if (!walk32_setup(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1, s)) return;
  int32_t u = s.u, w = s.w;
  uint32_t rem = s.rem;
  for (int32_t i = 0; i < s.cnt; ++i) {
    if (s.xmaj) emit(u, w);
    else emit(w, u);
    u += s.su;
    if (rem >= s.gap) {
      rem -= s.gap;
      w += s.sw;
    } else {
      rem += s.tm;
    }
  }
}

// walk_window32 giving each cell as its index (y - Y0) * stride + (x - X0)
// in the window (stride >= X1 - X0), stepped along the walk.
template <typename F>
__host__ __device__ inline void walk_window32_idx(int32_t ix0, int32_t iy0, int32_t ix1, int32_t iy1, int32_t X0,
                                                  int32_t X1, int32_t Y0, int32_t Y1, int32_t stride, F&& emit) {
  Walk32 s;
  if (!walk32_setup(ix0, iy0, ix1, iy1, X0, X1, Y0, Y1, s)) return;
  const int32_t x = s.xmaj ? s.u : s.w, y = s.xmaj ? s.w : s.u;
  uint32_t idx = (uint32_t)(y - Y0) * (uint32_t)stride + (uint32_t)(x - X0);
  const uint32_t du = s.xmaj ? (uint32_t)s.su : (uint32_t)(s.su * stride);
  const uint32_t dw = s.xmaj ? (uint32_t)(s.sw * stride) : (uint32_t)s.sw;
  uint32_t rem = s.rem;
#pragma unroll 2  // (2 % faster on long line runs than no unrolling; 4 measured the same as 2)
  for (int32_t i = 0; i < s.cnt; ++i) {
    emit(idx);
    idx += du;
    if (rem >= s.gap) {
      rem -= s.gap;
      idx += dw;
    } else {
      rem += s.tm;
    }
  }
}

// `v as isize` for f64 (saturating, NaN -> 0), for host and device.
__host__ __device__ inline int64_t as_isize(double v) {
  if (v != v) return 0;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  return (int64_t)v;
}

// AllTouchedBase::burn_line for one edge (rust/src/rasterization/
// burners.rs:115-246, the GDAL all-touched walker), in its order, with its
// f64 state stepped exactly as written (every expression compiled without
// contraction): cells inside the raster and inside the window
// [X0, X1) x [Y0, Y1) are emitted.  The walk moves left to right and, when
// sloped, monotonically down (up), so it stops once its column passes X1 or
// its row passes the window in that direction.  Each step reaches the next column or row
// boundary (or moves 1e-9 off one), so a walk takes at most about
// 2 (ncols + nrows) steps; past 4 (ncols + nrows) + 64 it stops and returns
// false (the reference would not end).
template <typename F>
__host__ __device__ inline bool walk_all_touched(double dfx, double dfy, double dfxe, double dfye, int64_t ncols,
                                                 int64_t nrows, int64_t X0, int64_t X1, int64_t Y0, int64_t Y1,
                                                 F&& emit) {
  const double EPSILON_INTERSECT = 1e-4, TOLERANCE = 1e-9;
  const double ncols_f = (double)ncols, nrows_f = (double)nrows;
  double t;
  if (dfx > dfxe) {  // proceed left to right
    t = dfx; dfx = dfxe; dfxe = t;
    t = dfy; dfy = dfye; dfye = t;
  }
  if (fabs(dfx - dfxe) < 0.01) {  // vertical (burners.rs:124-148)
    if (dfye < dfy) { t = dfy; dfy = dfye; dfye = t; }
    const int64_t ix = as_isize(floor(dfxe));
    int64_t iy = as_isize(floor(dfy));
    int64_t iy_end = as_isize(floor(dfye - EPSILON_INTERSECT));
    if (ix < 0 || ix >= ncols || ix < X0 || ix >= X1) return true;
    iy = iy < 0 ? 0 : iy;
    iy_end = iy_end > nrows - 1 ? nrows - 1 : iy_end;
    iy = iy < Y0 ? Y0 : iy;
    iy_end = iy_end > Y1 - 1 ? Y1 - 1 : iy_end;
    for (int64_t y = iy; y <= iy_end; ++y) emit(ix, y);
    return true;
  }
  if (fabs(dfy - dfye) < 0.01) {  // horizontal (burners.rs:150-176)
    if (dfxe < dfx) { t = dfx; dfx = dfxe; dfxe = t; }
    int64_t ix = as_isize(floor(dfx));
    const int64_t iy = as_isize(floor(dfy));
    int64_t ix_end = as_isize(floor(dfxe - EPSILON_INTERSECT));
    if (iy < 0 || iy >= nrows || iy < Y0 || iy >= Y1) return true;
    ix = ix < 0 ? 0 : ix;
    ix_end = ix_end > ncols - 1 ? ncols - 1 : ix_end;
    ix = ix < X0 ? X0 : ix;
    ix_end = ix_end > X1 - 1 ? X1 - 1 : ix_end;
    for (int64_t x = ix; x <= ix_end; ++x) emit(x, iy);
    return true;
  }
  // sloped (burners.rs:178-245)
  const double slope = (dfye - dfy) / (dfxe - dfx);
  const double inv_slope = 1.0 / slope;
  if (dfx < 0.0) {  // clip along x
    dfy += (0.0 - dfx) * slope;
    dfx = 0.0;
  }
  if (dfxe > ncols_f) {
    dfye += (ncols_f - dfxe) * slope;
    dfxe = ncols_f;
  }
  if (dfy < 0.0) {  // clip along y
    dfx += (0.0 - dfy) * inv_slope;
    dfy = 0.0;
  } else if (dfy > nrows_f) {
    dfx += (nrows_f - dfy) * inv_slope;
    dfy = nrows_f;
  }
  if (dfye < 0.0) {
    dfxe += (0.0 - dfye) * inv_slope;
  } else if (dfye > nrows_f) {
    dfxe += (nrows_f - dfye) * inv_slope;
  }
  // f64::clamp (a NaN stays NaN)
  if (dfx < 0.0) dfx = 0.0;
  if (dfx > ncols_f) dfx = ncols_f;
  if (dfxe < 0.0) dfxe = 0.0;
  if (dfxe > ncols_f) dfxe = ncols_f;
  int64_t steps = 4 * (ncols + nrows) + 64;
  while (dfx >= 0.0 && dfx < dfxe) {
    const int64_t ix = as_isize(floor(dfx));
    const int64_t iy = as_isize(floor(dfy));
    if (ix >= X1) return true;  // every later cell lies further right
    if (slope > 0.0 ? iy >= Y1 : iy < Y0) return true;  // ... and further down (up)
    if (ix >= 0 && ix < ncols && iy >= 0 && iy < nrows && ix >= X0 && iy >= Y0 && iy < Y1) emit(ix, iy);
    if (--steps < 0) return false;
    double sx = floor(dfx + 1.0) - dfx;
    double sy = sx * slope;
    if (as_isize(floor(dfy + sy)) == iy) {
      dfx += sx;
      dfy += sy;
    } else if (slope < 0.0) {
      sy = (double)iy - dfy;
      if (sy > -TOLERANCE) sy = -TOLERANCE;
      sx = sy / slope;
      dfx += sx;
      dfy += sy;
    } else {
      sy = (double)(int64_t)((uint64_t)iy + 1u) - dfy;
      if (sy < TOLERANCE) sy = TOLERANCE;
      sx = sy / slope;
      dfx += sx;
      dfy += sy;
    }
  }
  return true;
}

// The outline segment from local vertex lv to the next vertex of its ring
// in pixel coordinates, when extract_line keeps it (edges.rs:112-134: its
// box meets the raster).
template <typename JobRef>
__device__ inline bool kept_segment(const JobRef& J, int64_t lv, double& x0, double& y0, double& x1,
                                    double& y1) {
  if (!(J.vert_info[lv] & VI_HAS_NEXT)) return false;
  auto px_x = [&](double X) { return isnan(X) ? X : (X - J.xmin) / J.xres; };
  auto px_y = [&](double Y) { return isnan(Y) ? Y : (J.ymax - Y) / J.yres; };
  x0 = px_x(J.xy[2 * lv]);
  y0 = px_y(J.xy[2 * lv + 1]);
  x1 = px_x(J.xy[2 * lv + 2]);
  y1 = px_y(J.xy[2 * lv + 3]);
  return fmin(x0, x1) < (double)J.ncols && fmax(x0, x1) >= 0.0 && fmin(y0, y1) < (double)J.nrows &&
         fmax(y0, y1) >= 0.0;
}

}  // namespace rzk
