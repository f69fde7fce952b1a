// rz_sparse.hip -- encoding="sparse": every write of the reference's
// SparseArrayWriter (rust/src/encoding/writers.rs:85-142) as a (row, col,
// value) triplet, in the reference's order, built on the device; and
// SparseArray::build_array (rust/src/encoding/arrays.rs:102-141), the
// triplets folded back into a dense raster.  Included by rz_runtime.hip.
//
// Order of the triplets (rust/src/rasterize.rs:159-196, burners.rs): bands in
// order; in a band, its geometries in order (collection members in order);
// in a polygon item its rows top to bottom and each row's covered columns
// left to right (the sorted, paired crossings of burn_polygon give disjoint
// spans in increasing column order, so a row's writes are its coverage mask
// read left to right); in a line item its kept segments in order, each
// segment's cells in Bresenham order and the end cell last; points in order.
//
// Work units, in that order: a polygon item's 16-row strips (its coverage
// masks come from the wide pre-pass, run for every polygon item over its
// whole box), a line or point item's vertices (one segment or point each),
// and a line of a job with xres != yres as a whole (LineWriter + PixelCache,
// burn_geometry.rs:179-185, 202-208: the first visit of each pixel only, in
// write order, so such a line is walked sequentially).  A count pass sizes
// every unit, an exclusive scan places it, and the emit pass writes it.

namespace rzk {

// Unit of global index u: the position p of its item in band order
// (ubase[p] <= u < ubase[p + 1]).
__device__ inline uint32_t sp_unit_pos(const uint64_t* ubase, uint32_t n_pos, uint64_t u) {
  uint32_t lo = 0, hi = n_pos;  // ubase[lo] <= u < ubase[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ubase[mid] <= u) lo = mid;
    else hi = mid;
  }
  return lo;
}

struct SparseArgs {
  const JobDev* job;           // the call's job record
  const ItemDesc* desc;        // [n_items]
  const uint32_t* perm;        // [n_items] items in band order
  int64_t n_items;
  const uint64_t* ubase;       // [n_items + 1] first unit of each position
  const uint4* lseg;           // line / point vertex records
  const uint64_t* wide_masks;
  uint32_t* counts;            // [units] triplets of each unit (count pass)
  const uint64_t* toff;        // [units + 1] first triplet of each unit (emit pass)
  uint64_t* rows;              // [triplets] (emit pass)
  uint64_t* cols;
  uint8_t* values;             // [triplets] of esz bytes
  int esz;
  int dedup;                   // xres != yres: lines are walked whole
  int at;                      // all-touched: the GDAL walker, polygons burn their outline first
  int at_dedup;                // all-touched sum / count: LineWriter + FillWriter over each item's PixelCache
  const unsigned long long* item_win;  // all-touched dedup polygons' cache windows (by item)
  uint64_t* bits;              // first-visit bits of the whole-walked items
  const uint64_t* bit_base;    // [n_items + 1] bit offsets of those items (by position)
  uint32_t* status;
};

// A line item walked whole (one unit, first visit per pixel): xres != yres,
// or all-touched sum / count (burn_geometry.rs:168-210).
__device__ inline bool sp_line_whole(const SparseArgs& a) { return a.dedup || a.at_dedup; }

// Outline units of a polygon item before its strips: all-touched polygons
// burn their outline first, segment by segment (one unit per vertex), or
// as one unit walked whole through the PixelCache (sum / count).
__device__ inline uint32_t sp_outline_units(const SparseArgs& a, const ItemDesc& d) {
  if (!a.at) return 0u;
  return a.at_dedup ? 1u : (uint32_t)d.nv;
}

// Units of the item at position p: a polygon's strips, a line's or a point
// item's vertices (1 for a line walked whole).
__global__ void k_sp_units(SparseArgs a, uint32_t* n_units, uint64_t* nbits) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_items) return;
  const ItemDesc d = a.desc[a.perm[p]];
  const uint32_t kind = d.kind_layer & 3u;
  const bool boxed = d.r0 < d.r1 && d.c0 < d.c1;
  uint32_t n = 0;
  uint64_t b = 0;
  if (kind == 0) {
    n = boxed && (d.flags & kItemWide) ? (uint32_t)wide_strip_count(d.r0, d.r1) : 0u;
    if (boxed) n += sp_outline_units(a, d);
    CacheWin cw;
    if (boxed && a.at && a.at_dedup && item_cache(a.item_win, (int64_t)d.last, cw)) {
      if (cw.cap > ((uint64_t)1 << 34)) atomicOr(a.status, 32u);  // more than 2 GiB of cache bits
      else b = cw.cap;
    }
  } else if (kind == 1 && sp_line_whole(a)) {
    n = boxed ? 1u : 0u;
    b = boxed ? (uint64_t)(d.r1 - d.r0) * (uint64_t)(d.c1 - d.c0) : 0u;
  } else {
    n = (uint32_t)d.nv;
  }
  n_units[p] = n;
  if (nbits) nbits[p] = (b + 63) / 64;  // whole words per line
}

// The cells of one line vertex's segment inside the raster, in Bresenham
// order, then its end cell when it is the last kept edge of an open line.
template <typename F>
__device__ inline void sp_segment(const SparseArgs& a, const RZ_CONST JobDev& J, const ItemDesc& d, int64_t s,
                                  F&& emit) {
  const uint4 r = a.lseg[s];
  if (r.x == kLsNone) return;
  const bool end = (uint64_t)(s + 1) == d.last && !(r.w & kLsFlag);
  if (r.x != kLsFar) {
    const int32_t ix0 = (int32_t)((int64_t)r.x - kLsBias), iy0 = (int32_t)((int64_t)r.y - kLsBias);
    const int32_t ix1 = (int32_t)((int64_t)r.z - kLsBias), iy1 = (int32_t)((int64_t)(r.w & ~kLsFlag) - kLsBias);
    walk_window32(ix0, iy0, ix1, iy1, 0, J.ncols, 0, J.nrows, emit);
    if (end) emit(ix1, iy1);
  } else {
    const int64_t lv = s - J.v_base;
    auto px_x = [&](double X) { return isnan(X) ? X : (X - J.xmin) / J.xres; };
    auto px_y = [&](double Y) { return isnan(Y) ? Y : (J.ymax - Y) / J.yres; };
    const int64_t ix0 = sat_isize(floor(px_x(J.xy[2 * lv]))), iy0 = sat_isize(floor(px_y(J.xy[2 * lv + 1])));
    const int64_t ix1 = sat_isize(floor(px_x(J.xy[2 * lv + 2]))), iy1 = sat_isize(floor(px_y(J.xy[2 * lv + 3])));
    const int64_t lim = (int64_t)1 << 60;
    if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim && iy1 > -lim && iy1 < lim)
      walk_window<__int128>(ix0, iy0, ix1, iy1, 0, J.ncols, 0, J.nrows, emit);
    else
      atomicOr(a.status, 1u);
    if (end) emit(ix1, iy1);
  }
}

__device__ inline void sp_put(const SparseArgs& a, uint64_t o, int64_t row, int64_t col, uint64_t vbits) {
  a.rows[o] = (uint64_t)row;
  a.cols[o] = (uint64_t)col;
  uint8_t* dst = a.values + o * (uint64_t)a.esz;
  for (int k = 0; k < a.esz; ++k) dst[k] = (uint8_t)(vbits >> (8 * k));
}

// Line and point units, one thread each (EMIT: write, else count).
template <bool EMIT>
__global__ void k_sp_lines(SparseArgs a, uint64_t n_units) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n_units) return;
  const uint32_t p = sp_unit_pos(a.ubase, (uint32_t)a.n_items, u);
  const ItemDesc d = a.desc[a.perm[p]];
  const uint32_t kind = d.kind_layer & 3u;
  if (kind == 1 && sp_line_whole(a)) return;
  const uint64_t local = u - a.ubase[p];
  if (kind == 0 && (a.at_dedup || local >= (uint64_t)sp_outline_units(a, d))) return;  // strips, whole outlines
  const RZ_CONST JobDev& J = *as_const(a.job);
  const int64_t s = d.vbeg + (int64_t)local;
  uint64_t n = 0;
  const uint64_t o = EMIT ? a.toff[u] : 0;
  auto emit = [&](auto cx, auto cy) {
    if (cx < 0 || cx >= J.ncols || cy < 0 || cy >= J.nrows) return;
    if (EMIT) sp_put(a, o + n, (int64_t)cy, (int64_t)cx, d.value);
    ++n;
  };
  if (kind == 2) {  // a point: its cell (extract_point, edges.rs:79-88)
    const uint4 r = a.lseg[s];
    if (r.x != kLsNone) emit((int32_t)r.x, (int32_t)r.y);
  } else if (a.at) {  // an all-touched segment (burners.rs:115-246): a line's, or a polygon's outline
    double x0, y0, x1, y1;
    if (kept_segment(J, s - J.v_base, x0, y0, x1, y1) &&
        !walk_all_touched(x0, y0, x1, y1, J.ncols, J.nrows, 0, J.ncols, 0, J.nrows,
                          [&](int64_t x, int64_t y) { emit(x, y); }))
      atomicOr(a.status, 64u);
  } else {
    sp_segment(a, J, d, s, emit);
  }
  if (!EMIT) a.counts[u] = (uint32_t)n;
}

// Lines walked whole (xres != yres), one thread each: a pixel is written at
// its first visit only (PixelCache over the line's pixel box).  The count
// pass leaves the bits set; the emit pass runs on cleared bits.
// An all-touched sum / count polygon's outline is walked whole the same way
// (LineWriter over its PixelCache, burn_geometry.rs:225-234): a cell is
// written at the first insert of its cache index, and the bits then hold the
// cache its fill is filtered by.
template <bool EMIT>
__global__ void k_sp_dedup(SparseArgs a) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.n_items) return;
  const ItemDesc d = a.desc[a.perm[p]];
  const uint32_t kind = d.kind_layer & 3u;
  if (!(d.r0 < d.r1 && d.c0 < d.c1)) return;
  const bool outline = kind == 0 && a.at && a.at_dedup;
  if (!(kind == 1 && sp_line_whole(a)) && !outline) return;
  const RZ_CONST JobDev& J = *as_const(a.job);
  uint64_t* bits = a.bits + a.bit_base[p];
  const int64_t w = d.c1 - d.c0;
  CacheWin cw{0, 0, 0, 0};
  // (no kept segment: no cache; one past 2^34 bits was refused by k_sp_units)
  if (outline && (!item_cache(a.item_win, (int64_t)d.last, cw) || cw.cap > ((uint64_t)1 << 34))) return;
  const uint64_t u = a.ubase[p];
  uint64_t n = 0;
  const uint64_t o = EMIT ? a.toff[u] : 0;
  auto emit = [&](auto cx, auto cy) {
    if (cx < 0 || cx >= J.ncols || cy < 0 || cy >= J.nrows) return;
    if (!outline && (cy < d.r0 || cy >= d.r1 || cx < d.c0 || cx >= d.c1)) return;  // (cells lie in the box)
    // a line's first visits over its box; an outline's PixelCache index
    // (one past the capacity, which the reference rejects, is skipped)
    const uint64_t k = outline ? cache_index(cw, (int64_t)cx, (int64_t)cy)
                               : (uint64_t)((int64_t)cy - d.r0) * (uint64_t)w + (uint64_t)((int64_t)cx - d.c0);
    if (outline && k >= cw.cap) return;
    const uint64_t bit = 1ull << (k & 63);
    if (bits[k >> 6] & bit) return;
    bits[k >> 6] |= bit;
    if (EMIT) sp_put(a, o + n, (int64_t)cy, (int64_t)cx, d.value);
    ++n;
  };
  for (int64_t s = d.vbeg; s < d.vbeg + d.nv; ++s) {
    if (a.at) {
      double x0, y0, x1, y1;
      if (kept_segment(J, s - J.v_base, x0, y0, x1, y1) &&
          !walk_all_touched(x0, y0, x1, y1, J.ncols, J.nrows, 0, J.ncols, 0, J.nrows,
                            [&](int64_t x, int64_t y) { emit(x, y); }))
        atomicOr(a.status, 64u);
    } else {
      sp_segment(a, J, d, s, emit);
    }
  }
  if (!EMIT) a.counts[u] = (uint32_t)n;
}

// Polygon units, one wave each (grid-stride): the strip's 16 mask rows, each
// read left to right; the wave takes 64 words of a row at a time and places
// each word's bits after the popcounts of the words before it.
template <bool EMIT>
__global__ __launch_bounds__(256) void k_sp_polys(SparseArgs a, uint64_t n_units) {
  const int lane = threadIdx.x & 63;
  const uint64_t wstride = (uint64_t)gridDim.x * 4u;
  for (uint64_t u = (uint64_t)blockIdx.x * 4u + (threadIdx.x >> 6); u < n_units; u += wstride) {
    const uint32_t p = sp_unit_pos(a.ubase, (uint32_t)a.n_items, u);
    const ItemDesc d = load_desc(as_const(a.desc) + a.perm[p]);
    if ((d.kind_layer & 3u) != 0u) continue;
    const uint32_t n_out = sp_outline_units(a, d);
    if (u - a.ubase[p] < (uint64_t)n_out) continue;  // an outline unit
    const int32_t k = (int32_t)(u - a.ubase[p] - n_out);
    const int32_t nw = wide_row_words(d.c0, d.c1);
    // all-touched sum / count: the fill skips pixels the outline's cache
    // contains (FillWriter, writers.rs:38-60): inside the cache window the
    // pixel's own bit, outside it the bit its wrapped index lands on
    CacheWin cw{0, 0, 0, 0};
    const bool filter = a.at && a.at_dedup && item_cache(a.item_win, (int64_t)d.last, cw) &&
                        cw.cap <= ((uint64_t)1 << 34);
    const uint64_t* cbits = filter ? a.bits + a.bit_base[p] : nullptr;
    const uint64_t* m = a.wide_masks + (uint64_t)d.wide * 16u + (uint64_t)k * 16u * (uint64_t)nw;
    const int32_t Rw = ((d.r0 >> 4) + k) * WAVE_ROWS, C = (d.c0 >> 6) * 64;
    uint64_t o = EMIT ? a.toff[u] : 0, n = 0;
    for (int32_t r = 0; r < WAVE_ROWS; ++r) {
      for (int32_t w0 = 0; w0 < nw; w0 += 64) {
        const int32_t w = w0 + lane;
        uint64_t word = w < nw ? m[(uint64_t)r * nw + (uint64_t)w] : 0ull;
        if (filter) {
          for (uint64_t b = word; b; b &= b - 1) {
            const int bit = __builtin_ctzll(b);
            const uint64_t ci = cache_index(cw, (int64_t)(C + w * 64 + bit), (int64_t)(Rw + r));
            if (ci < cw.cap && ((cbits[ci >> 6] >> (ci & 63)) & 1ull)) word &= ~(1ull << bit);
          }
        }
        uint32_t c = (uint32_t)__popcll(word), x = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const uint32_t y = (uint32_t)__shfl_up((int)x, off, 64);
          if (lane >= off) x += y;
        }
        if (EMIT && word) {
          uint64_t q = o + n + (x - c);
          for (uint64_t b = word; b; b &= b - 1) sp_put(a, q++, Rw + r, C + w * 64 + __builtin_ctzll(b), d.value);
        }
        n += (uint32_t)__shfl((int)x, 63, 64);
      }
    }
    if (!EMIT && lane == 0) a.counts[u] = (uint32_t)n;
  }
}

// Sort keys of the items' band order: (band, item).
__global__ void k_band_keys(const ItemDesc* desc, int64_t n_items, uint32_t* keys, uint32_t* vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_items) return;
  keys[i] = desc[i].kind_layer >> 2;
  vals[i] = (uint32_t)i;
}

// First triplet of every band: the unit base of the band's first position.
__global__ void k_sp_bands(const ItemDesc* desc, const uint32_t* perm, int64_t n_items, const uint64_t* ubase,
                           const uint64_t* toff, uint32_t n_layers, uint64_t* band_first) {
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > n_layers) return;
  // first position whose item's band is >= b
  int64_t lo = 0, hi = n_items;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((desc[perm[mid]].kind_layer >> 2) < b) lo = mid + 1;
    else hi = mid;
  }
  band_first[b] = toff[ubase[lo]];
}

// SparseArray::build_array: the triplets of one band sorted by pixel
// (stable, so each pixel's writes stay in order) are folded from the
// background, one thread per pixel run.
template <typename T, int FUN>
__global__ void k_sp_fold(const uint32_t* keys, const uint32_t* idx, uint64_t n, const uint8_t* values,
                          uint64_t bg_bits, T* out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (i > 0 && keys[i - 1] == keys[i])) return;
  const T bg = from_bits<T>(bg_bits);
  T p = bg;
  for (uint64_t j = i; j < n && keys[j] == keys[i]; ++j) {
    T v;
    __builtin_memcpy(&v, values + (uint64_t)idx[j] * sizeof(T), sizeof(T));
    fold<T, FUN>(p, v, bg, 1u);
  }
  out[keys[i]] = p;
}

}  // namespace rzk

struct rz_sparse {
  rz_context* ctx = nullptr;
  uint32_t n_layers = 1;
  uint64_t nrows = 0, ncols = 0;
  rz_dtype dtype = RZ_F64;
  std::vector<uint64_t> band_first;  // [n_layers + 1]
  uint64_t* rows = nullptr;           // device
  uint64_t* cols = nullptr;
  uint8_t* values = nullptr;
};

extern "C" rz_status rz_sparse_free(rz_sparse* sp) {
  if (!sp) return RZ_OK;
  (void)hipSetDevice(sp->ctx->device);
  if (sp->rows) (void)hipFree(sp->rows);
  if (sp->cols) (void)hipFree(sp->cols);
  if (sp->values) (void)hipFree(sp->values);
  delete sp;
  return RZ_OK;
}

extern "C" rz_status rz_sparse_counts(const rz_sparse* sp, uint64_t* per_band) {
  if (!sp || !per_band) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_sparse_counts");
  for (uint32_t b = 0; b < sp->n_layers; ++b) per_band[b] = sp->band_first[b + 1] - sp->band_first[b];
  return RZ_OK;
}

extern "C" rz_status rz_sparse_copy(const rz_sparse* sp, uint64_t* rows, uint64_t* cols, void* values) {
  if (!sp) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_sparse_copy");
  const uint64_t n = sp->band_first[sp->n_layers];
  HIP_TRY(hipSetDevice(sp->ctx->device));
  hipStream_t s = sp->ctx->stream;
  if (n && rows) HIP_TRY(hipMemcpyAsync(rows, sp->rows, n * 8, hipMemcpyDeviceToHost, s));
  if (n && cols) HIP_TRY(hipMemcpyAsync(cols, sp->cols, n * 8, hipMemcpyDeviceToHost, s));
  if (n && values)
    HIP_TRY(hipMemcpyAsync(values, sp->values, n * rz_dtype_size(sp->dtype), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return RZ_OK;
}

namespace {

template <typename T>
void sp_fold_launch(int fun, const uint32_t* keys, const uint32_t* idx, uint64_t n, const uint8_t* values,
                    uint64_t bg, void* out, hipStream_t s) {
  const dim3 g(blocks_for((int64_t)n)), b(256);
  T* o = (T*)out;
  switch (fun) {
    case 0: hipLaunchKernelGGL((k_sp_fold<T, 0>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    case 1: hipLaunchKernelGGL((k_sp_fold<T, 1>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    case 2: hipLaunchKernelGGL((k_sp_fold<T, 2>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    case 3: hipLaunchKernelGGL((k_sp_fold<T, 3>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    case 4: hipLaunchKernelGGL((k_sp_fold<T, 4>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    case 5: hipLaunchKernelGGL((k_sp_fold<T, 5>), g, b, 0, s, keys, idx, n, values, bg, o); break;
    default: hipLaunchKernelGGL((k_sp_fold<T, 6>), g, b, 0, s, keys, idx, n, values, bg, o); break;
  }
}

// Pixel key of every triplet of a band, and its index (the sort's values).
__global__ void k_sp_keys(const uint64_t* rows, const uint64_t* cols, uint64_t n, uint64_t ncols, uint32_t* keys,
                          uint32_t* idx) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  keys[i] = (uint32_t)(rows[i] * ncols + cols[i]);
  idx[i] = (uint32_t)i;
}

// Fills n elements of esz bytes with the background's bits.
__global__ void k_fill_bits(uint8_t* out, uint64_t n, int esz, uint64_t bits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (esz) {
    case 1: out[i] = (uint8_t)bits; break;
    case 2: ((uint16_t*)out)[i] = (uint16_t)bits; break;
    case 4: ((uint32_t*)out)[i] = (uint32_t)bits; break;
    default: ((uint64_t*)out)[i] = bits; break;
  }
}

}  // namespace

extern "C" rz_status rz_sparse_to_dense(const rz_sparse* sp, rz_fun fun, const void* background, void* out) {
  if (!sp || !background || !out) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_sparse_to_dense");
  if ((unsigned)fun > 6) return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
  const uint64_t npix = sp->nrows * sp->ncols;
  if (npix >= ((uint64_t)1 << 32))
    return rz::fail(RZ_CAPACITY_ERROR, "a band of 2^32 pixels or more cannot be materialised");
  rz_context* c = sp->ctx;
  HIP_TRY(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int esz = (int)rz_dtype_size(sp->dtype);
  uint64_t bg = 0;
  std::memcpy(&bg, background, (size_t)esz);
  const uint64_t total_px = npix * sp->n_layers;
  void* d_out = nullptr;
  HIP_TRY(hipMalloc(&d_out, std::max<uint64_t>(total_px * esz, 1)));
  hipError_t e = hipSuccess;
  rz_status st = RZ_OK;
  if (total_px)
    hipLaunchKernelGGL(k_fill_bits, dim3(blocks_for((int64_t)total_px)), dim3(256), 0, s, (uint8_t*)d_out,
                       total_px, esz, bg);
  for (uint32_t b = 0; b < sp->n_layers && e == hipSuccess && st == RZ_OK; ++b) {
    const uint64_t t0 = sp->band_first[b], n = sp->band_first[b + 1] - t0;
    if (n == 0) continue;
    if (n >= ((uint64_t)1 << 32)) {
      st = rz::fail(RZ_CAPACITY_ERROR, "a band of 2^32 triplets or more cannot be materialised");
      break;
    }
    e = ensure(s, c->keys_a, n * 4);
    if (e == hipSuccess) e = ensure(s, c->vals_a, n * 4);
    if (e == hipSuccess) e = ensure(s, c->keys_b, n * 4);
    if (e == hipSuccess) e = ensure(s, c->vals_b, n * 4);
    const uint64_t nbk = sort_blocks(n);
    if (e == hipSuccess) e = ensure(s, c->sort_hist, nbk * kRadix * 4);
    if (e == hipSuccess) e = ensure(s, c->sort_off, nbk * kRadix * 4);
    if (e == hipSuccess) e = ensure(s, c->sort_tmp, scan_blocks(nbk * kRadix) * 4);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(k_sp_keys, dim3(blocks_for((int64_t)n)), dim3(256), 0, s, sp->rows + t0, sp->cols + t0, n,
                       sp->ncols, (uint32_t*)c->keys_a.p, (uint32_t*)c->vals_a.p);
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)1 << end_bit) <= npix) ++end_bit;
    const bool in_b = radix_sort_pairs((uint32_t*)c->keys_a.p, (uint32_t*)c->vals_a.p, (uint32_t*)c->keys_b.p,
                                       (uint32_t*)c->vals_b.p, n, end_bit, (uint32_t*)c->sort_hist.p,
                                       (uint32_t*)c->sort_off.p, (uint32_t*)c->sort_tmp.p, s);
    const uint32_t* keys = (const uint32_t*)(in_b ? c->keys_b.p : c->keys_a.p);
    const uint32_t* idx = (const uint32_t*)(in_b ? c->vals_b.p : c->vals_a.p);
    const uint8_t* vals = sp->values + t0 * (uint64_t)esz;
    void* band = (uint8_t*)d_out + (uint64_t)b * npix * esz;
    switch (sp->dtype) {
      case RZ_U8: sp_fold_launch<uint8_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_U16: sp_fold_launch<uint16_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_U32: sp_fold_launch<uint32_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_U64: sp_fold_launch<uint64_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_I8: sp_fold_launch<int8_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_I16: sp_fold_launch<int16_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_I32: sp_fold_launch<int32_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_I64: sp_fold_launch<int64_t>((int)fun, keys, idx, n, vals, bg, band, s); break;
      case RZ_F32: sp_fold_launch<float>((int)fun, keys, idx, n, vals, bg, band, s); break;
      default: sp_fold_launch<double>((int)fun, keys, idx, n, vals, bg, band, s); break;
    }
  }
  if (e == hipSuccess && st == RZ_OK) e = hipGetLastError();
  if (e == hipSuccess && st == RZ_OK) e = hipMemcpyAsync(out, d_out, total_px * esz, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_out);
  if (e != hipSuccess) return rz::fail(RZ_DEVICE_ERROR, std::string("HIP error: ") + hipGetErrorString(e));
  return st;
}

// The sparse triplets of one call: the call's set-up and coverage masks
// (run_batch, masks only), then units counted, placed and emitted.
static rz_status sparse_build(rz_context* c, const rz_job& j, rz_sparse* sp) {
  rz_status st = run_batch(c, &j, 1, nullptr, true);
  if (st) return st;
  const rz_device_geometry* dg = j.geometry;
  const int64_t I = dg->I;
  hipStream_t s = c->stream;
  const uint32_t nl = sp->n_layers;
  sp->band_first.assign((size_t)nl + 1, 0);
  if (I == 0) return RZ_OK;
  // 1. items in band order: a stable sort of (band, item)
  HIP_TRY(ensure(s, c->keys_a, (size_t)I * 4));
  HIP_TRY(ensure(s, c->vals_a, (size_t)I * 4));
  HIP_TRY(ensure(s, c->keys_b, (size_t)I * 4));
  HIP_TRY(ensure(s, c->vals_b, (size_t)I * 4));
  const uint64_t nbk = sort_blocks((uint64_t)I);
  HIP_TRY(ensure(s, c->sort_hist, nbk * kRadix * 4));
  HIP_TRY(ensure(s, c->sort_off, nbk * kRadix * 4));
  HIP_TRY(ensure(s, c->sort_tmp, scan_blocks(nbk * kRadix) * 4));
  hipLaunchKernelGGL(k_band_keys, dim3(blocks_for(I)), dim3(256), 0, s, (const ItemDesc*)c->desc.p, I,
                     (uint32_t*)c->keys_a.p, (uint32_t*)c->vals_a.p);
  int end_bit = 1;
  while (end_bit < 32 && ((uint64_t)1 << end_bit) <= (uint64_t)nl) ++end_bit;
  const bool in_b = radix_sort_pairs((uint32_t*)c->keys_a.p, (uint32_t*)c->vals_a.p, (uint32_t*)c->keys_b.p,
                                     (uint32_t*)c->vals_b.p, (uint64_t)I, end_bit, (uint32_t*)c->sort_hist.p,
                                     (uint32_t*)c->sort_off.p, (uint32_t*)c->sort_tmp.p, s);
  const uint32_t* perm = (const uint32_t*)(in_b ? c->vals_b.p : c->vals_a.p);
  // 2. units of each position, placed by a scan (and the dedup lines' bits)
  SparseArgs a{};
  std::memset(&a, 0, sizeof(a));
  a.job = (const JobDev*)c->jobtab.p;
  a.desc = (const ItemDesc*)c->desc.p;
  a.perm = perm;
  a.n_items = I;
  a.lseg = (const uint4*)c->lseg.p;
  a.wide_masks = (const uint64_t*)c->wide_masks.p;
  a.esz = (int)rz_dtype_size(j.dtype);
  a.dedup = j.info.xres != j.info.yres;
  a.at = j.all_touched ? 1 : 0;
  a.at_dedup = j.all_touched && (j.fun == RZ_SUM || j.fun == RZ_COUNT);  // prelude.rs:116-118
  a.item_win = (a.at_dedup && dg->has_poly) ? (const unsigned long long*)c->item_win.p : nullptr;
  a.status = (uint32_t*)c->misc.p;
  const bool whole = a.dedup || a.at_dedup;  // items walked whole, with first-visit bits
  HIP_TRY(ensure(s, c->counts, (size_t)(I + 1) * 4));
  HIP_TRY(ensure(s, c->offsets, (size_t)(I + 1) * 8));
  HIP_TRY(ensure(s, c->band_base, (size_t)(I + 1) * 8));  // dedup bits per position
  HIP_TRY(ensure(s, c->heavy_base, (size_t)(I + 1) * 8)); // their scan
  HIP_TRY(ensure(s, c->scan_tmp, scan_blocks((uint64_t)I + 1) * 8));
  InitArgs ia{};
  std::memset(&ia, 0, sizeof(ia));
  ia.clear((uint32_t*)c->counts.p + I, 1);
  ia.clear((uint64_t*)c->band_base.p + I, 2);
  hipLaunchKernelGGL(k_init, dim3(1), dim3(256), 0, s, ia);
  hipLaunchKernelGGL(k_sp_units, dim3(blocks_for(I)), dim3(256), 0, s, a, (uint32_t*)c->counts.p,
                     whole ? (uint64_t*)c->band_base.p : nullptr);
  exclusive_scan<uint32_t, uint64_t>((const uint32_t*)c->counts.p, (uint64_t*)c->offsets.p, (uint64_t)I + 1,
                                     (uint64_t*)c->scan_tmp.p, s);
  a.ubase = (const uint64_t*)c->offsets.p;
  uint64_t n_bits_words = 0;
  if (whole) {
    exclusive_scan<uint64_t, uint64_t>((const uint64_t*)c->band_base.p, (uint64_t*)c->heavy_base.p,
                                       (uint64_t)I + 1, (uint64_t*)c->scan_tmp.p, s);
    HIP_TRY(hipMemcpyAsync(&c->host_misc[1], (const uint64_t*)c->heavy_base.p + I, 8, hipMemcpyDeviceToHost, s));
  }
  HIP_TRY(hipMemcpyAsync(&c->host_misc[0], (const uint64_t*)c->offsets.p + I, 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const uint64_t U = c->host_misc[0];
  if (whole) n_bits_words = c->host_misc[1];
  if (U >= ((uint64_t)1 << 32)) return rz::fail(RZ_CAPACITY_ERROR, "more than 2^32 sparse work units");
  // 3. triplets per unit, then their places
  HIP_TRY(ensure(s, c->cover_list, (size_t)(U + 1) * 4));   // counts per unit
  HIP_TRY(ensure(s, c->heavy_list, (size_t)(U + 1) * 8));   // their scan
  HIP_TRY(ensure(s, c->scan_tmp2, scan_blocks(U + 1) * 8));
  a.counts = (uint32_t*)c->cover_list.p;
  if (whole) {
    HIP_TRY(ensure(s, c->wide_item, std::max<uint64_t>(2 * n_bits_words, 1) * 8));  // count + emit halves
    a.bits = (uint64_t*)c->wide_item.p;
    a.bit_base = (const uint64_t*)c->heavy_base.p;
  }
  InitArgs ib{};
  std::memset(&ib, 0, sizeof(ib));
  ib.clear(a.counts, U + 1);
  if (whole) ib.clear(a.bits, 4 * n_bits_words);  // (32-bit words)
  hipLaunchKernelGGL(k_init, dim3(blocks_for((int64_t)std::max<uint64_t>(ib.words(), 1))), dim3(256), 0, s, ib);
  // (the whole walks first: an all-touched outline's bits filter its fill)
  if (U > 0) {
    if (whole) hipLaunchKernelGGL(k_sp_dedup<false>, dim3(blocks_for(I)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_sp_lines<false>, dim3(blocks_for((int64_t)U)), dim3(256), 0, s, a, U);
    hipLaunchKernelGGL(k_sp_polys<false>, dim3((unsigned)std::min<uint64_t>((U + 3) / 4, 4096)), dim3(256), 0, s,
                       a, U);
  }
  exclusive_scan<uint32_t, uint64_t>(a.counts, (uint64_t*)c->heavy_list.p, U + 1, (uint64_t*)c->scan_tmp2.p, s);
  a.toff = (const uint64_t*)c->heavy_list.p;
  HIP_TRY(ensure(s, c->tile_off, ((size_t)nl + 1) * 8));
  hipLaunchKernelGGL(k_sp_bands, dim3(blocks_for((int64_t)nl + 1)), dim3(256), 0, s, a.desc, perm, I, a.ubase,
                     a.toff, nl, (uint64_t*)c->tile_off.p);
  HIP_TRY(hipMemcpyAsync(sp->band_first.data(), c->tile_off.p, ((size_t)nl + 1) * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  HIP_TRY(hipGetLastError());
  const uint64_t T = sp->band_first[nl];
  // 4. the triplets
  HIP_TRY(hipMalloc((void**)&sp->rows, std::max<uint64_t>(T, 1) * 8));
  HIP_TRY(hipMalloc((void**)&sp->cols, std::max<uint64_t>(T, 1) * 8));
  HIP_TRY(hipMalloc((void**)&sp->values, std::max<uint64_t>(T, 1) * (uint64_t)a.esz));
  a.rows = sp->rows;
  a.cols = sp->cols;
  a.values = sp->values;
  if (whole) a.bits += n_bits_words;  // the second (clear) half
  if (U > 0 && T > 0) {
    if (whole) hipLaunchKernelGGL(k_sp_dedup<true>, dim3(blocks_for(I)), dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_sp_lines<true>, dim3(blocks_for((int64_t)U)), dim3(256), 0, s, a, U);
    hipLaunchKernelGGL(k_sp_polys<true>, dim3((unsigned)std::min<uint64_t>((U + 3) / 4, 4096)), dim3(256), 0, s,
                       a, U);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(s));
  return check_status_word(c);
}

extern "C" rz_status rz_rasterize_sparse(const rz_geometry* geoms, const rz_raster_info* info, rz_dtype dtype,
                                         rz_fun fun, const void* field, const void* scalar,
                                         const uint32_t* layer_of_geom, uint32_t n_layers, int all_touched,
                                         rz_sparse** out) {
  if (!geoms || !info || !out) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_rasterize_sparse");
  *out = nullptr;
  if ((unsigned)dtype > 9) return rz::fail(RZ_TYPE_ERROR, "Invalid dtype or encoding provided.");
  if (!field && !scalar) return rz::fail(RZ_VALUE_ERROR, "no field and no scalar value");
  if ((unsigned)fun > 6) return rz::fail(RZ_VALUE_ERROR, "Unknown pixel function");
  if (!layer_of_geom) n_layers = 1;
  if (n_layers == 0) return rz::fail(RZ_VALUE_ERROR, "n_layers must be at least 1");
  for (int64_t i = 0; layer_of_geom && i < geoms->n_geoms; ++i)
    if (layer_of_geom[i] >= n_layers) return rz::fail(RZ_VALUE_ERROR, "layer index out of range");
  std::lock_guard<std::mutex> lock(g_default_mu);
  rz_context* c = nullptr;
  rz_status st = default_context(&c);
  if (st) return st;
  HIP_TRY(hipSetDevice(c->device));
  auto* sp = new (std::nothrow) rz_sparse();
  if (!sp) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  sp->ctx = c;
  sp->n_layers = n_layers;
  sp->nrows = info->nrows;
  sp->ncols = info->ncols;
  sp->dtype = dtype;
  const size_t esz = rz_dtype_size(dtype);
  rz_device_geometry* dg = nullptr;
  st = rz_device_geometry_upload(c, geoms, &dg);
  if (st) {
    delete sp;
    return st;
  }
  void *d_field = nullptr, *d_layer = nullptr;
  hipError_t e = hipSuccess;
  if (field) e = upload(c, (uint8_t**)&d_field, (const uint8_t*)field, (int64_t)(geoms->n_geoms * esz));
  if (e == hipSuccess && layer_of_geom) e = upload(c, (uint32_t**)&d_layer, layer_of_geom, geoms->n_geoms);
  if (e == hipSuccess && info->nrows > 0 && info->ncols > 0) {
    rz_job j;
    std::memset(&j, 0, sizeof(j));
    j.geometry = dg;
    j.info = *info;
    j.dtype = dtype;
    j.fun = fun;  // (decides the all-touched dedup, prelude.rs:116-118)
    j.all_touched = all_touched ? 1 : 0;
    j.field_dev = d_field;
    if (!field) std::memcpy(&j.scalar_bits, scalar, esz);
    j.layer_dev = (const uint32_t*)d_layer;
    j.n_layers = layer_of_geom ? n_layers : 1u;
    j.out_dev = nullptr;
    st = sparse_build(c, j, sp);
  } else if (e == hipSuccess) {
    sp->band_first.assign((size_t)n_layers + 1, 0);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_field) (void)hipFree(d_field);
  if (d_layer) (void)hipFree(d_layer);
  rz_device_geometry_free(dg);
  if (e != hipSuccess) st = rz::fail(RZ_DEVICE_ERROR, std::string("HIP error: ") + hipGetErrorString(e));
  if (st) {
    rz_sparse_free(sp);
    return st;
  }
  *out = sp;
  return RZ_OK;
}
