// rz_kernels.hip -- gfx950 kernels: device-side edge construction, binning
// and the tile scanline / fold kernel.  See rz_kernels.h for the pipeline.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include "rz_kernels.h"

namespace rzk {

// ------------------------------------------------------ Rust scalar rules --

// `v as usize` (saturating, NaN -> 0), then clamped to `cap`.
__device__ inline int32_t sat_usize_cap(double v, int32_t cap) {
  if (!(v > 0.0)) return 0;
  if (v >= (double)cap) return cap;
  return (int32_t)v;
}
// `v as isize`: saturating, NaN -> 0.
__device__ inline int64_t sat_isize(double v) {
  if (v != v) return 0;
  if (v <= -9223372036854775808.0) return INT64_MIN;
  if (v >= 9223372036854775808.0) return INT64_MAX;
  return (int64_t)v;
}
// PolyEdge::intersect_at (rust/src/geo/edges.rs:50-55): x0 + ((r+0.5) - y0) * dxdy.
// A NaN produced by an invalid operation (0 * inf, inf - inf) is given the
// sign x86 SSE gives its default NaN (negative), so total_cmp places it like
// the reference does; propagated NaNs keep their operand's sign.
__device__ inline double crossing(const PolyEdge& e, int32_t r) {
  double t = ((double)r + 0.5) - e.y0;
  double m = t * e.dxdy;
  if (isnan(m) && !isnan(t) && !isnan(e.dxdy)) m = -__builtin_nan("");
  double x = e.x0 + m;
  if (isnan(x) && !isnan(e.x0) && !isnan(m)) x = -__builtin_nan("");
  return x;
}

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

__global__ void k_vert_ring(const int64_t* ring_voff, int64_t n_rings, int32_t* vert_ring) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rings) return;
  for (int64_t v = ring_voff[r]; v < ring_voff[r + 1]; ++v) vert_ring[v] = (int32_t)r;
}

__global__ void k_init_items(ItemBox* box, unsigned long long* last, int64_t n_items) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_items) return;
  box[i] = ItemBox{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  last[i] = 0ull;
}

// ------------------------------------------------------------------ prep --

// Segmented (by item) min/max reduction across the wave, then one set of
// atomics per item segment.  Items own contiguous vertex ranges, so lanes of
// one item are contiguous.
__device__ inline void reduce_box(int32_t item, ItemBox b, ItemBox* out, bool active) {
  const int lane = threadIdx.x & 63;
  int32_t it = active ? item : -1;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int32_t oit = __shfl_up(it, off, 64);
    int32_t r0 = __shfl_up(b.r0, off, 64), r1 = __shfl_up(b.r1, off, 64);
    int32_t c0 = __shfl_up(b.c0, off, 64), c1 = __shfl_up(b.c1, off, 64);
    if (lane >= off && oit == it) {
      b.r0 = min(b.r0, r0);
      b.r1 = max(b.r1, r1);
      b.c0 = min(b.c0, c0);
      b.c1 = max(b.c1, c1);
    }
  }
  int32_t nit = __shfl_down(it, 1, 64);
  bool tail = (lane == 63) || (nit != it);
  if (active && tail && b.r0 < b.r1 && b.c0 < b.c1) {
    atomicMin(&out[item].r0, b.r0);
    atomicMax(&out[item].r1, b.r1);
    atomicMin(&out[item].c0, b.c0);
    atomicMax(&out[item].c1, b.c1);
  }
}

__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = v < a.n_vertices;
  int32_t item = -1;
  ItemBox box{INT_MAX, INT_MIN, INT_MAX, INT_MIN};
  if (active) {
    const int32_t ring = a.vert_ring[v];
    item = a.ring_item[ring];
    const int32_t kind = a.item_kind[item];
    const int64_t rbeg = a.ring_voff[ring], rend = a.ring_voff[ring + 1];
    const bool has_next = v + 1 < rend;
    const double rows = (double)a.nrows, cols = (double)a.ncols;
    // world -> pixel, rust/src/geo/edges.rs:94-97
    const double x0 = (a.xy[2 * v] - a.xmin) / a.xres;
    const double y0 = (a.ymax - a.xy[2 * v + 1]) / a.yres;
    if (kind == 0) {  // polygon rings: extract_ring + PolyEdge::new
      PolyEdge e{0.0, 0.0, 0.0, 0, 0};
      if (has_next) {
        const double x1 = (a.xy[2 * v + 2] - a.xmin) / a.xres;
        const double y1 = (a.ymax - a.xy[2 * v + 3]) / a.yres;
        if (fabs(y0 - y1) >= 2.220446049250313e-16) {  // edges.rs:100
          const double min_y = fmin(y0, y1), max_y = fmax(y0, y1);
          if (min_y < rows && max_y >= 0.0) {  // edges.rs:105
            double xt, yt, xb, yb;
            if (y0 < y1) { xt = x0; yt = y0; xb = x1; yb = y1; }
            else { xt = x1; yt = y1; xb = x0; yb = y0; }
            e.ystart = sat_usize_cap(ceil(yt - 0.5), a.nrows);
            e.yend = sat_usize_cap(ceil(yb - 0.5), a.nrows);
            e.dxdy = (xb - xt) / (yb - yt);
            e.x0 = xt;
            e.y0 = yt;
            if (e.ystart < e.yend) {
              box.r0 = e.ystart;
              box.r1 = e.yend;
              // Crossings of this edge lie between its end points up to a few
              // ulps, so their span columns are within one pixel of these.
              const double ax = fabs(xt), bx = fabs(xb);
              if (isnan(xt) || isnan(xb) || ax > 1.0e12 || bx > 1.0e12) {
                box.c0 = 0;
                box.c1 = a.ncols;
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
    } else if (kind == 1) {  // line rings: extract_line
      a.px[v] = make_double2(x0, y0);
      uint8_t f = has_next ? VF_HAS_NEXT : 0;
      // geo-types LineString::is_closed: first coordinate == last coordinate
      if (a.xy[2 * rbeg] == a.xy[2 * (rend - 1)] && a.xy[2 * rbeg + 1] == a.xy[2 * (rend - 1) + 1])
        f |= VF_CLOSED;
      a.vflags[v] = f;
      if (has_next) {
        const double x1 = (a.xy[2 * v + 2] - a.xmin) / a.xres;
        const double y1 = (a.ymax - a.xy[2 * v + 3]) / a.yres;
        const double mnx = fmin(x0, x1), mxx = fmax(x0, x1);
        const double mny = fmin(y0, y1), mxy = fmax(y0, y1);
        if (mnx < cols && mxx >= 0.0 && mny < rows && mxy >= 0.0) {  // edges.rs:130
          atomicMax(&a.item_last[item], (unsigned long long)(v + 1));
          // Bresenham visits cells between the floored end points only.
          const int64_t fx0 = sat_isize(floor(mnx)), fx1 = sat_isize(floor(mxx));
          const int64_t fy0 = sat_isize(floor(mny)), fy1 = sat_isize(floor(mxy));
          box.c0 = (int32_t)max((int64_t)0, min(fx0, (int64_t)a.ncols));
          box.c1 = (int32_t)max((int64_t)0, min(fx1 + 1, (int64_t)a.ncols));
          box.r0 = (int32_t)max((int64_t)0, min(fy0, (int64_t)a.nrows));
          box.r1 = (int32_t)max((int64_t)0, min(fy1 + 1, (int64_t)a.nrows));
        }
      }
    } else {  // points: extract_point
      a.px[v] = make_double2(x0, y0);
      a.vflags[v] = 0;
      if (x0 >= 0.0 && x0 < cols && y0 >= 0.0 && y0 < rows) {  // edges.rs:85
        const int32_t cx = (int32_t)x0, cy = (int32_t)y0;
        box = ItemBox{cy, cy + 1, cx, cx + 1};
      }
    }
  }
  reduce_box(item, box, a.item_box, active);
}

// --------------------------------------------------------------- binning --

__device__ inline bool tile_rect(const ItemBox& b, int32_t tw, int32_t& tx0, int32_t& tx1,
                                 int32_t& ty0, int32_t& ty1) {
  if (!(b.r0 < b.r1 && b.c0 < b.c1)) return false;
  tx0 = b.c0 / tw;
  tx1 = (b.c1 - 1) / tw;
  ty0 = b.r0 / TILE_H;
  ty1 = (b.r1 - 1) / TILE_H;
  return true;
}

__global__ void k_bin_count(BinArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  int32_t tx0, tx1, ty0, ty1;
  uint32_t c = 0;
  if (tile_rect(a.item_box[i], a.tile_w, tx0, tx1, ty0, ty1))
    c = (uint32_t)(tx1 - tx0 + 1) * (uint32_t)(ty1 - ty0 + 1);
  a.counts[i] = c;
}

__global__ void k_bin_emit(BinArgs a) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_items) return;
  int32_t tx0, tx1, ty0, ty1;
  if (!tile_rect(a.item_box[i], a.tile_w, tx0, tx1, ty0, ty1)) return;
  uint32_t layer = a.layer_of_geom ? a.layer_of_geom[a.item_geom[i]] : 0u;
  uint64_t o = a.offsets[i];
  uint32_t base = layer * a.tiles_per_layer;
  for (int32_t ty = ty0; ty <= ty1; ++ty)
    for (int32_t tx = tx0; tx <= tx1; ++tx) {
      a.keys[o] = base + (uint32_t)ty * (uint32_t)a.ntx + (uint32_t)tx;
      a.vals[o] = (uint32_t)i;
      ++o;
    }
}

// tile_off[t] = first pair index whose key >= t, for t in [0, n_tiles].
__global__ void k_tile_offsets(const uint32_t* keys, uint32_t n_pairs, uint32_t n_tiles,
                               uint32_t* tile_off) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  uint32_t lo = 0, hi = n_pairs;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < (uint32_t)t) lo = mid + 1;
    else hi = mid;
  }
  tile_off[t] = lo;
}

// ----------------------------------------------------------- pixel folds --

template <typename T> __device__ inline bool isnan_t(T) { return false; }
template <> __device__ inline bool isnan_t<float>(float v) { return v != v; }
template <> __device__ inline bool isnan_t<double>(double v) { return v != v; }

template <typename T> struct Unsigned { using type = T; };
template <> struct Unsigned<int8_t> { using type = uint8_t; };
template <> struct Unsigned<int16_t> { using type = uint16_t; };
template <> struct Unsigned<int32_t> { using type = uint32_t; };
template <> struct Unsigned<int64_t> { using type = uint64_t; };

// Integer `+=` wraps (release profile); floats add in their own precision.
template <typename T> __device__ inline T add_wrap(T a, T b) {
  using U = typename Unsigned<T>::type;
  return (T)(U)((U)a + (U)b);
}
template <> __device__ inline float add_wrap<float>(float a, float b) { return a + b; }
template <> __device__ inline double add_wrap<double>(double a, double b) { return a + b; }

// The seven transitions of rust/src/rasterization/pixel_functions.rs:55-123,
// applied k >= 1 times with the same value (all writes of one geometry
// carry one value, so only their number matters; sum and count are the
// only non-idempotent transitions).
template <typename T, int FUN>
__device__ inline void fold(T& p, T v, T bg, uint32_t k) {
  if (FUN == 0) {  // sum
    for (uint32_t i = 0; i < k; ++i)
      p = (p == bg || isnan_t(p) || isnan_t(v)) ? v : add_wrap(p, v);
  } else if (FUN == 1) {  // first
    if (p == bg || isnan_t(p)) p = v;
  } else if (FUN == 2) {  // last
    p = v;
  } else if (FUN == 3) {  // min
    if (p == bg || isnan_t(p) || p > v) p = v;
  } else if (FUN == 4) {  // max
    if (p == bg || isnan_t(p) || p < v) p = v;
  } else if (FUN == 5) {  // count
    for (uint32_t i = 0; i < k; ++i) p = (p == bg || isnan_t(p)) ? (T)1 : add_wrap(p, (T)1);
  } else {  // any
    p = (T)1;
  }
}

template <typename T> __device__ inline T from_bits(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}

// ------------------------------------------------------------- tile kernel --

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint32_t xcd_remap(uint32_t b, uint32_t n) {
  // Workgroup b runs on XCD b % 8; give each XCD a contiguous run of tiles so
  // neighbouring tiles (shared cache lines, shared items) meet in one L2.
  uint32_t q = n / NUM_XCD, r = n % NUM_XCD;
  uint32_t x = b % NUM_XCD, i = b / NUM_XCD;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <typename T, int FUN>
__global__ __launch_bounds__(256) void k_tile(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int PPT = TW / 4;      // pixels per lane (one quarter of a tile row)
  constexpr int NWORD = TW / 64;   // u64 words of a row mask
  constexpr int CNT_WORDS = WAVE_ROWS * TW / 2;  // u16 line counters, 2 per word
  __shared__ uint32_t s_cnt[4][CNT_WORDS];

  const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t layer = tile / a.tiles_per_layer;
  const uint32_t tl = tile - layer * a.tiles_per_layer;
  const int32_t ty = (int32_t)(tl / (uint32_t)a.ntx), tx = (int32_t)(tl % (uint32_t)a.ntx);
  const int32_t R0 = ty * TILE_H, C0 = tx * TW;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;          // row in the wave strip, quarter
  const int32_t Rw = R0 + wave * WAVE_ROWS;         // first row of this wave's strip
  const int32_t r = Rw + rl;                        // this lane's row
  const int32_t seg_lo = q * PPT;                   // first local column of this lane
  const double ncols_f = (double)a.ncols;

  const T bg = from_bits<T>(a.bg_bits);
  const T scalar = from_bits<T>(a.scalar_bits);
  const T* field = (const T*)a.field;

  T acc[PPT];
#pragma unroll
  for (int j = 0; j < PPT; ++j) acc[j] = bg;
  unsigned long long burns = 0;

  uint32_t* cnt = s_cnt[wave];
  // counters start zeroed; every apply pass zeroes what it read
  for (int w = lane; w < CNT_WORDS; w += 64) cnt[w] = 0u;
  wave_sync();

  const int32_t strip_r1 = min(Rw + WAVE_ROWS, a.nrows);
  const int32_t tile_c1 = min(C0 + TW, a.ncols);

  auto process_item = [&](uint32_t item) {
    const ItemBox bx = a.item_box[item];
    // wave-uniform skip: item rows do not meet this wave's strip
    if (bx.r1 <= Rw || bx.r0 >= strip_r1 || bx.c1 <= C0 || bx.c0 >= tile_c1) return;
    const int32_t kind = a.item_kind[item];
    const int32_t g = a.item_geom[item];
    const T v = field ? field[g] : scalar;
    const int64_t vb = a.item_vbeg[item], ve = a.item_vbeg[item + 1];
    if (kind == 0) {
      // ---- polygon: even-odd coverage of this row from crossing counts
      // (segment_coverage in rz_kernels.h states why this equals the spans).
      uint64_t tog[NWORD];
#pragma unroll
      for (int w = 0; w < NWORD; ++w) tog[w] = 0ull;
      uint32_t lp = 0;
      int32_t amax = -1;
      if (r < strip_r1 && r >= bx.r0 && r < bx.r1) {
        for (int64_t e = vb + q; e < ve; e += 4) {
          const PolyEdge ed = a.edges[e];
          if (ed.ystart <= r && r < ed.yend) {
            const double x = crossing(ed, r);
            if (!is_pos_nan(x)) toggle_crossing(span_col(x, ncols_f), C0, TW, tog, lp, amax);
          }
        }
      }
      // combine the four lanes of this row (lanes 4*rl .. 4*rl+3)
#pragma unroll
      for (int msk = 1; msk <= 2; msk <<= 1) {
        lp ^= (uint32_t)__shfl_xor((int)lp, msk, 64);
        amax = max(amax, __shfl_xor(amax, msk, 64));
#pragma unroll
        for (int w = 0; w < NWORD; ++w)
          tog[w] ^= (uint64_t)__shfl_xor((unsigned long long)tog[w], msk, 64);
      }
      const uint64_t cov = segment_coverage<PPT, NWORD>(tog, lp, amax, C0, seg_lo);
      if (cov) {
        burns += (unsigned long long)__popcll(cov);
#pragma unroll
        for (int j = 0; j < PPT; ++j)
          if ((cov >> j) & 1ull) fold<T, FUN>(acc[j], v, bg, 1u);
      }
      return;
    }
    // ---- lines and points: multiplicities in wave-private u16 counters
    auto emit = [&](int64_t cx, int64_t cy) {
      if (cx < C0 || cx >= tile_c1 || cy < Rw || cy >= strip_r1) return;
      const uint32_t idx = (uint32_t)(cy - Rw) * TW + (uint32_t)(cx - C0);
      atomicAdd(&cnt[idx >> 1], 1u << ((idx & 1u) * 16u));
    };
    const bool dedup = kind == 1 && a.dedup_lines;
    const unsigned long long last = kind == 1 ? a.item_last[item] : 0ull;
    uint32_t seen = 0u;  // PixelCache bits of this lane's pixels (PPT <= 32)
    for (int64_t cs = vb; cs < ve; cs += 65535) {  // <= 65535 writes per pixel per chunk
      const int64_t ce = min(ve, cs + 65535);
      for (int64_t s = cs + lane; s < ce; s += 64) {
        const double2 p0 = a.px[s];
        if (kind == 2) {  // point: extract_point + burn_point
          if (p0.x >= 0.0 && p0.x < ncols_f && p0.y >= 0.0 && p0.y < (double)a.nrows)
            emit((int64_t)p0.x, (int64_t)p0.y);
          continue;
        }
        const uint8_t f = a.vflags[s];
        if (!(f & VF_HAS_NEXT)) continue;
        const double2 p1 = a.px[s + 1];
        if (!(fmin(p0.x, p1.x) < ncols_f && fmax(p0.x, p1.x) >= 0.0 &&
              fmin(p0.y, p1.y) < (double)a.nrows && fmax(p0.y, p1.y) >= 0.0))
          continue;  // extract_line bbox filter, edges.rs:130
        const int64_t ix0 = sat_isize(floor(p0.x)), iy0 = sat_isize(floor(p0.y));
        const int64_t ix1 = sat_isize(floor(p1.x)), iy1 = sat_isize(floor(p1.y));
        const int64_t lim = (int64_t)1 << 30;
        if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
            iy1 > -lim && iy1 < lim) {
          walk_window<int64_t>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
        } else if (ix0 > -(lim << 30) && ix0 < (lim << 30) && ix1 > -(lim << 30) &&
                   ix1 < (lim << 30) && iy0 > -(lim << 30) && iy0 < (lim << 30) &&
                   iy1 > -(lim << 30) && iy1 < (lim << 30)) {
          walk_window<__int128>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
        } else {
          atomicOr(a.status, 1u);  // pixel coordinates beyond +-2^60: unsupported
        }
        // end cell of the last kept edge of an open line (burners.rs:86-89)
        if ((unsigned long long)(s + 1) == last && !(f & VF_CLOSED)) emit(ix1, iy1);
      }
      wave_sync();
      // apply the multiplicities of this lane's pixels, then clear them
      uint32_t* crow = cnt + (rl * TW + seg_lo) / 2;
#pragma unroll
      for (int j2 = 0; j2 < PPT / 2; ++j2) {
        const uint32_t wv = crow[j2];
        if (wv) {
          crow[j2] = 0u;
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            uint32_t k = (wv >> (16 * h)) & 0xFFFFu;
            const int j = 2 * j2 + h;
            if (k && dedup) {  // LineWriter + PixelCache: one write per pixel per item
              const uint32_t bit = 1u << j;
              if (seen & bit) k = 0;
              else { seen |= bit; k = 1; }
            }
            if (k) {
              burns += k;
              fold<T, FUN>(acc[j], v, bg, k);
            }
          }
        }
      }
      wave_sync();
    }
  };

  if (a.pair_item) {
    const uint32_t b = a.tile_off[tile], e = a.tile_off[tile + 1];
    for (uint32_t i = b; i < e; ++i) process_item(a.pair_item[i]);
  } else {
    // scan mode: test every item's box in order (small item counts)
    const int32_t tile_r1 = min(R0 + TILE_H, a.nrows);
    for (int64_t base = 0; base < a.n_items; base += 64) {
      const int64_t it = base + lane;
      bool hit = false;
      if (it < a.n_items) {
        const ItemBox bx = a.item_box[it];
        const uint32_t il = a.layer_of_geom ? a.layer_of_geom[a.item_geom[it]] : 0u;
        hit = il == layer && bx.r0 < tile_r1 && bx.r1 > R0 && bx.c0 < tile_c1 && bx.c1 > C0 &&
              bx.r0 < bx.r1 && bx.c0 < bx.c1;
      }
      uint64_t m = __ballot(hit);
      while (m) {
        const int bit = __builtin_ctzll(m);
        m &= m - 1;
        process_item((uint32_t)(base + bit));
      }
    }
  }

  // ---- write the tile once (background + burns), 16-byte stores when aligned
  if (r < a.nrows) {
    T* orow = (T*)a.out + ((uint64_t)layer * (uint64_t)a.nrows + (uint64_t)r) * (uint64_t)a.ncols;
    const int32_t c = C0 + seg_lo;
    constexpr int BYTES = PPT * (int)sizeof(T);
    if (c + PPT <= a.ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
      u32x4 vec[BYTES / 16];
      __builtin_memcpy(vec, acc, BYTES);
      u32x4* dst = (u32x4*)(orow + c);
#pragma unroll
      for (int k = 0; k < BYTES / 16; ++k) __builtin_nontemporal_store(vec[k], dst + k);
    } else {
#pragma unroll
      for (int j = 0; j < PPT; ++j)
        if (c + j < a.ncols) orow[c + j] = acc[j];
    }
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(a.burns, burns);
  }
}

}  // namespace rzk
