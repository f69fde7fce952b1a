// rz_tile.hip -- the tile kernel: scanline coverage, burn folds and the
// write-once store of every output tile.
#pragma once

#include "rz_kernels.h"

namespace rzk {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// XOR-combine a value across the four lanes of a quad with DPP quad
// permutations ([1,0,3,2] then [2,3,0,1]): VALU data movement, no LDS.
__device__ inline uint32_t quad_xor_all(uint32_t x) {
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);
  x ^= (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);
  return x;
}
__device__ inline uint64_t quad_xor_all64(uint64_t x) {
  return (uint64_t)quad_xor_all((uint32_t)x) | ((uint64_t)quad_xor_all((uint32_t)(x >> 32)) << 32);
}
__device__ inline int32_t quad_max_all(int32_t x) {
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false));
  x = max(x, __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false));
  return x;
}

// One 64-byte descriptor from a read-only table (four 16-byte loads).
__device__ inline ItemDesc load_desc(const RZ_CONST ItemDesc* p) {
  const RZ_CONST u32x4* q = (const RZ_CONST u32x4*)p;
  const u32x4 w[4] = {q[0], q[1], q[2], q[3]};
  ItemDesc d;
  __builtin_memcpy(&d, w, sizeof(d));
  return d;
}

// First index of XCD x's contiguous share of n.
__device__ inline uint32_t xcd_unit_lo(uint32_t x, uint32_t n) {
  const uint32_t per = n / NUM_XCD, rem = n % NUM_XCD;
  return x * per + min(x, rem);
}

// Coverage of one polygon item over the 16 rows [Rw, Rw + 16) (rows from
// strip_r1 on excluded) and the columns [C0, C0 + 64 * NWORD): lane
// (rl = lane / 4, q = lane % 4) returns word q of row Rw + rl's mask (lanes
// with q >= NWORD return 0).  Lane layout: the four lanes of a row take every
// 4th active edge; their crossing toggles are XOR-combined with DPP and the
// row mask follows from the note above row_coverage_word in rz_kernels.h.
// `stage` is 64 edges of wave-private LDS.
template <int NWORD>
__device__ inline uint64_t strip_coverage(const PolyEdge* ebase, uint32_t n_e, int32_t Rw,
                                          int32_t strip_r1, int32_t C0, double ncols_f,
                                          PolyEdge* stage, int lane) {
  const int rl = lane >> 2, q = lane & 3;
  const int32_t r = Rw + rl;
  const bool row_ok = r < strip_r1;
  uint64_t tog[NWORD];
#pragma unroll
  for (int w = 0; w < NWORD; ++w) tog[w] = 0ull;
  uint32_t lp = 0;
  int32_t amax = -1;
  for (uint32_t cb = 0; cb < n_e; cb += 64) {
    // stage the edges of this chunk that reach the strip, compacted
    // (two 16-byte vector loads and stores: no struct copy through scratch)
    const uint32_t e = cb + (uint32_t)lane;
    uint4 e_lo = make_uint4(0u, 0u, 0u, 0u), e_hi = make_uint4(0u, 0u, 0u, 0u);
    bool act = false;
    if (e < n_e) {
      const uint4* src = reinterpret_cast<const uint4*>(ebase + e);
      e_lo = src[0];
      e_hi = src[1];  // dxdy, ystart (z), yend (w)
      const int32_t ys = (int32_t)e_hi.z, ye = (int32_t)e_hi.w;
      act = ys < strip_r1 && ye > Rw && ys < ye;
    }
    const uint64_t m = __ballot(act);
    if (act) {
      uint4* dst = reinterpret_cast<uint4*>(stage + __popcll(m & ((1ull << lane) - 1ull)));
      dst[0] = e_lo;
      dst[1] = e_hi;
    }
    wave_sync();
    const int n = __popcll(m);
    for (int k0 = 0; k0 < n; k0 += 4) {  // uniform trip count, predicated lanes
      const int k = k0 + q;
      const PolyEdge ee = stage[k < n ? k : n - 1];
      const bool on = k < n && row_ok && ee.ystart <= r && r < ee.yend;
      const double x = crossing(ee, r);
      toggle_crossing_if<NWORD>(on, span_col(x, ncols_f), C0, tog, lp, amax);
    }
    wave_sync();
  }
  lp = quad_xor_all(lp);
  amax = quad_max_all(amax);
#pragma unroll
  for (int w = 0; w < NWORD; ++w) tog[w] = quad_xor_all64(tog[w]);
  return q < NWORD ? row_coverage_word<NWORD>(tog, lp, amax, C0, q) : 0ull;
}

// Largest value of v (0..16) over the lanes of the wave, from five ballots.
__device__ inline int32_t wave_max16(int32_t v) {
  int32_t m = 0;
#pragma unroll
  for (int step = 16; step >= 1; step >>= 1)
    if (__any(v >= m + step)) m += step;
  return m;
}

// Persistent work-queue tile kernel, one launch per pixel-function group of
// a batch (a single call is a batch of one job).
//
// Scheduling: a unit of work is one 16-row strip of one output tile, and
// every wave works alone (no workgroup barrier).  Workgroup b runs on XCD
// b % 8; each XCD owns a contiguous eighth of the launch's tiles, so
// neighbouring tiles share cache lines and items in that XCD's L2.  Waves
// take blocks of tiles from per-(XCD, strip) counters and, when their XCD's
// range is used up, from the other XCDs' counters, so jobs of very uneven
// size still end together.  Every wave leaves when all eight ranges of its
// strip are used up.
//
// Per unit (16 rows x TW columns of the job's output): the background fill
// of RasterInfo::build_raster (rust/src/geo/raster.rs:23-28) and every write
// of every item meeting the strip, folded in geometry order in registers,
// then written to HBM once.
//
// Two lane layouts per wave:
//  * coverage: lane (rl = lane/4, q = lane%4) takes every 4th active edge of
//    row rl; the four lanes of a row XOR-combine their crossing toggles and
//    publish the row's coverage mask in LDS;
//  * fold / store: lane l owns the 16-byte chunk l % LPR of rows
//    l / LPR + k * RSTEP, so every store instruction writes whole tile rows
//    (full 128-byte lines, contiguous across lanes).
// The polygon path is branch-free per lane (predicated crossings and folds).
// Waves per SIMD the register allocation targets: 7 for 4-byte outputs, 6
// for 1- and 2-byte outputs (measured: the latency-bound strip loop gains more
// from the extra wave than the few spilled registers cost), the compiler's
// choice for 8-byte outputs and the line / point variants.
template <typename T, bool LINES>
struct TileOccupancy {
  static constexpr int value = LINES ? 4 : (sizeof(T) == 4 ? 7 : (sizeof(T) == 8 ? 5 : 6));
};

template <typename T, int FUN, bool LINES>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(TileOccupancy<T, LINES>::value))) void k_tile(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;                    // u64 words of a row mask
  constexpr int PPC = 16 / (int)sizeof(T);          // pixels per 16-byte chunk
  constexpr int LPR = TW / PPC;                     // lanes per row (store layout)
  constexpr int RSTEP = 64 / LPR;                   // rows per store instruction
  constexpr int RPL = WAVE_ROWS / RSTEP;            // rows per lane
  constexpr int NPIX = RPL * PPC;                   // pixels per lane
  // 1- and 2-byte outputs of the polygon variants keep their pixels packed,
  // four (two) per 32-bit word, and fold them lane-wise (fold_word)
  constexpr bool PACKED = !LINES && sizeof(T) <= 2;
  constexpr int NW = PACKED ? NPIX * (int)sizeof(T) / 4 : 1;  // words per lane
  constexpr int PPW = 4 / (int)sizeof(T);                     // pixels per word (packed)
  constexpr int CNT_WORDS = LINES ? WAVE_ROWS * TW / 2 : 1;  // u16 counters, 2 per word
  static_assert(NPIX <= 64, "dedup bits assume <= 64 pixels per lane");
  __shared__ PolyEdge s_edge[4][64];
  __shared__ uint64_t s_mask[4][2][WAVE_ROWS][NWORD];
  __shared__ uint32_t s_cnt[4][CNT_WORDS];

  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;            // coverage layout
  const int srow = lane / LPR, ch = lane % LPR;      // store layout
  const int ccol = ch * PPC;                         // first local column of the chunk
  PolyEdge* stage = s_edge[wave];
  uint32_t* cnt = s_cnt[wave];
  if (LINES) {
    for (int w = lane; w < CNT_WORDS; w += 64) cnt[w] = 0u;  // every apply re-zeroes
    wave_sync();
  }
  unsigned long long burns = 0;
  int mbuf = 0;  // alternating row-mask buffer: one wave_sync per polygon item

  // ---- work queue.  Wave w of every workgroup rasterises strip w (rows
  // 16w..16w+15) of its tiles.  Each XCD owns a contiguous eighth of the
  // launch's tiles.  A wave's first tile is its workgroup's slot in that
  // range (no atomic); after the first `slots` tiles, queue (x, w) hands out
  // blocks of a.chunk consecutive tiles, so the waves working on one strip
  // index walk neighbouring tiles together (their stores fill neighbouring
  // lines).  A grab is issued one block ahead, so its latency overlaps the
  // current block.  A wave whose own XCD range is used up reads all eight
  // counters of its strip with one relaxed load (counters only grow: a range
  // that reads as used up is used up) and continues on the next XCD range
  // with work left.
  const uint32_t n_loc = a.t_hi - a.t_lo;
  const uint32_t xcd = blockIdx.x % NUM_XCD, slot = blockIdx.x / NUM_XCD;
  const uint32_t slots = gridDim.x / NUM_XCD;
  const uint32_t B = a.chunk;
  uint32_t* const qbase = a.queue + (uint32_t)wave * kQueueStride;  // queue (x, w) at qbase + 4x*stride
  uint32_t cur_x = xcd;
  uint32_t pending = 0;  // lane 0: result of the grab in flight on cur_x
  auto issue = [&](uint32_t x) {
    if (lane == 0) pending = atomicAdd(qbase + x * 4u * kQueueStride, 1u);
  };
  // first tile handed out by queue x's grab number c
  auto grab_tile = [&](uint32_t x, uint32_t c) -> uint64_t {
    return (uint64_t)xcd_unit_lo(x, n_loc) + slots + (uint64_t)c * B;
  };
  uint32_t t_cur = xcd_unit_lo(xcd, n_loc) + slot;  // current block [t_cur, t_end)
  uint32_t t_end = t_cur + 1u;
  auto take = [&](uint32_t x) -> bool {
    const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)pending, 0);
    const uint64_t s = grab_tile(x, c), xhi = xcd_unit_lo(x + 1u, n_loc);
    if (s >= xhi) return false;
    t_cur = (uint32_t)s;
    t_end = (uint32_t)min(s + (uint64_t)B, xhi);
    return true;
  };
  // (a grab on cur_x is always in flight when next_block runs)
  auto next_block = [&]() -> bool {
    if (take(cur_x)) return true;
    for (;;) {
      bool avail = false;
      if (lane < NUM_XCD) {
        const uint32_t seen = __hip_atomic_load(qbase + (uint32_t)lane * 4u * kQueueStride,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        avail = grab_tile((uint32_t)lane, seen) < (uint64_t)xcd_unit_lo((uint32_t)lane + 1u, n_loc);
      }
      const uint32_t m = (uint32_t)__ballot(avail) & 0xFFu;
      if (m == 0u) return false;
      const uint32_t rot = ((m >> xcd) | (m << (NUM_XCD - xcd))) & 0xFFu;
      const uint32_t x = (xcd + (uint32_t)__builtin_ctz(rot)) % NUM_XCD;
      issue(x);
      if (take(x)) {
        cur_x = x;
        return true;
      }
    }
  };
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST int64_t* tend_c = as_const(a.job_tend);
  const RZ_CONST uint32_t* toff_c = as_const(a.tile_off);
  const RZ_CONST ItemDesc* pdesc_c = as_const(a.pair_desc);
  const RZ_CONST ItemDesc* desc_c = as_const(a.desc);
  const RZ_CONST uint64_t* band_c = as_const(a.band_off);
  const RZ_CONST uint64_t* heavy_c = as_const(a.heavy_base);
  const RZ_CONST uint64_t* mask_c = as_const(a.masks);
  const RZ_CONST uint64_t* wide_c = as_const(a.wide_masks);
  uint32_t pb = 0, pe = 0;
  auto fetch_range = [&](uint32_t tl) {
    if (a.tile_off) {
      pb = toff_c[a.t_lo + tl];
      pe = toff_c[a.t_lo + tl + 1u];
    }
  };
  bool have = t_cur < xcd_unit_lo(xcd + 1u, n_loc);
  issue(xcd);  // the grab after the first tile (or the first block, if there is no first tile)
  if (!have) {
    have = next_block();
    if (have) issue(cur_x);
  }
  if (have) fetch_range(t_cur);
  const int32_t strip = wave;

  while (have) {
    const uint32_t t = a.t_lo + t_cur++;
    const RZ_CONST JobDev& J = jobs_c[find_job(tend_c, a.n_jobs, (int64_t)t)];
    const int32_t nrows = J.nrows, ncols = J.ncols;
    const double ncols_f = (double)ncols, nrows_f = (double)nrows;
    const T bg = from_bits<T>(J.bg_bits);
    const uint32_t tj = t - J.tile_base;
    const uint32_t layer = tj / J.tiles_per_layer;
    const uint32_t tl = tj - layer * J.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
    const int32_t C0 = tx * TW;
    const int32_t Rw = ty * TILE_H + strip * WAVE_ROWS;  // first row of the strip
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, nrows);
    const int32_t tile_c1 = min(C0 + TW, ncols);
    const bool dedup_job = J.dedup_lines != 0;
    // heavy tile (many items): every polygon mask comes from the pre-pass
    // (a tile is heavy iff its pair range holds at least heavy_min items, so
    // the slot base is only loaded for heavy tiles)
    uint64_t hbase = 0;
    const bool hb_ok = a.heavy_base && !J.scan && pe - pb >= a.heavy_min;
    if (hb_ok) hbase = heavy_c[t];

    T acc[PACKED ? 1 : NPIX];  // acc[k * PPC + j]: row srow + k*RSTEP, column ccol + j
    uint32_t accw[NW];         // packed: pixel k * PPC + j in word (k * PPC + j) / PPW
    const uint32_t bgw = PACKED ? lane_splat<T>(J.bg_bits) : 0u;
    if constexpr (PACKED) {
#pragma unroll
      for (int i = 0; i < NW; ++i) accw[i] = bgw;
    } else {
#pragma unroll
      for (int j = 0; j < NPIX; ++j) acc[j] = bg;
    }

    auto process_item = [&](const ItemDesc d, uint32_t pi) {
      // wave-uniform skip: the item box misses this strip
      if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) return;
      const int32_t kind = (int32_t)(d.kind_layer & 3u);
      const T v = from_bits<T>(d.value);
      const uint64_t vbits = d.value;
      const int64_t vb = d.vbeg, ve = d.vbeg + d.nv;
      if (kind == 0) {
        uint64_t cov = 0;
        if (d.flags & kItemWide) {  // the wide pre-pass wrote the item's masks
          const int32_t nw = wide_row_words(d.c0, d.c1), wi = (C0 >> 6) + q - (d.c0 >> 6);
          if (q < NWORD && wi >= 0 && wi < nw)
            cov = wide_c[(uint64_t)d.wide * 16u +
                         (uint64_t)(((Rw >> 4) - (d.r0 >> 4)) * 16 + rl) * (uint64_t)nw + (uint64_t)wi];
        } else if (hb_ok) {  // heavy tile: the coverage pre-pass wrote this mask
          if (q < NWORD)
            cov = mask_c[(hbase * 4u + (uint64_t)strip * (pe - pb) + (pi - pb)) * a.slot_words +
                         (uint32_t)(rl * NWORD + q)];
        } else {
          // the item's edges, or for a banded item only those of this tile row
          const PolyEdge* ebase = a.edges + vb;
          uint32_t n_e = (uint32_t)d.nv;
          if (d.band) {
            const uint32_t bi = d.band - 1u + (uint32_t)(Rw / kBandRows - d.r0 / kBandRows);
            const uint64_t b0 = band_c[bi];
            ebase = a.band_edges + b0;
            n_e = (uint32_t)(band_c[bi + 1u] - b0);
          }
          cov = strip_coverage<NWORD>(ebase, n_e, Rw, strip_r1, C0, ncols_f, stage, lane);
        }
        if (q < NWORD) s_mask[wave][mbuf][rl][q] = cov;
        wave_sync();
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const uint64_t row_word = s_mask[wave][mbuf][srow + k * RSTEP][ccol >> 6];
          const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
          if (__any(bits != 0u)) {  // wave-uniform skip of uncovered row groups
            burns += (unsigned long long)__popc(bits);
            if constexpr (PACKED) {
              const uint32_t vw = lane_splat<T>(vbits);
#pragma unroll
              for (int i = 0; i < PPC / PPW; ++i)
                accw[k * (PPC / PPW) + i] = fold_word<T, FUN>(
                    accw[k * (PPC / PPW) + i], vw, bgw,
                    lane_mask<(int)sizeof(T)>((bits >> (i * PPW)) & ((1u << PPW) - 1u)));
            } else {
#pragma unroll
              for (int j = 0; j < PPC; ++j)
                acc[k * PPC + j] = fold_select<T, FUN>(acc[k * PPC + j], v, bg, (bits >> j) & 1u);
            }
          }
        }
        mbuf ^= 1;  // the next polygon writes the other buffer
        return;
      }
      if (!LINES) return;  // the host only picks this variant without lines / points
      // ---- lines and points: multiplicities in wave-private u16 counters
      bool hit = false;  // this lane counted a cell of the strip
      auto emit = [&](int64_t cx, int64_t cy) {
        if (cx < C0 || cx >= tile_c1 || cy < Rw || cy >= strip_r1) return;
        const uint32_t idx = (uint32_t)(cy - Rw) * TW + (uint32_t)(cx - C0);
        atomicAdd(&cnt[idx >> 1], 1u << ((idx & 1u) * 16u));
        hit = true;
      };
      const bool dedup = kind == 1 && dedup_job;
      uint64_t seen = 0u;  // PixelCache bits of this lane's pixels
      for (int64_t cs = vb; cs < ve; cs += 65535) {  // <= 65535 writes per pixel per chunk
        const int64_t ce = min(ve, cs + 65535);
        for (int64_t s = cs + lane; s < ce; s += 64) {
          const double2 p0 = a.px[s];
          if (kind == 2) {  // point: extract_point + burn_point
            if (p0.x >= 0.0 && p0.x < ncols_f && p0.y >= 0.0 && p0.y < nrows_f)
              emit((int64_t)p0.x, (int64_t)p0.y);
            continue;
          }
          const uint8_t f = a.vflags[s];
          if (!(f & VF_HAS_NEXT)) continue;
          const double2 p1 = a.px[s + 1];
          if (!(fmin(p0.x, p1.x) < ncols_f && fmax(p0.x, p1.x) >= 0.0 &&
                fmin(p0.y, p1.y) < nrows_f && fmax(p0.y, p1.y) >= 0.0))
            continue;  // extract_line bbox filter, edges.rs:130
          const int64_t ix0 = sat_isize(floor(p0.x)), iy0 = sat_isize(floor(p0.y));
          const int64_t ix1 = sat_isize(floor(p1.x)), iy1 = sat_isize(floor(p1.y));
          const int64_t lim = (int64_t)1 << 30, lim2 = (int64_t)1 << 60;
          if (ix0 > -lim && ix0 < lim && ix1 > -lim && ix1 < lim && iy0 > -lim && iy0 < lim &&
              iy1 > -lim && iy1 < lim) {
            walk_window<int64_t>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
          } else if (ix0 > -lim2 && ix0 < lim2 && ix1 > -lim2 && ix1 < lim2 && iy0 > -lim2 &&
                     iy0 < lim2 && iy1 > -lim2 && iy1 < lim2) {
            walk_window<__int128>(ix0, iy0, ix1, iy1, C0, tile_c1, Rw, strip_r1, emit);
          } else {
            atomicOr(a.status, 1u);  // pixel coordinates beyond +-2^60: unsupported
          }
          // end cell of the last kept edge of an open line (burners.rs:86-89)
          if ((uint64_t)(s + 1) == d.last && !(f & VF_CLOSED)) emit(ix1, iy1);
        }
        // nothing counted in this strip: the counters are still zero
        if (!__any(hit)) continue;
        hit = false;
        wave_sync();
        // apply the multiplicities of this lane's pixels, then clear them
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          uint32_t* crow = cnt + ((srow + k * RSTEP) * TW + ccol) / 2;
#pragma unroll
          for (int j2 = 0; j2 < PPC / 2; ++j2) {
            const uint32_t wv = crow[j2];
            if (wv) {
              crow[j2] = 0u;
#pragma unroll
              for (int h = 0; h < 2; ++h) {
                uint32_t kk = (wv >> (16 * h)) & 0xFFFFu;
                const int j = k * PPC + 2 * j2 + h;
                if (kk && dedup) {  // LineWriter + PixelCache: one write per pixel per item
                  const uint64_t bit = 1ull << j;
                  if (seen & bit) kk = 0;
                  else {
                    seen |= bit;
                    kk = 1;
                  }
                }
                if (kk) {
                  burns += kk;
                  fold<T, FUN>(acc[j], v, bg, kk);
                }
              }
            }
          }
        }
        wave_sync();
      }
    };

    if (!J.scan) {
      // (a heavy tile of a polygon-only group is folded by k_tile_heavy,
      // which runs after this kernel and stores over this strip)
      const uint32_t pend = (hb_ok && a.heavy_split) ? pb : pe;
      for (uint32_t i = pb; i < pend; ++i) process_item(load_desc(pdesc_c + i), i);
    } else {
      // scan job: test every item's box in order (small item counts)
      const int64_t ib = J.i_base, ie = J.i_base + J.ni;
      for (int64_t base = ib; base < ie; base += 64) {
        const int64_t it = base + lane;
        bool hit = false;
        if (it < ie) {
          const ItemDesc d = load_desc(desc_c + it);
          hit = (d.kind_layer >> 2) == layer && d.r0 < strip_r1 && d.r1 > Rw && d.c0 < tile_c1 &&
                d.c1 > C0 && d.r0 < d.r1 && d.c0 < d.c1;
        }
        uint64_t m = __ballot(hit);
        while (m) {
          const int bit = __builtin_ctzll(m);
          m &= m - 1;
          process_item(load_desc(desc_c + base + bit), 0u);
        }
      }
    }

    // take the next tile (and its item range) before this strip's stores
    if (t_cur == t_end) {
      have = next_block();
      if (have) issue(cur_x);
    }
    if (have) fetch_range(t_cur);

    // ---- write the strip once (background + burns): each store instruction
    // covers RSTEP whole rows of the tile
    const int32_t c = C0 + ccol;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int32_t row = Rw + srow + k * RSTEP;
      if (row < nrows) {
        T* orow = (T*)J.out + ((uint64_t)layer * (uint64_t)nrows + (uint64_t)row) * (uint64_t)ncols;
        if (c + PPC <= ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
          u32x4 vec;
          if constexpr (PACKED) vec = u32x4{accw[k * 4], accw[k * 4 + 1], accw[k * 4 + 2], accw[k * 4 + 3]};
          else __builtin_memcpy(&vec, &acc[k * PPC], 16);
          __builtin_nontemporal_store(vec, (u32x4*)(orow + c));
        } else {
#pragma unroll
          for (int j = 0; j < PPC; ++j)
            if (c + j < ncols) {
              if constexpr (PACKED) {
                T x;
                __builtin_memcpy(&x, reinterpret_cast<const char*>(accw) + (k * PPC + j) * sizeof(T), sizeof(T));
                orow[c + j] = x;
              } else {
                orow[c + j] = acc[k * PPC + j];
              }
            }
        }
      }
    }
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(&a.burns[blockIdx.x % kBurnSlots], burns);
  }
}

// ---------------------------------------------------------- heavy tiles --
//
// A tile with at least kHeavyItems items would serialise all their coverage
// work in its four strip waves.  For such tiles the coverage masks of every
// (polygon item, strip) are computed beforehand by k_cover, one wave per
// (item, strip), in parallel; the tile kernel then only folds the masks in
// geometry order (every pixel function, ordered sum included, stays exact).

// Items of every heavy tile (0 for the others) and its heavy flag.
__global__ void k_heavy_count(const uint32_t* tile_off, uint32_t n_tiles, uint32_t min_items,
                              uint32_t* heavy_cnt, uint32_t* heavy_flag) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n_tiles) return;
  uint32_t c = 0;
  if (t < n_tiles) {
    c = tile_off[t + 1] - tile_off[t];
    if (c < min_items) c = 0;
  }
  heavy_cnt[t] = c;
  heavy_flag[t] = c > 0 ? 1u : 0u;
}

// The heavy tiles in tile order: heavy_list[heavy_rank[t]] = t, so the heavy
// tiles of tiles [t_lo, t_hi) are heavy_list[heavy_rank[t_lo] .. heavy_rank[t_hi]).
__global__ void k_heavy_list(const uint32_t* heavy_flag, const uint32_t* heavy_rank, uint32_t n_tiles,
                             uint32_t* heavy_list) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tiles && heavy_flag[t]) heavy_list[heavy_rank[t]] = t;
}

// The compact list of the pairs of heavy tiles that hold polygon items other
// than wide ones (the only ones k_cover writes; k_wide_cover writes the wide
// ones), appended with one atomic per wave.  For a wide item's pair it also
// records the item's place in the tile's pair list, at the pair's emission
// index (k_bin_emit's order over the item's tile rectangle), where
// k_wide_cover finds it with one load.
__global__ void k_heavy_pairs(const JobDev* jobs, const uint32_t* pair_tile, const uint32_t* pair_item,
                              const ItemDesc* pair_desc, const uint64_t* n_pairs, uint64_t cap,
                              const uint32_t* tile_off, const uint64_t* heavy_base, const uint64_t* pair_off,
                              uint32_t* wide_j, uint32_t* cover_list, uint32_t* cover_count) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool poly = false;
  if (i < min(*n_pairs, cap)) {
    const uint32_t t = pair_tile[i];
    if (heavy_base[t + 1] > heavy_base[t]) {
      const ItemDesc& d = pair_desc[i];
      poly = (d.kind_layer & 3u) == 0u && !(d.flags & kItemWide);  // k_wide_cover writes those
      int32_t tx0, tx1, ty0, ty1;
      if ((d.flags & kItemWide) && tile_rect(d.r0, d.r1, d.c0, d.c1, jobs[d.job].tile_w, tx0, tx1, ty0, ty1)) {
        const JobDev& J = jobs[d.job];
        const uint32_t tl = (t - J.tile_base) % J.tiles_per_layer;
        const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
        wide_j[pair_off[pair_item[i]] + (uint64_t)((ty - ty0) * (tx1 - tx0 + 1) + (tx - tx0))] =
            (uint32_t)(i - tile_off[t]);
      }
    }
  }
  const uint64_t m = __ballot(poly);
  if (m == 0) return;
  uint32_t base = 0;
  const int leader = __builtin_ctzll(m);
  if (lane == leader) base = atomicAdd(cover_count, (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, leader, 64);
  if (poly) cover_list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
}

template <int NWORD>
__global__ __launch_bounds__(256) void k_cover(CoverArgs a) {
  constexpr int TW = 64 * NWORD;
  __shared__ PolyEdge s_edge[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST uint64_t* band_c = as_const(a.band_off);
  const uint64_t units = 4ull * *a.cover_count;
  const uint64_t ustride = (uint64_t)gridDim.x * 4u;
  uint64_t u = (uint64_t)blockIdx.x * 4u + (uint64_t)wave;
  // the next unit's pair is loaded one unit ahead
  uint32_t i_next = u < units ? a.cover_list[u >> 2] : 0u;
  for (; u < units; u += ustride) {
    const uint32_t i = (uint32_t)__builtin_amdgcn_readfirstlane((int)i_next);  // wave-uniform
    if (u + ustride < units) i_next = a.cover_list[(u + ustride) >> 2];
    const int32_t strip = (int32_t)(u & 3u);
    const uint32_t t = a.pair_tile[i];
    const ItemDesc d = load_desc(as_const(a.pair_desc) + i);
    const RZ_CONST JobDev& J = jobs_c[d.job];
    if (J.tile_w != TW) continue;
    const uint32_t tj = t - J.tile_base;
    const uint32_t layer = tj / J.tiles_per_layer;
    const uint32_t tl = tj - layer * J.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
    const int32_t C0 = tx * TW, Rw = ty * TILE_H + strip * WAVE_ROWS;
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, J.nrows), tile_c1 = min(C0 + TW, J.ncols);
    if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) continue;
    const PolyEdge* ebase = a.edges + d.vbeg;
    uint32_t n_e = (uint32_t)d.nv;
    if (d.band) {
      const uint32_t bi = d.band - 1u + (uint32_t)(Rw / kBandRows - d.r0 / kBandRows);
      const uint64_t b0 = band_c[bi];
      ebase = a.band_edges + b0;
      n_e = (uint32_t)(band_c[bi + 1u] - b0);
    }
    const uint64_t cov = strip_coverage<NWORD>(ebase, n_e, Rw, strip_r1, C0, (double)J.ncols, s_edge[wave], lane);
    // strip-major slots: item j of heavy tile t, strip s at heavy_base[t] * 4 + s * cnt + j
    const uint32_t p0 = a.tile_off[t], cnt = a.tile_off[t + 1] - p0;
    const uint64_t slot = a.heavy_base[t] * 4u + (uint64_t)strip * cnt + (i - p0);
    if (q < NWORD) a.masks[slot * a.slot_words + (uint64_t)(rl * NWORD + q)] = cov;
  }
}

// Heavy tiles of a polygon-only group (k_tile skips them): one wave per
// (heavy tile, strip) unit, taken from a work counter.  Every item's strip
// mask was written by the pre-pass (k_cover), strip-major per tile.  Masks
// and descriptors go through LDS kBlk items at a time: the next block is
// loaded into registers while the current block is folded from LDS, then
// stored to LDS, so the per-item loads of a tile with thousands of items
// overlap instead of serialising.  Same folds, geometry order and stores as
// k_tile.
template <typename T, int FUN>
__global__ __launch_bounds__(256) void k_tile_heavy(TileArgs a) {
  constexpr int TW = tile_w(sizeof(T));
  constexpr int NWORD = TW / 64;
  constexpr int PPC = 16 / (int)sizeof(T);
  constexpr int LPR = TW / PPC;
  constexpr int RSTEP = 64 / LPR;
  constexpr int RPL = WAVE_ROWS / RSTEP;
  constexpr int NPIX = RPL * PPC;
  constexpr bool PACKED = sizeof(T) <= 2;  // pixels packed in 32-bit words (fold_word)
  constexpr int NW = PACKED ? NPIX * (int)sizeof(T) / 4 : 1;
  constexpr int PPW = 4 / (int)sizeof(T);
  constexpr int SW = 16 * NWORD;          // words of one item's strip mask
  constexpr uint32_t kBlk = 256 / SW;     // items per block (16 or 4)
  __shared__ uint64_t s_mk[4][256];
  __shared__ ItemDesc s_ds[4][16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int srow = lane / LPR, ch = lane % LPR, ccol = ch * PPC;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST int64_t* tend_c = as_const(a.job_tend);
  // this launch's heavy tiles: heavy_list[h0 .. h1), four units (strips) each
  const uint32_t h0 = a.heavy_rank[a.t_lo], units = 4u * (a.heavy_rank[a.t_hi] - h0);
  // units per counter grab: a few per wave of the grid, at most 16
  const uint32_t grab = max(1u, min(16u, units / (gridDim.x * 4u * 4u)));
  uint64_t* mk = s_mk[wave];
  ItemDesc* ds = s_ds[wave];
  unsigned long long burns = 0;
  uint32_t u = 0, u_end = 0;
  for (;;) {
    if (u == u_end) {
      uint32_t g = 0;
      if (lane == 0) g = atomicAdd(a.heavy_ctr, grab);
      u = (uint32_t)__builtin_amdgcn_readfirstlane((int)g);
      u_end = min(u + grab, units);
      if (u >= units) break;
    }
    const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)a.heavy_list[h0 + (u >> 2)]);
    const int32_t strip = (int32_t)(u & 3u);
    ++u;
    const RZ_CONST JobDev& J = jobs_c[find_job(tend_c, a.n_jobs, (int64_t)t)];
    const int32_t nrows = J.nrows, ncols = J.ncols;
    const T bg = from_bits<T>(J.bg_bits);
    const uint32_t tj = t - J.tile_base;
    const uint32_t layer = tj / J.tiles_per_layer;
    const uint32_t tl = tj - layer * J.tiles_per_layer;
    const int32_t ty = (int32_t)(tl / (uint32_t)J.ntx), tx = (int32_t)(tl % (uint32_t)J.ntx);
    const int32_t C0 = tx * TW, Rw = ty * TILE_H + strip * WAVE_ROWS;
    const int32_t strip_r1 = min(Rw + WAVE_ROWS, nrows), tile_c1 = min(C0 + TW, ncols);
    const uint32_t pb = a.tile_off[t], cnt = a.tile_off[t + 1] - pb;
    const uint64_t* sbase = a.masks + (a.heavy_base[t] * 4u + (uint64_t)strip * cnt) * a.slot_words;
    const ItemDesc* dbase = a.pair_desc + pb;

    T acc[PACKED ? 1 : NPIX];
    uint32_t accw[NW];
    const uint32_t bgw = PACKED ? lane_splat<T>(J.bg_bits) : 0u;
    if constexpr (PACKED) {
#pragma unroll
      for (int i = 0; i < NW; ++i) accw[i] = bgw;
    } else {
#pragma unroll
      for (int j = 0; j < NPIX; ++j) acc[j] = bg;
    }

    // block j0 into registers: masks (lane L: words 4L .. 4L + 3 of the
    // block) and descriptors (lane L: bytes 16 (L % 4) .. of descriptor L / 4)
    const uint32_t jm = (uint32_t)(lane * 4) / SW, om = (uint32_t)(lane * 4) % SW, jd = (uint32_t)lane >> 2;
    uint4 m_lo = make_uint4(0u, 0u, 0u, 0u), m_hi = m_lo, d_q = m_lo;
    auto load_block = [&](uint32_t j0) {
      if (j0 + jm < cnt) {
        const uint4* src = reinterpret_cast<const uint4*>(sbase + (uint64_t)(j0 + jm) * a.slot_words + om);
        m_lo = src[0];
        m_hi = src[1];
      }
      if (jd < kBlk && j0 + jd < cnt)
        d_q = reinterpret_cast<const uint4*>(dbase + j0 + jd)[lane & 3];
    };
    auto stage_block = [&]() {
      wave_sync();  // the previous block's reads are done
      uint4* dm = reinterpret_cast<uint4*>(mk + lane * 4);
      dm[0] = m_lo;
      dm[1] = m_hi;
      if (jd < kBlk) reinterpret_cast<uint4*>(ds + jd)[lane & 3] = d_q;
      wave_sync();
    };
    load_block(0u);
    for (uint32_t j0 = 0; j0 < cnt; j0 += kBlk) {
      stage_block();
      if (j0 + kBlk < cnt) load_block(j0 + kBlk);  // in flight during the folds below
      const uint32_t gn = min(kBlk, cnt - j0);
      for (uint32_t g = 0; g < gn; ++g) {
        const ItemDesc& d = ds[g];
        if (d.r1 <= Rw || d.r0 >= strip_r1 || d.c1 <= C0 || d.c0 >= tile_c1) continue;
        const T v = from_bits<T>(d.value);
        const uint64_t vbits = d.value;
        const uint64_t* m = mk + g * SW;
#pragma unroll
        for (int k = 0; k < RPL; ++k) {
          const uint64_t row_word = m[(srow + k * RSTEP) * NWORD + (ccol >> 6)];
          const uint32_t bits = (uint32_t)(row_word >> (ccol & 63)) & ((1u << PPC) - 1u);
          if (__any(bits != 0u)) {
            burns += (unsigned long long)__popc(bits);
            if constexpr (PACKED) {
              const uint32_t vw = lane_splat<T>(vbits);
#pragma unroll
              for (int i = 0; i < PPC / PPW; ++i)
                accw[k * (PPC / PPW) + i] = fold_word<T, FUN>(
                    accw[k * (PPC / PPW) + i], vw, bgw,
                    lane_mask<(int)sizeof(T)>((bits >> (i * PPW)) & ((1u << PPW) - 1u)));
            } else {
#pragma unroll
              for (int j = 0; j < PPC; ++j)
                acc[k * PPC + j] = fold_select<T, FUN>(acc[k * PPC + j], v, bg, (bits >> j) & 1u);
            }
          }
        }
      }
    }

    // write the strip once, as k_tile does
    const int32_t c = C0 + ccol;
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int32_t row = Rw + srow + k * RSTEP;
      if (row < nrows) {
        T* orow = (T*)J.out + ((uint64_t)layer * (uint64_t)nrows + (uint64_t)row) * (uint64_t)ncols;
        if (c + PPC <= ncols && ((uintptr_t)(orow + c) & 15u) == 0) {
          u32x4 vec;
          if constexpr (PACKED) vec = u32x4{accw[k * 4], accw[k * 4 + 1], accw[k * 4 + 2], accw[k * 4 + 3]};
          else __builtin_memcpy(&vec, &acc[k * PPC], 16);
          __builtin_nontemporal_store(vec, (u32x4*)(orow + c));
        } else {
#pragma unroll
          for (int j = 0; j < PPC; ++j)
            if (c + j < ncols) {
              if constexpr (PACKED) {
                T x;
                __builtin_memcpy(&x, reinterpret_cast<const char*>(accw) + (k * PPC + j) * sizeof(T), sizeof(T));
                orow[c + j] = x;
              } else {
                orow[c + j] = acc[k * PPC + j];
              }
            }
        }
      }
    }
  }
  if (a.burns) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) burns += __shfl_xor(burns, off, 64);
    if (lane == 0 && burns) atomicAdd(&a.burns[blockIdx.x % kBurnSlots], burns);
  }
}

// The wide pre-pass: one wave per unit = (kItemWide item, 16-row strip,
// window of at most kWideWords mask words), grid-stride over the unit list.
// The strip's edges (its band, or the whole edge list) are visited once for
// the window, in the cheaper of the two layouts per chunk of 64 edges
// (row-major: four lanes per row over the staged active edges; edge-major:
// one lane per edge over the rows it crosses), and every crossing updates
// the row's LDS state with an atomic (XOR of its toggle bit or of the left
// parity, max of the column; order-free).  Then two rows at a time, lane
// (half, w) turns word w of the toggles of its row into coverage (prefix
// parity across the words from one ballot, right limit from the rightmost
// crossing) and stores it.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_wide_cover(WideCoverArgs a) {
  __shared__ PolyEdge s_edge[4][64];
  // rows padded by one word, so the same word of different rows falls in
  // different LDS banks
  __shared__ uint64_t s_tog[4][WAVE_ROWS][kWideWords + 1];
  __shared__ uint32_t s_lp[4][WAVE_ROWS];
  __shared__ int32_t s_amax[4][WAVE_ROWS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rl = lane >> 2, q = lane & 3;
  const RZ_CONST JobDev* jobs_c = as_const(a.jobs);
  const RZ_CONST uint64_t* band_c = as_const(a.band_off);
  const uint64_t units = min(*a.n_units, a.unit_cap);
  PolyEdge* stage = s_edge[wave];
  uint64_t(*tog)[kWideWords + 1] = s_tog[wave];
  uint32_t* lp = s_lp[wave];
  int32_t* amax = s_amax[wave];
  if (a.n_items <= 0) return;
  // the toggle rows start at zero, and each unit's finalize step clears the
  // words it read (no crossing touches a word past the unit's window)
  for (int idx = lane; idx < WAVE_ROWS * (kWideWords + 1); idx += 64) (&tog[0][0])[idx] = 0ull;
  const uint64_t ustride = (uint64_t)gridDim.x * 4u;
  uint64_t u = (uint64_t)blockIdx.x * 4u + (uint64_t)wave;
  // the next unit's item is loaded one unit ahead, and a unit's range and
  // descriptor are loaded together (clamped to a valid item) and checked
  // afterwards, so a unit starts after one round trip instead of three
  uint32_t it_next = u < units ? a.unit_item[u] : 0u;
  for (; u < units; u += ustride) {
    // (wave-uniform: the unit's header and descriptor go to scalar registers)
    const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane((int)it_next);
    if (u + ustride < units) it_next = a.unit_item[u + ustride];
    const uint32_t itc = (int64_t)it < a.n_items ? it : (uint32_t)(a.n_items - 1);
    const uint64_t ub = a.unit_base[itc], ue = a.unit_base[itc + 1];
    const ItemDesc d = load_desc(as_const(a.desc) + itc);
    // a unit of an item past the capacities, or a stale one
    if ((int64_t)it >= a.n_items || u < ub || u >= ue || !(d.flags & kItemWide)) continue;
    const RZ_CONST JobDev& J = jobs_c[d.job];
    const int32_t nw = wide_row_words(d.c0, d.c1);
    const int32_t nwin = (nw + kWideWords - 1) / kWideWords;
    const int32_t local = (int32_t)(u - ub);
    const int32_t k = local / nwin, w0 = (local - k * nwin) * kWideWords;
    const int32_t ww = min(kWideWords, nw - w0);
    const int32_t Rw = ((d.r0 >> 4) + k) * WAVE_ROWS, strip_r1 = min(Rw + WAVE_ROWS, J.nrows);
    const int32_t W0 = ((d.c0 >> 6) + w0) * 64, WEND = W0 + ww * 64;
    const double ncols_f = (double)J.ncols;
    const PolyEdge* ebase = a.edges + d.vbeg;
    uint32_t n_e = (uint32_t)d.nv;
    if (d.band) {
      const uint32_t bi = d.band - 1u + (uint32_t)k;
      const uint64_t b0 = band_c[bi];
      ebase = a.band_edges + b0;
      n_e = (uint32_t)(band_c[bi + 1u] - b0);
    }
    // the rightmost crossing of a row matters only with an odd count
    const bool odd_rows = (d.flags & kItemOddRows) != 0;
    if (lane < WAVE_ROWS) {
      lp[lane] = 0u;
      amax[lane] = -1;
    }
    wave_sync();
    // one crossing of row `row` (0..15) at x
    auto put = [&](int32_t row, double x) {
      const int32_t col = span_col(x, ncols_f);
      const int32_t rel = col - W0;
      if (rel <= 0) {
        __hip_atomic_fetch_xor(&lp[row], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        return;
      }
      if (odd_rows) __hip_atomic_fetch_max(&amax[row], col, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      if (col < WEND)
        __hip_atomic_fetch_xor(&tog[row][rel >> 6], 1ull << (rel & 63), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WAVEFRONT);
    };
    for (uint32_t cb = 0; cb < n_e; cb += 64) {
      const uint32_t e = cb + (uint32_t)lane;
      uint4 e_lo = make_uint4(0u, 0u, 0u, 0u), e_hi = make_uint4(0u, 0u, 0u, 0u);
      int32_t a0 = 0, span = 0;  // rows [a0, a0 + span) of the strip the edge crosses
      if (e < n_e) {
        const uint4* src = reinterpret_cast<const uint4*>(ebase + e);
        e_lo = src[0];
        e_hi = src[1];  // dxdy, ystart (z), yend (w)
        a0 = max((int32_t)e_hi.z, Rw);
        span = max(0, min((int32_t)e_hi.w, strip_r1) - a0);
      }
      const uint64_t m = __ballot(span > 0);
      const int n = __popcll(m);
      const int32_t steps = wave_max16(span);
      if (a.cov_mode == 2 || (a.cov_mode == 0 && steps <= (n + 3) / 4)) {  // edge-major
        PolyEdge ee;
        __builtin_memcpy(&ee, &e_lo, 16);
        __builtin_memcpy(reinterpret_cast<char*>(&ee) + 16, &e_hi, 16);
        for (int32_t j = 0; j < steps; ++j) {
          const int32_t rr = a0 + j;
          const double x = crossing(ee, rr);
          if (j < span) put(rr - Rw, x);
        }
        continue;
      }
      // row-major over the staged active edges
      if (span > 0) {
        uint4* dst = reinterpret_cast<uint4*>(stage + __popcll(m & ((1ull << lane) - 1ull)));
        dst[0] = e_lo;
        dst[1] = e_hi;
      }
      wave_sync();
      const int32_t r = Rw + rl;
      for (int k0 = 0; k0 < n; k0 += 4) {
        const int kk = k0 + q;
        const PolyEdge ee = stage[kk < n ? kk : n - 1];
        const double x = crossing(ee, r);
        if (kk < n && r < strip_r1 && ee.ystart <= r && r < ee.yend) put(rl, x);
      }
      wave_sync();
    }
    wave_sync();
    // two rows per step: lanes 0..31 row 2i, lanes 32..63 row 2i + 1
    static_assert(kWideWords == 32, "two rows of window words per wave step");
    uint64_t* out = a.wide_masks + (uint64_t)d.wide * 16u + (uint64_t)k * 16u * (uint64_t)nw + (uint64_t)w0;
    const int half = lane >> 5, w = lane & 31;
    // A heavy tile's fold reads its items' masks from its pre-pass slots, so
    // each word also goes to the slot of the tile that holds it when that
    // tile is heavy: tile column gw / (TW / 64), the item's place j in the
    // tile's (item-ordered) pair list found by binary search.
    uint64_t* slot = nullptr;
    if (a.heavy_cnt && !J.scan && w < ww) {
      const int32_t nwt = J.tile_w >> 6;  // words per tile row
      const int32_t gw = (W0 >> 6) + w, tx = gw / nwt;
      const uint32_t t = J.tile_base + (d.kind_layer >> 2) * J.tiles_per_layer +
                         (uint32_t)(Rw / TILE_H) * (uint32_t)J.ntx + (uint32_t)tx;
      const uint32_t cnt = a.heavy_cnt[t];
      // the pair (it, t) in k_bin_emit's order over the item's tile
      // rectangle: row Rw / TILE_H - ty0, column tx - tx0 (uniform part first)
      int32_t tx0, tx1, ty0, ty1;
      uint32_t j = ~0u;
      if (cnt && a.pair_off[it + 1] <= a.pair_cap &&
          tile_rect(d.r0, d.r1, d.c0, d.c1, J.tile_w, tx0, tx1, ty0, ty1)) {
        const int64_t jrow = (int64_t)a.pair_off[it] + (int64_t)(Rw / TILE_H - ty0) * (tx1 - tx0 + 1) - tx0;
        j = a.wide_j[jrow + tx];
      }
      if (j < cnt) {
        const uint32_t s = (uint32_t)((Rw % TILE_H) / WAVE_ROWS);
        slot = a.masks + (a.heavy_base[t] * 4u + (uint64_t)s * cnt + j) * a.slot_words + (uint32_t)(gw % nwt);
      }
    }
    const uint64_t below = ((1ull << lane) - 1ull) & (half ? 0xFFFFFFFF00000000ull : 0xFFFFFFFFull);
    for (int r2 = 0; r2 < WAVE_ROWS; r2 += 2) {
      const int row = r2 + half;
      uint64_t t = 0ull;
      if (w < ww) {
        t = tog[row][w];
        tog[row][w] = 0ull;
      }
      const uint64_t odd = __ballot(__popcll(t) & 1);
      const uint32_t carry = lp[row] ^ ((uint32_t)__popcll(odd & below) & 1u);
      uint64_t x = prefix_xor64(t) ^ (carry ? ~0ull : 0ull);
      const int32_t lim = odd_rows ? amax[row] - (W0 + 64 * w) : 64;
      x &= lim <= 0 ? 0ull : (lim >= 64 ? ~0ull : ((1ull << lim) - 1ull));
      if (w < ww) out[(uint64_t)row * (uint64_t)nw + (uint64_t)w] = x;
      if (slot) {
        // the tile's words left of the item's first word or right of its
        // last one are empty: written as such (the fold reads every word)
        const int32_t nwt = J.tile_w >> 6, gw = (W0 >> 6) + w, qq = gw % nwt;
        uint64_t* srow_p = slot + row * nwt - qq;
        srow_p[qq] = x;
        if (gw == (d.c0 >> 6))
          for (int z = 0; z < qq; ++z) srow_p[z] = 0ull;
        if (gw == ((d.c1 - 1) >> 6))
          for (int z = qq + 1; z < nwt; ++z) srow_p[z] = 0ull;
      }
    }
    wave_sync();  // the rows are read and cleared before the next unit's crossings
  }
}

}  // namespace rzk
