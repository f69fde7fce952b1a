// rz_tile_t.h -- the tile kernel templates (k_tile, k_tile_heavy): scanline
// coverage, burn folds and the write-once store of every output tile.
// Instantiated per output dtype in rz_tile_inst.hip.
#pragma once

#include "rz_device.h"

namespace rzk {

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

}  // namespace rzk
